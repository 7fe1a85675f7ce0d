#!/usr/bin/env python
"""Benchmark of the VPFP time step on one MI355X.

Headline (`value`): the batched ragged parameter scan in instances/s at the
damping-rate bar.  4096 instances, k ~ U[0.25, 0.45], amplitude
10^U[-4, -1], nu 10^U[-5, -2] (Dougherty), nx in {32, 64, 128}, nv in
{256, 512, 1024, 2048} drawn from default_rng(20261202); v in [-6.4, 6.4],
f0 = (1 + amp cos kx) exp(-v^2/2)/sqrt(2 pi), e0 = 0, dt = 0.1, 2000 leapfrog
+ collision steps, the <E^2>(t) history of every instance recorded.  The
timed configuration stores f - M in FP32 (delta-f storage) with FP64
moments, Poisson and tridiagonal; the same scan in FP64 is timed beside it
and is the reference: `damping_rate` gives the fraction of fitted instances
whose damping rate agrees with the FP64 scan's within 1 %.  Plain FP32
storage (f itself) is timed as well.

Beside it: the collisional nonlinear electron plasma wave (`epw`, nx=64,
nv=1024, Dougherty nu=1e-3, dt=0.05, 4000 steps with the storage moments) in
FP32, FP64 and FP32 delta-f, with the row kernel's roofline; the
high-resolution grid (`large_grid`, 1024 x 8192); the learned closure
(`closure`).  CPU baselines: the NumPy FP64 oracle on one core and on every
core of the host's share (one process per core), on a sample of the scan.

    python bench.py [--scan-steps 2000]

Prints one JSON line.
"""

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CLOSURE_NPZ = os.path.join(HERE, "vlapy_amd", "data", "closure_c5.npz")  # the trained closure
METRIC = "VPFP cell-updates/s + % HBM roofline; scan runs/s at ≤1% damping-rate error"
HBM_PEAK_GBS = 8000.0

NX, NV, K0, VMAX, AMP, NU, DT, NSTEPS = 64, 1024, 0.3, 6.4, 0.05, 1e-3, 0.05, 4000
WORKLOAD = ("collisional nonlinear EPW: nx=64, nv=1024, v in [-6.4, 6.4], k=0.3, density 1+0.05cos(kx) "
            "on a unit Maxwellian, Dougherty nu=1e-3 implicit per x every step, dt=0.05, 4000 steps")


def initial_state(np):
    xmax = 2 * np.pi / K0
    dx = xmax / NX
    x = np.linspace(dx / 2.0, xmax - dx / 2.0, NX)
    kx = np.fft.fftfreq(NX, d=dx) * 2.0 * np.pi
    okx = np.zeros_like(kx)
    okx[1:] = 1.0 / kx[1:]
    dv = 2 * VMAX / NV
    v = np.linspace(-VMAX + dv / 2.0, VMAX - dv / 2.0, NV)
    kv = np.fft.fftfreq(NV, d=dv) * 2.0 * np.pi
    f0 = (1.0 + AMP * np.cos(K0 * x))[:, None] * (np.exp(-v ** 2 / 2.0) / np.sqrt(2 * np.pi))[None, :]
    return x, kx, okx, v, kv, dv, f0


# --------------------------------------------------------------------------- CPU baseline


def cpu_child(target_s):
    """Time the NumPy FP64 oracle on one core (run in a child with 1 BLAS/OMP thread)."""
    import numpy as np

    sys.path.insert(0, HERE)
    from oracle import vpfp_oracle as O

    x, kx, okx, v, kv, dv, f = initial_state(np)
    g = O.Grid(x, kx, okx, v, kv, dv)
    e = np.zeros(NX)
    steps = 0
    t0 = time.perf_counter()
    while True:
        e, f = O.leapfrog_step(g, e, f, steps * DT, DT)
        f = O.collide(f, v, NU, DT, dv, "dg")
        mom = O.step_moments(f, v, dv)
        O.step_series(mom, e, np.zeros(NX), f, dv)
        steps += 1
        el = time.perf_counter() - t0
        if (el >= target_s and steps >= 10) or steps >= NSTEPS:
            break
    print(json.dumps({"steps": steps, "seconds": el}))


def oracle_child(arg):
    """Time `steps` steps of the NumPy FP64 oracle on one configuration (one
    thread): "c1:<steps>" the collisionless Landau run, "c4:<steps>" the
    1024 x 8192 run with the Dougherty collision."""
    import numpy as np

    sys.path.insert(0, HERE)
    from oracle import vpfp_oracle as O

    cfg, steps = arg.split(":")
    steps = int(steps)
    if cfg == "c1":
        x, kx, okx, v, kv, dv, f = c1_state(np)
        nu, dt = 0.0, C1_DT
    else:
        x, kx, okx, v, kv, dv, f = large_state(np)
        nu, dt = LG_NU, LG_DT
    g = O.Grid(x, kx, okx, v, kv, dv)
    e = np.zeros(x.size)
    t0 = time.perf_counter()
    for i in range(steps):
        e, f = O.leapfrog_step(g, e, f, i * dt, dt)
        if nu > 0.0:
            f = O.collide(f, v, nu, dt, dv, "dg")
    print(json.dumps({"steps": steps, "seconds": time.perf_counter() - t0}))


def oracle_time(cfg, steps):
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--oracle-child", "%s:%d" % (cfg, steps)],
                         capture_output=True, text=True, env=env, timeout=600)
    if res.returncode != 0:
        raise RuntimeError("oracle timing failed: " + res.stderr[-2000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


def cpu_baseline(target_s):
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", str(target_s)],
                         capture_output=True, text=True, env=env, timeout=600)
    if res.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + res.stderr[-2000:])
    r = json.loads(res.stdout.strip().splitlines()[-1])
    return {
        "value": NX * NV * r["steps"] / r["seconds"],
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": "first %d of the 4000 steps of the same workload on the NumPy FP64 oracle "
                  "(leapfrog + Dougherty collision + storage moments), %.1f s, one thread" % (r["steps"], r["seconds"]),
    }


# --------------------------------------------------------------------------- GPU


def gpu_run(precision, steps, warmup, profile_steps, autotune=True, delta=False):
    import numpy as np
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import NMOMENTS, Plan

    dev = torch.device("cuda:0")
    x, kx, okx, v, kv, dv, f0 = initial_state(np)
    plan = Plan(NX, NV, precision, kx, okx, v, kv, dv, NU, DT, "dg")
    if delta:
        # f - M stored (vpfp_plan_set_delta), formed in FP64
        plan.set_delta(True)
        f0 = f0 - (np.exp(-v ** 2 / 2.0) / np.sqrt(2 * np.pi))[None, :]
    fdt = ops.f_dtype(plan)
    # launch shapes picked by timing them on scratch state (untimed setup)
    tuning = plan.autotune(50) if autotune else plan.tuning()

    def fresh():
        return (torch.as_tensor(f0, dtype=fdt, device=dev).contiguous(),
                torch.zeros(NX, dtype=torch.float64, device=dev))

    def outputs(n):
        return (torch.zeros(n, NX, dtype=torch.float64, device=dev),
                torch.zeros(n, NMOMENTS, NX, dtype=torch.float64, device=dev),
                torch.zeros(n, 2, dtype=torch.float64, device=dev))

    # warmup (untimed)
    f, e = fresh()
    if warmup > 0:
        eh, mo, se = outputs(warmup)
        ops.step_n(plan, f, e, warmup, None, eh, mo, se)
    torch.cuda.synchronize()

    # timed region: `steps` full timesteps from the initial state
    f, e = fresh()
    eh, mo, se = outputs(steps)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    ops.step_n(plan, f, e, steps, None, eh, mo, se)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    if not torch.isfinite(eh).all():
        raise RuntimeError("non-finite field history")
    if delta:
        plan.background(f, 1)
    result = {"ms": ms, "wall_s": wall, "e_hist": eh.cpu().numpy(), "series": se.cpu().numpy(),
              "f": f.double().cpu().numpy()}

    # per-kernel durations from HIP events around every launch (separate pass)
    f, e = fresh()
    eh, mo, se = outputs(profile_steps)
    plan.set_profiling(True)
    ops.step_n(plan, f, e, profile_steps, None, eh, mo, se)
    torch.cuda.synchronize()
    result["kernels"] = plan.kernel_times()
    plan.set_profiling(False)
    result["bytes_f"] = 8 if precision == "fp64" else 4
    result["ntiles"] = NV // tuning["drift_tile"]
    result["tuning"] = tuning
    return result


def roofline(run, cells, ms_per_step, traffic_file=None):
    """Dominant kernel: achieved algorithmic GB/s vs the 8 TB/s HBM peak.

    Algorithmic bytes are one read and one write of f per split operator
    (moments, Poisson and the storage epilogue fused: no extra bytes), s = 4
    (FP32 f) or 8 (FP64 f) bytes per cell:
      row kernel   kick(dt/2) + collision + kick(dt/2) = 3 operators: 6 s cells
      drift        one operator: 2 s cells
      field        the partial densities it reads and E: 8 (ntiles + 1) nx
    and the whole step (leapfrog + collision, 4 operators) 8 s per cell-update.
    hbm_min_bytes_per_launch is what the fused kernel itself must move."""
    s = run["bytes_f"]
    kern = run["kernels"]
    dom = max(("row", "drift", "field"), key=lambda k: kern[k][0])
    total_ms, launches = kern[dom]
    avg_s = total_ms / launches / 1e3
    nx = run.get("nx", NX)
    alg = {"row": 6 * s * cells, "drift": 2 * s * cells, "field": 8 * (run["ntiles"] + 1) * nx}[dom]
    hbm_min = {"row": 2 * s * cells, "drift": 2 * s * cells + 8 * nx * run["ntiles"],
               "field": 8 * (run["ntiles"] + 1) * nx}[dom]
    achieved = alg / avg_s / 1e9
    traffic = None
    if traffic_file:
        prof = os.path.join(HERE, "profiles", traffic_file)
        if os.path.exists(prof):
            with open(prof) as fh:
                traffic = json.load(fh).get(dom)
    step_gbs = 8 * s * cells / (ms_per_step / 1e3) / 1e9
    return {
        "bound": "hbm",
        "kernel": dom,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "alg_bytes_per_launch": alg,
        "hbm_min_bytes_per_launch": hbm_min,
        "hbm_min_frac": hbm_min / avg_s / 1e9 / HBM_PEAK_GBS,
        "avg_launch_us": avg_s * 1e6,
        "kernel_us": {k: (v[0] / v[1] * 1e3 if v[1] else None) for k, v in kern.items()},
        "step_alg_GBps": step_gbs,
        "step_frac": step_gbs / HBM_PEAK_GBS,
    }


# --------------------------------------------------------------------------- scan

SCAN_N, SCAN_SEED, SCAN_STEPS, SCAN_DT = 4096, 20261202, 2000, 0.1
SCAN_GRAPH_STEPS = 100


def scan_window(p, i, e2, t):
    """Fit window for instance i: after one period, before half the trapping
    time (pi / sqrt(amp)), half the recurrence time 2 pi / (k dv) and the
    point where <E^2> has fallen 1e8 below its start (FP32 noise)."""
    import numpy as np

    dv = 12.8 / p["nv"][i]
    t_rec = 2 * np.pi / (p["k"][i] * dv)
    t_trap = 2 * np.pi / np.sqrt(p["amp"][i])
    y = e2[:, i]
    floor = np.nonzero(y < 1e-8 * y[:20].max())[0]
    t_noise = t[floor[0]] if floor.size else t[-1]
    return 6.0, min(0.5 * t_trap, 0.5 * t_rec, t_noise, t[-1])


SCAN_PROFILE_STEPS = 100


SCAN_TUNING = {"in": None, "out": None}  # --scan-tuning-in / --scan-tuning-out (JSON files)


def _tuning_key(prec, delta):
    return "%s%s" % (prec, "_delta" if delta else "")


def scan_run(precision, steps, delta=False, graph=False, profile_steps=0, warmup=2):
    """One timed scan of `steps` steps from the initial state (launch shapes
    autotuned first, untimed), <E^2> recorded every step; then optionally
    `profile_steps` more steps with per-kernel HIP events, and the same steps
    replayed from one HIP graph of SCAN_GRAPH_STEPS steps."""
    import torch

    from vlapy_amd.scan import Scan, draw_parameters

    p = draw_parameters(SCAN_N, SCAN_SEED)
    sc = Scan(p, dt=SCAN_DT, op="dg", precision=precision, delta=delta)
    sc.run(max(1, warmup), record=False)  # warmup (also first-call kernel setup)
    torch.cuda.synchronize()
    del sc
    sc = Scan(p, dt=SCAN_DT, op="dg", precision=precision, delta=delta)
    key = _tuning_key(precision, delta)
    saved = None
    if SCAN_TUNING["in"]:
        with open(SCAN_TUNING["in"]) as fh:
            saved = json.load(fh).get(key)
    if saved:
        # the launch shapes of an earlier timed run (counter runs)
        tuning = {(c.nx, c.nv): saved["%dx%d" % (c.nx, c.nv)] for c in sc.classes}
        sc.set_tuning(tuning)
    else:
        tuning = sc.autotune(3)  # launch shapes per size class (untimed setup)
    if SCAN_TUNING["out"]:
        old = {}
        if os.path.exists(SCAN_TUNING["out"]):
            with open(SCAN_TUNING["out"]) as fh:
                old = json.load(fh)
        old[key] = {"%dx%d" % k: v for k, v in tuning.items()}
        with open(SCAN_TUNING["out"], "w") as fh:
            json.dump(old, fh, indent=1)
    n0, w0 = sc.density(), sc.energy(self_consistent=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e2 = sc.run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import numpy as np

    r = {"p": p, "cells": sc.cells, "e2": e2, "wall": wall, "tuning": tuning,
         "density_drift": np.abs(sc.density() / n0 - 1.0), "energy_drift": np.abs(sc.energy() / w0 - 1.0)}
    if profile_steps:
        sc.set_profiling(True)
        sc.run(profile_steps, record=False)
        torch.cuda.synchronize()
        r["kernels"] = sc.kernel_times()
        r["profile_steps"] = profile_steps
        r["launch_bytes"] = launch_bytes(sc, tuning, 8 if precision == "fp64" else 4)
        sc.set_profiling(False)
    r["graph_wall"] = None
    if graph and steps % SCAN_GRAPH_STEPS == 0:
        del sc
        sc = Scan(p, dt=SCAN_DT, op="dg", precision=precision, delta=delta)
        sc.set_tuning(tuning)
        sc.capture(SCAN_GRAPH_STEPS)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sc.replay(steps // SCAN_GRAPH_STEPS)
        torch.cuda.synchronize()
        r["graph_wall"] = time.perf_counter() - t0
    return r


def launch_bytes(sc, tuning, s):
    """Algorithmic bytes of one launch of each kernel of the scan's step, keyed
    as Scan.kernel_times(): per grouped launch ("nv", nv) / ("nx", nx), or per
    size class.  s bytes per stored value, cells of the launch's classes:
    row kernel 3 split operators (kick, collision, kick) = 6 s cells per fused
    launch; drift 2 s cells; field the partial densities and E,
    8 (drift workgroups + 1) nx per instance."""
    out = {}
    for c in sc.classes:
        t = tuning[(c.nx, c.nv)]
        gx = -(-(c.nv // t["drift_tile"]) // max(1, t["drift_tiles"]))
        cells = c.n * c.nx * c.nv
        rk, xk = (("nv", c.nv), ("nx", c.nx)) if sc.group is not None else ((c.nx, c.nv), (c.nx, c.nv))
        out.setdefault(rk, {}).setdefault("row", 0)
        out.setdefault(xk, {}).setdefault("drift", 0)
        out[xk].setdefault("field", 0)
        out[rk]["row"] += 6 * s * cells
        out[xk]["drift"] += 2 * s * cells
        out[xk]["field"] += 8 * (gx + 1) * c.nx * c.n
    return out


def _launch_name(key):
    return "%s=%d" % key if isinstance(key[0], str) else "nx=%d, nv=%d" % key


def scan_roofline(run, traffic_file=None):
    """Dominant kernel of the scan (the launch with the most event time over
    the profiled steps): its algorithmic bytes (launch_bytes) over its
    average launch duration, against the 8 TB/s HBM peak.  Over n profiled
    steps a row launch is made n + 1 times carrying 3n split operators in
    all (the first is a half kick alone, the last has no closing half kick),
    which is what `achieved` divides."""
    kern = run["kernels"]
    cands = [(k, kind) for k in kern for kind in ("row", "drift", "field") if kern[k].get(kind, (0, 0))[1]]
    key, kind = max(cands, key=lambda kk: kern[kk[0]][kk[1]][0])
    total_ms, launches = kern[key][kind]
    n = run["profile_steps"]
    per_launch = run["launch_bytes"][key][kind]
    achieved = per_launch * n / (total_ms / 1e3) / 1e9
    traffic = None
    if traffic_file:
        prof = os.path.join(HERE, "profiles", traffic_file)
        if os.path.exists(prof):
            with open(prof) as fh:
                t = json.load(fh)
            if t.get("launch") == _launch_name(key):
                traffic = t.get(kind)
    step_ms = sum(v[0] for d in kern.values() for v in d.values()) / n
    return {
        "bound": "hbm",
        "kernel": "%s (%s)" % (kind, _launch_name(key)),
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "alg_bytes_per_launch": per_launch,
        "avg_launch_us": total_ms / launches * 1e3,
        "launches": launches,
        "kernel_ms_per_step": {_launch_name(k): {kk: v[0] / n for kk, v in d.items() if v[1]} for k, d in kern.items()},
        "sum_of_kernel_ms_per_step": step_ms,
    }


def drift_mfma_section(tuning, steps=20):
    """The x-advection of the nx = 32 and nx = 64 launches of the FP32
    delta-f scan as Stockham transforms (the tuned shapes) and as dense
    real-DFT products on MFMA (drift_dft, 16 and 32 columns per pass, the
    same tiles per workgroup): microseconds per launch from HIP events over
    `steps` steps, the faster per nx, and the MFMA instruction rate."""
    import copy

    import torch

    from vlapy_amd.scan import Scan, draw_parameters

    p = draw_parameters(SCAN_N, SCAN_SEED)
    sc = Scan(p, dt=SCAN_DT, op="dg", precision="fp32", delta=True)
    out = {}
    for name, dft, C in (("stockham", 0, None), ("mfma_16", 1, 16), ("mfma_32", 1, 32)):
        t = copy.deepcopy(tuning)
        for (nx, nv), tt in t.items():
            if nx in (32, 64):
                tt["drift_dft"] = dft
                if C:
                    tt["drift_tile"] = C
        sc.set_tuning(t)
        sc.run(2, record=False)
        sc.set_profiling(True)
        sc.run(steps, record=False)
        torch.cuda.synchronize()
        kt = sc.kernel_times()
        sc.set_profiling(False)
        for nx in (32, 64):
            ms, n = kt[("nx", nx)]["drift"]
            out.setdefault("nx=%d" % nx, {})[name] = ms / n * 1e3
    for nx in (32, 64):
        d = out["nx=%d" % nx]
        d["faster"] = min(d, key=d.get)
    return {"workload": "x-advection launches of the FP32 delta-f scan (grouped), %d steps" % steps,
            "us_per_launch": out}


def scan_cpu(p, steps_sample, n_sample, procs):
    """Oracle time on a sample of the scan's instances, extrapolated by
    cell-steps to the whole scan (run in a child process with one BLAS/OMP
    thread per process): procs = 1 is one core; procs > 1 runs the sample
    as one process per core over the instance list."""
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    res = subprocess.run([sys.executable, os.path.abspath(__file__), "--scan-cpu-child",
                          "%d,%d,%d" % (steps_sample, n_sample, procs)], capture_output=True, text=True, env=env,
                         timeout=900)
    if res.returncode != 0:
        raise RuntimeError("scan cpu baseline failed: " + res.stderr[-2000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


def _oracle_instance(args):
    """Run instance i of the scan draw for `steps` steps on the oracle;
    returns (seconds, cell-steps)."""
    i, steps = args
    import numpy as np

    sys.path.insert(0, HERE)
    from oracle import vpfp_oracle as O
    from vlapy_amd.scan import draw_parameters, instance_grids

    p = draw_parameters(SCAN_N, SCAN_SEED)
    x, kx, okx, v, kv, dv = instance_grids(int(p["nx"][i]), p["k"][i], int(p["nv"][i]), 6.4)
    g = O.Grid(x, kx, okx, v, kv, dv)
    f = O.perturbed_maxwellian(x, v, p["k"][i], p["amp"][i])
    t0 = time.perf_counter()
    O.run_steps(g, np.zeros(x.size), f, steps, SCAN_DT, nu=p["nu"][i], fp_type="dg")
    return time.perf_counter() - t0, steps * x.size * v.size


def scan_cpu_child(arg):
    import numpy as np

    from vlapy_amd.scan import draw_parameters

    steps, n_sample, procs = (int(a) for a in arg.split(","))
    p = draw_parameters(SCAN_N, SCAN_SEED)
    # a sample spread over the size classes: every (4096 / n_sample)-th instance
    idx = np.arange(0, SCAN_N, SCAN_N // n_sample)[:n_sample]
    t0 = time.perf_counter()
    if procs == 1:
        res = [_oracle_instance((int(i), steps)) for i in idx]
    else:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(procs) as pool:
            res = pool.map(_oracle_instance, [(int(i), steps) for i in idx], chunksize=1)
    wall = time.perf_counter() - t0
    work = float(sum(r[1] for r in res))
    cells_all = float(np.sum(p["nx"].astype(float) * p["nv"].astype(float)))
    # seconds for the whole scan at the sample's throughput in cell-steps/s
    est = wall * cells_all * SCAN_STEPS / work
    print(json.dumps({"sample_instances": int(len(idx)), "sample_steps": steps, "procs": procs,
                      "sample_wall_s": wall, "sample_cpu_s": float(sum(r[0] for r in res)),
                      "est_scan_seconds": est}))


def host_cores():
    """Cores of this host's share: the affinity mask, at most 16 (the CPU
    share of one GPU on the benchmark host)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return min(aff, 16), aff


def scan_section(timed_steps, warmup, with_cpu, check=True, profile_steps=SCAN_PROFILE_STEPS):
    """The headline: `timed_steps` steps of the FP32 delta-f scan from the
    initial state (value = completed runs of SCAN_STEPS steps per second,
    SCAN_N * timed_steps / SCAN_STEPS / wall), then the full-length runs that
    check it: FP32 delta-f (the same run when timed_steps == SCAN_STEPS),
    FP64 and plain FP32, SCAN_STEPS steps each."""
    import numpy as np

    from vlapy_amd.diagnostics.low_level_helpers import get_standing_wave_rate_and_frequency as rate

    steps = SCAN_STEPS
    head = scan_run("fp32", timed_steps, delta=True, warmup=warmup, profile_steps=max(1, profile_steps),
                    graph=(timed_steps == SCAN_STEPS))
    if not check:
        # profiling runs: the timed steps and the kernel roofline only
        return {"headline": {"timed_steps": timed_steps, "seconds": head["wall"],
                             "runs_per_s": SCAN_N * timed_steps / SCAN_STEPS / head["wall"],
                             "ms_per_step": head["wall"] / timed_steps * 1e3},
                "fp32_delta_f": {"damping_rate": None},
                "roofline": scan_roofline(head, "traffic_scan_fp32_delta.json")}
    rd = head if timed_steps == SCAN_STEPS else scan_run("fp32", steps, delta=True, graph=True)
    r64 = scan_run("fp64", steps)
    r32 = scan_run("fp32", steps, graph=True)
    p = rd["p"]
    t = SCAN_DT * (np.arange(steps) + 1)
    e2_64 = r64["e2"]

    def damping_errs(e2):
        errs = []
        for i in range(SCAN_N):
            lo, hi = scan_window(p, i, e2_64, t)
            try:
                g64, _ = rate(e2_64[:, i], t, lo, hi)
                g32, _ = rate(e2[:, i], t, lo, hi)
            except ValueError:
                continue
            errs.append(abs(g32 / g64 - 1.0))
        return np.asarray(errs)

    def summary(errs):
        return {"instances_fitted": int(errs.size),
                "within_1pct": float(np.mean(errs <= 0.01)) if errs.size else None,
                "median_rel_err": float(np.median(errs)) if errs.size else None,
                "p95_rel_err": float(np.percentile(errs, 95)) if errs.size else None}

    def drifts(r):
        return {"density_max": float(r["density_drift"].max()), "density_median": float(np.median(r["density_drift"])),
                "energy_max": float(r["energy_drift"].max()), "energy_median": float(np.median(r["energy_drift"]))}

    cells = rd["cells"]
    out = {
        "workload": "4096-instance ragged scan, seed 20261202, Dougherty, dt=0.1, %d steps" % steps,
        "cells": cells,
        "headline": {"timed_steps": timed_steps, "seconds": head["wall"],
                     "runs_per_s": SCAN_N * timed_steps / SCAN_STEPS / head["wall"],
                     "ms_per_step": head["wall"] / timed_steps * 1e3},
        "fp32_delta_f": {"instances_per_s": SCAN_N / rd["wall"], "seconds": rd["wall"],
                         "ms_per_step": rd["wall"] / steps * 1e3,
                         "cell_updates_per_s": cells * steps / rd["wall"],
                         "graph_instances_per_s": SCAN_N / rd["graph_wall"] if rd["graph_wall"] else None,
                         "damping_rate": summary(damping_errs(rd["e2"])), "drift": drifts(rd),
                         "tuning": {"%dx%d" % k: v for k, v in rd["tuning"].items()},
                         "tuning_raw": rd["tuning"]},
        "fp64": {"instances_per_s": SCAN_N / r64["wall"], "seconds": r64["wall"],
                 "cell_updates_per_s": cells * steps / r64["wall"], "drift": drifts(r64)},
        "fp32": {"instances_per_s": SCAN_N / r32["wall"], "seconds": r32["wall"],
                 "graph_instances_per_s": SCAN_N / r32["graph_wall"] if r32["graph_wall"] else None,
                 "damping_rate": summary(damping_errs(r32["e2"])), "drift": drifts(r32)},
        # algorithmic bytes of a leapfrog + collision step: 8 s per cell-update
        "step_alg_GBps": {"fp32_delta_f": 32.0 * cells * steps / rd["wall"] / 1e9,
                          "fp64": 64.0 * cells * steps / r64["wall"] / 1e9},
        "roofline": scan_roofline(head, "traffic_scan_fp32_delta.json"),
    }
    if with_cpu:
        procs, aff = host_cores()
        one = scan_cpu(p, 25, 24, 1)
        allc = scan_cpu(p, 30, 64, procs)
        est1 = one["est_scan_seconds"] * steps / SCAN_STEPS
        estn = allc["est_scan_seconds"] * steps / SCAN_STEPS
        out["cpu_baseline"] = {
            "one_core": {"instances_per_s": SCAN_N / est1, "cores": 1, "kind": "port",
                         "sample": "%d instances x %d steps of the scan on the NumPy FP64 oracle (%.1f s), "
                                   "extrapolated by cell-steps to the whole scan" % (
                                       one["sample_instances"], one["sample_steps"], one["sample_wall_s"])},
            "all_cores": {"instances_per_s": SCAN_N / estn, "cores": procs, "kind": "port",
                          "sample": "%d instances x %d steps of the scan on the NumPy FP64 oracle, one process per "
                                    "core over the instance list (%.1f s wall, %.1f s CPU), extrapolated by "
                                    "cell-steps to the whole scan" % (allc["sample_instances"], allc["sample_steps"],
                                                                     allc["sample_wall_s"], allc["sample_cpu_s"])},
            "os_cpu_count": os.cpu_count(), "affinity_cores": aff,
        }
    return out


# --------------------------------------------------------------------------- large grid

LG_NX, LG_NV, LG_K, LG_VMAX, LG_AMP, LG_NU, LG_DT = 1024, 8192, 0.35, 8.0, 0.1, 1e-4, 0.02


C1_NX, C1_NV, C1_K, C1_VMAX, C1_AMP, C1_DT, C1_STEPS = 64, 512, 0.3, 6.4, 1e-4, 0.1, 1000


def c1_state(np):
    """Linear Landau damping: nx=64 on [0, 2 pi/k), k=0.3; nv=512 on [-6.4, 6.4];
    f0 = (1 + 1e-4 cos kx) exp(-v^2/2)/sqrt(2 pi); e0 = 0; no collisions."""
    xmax = 2 * np.pi / C1_K
    dx = xmax / C1_NX
    x = np.linspace(dx / 2.0, xmax - dx / 2.0, C1_NX)
    kx = np.fft.fftfreq(C1_NX, d=dx) * 2.0 * np.pi
    okx = np.zeros_like(kx)
    okx[1:] = 1.0 / kx[1:]
    dv = 2 * C1_VMAX / C1_NV
    v = np.linspace(-C1_VMAX + dv / 2.0, C1_VMAX - dv / 2.0, C1_NV)
    kv = np.fft.fftfreq(C1_NV, d=dv) * 2.0 * np.pi
    f0 = (1.0 + C1_AMP * np.cos(C1_K * x))[:, None] * (np.exp(-v ** 2 / 2.0) / np.sqrt(2 * np.pi))[None, :]
    return x, kx, okx, v, kv, dv, f0


def c1_section(with_cpu):
    """The collisionless Landau run, FP64, 1000 leapfrog steps with the E
    history: ms per step (HIP events; eager and from a 100-step HIP graph),
    the damping rate and frequency of <E^2> against the Landau root, and the
    oracle's time for the same 1000 steps on one core."""
    import numpy as np
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import Plan
    from vlapy_amd.diagnostics.low_level_helpers import get_standing_wave_rate_and_frequency

    dev = torch.device("cuda:0")
    x, kx, okx, v, kv, dv, f0 = c1_state(np)
    plan = Plan(C1_NX, C1_NV, "fp64", kx, okx, v, kv, dv, 0.0, C1_DT, "lb")
    tuning = plan.autotune(50)
    f = torch.as_tensor(f0, device=dev).contiguous()
    e = torch.zeros(C1_NX, dtype=torch.float64, device=dev)
    eh = torch.zeros(C1_STEPS, C1_NX, dtype=torch.float64, device=dev)
    ops.step_n(plan, f, e, 20, None, eh[:20])  # warm-up
    f.copy_(torch.as_tensor(f0, device=dev))
    e.zero_()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    ops.step_n(plan, f, e, C1_STEPS, None, eh)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / C1_STEPS
    e2 = (eh ** 2).mean(dim=1).cpu().numpy()
    t = C1_DT * (np.arange(C1_STEPS) + 1)
    gamma, omega = get_standing_wave_rate_and_frequency(e2, t, t_min=8.0)
    # the same loop replayed from a HIP graph of 100 steps
    chunk = 100
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            ops.step_n(plan, f, e, chunk, None, eh[:chunk])
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    a.record()
    for _ in range(C1_STEPS // chunk):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    graph_ms = a.elapsed_time(b) / C1_STEPS
    out = {"workload": "collisionless Landau damping: nx=64, nv=512, v in [-6.4, 6.4], k=0.3, density "
                       "1+1e-4cos(kx), e0=0, dt=0.1, 1000 leapfrog steps with the E history, FP64",
           "ms_per_step": ms, "cell_updates_per_s": C1_NX * C1_NV / (ms / 1e3), "graph_ms_per_step": graph_ms,
           "tuning": tuning, "gamma": gamma, "omega": omega,
           "landau_root": {"gamma": -0.012623, "omega": 1.1598}}
    if with_cpu:
        r = oracle_time("c1", C1_STEPS)
        out["cpu_baseline"] = {"value": C1_NX * C1_NV * r["steps"] / r["seconds"], "unit": "cell-updates/s",
                               "cores": 1, "kind": "port",
                               "sample": "the same 1000 steps on the NumPy FP64 oracle, %.1f s, one thread"
                                         % r["seconds"]}
    return out


def large_state(np):
    """nx=1024 on [0, 2 pi/k), k=0.35; nv=8192 on [-8, 8]; f0 = (1 + 0.1 sum_{m=1..3}
    cos(m k x + phi_m)) exp(-v^2/2)/sqrt(2 pi), phi_m ~ U[0, 2 pi) from default_rng(4)."""
    xmax = 2 * np.pi / LG_K
    dx = xmax / LG_NX
    x = np.linspace(dx / 2.0, xmax - dx / 2.0, LG_NX)
    kx = np.fft.fftfreq(LG_NX, d=dx) * 2.0 * np.pi
    okx = np.zeros_like(kx)
    okx[1:] = 1.0 / kx[1:]
    dv = 2 * LG_VMAX / LG_NV
    v = np.linspace(-LG_VMAX + dv / 2.0, LG_VMAX - dv / 2.0, LG_NV)
    kv = np.fft.fftfreq(LG_NV, d=dv) * 2.0 * np.pi
    phi = np.random.default_rng(4).uniform(0.0, 2 * np.pi, 3)
    n = 1.0 + LG_AMP * sum(np.cos(m * LG_K * x + phi[m - 1]) for m in (1, 2, 3))
    f0 = n[:, None] * (np.exp(-v ** 2 / 2.0) / np.sqrt(2 * np.pi))[None, :]
    return x, kx, okx, v, kv, dv, f0


def large_section(steps, autotune=True, with_cpu=False):
    """Single high-resolution run (8.4M cells), Dougherty nu=1e-4, dt=0.02:
    `steps` leapfrog + collision steps with the E history, FP32 and FP64;
    with_cpu: the oracle's time for 2 steps on one core."""
    import numpy as np
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import Plan

    dev = torch.device("cuda:0")
    x, kx, okx, v, kv, dv, f0 = large_state(np)
    cells = LG_NX * LG_NV
    out = {"workload": "nx=1024, nv=8192, v in [-8, 8], k=0.35, three harmonics of amplitude 0.1 with phases from "
                       "default_rng(4), Dougherty nu=1e-4, dt=0.02; %d steps timed (leapfrog + collision + E "
                       "history)" % steps}
    for prec in ("fp32", "fp64"):
        plan = Plan(LG_NX, LG_NV, prec, kx, okx, v, kv, dv, LG_NU, LG_DT, "dg")
        tuning = plan.autotune(3) if autotune else plan.tuning()
        fdt = ops.f_dtype(plan)
        f = torch.as_tensor(f0, dtype=fdt, device=dev).contiguous()
        e = torch.zeros(LG_NX, dtype=torch.float64, device=dev)
        eh = torch.zeros(steps, LG_NX, dtype=torch.float64, device=dev)
        warm, prof_n = min(5, steps), min(20, steps)
        ops.step_n(plan, f, e, warm, None, eh[:warm])
        f = torch.as_tensor(f0, dtype=fdt, device=dev).contiguous()
        e.zero_()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ops.step_n(plan, f, e, steps, None, eh)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        if not torch.isfinite(eh).all():
            raise RuntimeError("non-finite field history (large grid)")
        plan.set_profiling(True)
        ops.step_n(plan, f, e, prof_n, None, eh[:prof_n])
        torch.cuda.synchronize()
        run = {"kernels": plan.kernel_times(), "bytes_f": 8 if prec == "fp64" else 4,
               "ntiles": min(LG_NV // tuning["drift_tile"], 1024), "nx": LG_NX}
        plan.set_profiling(False)
        out[prec] = {"cell_updates_per_s": cells / (ms / 1e3), "ms_per_step": ms, "tuning": tuning,
                     "roofline": roofline(run, cells, ms, "traffic_large_%s.json" % prec)}
        del plan, f
    if with_cpu:
        r = oracle_time("c4", 2)
        out["cpu_baseline"] = {"value": cells * r["steps"] / r["seconds"], "unit": "cell-updates/s", "cores": 1,
                               "kind": "port",
                               "sample": "the first 2 steps on the NumPy FP64 oracle (leapfrog + Dougherty "
                                         "collision), %.1f s, one thread" % r["seconds"]}
    return out


# --------------------------------------------------------------------------- closure

CLOSURE_N, CLOSURE_SEED, CLOSURE_STEPS = 512, 12, 4000


def closure_section():
    """The learned-closure configuration: CLOSURE_N held-out runs (nx=64, dt=0.05,
    4000 steps, parameters from vlapy_amd.closure.draw_instances) on the fine grid
    (nv=4096, FP64: the plain FP64 path), the coarse grid (nv=256, FP32) and the
    coarse grid with the trained closure; instances/s and errors against the fine
    grid."""
    import numpy as np
    import torch

    from vlapy_amd import closure as CL

    path = CLOSURE_NPZ
    if not os.path.exists(path):
        return None
    net = CL.Closure.load(path, max_rows=64)
    inst = CL.draw_instances(CLOSURE_N, CLOSURE_SEED)
    t = CL.DT * (np.arange(CLOSURE_STEPS) + 1)
    out = {"workload": "%d runs, k ~ U[0.25, 0.45], amplitude 10^U[-4,-1], nu 10^U[-5,-2] (Dougherty), nx=64, "
                       "dt=0.05, %d steps; fine grid nv=4096 FP64 vs coarse nv=256 FP32 with and without the "
                       "closure" % (CLOSURE_N, CLOSURE_STEPS)}
    runs = (("fine_fp64", CL.NV_FINE, "fp64", None, False), ("coarse_fp32", CL.NV, "fp32", None, False),
            ("coarse_fp32_closure", CL.NV, "fp32", net, False), ("coarse_fp32_delta_f", CL.NV, "fp32", None, True))
    ref = None
    for name, nv, prec, n, delta in runs:
        CL.rollout(CL.draw_instances(4, 99), nv, prec, 2, n, delta)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e2 = CL.rollout(inst, nv, prec, CLOSURE_STEPS, n, delta)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        r = {"instances_per_s": CLOSURE_N / sec, "seconds": sec,
             "density_drift_max": float(np.max(CL.rollout.density_drift))}
        if ref is None:
            ref = e2
        else:
            h = CL.history_errors(ref, e2)
            d = CL.damping_errors(inst, ref, e2, t)
            ok = ~np.isnan(d)
            r.update({"history_err_median": float(np.median(h)), "damping_fitted": int(ok.sum()),
                      "damping_err_median": float(np.median(d[ok])),
                      "damping_within_1pct": float(np.mean(d[ok] <= 0.01))})
        out[name] = r
    out["speedup_vs_fine_fp64"] = {k: out[k]["instances_per_s"] / out["fine_fp64"]["instances_per_s"]
                                   for k in ("coarse_fp32", "coarse_fp32_closure", "coarse_fp32_delta_f")}
    out["network"] = closure_network_rate(path)
    return out


def closure_network_rate(path, rows=262144, reps=5):
    """The closure's inference kernel alone: `rows` coarse rows (the
    4096-instance scan at nx = 64) through Closure.apply, HIP events over
    `reps` calls; TFLOP/s of its convolutions (2 * 5 * (64 + 64 * 64 * 2 + 64)
    * 256 flops per row) against the 2.5 PFLOP/s dense BF16 peak."""
    import numpy as np
    import torch

    from vlapy_amd import closure as CL

    net = CL.Closure.load(path, max_rows=rows)
    v = np.linspace(-6.4 + 0.025, 6.4 - 0.025, 256)
    rng = np.random.default_rng(7)
    f = (np.exp(-v ** 2 / 2)[None, :] / np.sqrt(2 * np.pi)) * (1 + 1e-2 * rng.standard_normal((rows, 256)))
    x = torch.as_tensor(f.astype(np.float32), device="cuda:0")
    net.apply(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        net.apply(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    flops = 2.0 * 5 * (64 + 64 * 64 * 2 + 64) * 256 * rows
    tf = flops / (ms / 1e3) / 1e12
    return {"rows": rows, "ms_per_call": ms, "tflops": tf, "bf16_peak_tflops": 2500.0, "frac": tf / 2500.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=SCAN_STEPS,
                    help="timed steps of the headline scan (a full run is %d)" % SCAN_STEPS)
    ap.add_argument("--warmup", type=int, default=2, help="untimed warmup steps of the headline scan")
    ap.add_argument("--epw-steps", type=int, default=NSTEPS)
    ap.add_argument("--epw-warmup", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", "--no-judge", dest="no_check", action="store_true",
                    help="skip the full-length FP64 / FP32 scans that check the damping rates (profiling runs)")
    ap.add_argument("--no-epw", action="store_true")
    ap.add_argument("--large-steps", type=int, default=200)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the reference inner-loop timing of the EPW run")
    ap.add_argument("--closure-network", action="store_true",
                    help="only time the closure's inference kernel (262144 rows) and print its line")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (for counter runs that should see only the timed shapes)")
    ap.add_argument("--copy-bytes", type=int, default=None,
                    help="only copy a device buffer of this many bytes 5 times (HBM counter calibration)")
    ap.add_argument("--scan-tuning-out", default=None, help="write the scan's launch shapes to this JSON file")
    ap.add_argument("--scan-tuning-in", default=None,
                    help="use the scan launch shapes of this JSON file instead of autotuning (counter runs)")
    ap.add_argument("--scan-profile-steps", type=int, default=SCAN_PROFILE_STEPS)
    ap.add_argument("--cpu-child", type=float, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--scan-cpu-child", type=str, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--oracle-child", type=str, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_child is not None:
        cpu_child(args.cpu_child)
        return 0
    if args.oracle_child is not None:
        oracle_child(args.oracle_child)
        return 0
    if args.scan_cpu_child is not None:
        scan_cpu_child(args.scan_cpu_child)
        return 0
    if args.closure_network:
        print(json.dumps({"closure_network": closure_network_rate(CLOSURE_NPZ)}))
        return 0
    if args.copy_bytes is not None:
        import torch

        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4, "copies": 5}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: the build has no multi-GPU path (the parameter scan "
              "would shard its instances over GPUs with no exchange)", file=sys.stderr)
        return 2
    if args.steps < 1:
        print("--steps must be positive", file=sys.stderr)
        return 2

    import numpy as np

    SCAN_TUNING["in"], SCAN_TUNING["out"] = args.scan_tuning_in, args.scan_tuning_out
    scan = scan_section(args.steps, args.warmup, not args.no_cpu, check=not args.no_check,
                        profile_steps=args.scan_profile_steps)
    head = scan["headline"]
    dr = scan["fp32_delta_f"]["damping_rate"]
    out = {
        "metric": METRIC,
        "value": head["runs_per_s"],
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "ragged parameter scan: 4096 instances, k ~ U[0.25, 0.45], amplitude 10^U[-4,-1], "
                        "nu 10^U[-5,-2] (Dougherty), nx in {32,64,128}, nv in {256,512,1024,2048} from "
                        "default_rng(20261202), v in [-6.4, 6.4], dt=0.1, %d steps per run, <E^2>(t) recorded; "
                        "value = completed runs per second over the timed steps" % SCAN_STEPS,
            "precision": "f32 storage of f - M(v) (delta-f) and f32 advection FFTs; f64 moments, Poisson and "
                         "tridiagonal",
            "damping_rate_vs_fp64": dr,
            "parallelism": "single GPU",
        },
        "roofline": scan["roofline"],
        "cpu_baseline": None,
        "scan": scan,
    }
    if not args.no_check:
        scan["drift_mfma"] = drift_mfma_section(scan["fp32_delta_f"]["tuning_raw"])
    if "cpu_baseline" in scan:
        a = scan["cpu_baseline"]["all_cores"]
        out["cpu_baseline"] = {"value": a["instances_per_s"], "unit": "instances/s", "cores": a["cores"],
                               "kind": a["kind"], "sample": a["sample"],
                               "one_core_value": scan["cpu_baseline"]["one_core"]["instances_per_s"]}
    if not args.no_epw:
        out["epw"] = epw_section(args.epw_steps, args.epw_warmup, min(args.profile_steps, args.epw_steps),
                                 not args.no_autotune, None if args.no_cpu else args.cpu_seconds,
                                 dropin=not args.no_dropin)
    if not args.no_large:
        out["large_grid"] = large_section(args.large_steps, not args.no_autotune, not args.no_cpu)
    if not args.no_epw:
        out["landau_c1"] = c1_section(not args.no_cpu)
    if not args.no_closure:
        out["closure"] = closure_section()
    out["scan"]["fp32_delta_f"].pop("tuning_raw", None)
    print(json.dumps(out))
    return 0


def epw_graph(precision, steps, chunk=100):
    """The EPW step loop (step_n with the E history, moments and series)
    captured for `chunk` steps in one HIP graph and replayed steps / chunk
    times: ms per step.  Every replay writes the same history rows."""
    import numpy as np
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import NMOMENTS, Plan

    dev = torch.device("cuda:0")
    x, kx, okx, v, kv, dv, f0 = initial_state(np)
    plan = Plan(NX, NV, precision, kx, okx, v, kv, dv, NU, DT, "dg")
    plan.autotune(20)
    f = torch.as_tensor(f0, dtype=ops.f_dtype(plan), device=dev).contiguous()
    e = torch.zeros(NX, dtype=torch.float64, device=dev)
    eh = torch.zeros(chunk, NX, dtype=torch.float64, device=dev)
    mo = torch.zeros(chunk, NMOMENTS, NX, dtype=torch.float64, device=dev)
    se = torch.zeros(chunk, 2, dtype=torch.float64, device=dev)
    ops.step_n(plan, f, e, chunk, None, eh, mo, se)  # warm-up
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            ops.step_n(plan, f, e, chunk, None, eh, mo, se)
    torch.cuda.current_stream().wait_stream(cap)
    torch.cuda.synchronize()
    reps = max(1, steps // chunk)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / (reps * chunk)


def dropin_section(steps=NSTEPS, steps_in_loop=1000, scheme="leapfrog"):
    """The EPW workload through the reference's inner-loop interface
    (outer_loop.get_inner_loop_stepper with backend "hip", FP64): batches of
    steps_in_loop steps, each filling the reference's temp_storage (fields,
    series, the k0/k1 f store) on the host.  ms per step, beside the device
    loop alone (step_n_store, or step_n_scheme for "pefrl" / "h-sixth")."""
    steps_in_loop = min(steps_in_loop, steps)
    import numpy as np
    import torch

    from vlapy_amd import initializers, outer_loop

    ap = initializers.make_default_params_dictionary()
    ap.update({"nx": NX, "nv": NV, "nu": NU, "xmin": 0.0, "xmax": 2 * np.pi / K0, "vmax": VMAX,
               "tmax": DT * steps, "nt": steps + 1})
    ap["fokker-planck"]["type"] = "dg"
    ap["vlasov-poisson"]["time"] = scheme

    class D:
        rules_to_store_f = {"space": ["k0", "k1"], "time": "all"}

    stuff = outer_loop.get_everything_ready_for_outer_loop(D, ap, {}, steps)
    x = stuff["x"]
    stuff["f"] = stuff["f"] * (1 + AMP * np.cos(K0 * x))[:, None]
    stepper = outer_loop.get_inner_loop_stepper(ap, stuff, steps_in_loop)
    ts = outer_loop.get_arrays_for_inner_loop(stuff, steps_in_loop, D.rules_to_store_f)
    stepper(stuff["t"][:steps_in_loop], stuff["driver"][:steps_in_loop], ts)  # warm-up batch
    torch.cuda.synchronize()
    ts = outer_loop.get_arrays_for_inner_loop(stuff, steps_in_loop, D.rules_to_store_f)
    t0 = time.perf_counter()
    for b in range(steps // steps_in_loop):
        sl = slice(b * steps_in_loop, (b + 1) * steps_in_loop)
        ts = stepper(stuff["t"][sl], stuff["driver"][sl], ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the device work of the same batches alone: step_n_store with the E
    # history, moments and the k0/k1 store on the stepper's plan
    from vlapy_amd import ops
    from vlapy_amd._lib import NMOMENTS
    dev = torch.device("cuda:0")
    f = torch.as_tensor(stuff["f"], device=dev)
    e = torch.zeros(NX, dtype=torch.float64, device=dev)
    eh = torch.empty(steps_in_loop, NX, dtype=torch.float64, device=dev)
    mo = torch.empty(steps_in_loop, NMOMENTS, NX, dtype=torch.float64, device=dev)
    st = torch.empty(steps_in_loop, 2, NV, dtype=torch.complex128, device=dev)

    def device_batch():
        if scheme == "leapfrog":
            ops.step_n_store(stepper.plan, f, e, steps_in_loop, st, 2, None, eh, mo)
        else:
            ops.step_n_scheme(stepper.plan, scheme, f, e, steps_in_loop, None, eh, mo, None, st, 2)

    device_batch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for b in range(steps // steps_in_loop):
        device_batch()
    torch.cuda.synchronize()
    dev_wall = time.perf_counter() - t1
    return {"workload": WORKLOAD + "; %s; reference inner-loop interface, %d-step batches, k0/k1 f store"
                                   % (scheme, steps_in_loop),
            "ms_per_step": wall / steps * 1e3, "steps": steps,
            "device_loop_ms_per_step": dev_wall / steps * 1e3, "vs_device_loop": wall / dev_wall}


def epw_section(steps, warmup, profile_steps, autotune, cpu_seconds, dropin=True):
    """The collisional EPW (nx=64, nv=1024) in FP32, FP64 and FP32 delta-f,
    each `steps` fused steps with the storage moments."""
    import numpy as np

    cpu = cpu_baseline(cpu_seconds) if cpu_seconds else None
    r32 = gpu_run("fp32", steps, warmup, profile_steps, autotune)
    r64 = gpu_run("fp64", steps, warmup, profile_steps, autotune)
    r32d = gpu_run("fp32", steps, warmup, profile_steps, autotune, delta=True)
    cells = NX * NV
    e64 = r64["e_hist"]

    def err(r):
        return {"E_max_rel": float(np.abs(r["e_hist"] - e64).max() / np.abs(e64).max()),
                "f_max_rel": float(np.abs(r["f"] - r64["f"]).max() / np.abs(r64["f"]).max())}

    def line(r, traffic):
        return {"cell_updates_per_s": cells * steps / (r["ms"] / 1e3), "ms_per_step": r["ms"] / steps,
                "roofline": roofline(r, cells, r["ms"] / steps, traffic), "tuning": r["tuning"]}

    out = {"workload": WORKLOAD, "steps": steps,
           "fp32": dict(line(r32, "traffic_c2_fp32.json"), vs_fp64_after_run=err(r32)),
           "fp64": line(r64, "traffic_c2_fp64.json"),
           "fp32_delta_f": dict(line(r32d, None), vs_fp64_after_run=err(r32d)),
           "cpu_baseline": cpu}
    # launch overhead removed: the same loop replayed from a HIP graph
    out["fp64"]["graph_ms_per_step"] = epw_graph("fp64", steps)
    out["fp32"]["graph_ms_per_step"] = epw_graph("fp32", steps)
    # the reference's own inner-loop interface on the same workload (FP64)
    if dropin:
        d = dropin_section(steps)
        d["vs_step_n"] = d["ms_per_step"] / out["fp64"]["ms_per_step"]
        out["dropin_fp64"] = d
        # the 4th-order PEFRL and 6th-order splittings through the same interface
        for scheme in ("pefrl", "h-sixth"):
            out["dropin_%s_fp64" % scheme] = dropin_section(min(steps, 1000), scheme=scheme)
    return out


if __name__ == "__main__":
    sys.exit(main())
