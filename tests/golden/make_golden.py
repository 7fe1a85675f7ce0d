"""Write golden input/output vectors for the parts of the path that closed
forms and invariants do not pin: the compositions of the split integrators
(leapfrog, PEFRL, h-sixth, with a non-zero incoming field and a driver), the
full timestep with its storage epilogue, and the collision step on a
non-unit-density input.

It imports ``vlapy.core`` from a VlaPy source tree (default
/root/reference), which needs only numpy and scipy, and writes small .npz
fixtures next to this file.  Run once, in a container that holds the
reference; the fixtures are committed and the reference is not needed after.

    python tests/golden/make_golden.py [path-to-vlapy-tree]
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def grid(nx, nv, k, vmax):
    xmax = 2.0 * np.pi / k
    dx = xmax / nx
    x = np.linspace(dx / 2.0, xmax - dx / 2.0, nx)
    kx = np.fft.fftfreq(nx, d=dx) * 2.0 * np.pi
    okx = np.zeros_like(kx)
    okx[1:] = 1.0 / kx[1:]
    dv = 2 * vmax / nv
    v = np.linspace(-vmax + dv / 2.0, vmax - dv / 2.0, nv)
    kv = np.fft.fftfreq(nv, d=dv) * 2.0 * np.pi
    return x, kx, okx, v, kv, dv


def initial_f(x, v, k, rng):
    dens = 1.0 + 0.2 * np.cos(k * x) + 0.05 * np.sin(2 * k * x + 0.3)
    shape = np.exp(-((v - 0.4) ** 2) / 2.0) + 0.3 * np.exp(-((v + 1.5) ** 2) / 0.5)
    f = dens[:, None] * shape[None, :]
    return f * (1.0 + 1e-3 * rng.standard_normal(f.shape))


PULSE = {"p": {"k0": 0.35, "w0": 1.2, "t_L": 0.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 0.05}}


def main(tree):
    sys.path.insert(0, tree)
    from vlapy.core import field, step, vlasov, vlasov_poisson  # noqa: F401

    rng = np.random.default_rng(20261202)
    nx, nv, k, vmax = 16, 32, 0.35, 6.0
    x, kx, okx, v, kv, dv = grid(nx, nv, k, vmax)
    f0 = initial_f(x, v, k, rng)
    e0 = 0.02 * rng.standard_normal(nx)
    t0, dt = 1.3, 0.1

    def driver_function(t):
        tot = np.zeros(nx)
        for p in PULSE.values():
            env = 0.5 * (np.tanh((t - p["t_L"]) / p["t_wL"]) - np.tanh((t - p["t_R"]) / p["t_wR"]))
            tot += env * p["k0"] * p["a0"] * np.sin(p["k0"] * x - p["w0"] * t)
        return tot

    stuff = {
        "kx": kx, "x": x, "v": v, "kv": kv, "dv": dv, "one_over_kx": okx,
        "nx": nx, "nv": nv, "dt": dt, "driver_function": driver_function,
    }
    out = {"x": x, "kx": kx, "one_over_kx": okx, "v": v, "kv": kv, "dv": dv,
           "f0": f0, "e0": e0, "t0": t0, "dt": dt}
    vdfdx = vlasov.get_vdfdx(stuff, "exponential")
    edfdv = vlasov.get_edfdv(stuff, "exponential")
    fs = field.get_field_solver(stuff, "spectral")
    out["vdfdx_f"] = vdfdx(f0, dt)
    out["edfdv_f"] = edfdv(f0, e0, dt)
    out["field_e"] = fs(driver_function(t0), f0)
    for name in ("leapfrog", "pefrl", "h-sixth"):
        vp = vlasov_poisson.get_time_integrator(name, vdfdx, edfdv, fs, stuff)
        e1, f1 = vp(e=e0, f=f0, t=t0)
        out[name + "_e"], out[name + "_f"] = e1, f1

    # collision step on a non-unit density, shifted input
    for op in ("lb", "dg"):
        fc = 1.7 * np.exp(-((v - 0.8) ** 2) / 2.0 / 1.3)[None, :] * np.array([1.0, 0.6, 1.2])[:, None]
        st = {"f": fc, "v": v, "nv": nv, "nx": 3, "nu": 3e-2, "dt": 0.2, "dv": dv}
        ap = {"fokker-planck": {"type": op, "solver": "batched_tridiagonal"}, "nu": 3e-2}
        fp = step.get_collision_step(st, ap)
        out["coll_" + op + "_in"] = fc
        out["coll_" + op + "_out"] = fp(fc.copy())

    # three full timesteps (VP + DG collisions + storage epilogue)
    nt = 3
    stuff_ts = dict(stuff)
    stuff_ts.update({"nu": 1e-2, "rules_to_store_f": {"space": "all"}})
    ap = {
        "vlasov-poisson": {"time": "leapfrog", "vdfdx": "exponential", "edfdv": "exponential", "poisson": "spectral"},
        "fokker-planck": {"type": "dg", "solver": "batched_tridiagonal"},
        "nu": 1e-2,
    }
    ts = step.get_timestep(ap, stuff_ts)
    tb = t0 + dt * np.arange(nt)
    drv = np.stack([driver_function(t) for t in tb])
    temp = {
        "time_batch": tb, "driver_array_batch": drv, "e": e0.copy(), "f": f0.copy(),
        "stored_f": np.zeros((nt, nx, nv)),
        "series": {kk: np.zeros(nt) for kk in
                   ("mean_n", "mean_j", "mean_T", "mean_e2", "mean_de2", "mean_f2", "mean_flogf")},
        "fields": {kk: np.zeros((nt, nx)) for kk in ("e", "driver", "n", "j", "T", "q", "fv4", "vN")},
    }
    for i in range(nt):
        temp, _ = ts(temp, i)
    out["ts_driver"] = drv
    out["ts_f"] = temp["f"]
    out["ts_e"] = temp["e"]
    for kk, val in temp["fields"].items():
        out["ts_fields_" + kk] = val
    for kk, val in temp["series"].items():
        out["ts_series_" + kk] = val

    path = os.path.join(HERE, "vlapy_core_small.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "keys:", len(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
