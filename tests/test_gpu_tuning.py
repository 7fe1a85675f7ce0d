"""Launch shapes and the autotuner.

Every launch shape the autotuner may pick (row values per thread 4, 8 or 16,
drift tile of 4, 8, 16 or 32 columns in workgroup or XCD order, 1, 2 or 4
tiles per drift workgroup, field solve on its own, fused into the drift or
formed in the row kernel) must give the oracle's step to the FP64
bar (1e-12 in f, 1e-11 in E), and the autotuner must leave the plan on one of
them without touching the caller's state.
"""

import itertools

import numpy as np
import pytest

from oracle import vpfp_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _setup(nx, nv, k=0.35, nu=1e-2, dt=0.1, op="dg", precision="fp64"):
    import torch
    from vlapy_amd._lib import Plan

    g = O.Grid.make(nx, nv, 2 * np.pi / k, 6.0)
    f0 = O.perturbed_maxwellian(g.x, g.v, k, 0.2)
    e0 = 0.01 * np.random.default_rng(nx + nv).standard_normal(nx)
    p = Plan(nx, nv, precision, g.kx, g.one_over_kx, g.v, g.kv, g.dv, nu, dt, op)
    dev = torch.device("cuda:0")
    return g, f0, e0, p, dev


def _oracle(g, f0, e0, n, dt=0.1, nu=1e-2, op="dg"):
    e, f = e0.copy(), f0.copy()
    for i in range(n):
        e, f = O.leapfrog_step(g, e, f, i * dt, dt)
        f = O.collide(f, g.v, nu, dt, g.dv, op)
    return e, f


@pytest.mark.parametrize("nx,nv", [(16, 64), (256, 64), (64, 1024), (512, 32), (8, 2048)])
def test_every_launch_shape_matches_oracle(nx, nv):
    import torch
    from vlapy_amd import ops

    n = 3
    g, f0, e0, p, dev = _setup(nx, nv)
    e_ref, f_ref = _oracle(g, f0, e0, n)
    ran = 0
    for q, tile, fuse, order, per in itertools.product((4, 8, 16), (4, 8, 16, 32), (0, 1, 2), (0, 1), (1, 2, 4)):
        if tile > nv:
            continue
        print("shape", nx, nv, q, tile, fuse, order, per, flush=True)
        p.set_tuning(row_q=q, drift_tile=tile, fuse_field=fuse, drift_order=order, drift_tiles=per)
        f = torch.as_tensor(f0, device=dev).clone()
        e = torch.as_tensor(e0, device=dev).clone()
        try:
            ops.step_n(p, f, e, n)
        except NotImplementedError:
            # the shape does not fit: drift workgroup above 1024 threads, the
            # fused field tail asked for nx > 256 or the row prologue for nx > 128
            m = nx // 2
            wg = (m // (4 if m <= 64 else 8)) * tile
            assert (fuse == 1 and nx > 256) or (fuse == 2 and nx > 128) or wg > 1024, (q, tile, fuse, order)
            continue
        torch.cuda.synchronize()
        assert rel(f.cpu().numpy(), f_ref) < 1e-12, (q, tile, fuse, order)
        assert rel(e.cpu().numpy(), e_ref) < 1e-11, (q, tile, fuse, order)
        ran += 1
    assert ran >= 8


def test_autotune_picks_a_valid_shape_and_leaves_state_alone():
    import torch
    from vlapy_amd import ops

    nx, nv = 64, 1024
    g, f0, e0, p, dev = _setup(nx, nv, precision="fp32")
    f = torch.as_tensor(f0, dtype=torch.float32, device=dev).clone()
    e = torch.as_tensor(e0, device=dev).clone()
    t = p.autotune(n_steps=10)
    print("tuning", t, flush=True)
    assert t["row_q"] in (4, 8, 16) and t["drift_tile"] in (4, 8, 16, 32) and t["fuse_field"] in (0, 1, 2)
    assert t["drift_order"] in (0, 1) and t["drift_tiles"] in (1, 2, 4)
    assert t["us_per_step"] > 0.0
    assert p.tuning() == t
    assert torch.equal(f, torch.as_tensor(f0, dtype=torch.float32, device=dev))
    assert torch.equal(e, torch.as_tensor(e0, device=dev))
    # the tuned plan still steps to the FP32 bar of a single step
    ops.step_n(p, f, e, 1)
    e_ref, f_ref = _oracle(g, f0, e0, 1)
    assert rel(f.double().cpu().numpy(), f_ref) < 2e-6


@pytest.mark.parametrize("grouped", [True, False])
def test_batch_autotune_keeps_parity(grouped):
    from vlapy_amd.scan import Scan, instance_grids

    p = {"k": np.array([0.3, 0.4, 0.35]), "amp": np.array([1e-2, 1e-3, 5e-2]),
         "nu": np.array([1e-3, 1e-4, 1e-2]), "nx": np.array([32, 32, 64]), "nv": np.array([256, 256, 512])}
    sc = Scan(p, dt=0.1, op="dg", precision="fp64", grouped=grouped)
    tun = sc.autotune(n_steps=3)
    print("batch tuning", tun, flush=True)
    assert set(tun) == {(32, 256), (64, 512)}
    for t in tun.values():
        assert t["row_q"] in (4, 8, 16) and t["drift_tile"] in (4, 8, 16, 32) and t["us_per_step"] > 0.0
    n = 10
    e2 = sc.run(n)
    for i in range(3):
        x, kx, okx, v, kv, dv = instance_grids(int(p["nx"][i]), p["k"][i], int(p["nv"][i]), 6.4)
        gi = O.Grid(x, kx, okx, v, kv, dv)
        f0 = O.perturbed_maxwellian(x, v, p["k"][i], p["amp"][i])
        _, _, eh = O.run_steps(gi, np.zeros(x.size), f0, n, 0.1, nu=p["nu"][i], fp_type="dg")
        ref = np.mean(eh ** 2, axis=1)
        assert np.abs(e2[:, i] - ref).max() / np.abs(ref).max() < 1e-11, i


@pytest.mark.parametrize("nx,nv,C", [(32, 256, 16), (32, 512, 32), (64, 256, 16), (64, 1024, 32)])
def test_mfma_drift_matches_oracle(nx, nv, C):
    """The FP32 x-advection as dense real-DFT products on MFMA (drift_dft):
    one vdfdx call within the FP32 bound of the Stockham drift (2e-6 of
    max f), and 20 fused steps with the collision and field within the FP32
    path's bound of the oracle."""
    import torch

    from vlapy_amd import ops
    from vlapy_amd._lib import Plan

    k, dt = 0.35, 0.1
    print("mfma drift", nx, nv, C, flush=True)
    g = O.Grid.make(nx, nv, 2 * np.pi / k, 6.4)
    f0 = O.perturbed_maxwellian(g.x, g.v, k, 0.1)
    p = Plan(nx, nv, "fp32", g.kx, g.one_over_kx, g.v, g.kv, g.dv, 1e-3, dt, "dg")
    p.set_tuning(drift_tile=C, drift_dft=1, fuse_field=0)
    assert p.tuning()["drift_dft"] == 1
    f = torch.as_tensor(f0, dtype=torch.float32, device="cuda").contiguous()
    out = ops.vdfdx(p, f, torch.empty_like(f), dt)
    ref = O.vdfdx(f0, dt, g.kx, g.v)
    err = rel(out.double().cpu().numpy(), ref)
    print("  one call: %.2e" % err, flush=True)
    assert err < 2e-6
    n = 20
    e = torch.zeros(nx, dtype=torch.float64, device="cuda")
    eh = torch.zeros(n, nx, dtype=torch.float64, device="cuda")
    ops.step_n(p, f, e, n, None, eh)
    _, f_o, eh_o = O.run_steps(g, np.zeros(nx), f0, n, dt, nu=1e-3, fp_type="dg")
    ef, ee = rel(f.double().cpu().numpy(), f_o), rel(eh.cpu().numpy(), eh_o)
    print("  20 steps: f %.2e E %.2e" % (ef, ee), flush=True)
    assert ef < 2e-5 and ee < 2e-3


def test_mfma_drift_refuses_unsupported_plans():
    from vlapy_amd._lib import Plan

    g = O.Grid.make(128, 256, 2 * np.pi / 0.3, 6.4)
    p = Plan(128, 256, "fp32", g.kx, g.one_over_kx, g.v, g.kv, g.dv, 0.0, 0.1, "lb")
    with pytest.raises(NotImplementedError):
        p.set_tuning(drift_dft=1)
    g = O.Grid.make(32, 256, 2 * np.pi / 0.3, 6.4)
    p = Plan(32, 256, "fp64", g.kx, g.one_over_kx, g.v, g.kv, g.dv, 0.0, 0.1, "lb")
    with pytest.raises(NotImplementedError):
        p.set_tuning(drift_dft=1)
