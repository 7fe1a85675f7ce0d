"""Batched ragged parameter scan on the GPU against per-instance runs.

Each instance of a small ragged scan (several size classes, per-instance k,
amplitude and nu) must match the oracle's single-instance leapfrog +
collision loop (FP64 batch: 1e-11 in the <E^2> history and f after 20 steps)
and the FP32 batch must stay within its stated error: f is stored in FP32 and
each step makes four FP32 transform passes of <= 10 radix levels, so the
bound on f is 20 steps x 4 passes x ~4 ulp(f_max) ~ 2e-5 relative.
"""

import numpy as np
import pytest

from oracle import vpfp_oracle as O

pytestmark = pytest.mark.gpu


def _params():
    p = {"k": np.array([0.25, 0.31, 0.37, 0.45, 0.29, 0.41]),
         "amp": np.array([1e-4, 1e-3, 1e-2, 1e-1, 3e-2, 5e-3]),
         "nu": np.array([1e-5, 1e-4, 1e-3, 1e-2, 3e-3, 0.0]),
         "nx": np.array([32, 32, 64, 64, 32, 128]),
         "nv": np.array([256, 256, 512, 256, 1024, 256])}
    return p


def _oracle(p, i, n, dt, vmax=6.4, op="dg"):
    from vlapy_amd.scan import instance_grids

    x, kx, okx, v, kv, dv = instance_grids(int(p["nx"][i]), p["k"][i], int(p["nv"][i]), vmax)
    g = O.Grid(x, kx, okx, v, kv, dv)
    f0 = O.perturbed_maxwellian(x, v, p["k"][i], p["amp"][i])
    e, f, eh = O.run_steps(g, np.zeros(x.size), f0, n, dt, nu=p["nu"][i], fp_type=op)
    return f, np.mean(eh ** 2, axis=1)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-3)])
def test_scan_matches_single_instance_oracle(precision, tol):
    from vlapy_amd.scan import Scan

    p = _params()
    n, dt = 20, 0.1
    sc = Scan(p, dt=dt, op="dg", precision=precision)
    assert len(sc.classes) == 5
    e2 = sc.run(n)
    for i in range(p["k"].size):
        f_o, e2_o = _oracle(p, i, n, dt)
        err = np.abs(e2[:, i] - e2_o).max() / np.abs(e2_o).max()
        print(precision, "instance", i, "E^2 rel err", err)
        assert err < tol, (i, err)
        c = [c for c in sc.classes if i in c.idx][0]
        j = int(np.nonzero(c.idx == i)[0][0])
        f_g = c.f[j].double().cpu().numpy()
        assert np.abs(f_g - f_o).max() / np.abs(f_o).max() < (1e-12 if precision == "fp64" else 2e-5)


def test_scan_c3_draw_runs_and_conserves_density():
    """The 4096-instance ragged draw (seed 20261202), 5 steps, FP32: runs,
    finite, density drift per instance below 1e-6."""
    from vlapy_amd.scan import Scan, draw_parameters

    p = draw_parameters(4096, 20261202)
    sc = Scan(p, dt=0.1, op="dg", precision="fp32")
    n0 = sc.density()
    e2 = sc.run(5)
    assert np.all(np.isfinite(e2))
    n1 = sc.density()
    assert np.max(np.abs(n1 / n0 - 1)) < 1e-6


def test_scan_graph_replay_matches_eager():
    """Five steps captured in a HIP graph and replayed four times (with eager
    work and synchronisations between the replays) give the eager 20-step
    state bit for bit, after every replay."""
    import torch
    from vlapy_amd.scan import Scan

    p = _params()
    eager = Scan(p, dt=0.1, op="dg", precision="fp32")
    graph = Scan(p, dt=0.1, op="dg", precision="fp32")
    n = graph.capture(5)
    for r in range(4):
        eager.run(n, record=False)
        graph.replay()
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for ce, cg in zip(eager.classes, graph.classes):
            assert torch.equal(ce.f, cg.f), (r, ce.nx, ce.nv)
            assert torch.equal(ce.e, cg.e), (r, ce.nx, ce.nv)


def test_scan_lenard_bernstein_matches_oracle():
    """The scan with the Lenard-Bernstein operator (op "lb"), FP64: every
    instance's <E^2> history within 1e-11 of the oracle's single-instance run."""
    from vlapy_amd.scan import Scan

    p = _params()
    n, dt = 10, 0.1
    sc = Scan(p, dt=dt, op="lb", precision="fp64")
    e2 = sc.run(n)
    for i in range(p["k"].size):
        _, e2_o = _oracle(p, i, n, dt, op="lb")
        err = np.abs(e2[:, i] - e2_o).max() / np.abs(e2_o).max()
        assert err < 1e-11, (i, err)


@pytest.mark.parametrize("row_q", [0, 4, 8, 16])
@pytest.mark.parametrize("precision,tol_e2,tol_f", [("fp64", 1e-11, 1e-12), ("fp32", 2e-3, 2e-5)])
def test_scan_delta_f_matches_oracle(precision, tol_e2, tol_f, row_q):
    """Delta-f storage (f - M stored, vpfp_batch_set_delta) runs the same
    algorithm: FP64 within the plain scan's 1e-11 of the oracle, f (the
    deviation plus the background, vpfp_batch_background) within 1e-12; FP32
    within the plain FP32 scan's bounds.  Every row-kernel shape of the
    delta-f rows: row_q 0 (the row-count rule, Q = 4 at these sizes), 4, 8
    and 16 (which runs the Q = 8 kernel under delta-f storage)."""
    from vlapy_amd.scan import Scan

    p = _params()
    n, dt = 20, 0.1
    sc = Scan(p, dt=dt, op="dg", precision=precision, delta=True)
    tun = sc.shapes()
    for t in tun.values():
        t["row_q"] = row_q
    sc.set_tuning(tun)
    e2 = sc.run(n)
    for i in range(p["k"].size):
        f_o, e2_o = _oracle(p, i, n, dt)
        err = np.abs(e2[:, i] - e2_o).max() / np.abs(e2_o).max()
        print(precision, "delta-f instance", i, "E^2 rel err", err)
        assert err < tol_e2, (i, err)
        c = [c for c in sc.classes if i in c.idx][0]
        j = int(np.nonzero(c.idx == i)[0][0])
        f_g = c.full_f()[j].double().cpu().numpy()
        assert np.abs(f_g - f_o).max() / np.abs(f_o).max() < tol_f, i
    # the stored deviation is small next to f itself
    c = sc.classes[0]
    assert float(c.f.abs().max()) < 0.2 * float(c.full_f().abs().max())


def test_scan_delta_f_refuses_closure_and_round_trips_background():
    import ctypes

    import torch

    from vlapy_amd import _lib
    from vlapy_amd.closure import Closure
    from vlapy_amd.scan import Scan

    p = {key: val[:2] for key, val in _params().items()}
    sc = Scan(p, dt=0.1, precision="fp32", delta=True)
    c = sc.classes[0]
    net = Closure(max_rows=16)
    rc = _lib.library().vpfp_batch_set_closure(c.handle, net.handle)
    assert rc == _lib.VPFP_EUNSUPPORTED
    # FP64 batch: background added then removed leaves the array to rounding
    sc64 = Scan(p, dt=0.1, precision="fp64", delta=False)
    c64 = sc64.classes[0]
    f = c64.f.clone()
    lib = _lib.library()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.vpfp_batch_background(c64.handle, ctypes.c_void_p(f.data_ptr()), -1, st))
    assert float(f.abs().max()) < 0.01 * float(c64.f.abs().max()) + 1e-6
    _lib.check(lib.vpfp_batch_background(c64.handle, ctypes.c_void_p(f.data_ptr()), 1, st))
    torch.cuda.synchronize()
    assert float((f - c64.f).abs().max()) < 1e-15


@pytest.mark.parametrize("precision,delta,dft", [("fp64", False, 0), ("fp32", False, 0), ("fp32", True, 0),
                                               ("fp32", True, 1)])
def test_grouped_step_equals_per_class_batches(precision, delta, dft):
    """The grouped step (one row launch per nv, one drift and one field launch
    per nx, vpfp_group_step_n) equals the per-class batches (vpfp_batch_step_n)
    bit for bit when both use the same launch shapes: the kernels run the
    same arithmetic per row and per column tile."""
    import torch

    from vlapy_amd.scan import Scan

    p = _params()
    shape = {"row_q": 8, "drift_tile": 16, "fuse_field": 0, "drift_order": 0, "drift_tiles": 2, "us_per_step": 0.0}
    if dft:
        # the MFMA drift exists for nx 32 and 64 only: the nx = 128 class keeps
        # the Stockham drift
        p = {key: val[p["nx"] < 128] for key, val in p.items()}
        shape["drift_dft"] = 1
    runs = []
    for grouped in (True, False):
        sc = Scan(p, dt=0.1, op="dg", precision=precision, delta=delta, grouped=grouped)
        sc.set_tuning({(c.nx, c.nv): dict(shape) for c in sc.classes})
        e2 = sc.run(12)
        torch.cuda.synchronize()
        runs.append((e2, [c.f.clone() for c in sc.classes], [c.e.clone() for c in sc.classes]))
    (e2g, fg, eg), (e2c, fc, ec) = runs
    assert np.array_equal(e2g, e2c)
    for a, b in zip(fg, fc):
        assert torch.equal(a, b)
    for a, b in zip(eg, ec):
        assert torch.equal(a, b)


def test_grouped_scan_launches_per_step():
    """A grouped step makes one row launch per distinct nv and one drift and
    one field launch per distinct nx: 10 launches for the 4 x 3 size classes
    of the scan configuration (kernel_times counts them)."""
    from vlapy_amd.scan import Scan

    p = {"k": np.full(12, 0.3), "amp": np.full(12, 1e-3), "nu": np.full(12, 1e-4),
         "nx": np.repeat([32, 64, 128], 4), "nv": np.tile([256, 512, 1024, 2048], 3)}
    sc = Scan(p, dt=0.1, precision="fp32", delta=True)
    assert len(sc.classes) == 12
    sc.run(1, record=False)
    sc.set_profiling(True)
    sc.run(5, record=False)
    kt = sc.kernel_times()
    launches = sum(v[1] for d in kt.values() for v in d.values())
    # 5 steps: one opening half-kick row launch per nv, then 10 per step
    assert launches == 4 + 5 * 10, kt


def test_scan_mfma_drift_matches_oracle():
    """A grouped FP32 delta-f scan with the MFMA drift on its nx = 32 and 64
    classes matches the oracle within the FP32 scan's bounds."""
    from vlapy_amd.scan import Scan

    p = _params()
    n, dt = 20, 0.1
    sc = Scan(p, dt=dt, op="dg", precision="fp32", delta=True)
    tun = sc.shapes()
    for (nx, nv), t in tun.items():
        if nx in (32, 64):
            t.update(drift_tile=16, drift_dft=1)
    sc.set_tuning(tun)
    assert all(t["drift_dft"] == (1 if k[0] in (32, 64) else 0) for k, t in sc.shapes().items())
    e2 = sc.run(n)
    for i in range(p["k"].size):
        f_o, e2_o = _oracle(p, i, n, dt)
        err = np.abs(e2[:, i] - e2_o).max() / np.abs(e2_o).max()
        print("mfma drift scan instance", i, "E^2 rel err", err)
        assert err < 2e-3, (i, err)
        c = [c for c in sc.classes if i in c.idx][0]
        j = int(np.nonzero(c.idx == i)[0][0])
        f_g = c.full_f()[j].double().cpu().numpy()
        assert np.abs(f_g - f_o).max() / np.abs(f_o).max() < 2e-5, i
