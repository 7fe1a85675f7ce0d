"""GPU parity of each C-ABI entry point against the CPU oracle.

FP64 bar: max |gpu - oracle| / max |oracle| <= 1e-12 (per call).
FP32 storage/advection bar (moments and tridiagonal in FP64): 2e-6 per call.
"""

import numpy as np
import pytest

from oracle import vpfp_oracle as O

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-6


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def dev():
    import torch

    assert torch.cuda.is_available(), "no GPU visible"
    from vlapy_amd import _lib

    _lib.library()
    return torch.device("cuda:0")


def T(a, dev, dtype=None):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device=dev)


def make_plan(g, nx, nv, precision="fp64", nu=0.0, dt=0.0, op="lb"):
    from vlapy_amd._lib import Plan

    return Plan(nx, nv, precision, g.kx, g.one_over_kx, g.v, g.kv, g.dv, nu, dt, op)


SIZES = [(4, 8), (8, 16), (16, 64), (64, 512), (64, 1024), (32, 2048), (8, 8192), (1024, 64), (3, 256)]


def _fdtype(prec):
    import torch

    return torch.float64 if prec == "fp64" else torch.float32


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,nv", SIZES)
def test_edfdv(dev, nx, nv, prec):
    from vlapy_amd import ops

    print("edfdv", nx, nv, prec, flush=True)
    rng = np.random.default_rng(nx * 7 + nv)
    g = O.Grid.make(nx, nv, 2 * np.pi / 0.3, 6.4)
    f = 1 + 0.3 * rng.standard_normal((nx, nv))
    e = rng.standard_normal(nx)
    p = make_plan(g, nx, nv, prec)
    fi = T(f, dev, _fdtype(prec))
    out = ops.edfdv(p, fi, fi.clone(), T(e, dev), 0.17)
    want = O.edfdv(fi.double().cpu().numpy(), e, 0.17, g.kv)
    assert rel(out.double().cpu().numpy(), want) < (F64_TOL if prec == "fp64" else F32_TOL)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,nv", [s for s in SIZES if s[0] >= 4] + [(1024, 8192), (2048, 16)])
def test_vdfdx(dev, nx, nv, prec):
    from vlapy_amd import ops

    print("vdfdx", nx, nv, prec, flush=True)
    rng = np.random.default_rng(nx * 5 + nv)
    g = O.Grid.make(nx, nv, 2 * np.pi / 0.3, 6.4)
    f = 1 + 0.3 * rng.standard_normal((nx, nv))
    p = make_plan(g, nx, nv, prec)
    fi = T(f, dev, _fdtype(prec))
    out = ops.vdfdx(p, fi, fi.clone(), 0.37)
    want = O.vdfdx(fi.double().cpu().numpy(), 0.37, g.kx, g.v)
    assert rel(out.double().cpu().numpy(), want) < (F64_TOL if prec == "fp64" else F32_TOL)


@pytest.mark.parametrize("nx,nv", [(4, 8), (64, 512), (96, 32), (7, 16), (1024, 64)])
def test_field_and_charges(dev, nx, nv):
    import torch
    from vlapy_amd import ops

    print("field", nx, nv, flush=True)
    rng = np.random.default_rng(nx + nv)
    g = O.Grid.make(nx, nv, 2 * np.pi / 0.25, 5.0)
    f = np.abs(1 + 0.2 * rng.standard_normal((nx, nv))) / (nv * g.dv)
    drv = 0.01 * rng.standard_normal(nx)
    p = make_plan(g, nx, nv)
    n = ops.compute_charges(p, T(f, dev), torch.zeros(nx, dtype=torch.float64, device=dev))
    assert rel(n.cpu().numpy(), O.compute_charges(f, g.dv)) < F64_TOL
    e = ops.field(p, T(f, dev), T(drv, dev), torch.zeros(nx, dtype=torch.float64, device=dev))
    assert rel(e.cpu().numpy(), O.field_solve(drv, f, g.dv, g.one_over_kx)) < 1e-11
    rho = 1 + 0.5 * np.sin(0.25 * g.x) + 0.2 * np.cos(2 * 0.25 * g.x)
    e2 = ops.solve_for_field(p, T(rho, dev), torch.zeros(nx, dtype=torch.float64, device=dev))
    assert rel(e2.cpu().numpy(), O.solve_for_field(rho, g.one_over_kx)) < F64_TOL


def test_field_closed_forms_reference(dev):
    """tests/test_fieldsolver.py:27-54 of the reference on the GPU (nx=96)."""
    import torch
    from vlapy_amd import ops

    nx, kp = 96, 0.25
    _, x, kx, okx = O.spatial_grid(0.0, 2 * np.pi / kp, nx)
    from vlapy_amd._lib import Plan

    p = Plan(nx, 8, "fp64", kx, okx, np.zeros(8), np.fft.fftfreq(8) * 2 * np.pi, 1.0)
    dens = [1 + np.sin(kp * x), 1 + np.cos(2 * kp * x), 1 + np.sin(2 * kp * x) + np.cos(8 * kp * x)]
    want = [np.cos(kp * x) / kp, -np.sin(2 * kp * x) / 2 / kp, np.cos(2 * kp * x) / 2 / kp - np.sin(8 * kp * x) / 8 / kp]
    for n, w in zip(dens, want):
        e = ops.solve_for_field(p, T(1.0 - n, dev), torch.zeros(nx, dtype=torch.float64, device=dev))
        np.testing.assert_allclose(e.cpu().numpy(), w, atol=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("op", ["lb", "dg"])
@pytest.mark.parametrize("nx,nv", [(2, 1024), (8, 8), (64, 1024), (16, 8192), (5, 256)])
def test_collide(dev, nx, nv, op, prec):
    from vlapy_amd import ops

    print("collide", nx, nv, op, prec, flush=True)
    rng = np.random.default_rng(nx + 3 * nv)
    g = O.Grid.make(nx, nv, 2 * np.pi / 0.3, 6.0)
    f = O.perturbed_maxwellian(g.x, g.v, 0.3, 0.2) * (1 + 0.05 * rng.standard_normal((nx, nv)))
    f = np.abs(f)
    nu, dt = 1e-2, 0.1
    p = make_plan(g, nx, nv, prec, nu=nu, dt=dt, op=op)
    fi = T(f, dev, _fdtype(prec))
    out = ops.collide(p, fi, fi.clone())
    want = O.collide(fi.double().cpu().numpy(), g.v, nu, dt, g.dv, op)
    assert rel(out.double().cpu().numpy(), want) < (F64_TOL if prec == "fp64" else F32_TOL)


@pytest.mark.parametrize("op", ["lb", "dg"])
def test_diagonals_and_tridiag(dev, op):
    import torch
    from vlapy_amd import ops

    nx, nv = 6, 512
    g = O.Grid.make(nx, nv, 2 * np.pi / 0.3, 6.0)
    rng = np.random.default_rng(11)
    f = np.abs(O.perturbed_maxwellian(g.x, g.v, 0.3, 0.3) * (1 + 0.1 * rng.standard_normal((nx, nv))))
    p = make_plan(g, nx, nv, nu=3e-2, dt=0.2, op=op)
    z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=dev)  # noqa: E731
    a, b, c = ops.collision_diagonals(p, T(f, dev), z(nx, nv - 1), z(nx, nv), z(nx, nv - 1))
    wa, wb, wc = O.collision_diagonals(f, g.v, 3e-2, 0.2, g.dv, op)
    for got, want in ((a, wa), (b, wb), (c, wc)):
        assert rel(got.cpu().numpy(), want) < F64_TOL
    x = ops.tridiag_solve(p, a, b, c, T(f, dev), z(nx, nv))
    assert rel(x.cpu().numpy(), O.thomas(wa, wb, wc, f)) < F64_TOL


@pytest.mark.parametrize("nx,nv,nu,op,n", [(16, 64, 0.0, "lb", 12), (16, 64, 1e-2, "dg", 12), (16, 64, 1e-2, "lb", 12),
                                          (256, 64, 1e-2, "dg", 4), (1024, 1024, 1e-3, "dg", 2)])
def test_step_n_matches_oracle_loop(dev, nx, nv, nu, op, n):
    """Fused step vs the oracle's leapfrog + collision loop; nx >= 128 runs the
    field solve on several workgroups, nx = 1024 the drift on column-tile
    strides."""
    import torch
    from vlapy_amd import ops

    k, dt = 0.35, 0.1
    print("step_n", nx, nv, nu, op, flush=True)
    g = O.Grid.make(nx, nv, 2 * np.pi / k, 6.0)
    f0 = O.perturbed_maxwellian(g.x, g.v, k, 0.2)
    rng = np.random.default_rng(3)
    e0 = 0.01 * rng.standard_normal(nx)
    drv = 1e-3 * rng.standard_normal((n, nx))
    p = make_plan(g, nx, nv, nu=nu, dt=dt, op=op)
    f = T(f0, dev)
    e = T(e0, dev)
    eh = torch.zeros(n, nx, dtype=torch.float64, device=dev)
    mo = torch.zeros(n, 8, nx, dtype=torch.float64, device=dev)
    se = torch.zeros(n, 2, dtype=torch.float64, device=dev)
    ops.step_n(p, f, e, n, T(drv, dev), eh, mo, se)
    # oracle
    ef, ff = e0.copy(), f0.copy()
    for i in range(n):
        ef, ff = O.leapfrog_step(g, ef, ff, i * dt, dt, driver=lambda t, i=i: drv[i])
        ff = O.collide(ff, g.v, nu, dt, g.dv, op)
        assert rel(eh[i].cpu().numpy(), ef) < 1e-11, i
        m = O.step_moments(ff, g.v, g.dv)
        for q, key in enumerate(("n", "j", "T", "q", "fv4", "vN")):
            scale = O.trapz_v(np.abs(ff) * np.abs(g.v) ** q, g.dv).max()
            assert np.abs(mo[i, q].cpu().numpy() - m[key]).max() <= 1e-12 * scale, (i, key)
        assert abs(se[i, 0].item() - np.mean(ef ** 2)) <= 1e-11 * np.mean(ef ** 2)
        assert abs(se[i, 1].item() - np.mean(drv[i] ** 2)) <= 1e-12 * np.mean(drv[i] ** 2)
    assert rel(f.cpu().numpy(), ff) < F64_TOL
    assert rel(e.cpu().numpy(), ef) < 1e-11


@pytest.mark.parametrize("nu,delta,n_modes,nx,nv", [(1e-2, False, 0, 32, 256), (1e-2, False, 3, 32, 256),
                                                     (0.0, False, 0, 32, 256), (0.0, False, 2, 32, 256),
                                                     (1e-2, True, 0, 32, 256), (1e-2, True, 2, 32, 256),
                                                     (1e-3, False, 2, 64, 1024)])
def test_step_n_store_matches_stepwise_copies(dev, nu, delta, n_modes, nx, nv):
    """vpfp_step_n_store over n steps equals n single steps of vpfp_step_n
    with f (or its first x-modes from vpfp_x_modes) taken after each: the
    row kernel writes f between the collision and the next half kick (f
    itself under delta-f storage) and the next step's field launch, or one
    launch after the last step, sums its x-modes in vpfp_x_modes' order."""
    import torch
    from vlapy_amd import ops

    k, dt, n = 0.35, 0.1, 6
    print("step_n_store nu", nu, "delta", delta, "modes", n_modes, "grid", nx, nv, flush=True)
    g = O.Grid.make(nx, nv, 2 * np.pi / k, 6.0)
    f0 = O.perturbed_maxwellian(g.x, g.v, k, 0.2)
    m = np.exp(-g.v ** 2 / 2.0) / np.sqrt(2 * np.pi)
    ref = make_plan(g, nx, nv, nu=nu, dt=dt, op="dg")
    f = T(f0, dev)
    e = torch.zeros(nx, dtype=torch.float64, device=dev)
    want = []
    for i in range(n):
        ops.step_n(ref, f, e, 1)
        if n_modes:
            out = torch.zeros(n_modes, nv, dtype=torch.complex128, device=dev)
            want.append(ops.x_modes(ref, f, n_modes, out).cpu().numpy())
        else:
            want.append(f.cpu().numpy().copy())
    p = make_plan(g, nx, nv, nu=nu, dt=dt, op="dg")
    if delta:
        p.set_delta(True)
    f2 = T(f0 - m[None, :] if delta else f0, dev)
    e2 = torch.zeros(nx, dtype=torch.float64, device=dev)
    if n_modes:
        st = torch.zeros(n, n_modes, nv, dtype=torch.complex128, device=dev)
    else:
        st = torch.zeros(n, nx, nv, dtype=torch.float64, device=dev)
    ops.step_n_store(p, f2, e2, n, st, n_modes)
    got = st.cpu().numpy()
    tol = 1e-12 if delta else 1e-15
    for i in range(n):
        assert rel(got[i], want[i]) < tol, i


def test_unsupported_sizes_raise(dev):
    import torch
    from vlapy_amd import ops

    g = O.Grid.make(6, 12, 2 * np.pi / 0.3, 6.0)
    p = make_plan(g, 6, 12)
    f = torch.zeros(6, 12, dtype=torch.float64, device=dev)
    with pytest.raises(NotImplementedError):
        ops.vdfdx(p, f, f, 0.1)
    with pytest.raises(NotImplementedError):
        ops.edfdv(p, f, f, torch.zeros(6, dtype=torch.float64, device=dev), 0.1)
    with pytest.raises(ValueError):
        ops.edfdv(p, f, f, torch.zeros(5, dtype=torch.float64, device=dev), 0.1)


@pytest.mark.parametrize("scheme,nu,delta,n_modes", [("pefrl", 1e-2, False, 0), ("pefrl", 0.0, False, 2),
                                                      ("h-sixth", 1e-2, False, 0), ("h-sixth", 0.0, False, 0),
                                                      ("h-sixth", 1e-2, True, 2)])
def test_step_n_scheme_matches_oracle(dev, scheme, nu, delta, n_modes):
    """vpfp_step_n_scheme (PEFRL / 6th-order splitting, each step followed by
    the Dougherty collision) against the oracle's pefrl_step / hsixth_step and
    collide over 6 steps with a driver and a non-zero initial field: the E
    history, the moments after every step, the f store and f, FP64.  The
    fused kicks (last kick, collision, next step's first kick in one launch;
    one kick by the summed shift without a collision) change only rounding."""
    import torch
    from vlapy_amd import ops
    from vlapy_amd.outer_loop import _scheme_field_times

    nx, nv, k, dt, n = 32, 256, 0.35, 0.1, 6
    print("step_n_scheme", scheme, "nu", nu, "delta", delta, "modes", n_modes, flush=True)
    g = O.Grid.make(nx, nv, 2 * np.pi / k, 6.0)
    f0 = O.perturbed_maxwellian(g.x, g.v, k, 0.2)
    e0 = O._field(g, 0.0, f0)
    pulses = {"p": {"k0": k, "w0": 1.2, "t_L": 0.0, "t_R": 20.0, "t_wL": 2.5, "t_wR": 2.5, "a0": 1e-2}}
    drv_fn = lambda t: O.driver_field(g.x, pulses, t)
    step = O.INTEGRATORS[scheme]
    f, e = f0.copy(), e0.copy()
    eh, mo, st = [], [], []
    for i in range(n):
        e, f = step(g, e, f, i * dt, dt, drv_fn)
        f = O.collide(f, g.v, nu, dt, g.dv, "dg")
        eh.append(e.copy())
        m = O.step_moments(f, g.v, g.dv)
        mo.append(np.stack([m[key] for key in ("n", "j", "T", "q", "fv4", "vN")]))
        st.append(np.fft.fft(f, axis=0)[:n_modes] if n_modes else f.copy())
    times = _scheme_field_times(scheme, dt * np.arange(n), dt)
    drv = np.stack([[drv_fn(tt) for tt in row] for row in times])
    p = make_plan(g, nx, nv, nu=nu, dt=dt, op="dg")
    m_bg = np.exp(-g.v ** 2 / 2.0) / np.sqrt(2 * np.pi)
    if delta:
        p.set_delta(True)
    fd = T(f0 - m_bg[None, :] if delta else f0, dev)
    ed = T(e0, dev)
    ehd = torch.zeros(n, nx, dtype=torch.float64, device=dev)
    mod = torch.zeros(n, 8, nx, dtype=torch.float64, device=dev)
    std = (torch.zeros(n, n_modes, nv, dtype=torch.complex128, device=dev) if n_modes
           else torch.zeros(n, nx, nv, dtype=torch.float64, device=dev))
    ops.step_n_scheme(p, scheme, fd, ed, n, T(drv, dev), ehd, mod, None, std, n_modes)
    got_f = fd.cpu().numpy() + (m_bg[None, :] if delta else 0.0)
    tol = 1e-11 if delta else 1e-12
    assert rel(got_f, f) < tol
    assert rel(ehd.cpu().numpy(), np.array(eh)) < 1e-10
    assert rel(ed.cpu().numpy(), eh[-1]) < 1e-10
    assert rel(mod[:, :6].cpu().numpy(), np.array(mo)) < 1e-11
    assert rel(std.cpu().numpy(), np.array(st)) < tol


def test_step_n_scheme_refuses_leapfrog_and_bad_shapes(dev):
    import torch
    from vlapy_amd import ops

    g = O.Grid.make(16, 64, 2 * np.pi / 0.3, 6.0)
    p = make_plan(g, 16, 64, nu=1e-2, dt=0.1, op="dg")
    f = T(O.perturbed_maxwellian(g.x, g.v, 0.3, 0.1), dev)
    e = torch.zeros(16, dtype=torch.float64, device=dev)
    with pytest.raises(NotImplementedError):
        ops.step_n_scheme(p, "leapfrog", f, e, 2)
    with pytest.raises(ValueError):
        ops.step_n_scheme(p, "pefrl", f, e, 2, torch.zeros(2, 4, 16, dtype=torch.float64, device=dev))
