"""Device-tensor operations over a Plan: shape/dtype checks on the host, then
one C-ABI call on torch's current stream.

All tensors are torch CUDA (HIP) tensors.  f is (nx, nv) in the plan's
storage precision (float64 for fp64, float32 for fp32); fields, moments and
diagonals are float64.
"""

import torch

from . import _lib


def _stream():
    return torch.cuda.current_stream().cuda_stream


def f_dtype(plan):
    return torch.float64 if plan.precision == "fp64" else torch.float32


def _check(t, shape, dtype, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
    if t.dtype != dtype:
        raise ValueError("%s has dtype %s, expected %s" % (name, t.dtype, dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if not t.is_cuda:
        raise ValueError("%s must be a device tensor" % name)
    return t.data_ptr()


def _f(plan, f, name="f"):
    return _check(f, (plan.nx, plan.nv), f_dtype(plan), name)


def _x(plan, a, name):
    return _check(a, (plan.nx,), torch.float64, name)


def vdfdx(plan, f_in, f_out, dt):
    _lib.check(_lib.library().vpfp_vdfdx(plan.handle, _f(plan, f_in), _f(plan, f_out, "f_out"), float(dt), _stream()))
    return f_out


def edfdv(plan, f_in, f_out, e, dt):
    _lib.check(_lib.library().vpfp_edfdv(plan.handle, _f(plan, f_in), _f(plan, f_out, "f_out"), _x(plan, e, "e"),
                                         float(dt), _stream()))
    return f_out


def compute_charges(plan, f, n_out):
    _lib.check(_lib.library().vpfp_compute_charges(plan.handle, _f(plan, f), _x(plan, n_out, "n"), _stream()))
    return n_out


def solve_for_field(plan, charge_density, e_out):
    _lib.check(_lib.library().vpfp_solve_for_field(plan.handle, _x(plan, charge_density, "charge_density"),
                                                   _x(plan, e_out, "e"), _stream()))
    return e_out


def field(plan, f, driver, e_out):
    drv = _x(plan, driver, "driver_field") if driver is not None else None
    _lib.check(_lib.library().vpfp_field(plan.handle, _f(plan, f), drv, _x(plan, e_out, "e"), _stream()))
    return e_out


def collision_diagonals(plan, f, a, b, c):
    nx, nv = plan.nx, plan.nv
    _lib.check(_lib.library().vpfp_collision_diagonals(
        plan.handle, _f(plan, f), _check(a, (nx, nv - 1), torch.float64, "a"),
        _check(b, (nx, nv), torch.float64, "b"), _check(c, (nx, nv - 1), torch.float64, "c"), _stream()))
    return a, b, c


def tridiag_solve(plan, a, b, c, d, x):
    nx, nv = plan.nx, plan.nv
    _lib.check(_lib.library().vpfp_tridiag_solve(
        plan.handle, _check(a, (nx, nv - 1), torch.float64, "a"), _check(b, (nx, nv), torch.float64, "b"),
        _check(c, (nx, nv - 1), torch.float64, "c"), _f(plan, d, "d"), _f(plan, x, "x"), _stream()))
    return x


def collide(plan, f_in, f_out):
    _lib.check(_lib.library().vpfp_collide(plan.handle, _f(plan, f_in), _f(plan, f_out, "f_out"), _stream()))
    return f_out


def step_n(plan, f, e, n_steps, driver=None, e_hist=None, moments=None, series=None):
    """Fused leapfrog + collision steps; f and e updated in place."""
    nx = plan.nx
    n = int(n_steps)
    drv = _check(driver, (n, nx), torch.float64, "driver") if driver is not None else None
    eh = _check(e_hist, (n, nx), torch.float64, "e_hist") if e_hist is not None else None
    mo = _check(moments, (n, _lib.NMOMENTS, nx), torch.float64, "moments") if moments is not None else None
    se = _check(series, (n, 2), torch.float64, "series") if series is not None else None
    _lib.check(_lib.library().vpfp_step_n(plan.handle, _f(plan, f), _x(plan, e, "e"), n, drv, eh, mo, se,
                                          _stream()))
    return f, e


def step_n_store(plan, f, e, n_steps, f_store, n_modes=0, driver=None, e_hist=None, moments=None, series=None):
    """step_n with the per-step f store (vpfp_step_n_store): n_modes == 0:
    f_store (n, nx, nv) in the storage dtype receives f after each step;
    n_modes > 0: f_store (n, n_modes, nv) complex128 the first n_modes
    x-modes of f after each step."""
    nx, nv = plan.nx, plan.nv
    n = int(n_steps)
    nm = int(n_modes)
    drv = _check(driver, (n, nx), torch.float64, "driver") if driver is not None else None
    eh = _check(e_hist, (n, nx), torch.float64, "e_hist") if e_hist is not None else None
    mo = _check(moments, (n, _lib.NMOMENTS, nx), torch.float64, "moments") if moments is not None else None
    se = _check(series, (n, 2), torch.float64, "series") if series is not None else None
    if nm == 0:
        st = _check(f_store, (n, nx, nv), f_dtype(plan), "f_store")
    else:
        st = _check(f_store, (n, nm, nv), torch.complex128, "f_store")
    _lib.check(_lib.library().vpfp_step_n_store(plan.handle, _f(plan, f), _x(plan, e, "e"), n, drv, eh, mo, se, st,
                                                nm, _stream()))
    return f, e


SCHEMES = {"pefrl": 1, "h-sixth": 2}
SCHEME_FIELDS = 5


def step_n_scheme(plan, scheme, f, e, n_steps, driver=None, e_hist=None, moments=None, series=None, f_store=None,
                  n_modes=0):
    """n steps of the PEFRL ("pefrl") or 6th-order ("h-sixth") splitting, each
    followed by the collision step (vpfp_step_n_scheme): driver (n, 5, nx)
    float64 at the step's five field solves; e_hist, moments, series as
    step_n; f_store as step_n_store."""
    if scheme not in SCHEMES:
        raise NotImplementedError("df/dt : <" + str(scheme) + "> has not yet been implemented")
    nx, nv = plan.nx, plan.nv
    n = int(n_steps)
    nm = int(n_modes)
    drv = _check(driver, (n, SCHEME_FIELDS, nx), torch.float64, "driver") if driver is not None else None
    eh = _check(e_hist, (n, nx), torch.float64, "e_hist") if e_hist is not None else None
    mo = _check(moments, (n, _lib.NMOMENTS, nx), torch.float64, "moments") if moments is not None else None
    se = _check(series, (n, 2), torch.float64, "series") if series is not None else None
    st = None
    if f_store is not None:
        st = (_check(f_store, (n, nx, nv), f_dtype(plan), "f_store") if nm == 0
              else _check(f_store, (n, nm, nv), torch.complex128, "f_store"))
    elif nm:
        raise ValueError("n_modes > 0 needs f_store")
    _lib.check(_lib.library().vpfp_step_n_scheme(plan.handle, SCHEMES[scheme], _f(plan, f), _x(plan, e, "e"), n, drv,
                                                 eh, mo, se, st, nm, _stream()))
    return f, e


def moments(plan, f, out):
    """Trapz moments [8][nx]: n, j, T, q, fv4, vN, f^2, f log f."""
    _lib.check(_lib.library().vpfp_moments(plan.handle, _f(plan, f),
                                           _check(out, (_lib.NMOMENTS, plan.nx), torch.float64, "moments"),
                                           _stream()))
    return out


def x_modes(plan, f, n_modes, out):
    """First n_modes Fourier modes along x, complex128 [n_modes][nv]."""
    _lib.check(_lib.library().vpfp_x_modes(plan.handle, _f(plan, f), int(n_modes),
                                           _check(out, (int(n_modes), plan.nv), torch.complex128, "modes"),
                                           _stream()))
    return out


def field_noalloc(nx, nv, precision, dv, one_over_kx, f, driver, e_out):
    """E = driver + field of f with no plan and no device allocation
    (vpfp_field_noalloc): one_over_kx (nx,) float64, f (nx, nv) float64 or
    float32 (precision "fp64" / "fp32"), driver (nx,) or None, e_out (nx,)
    float64, all device tensors.  Shares one scratch array per GPU with every
    other call: do not overlap calls on one GPU."""
    nx, nv = int(nx), int(nv)
    fdt = torch.float64 if precision == "fp64" else torch.float32
    drv = _check(driver, (nx,), torch.float64, "driver_field") if driver is not None else None
    _lib.check(_lib.library().vpfp_field_noalloc(
        nx, nv, _lib.VPFP_F64 if precision == "fp64" else _lib.VPFP_F32, float(dv),
        _check(one_over_kx, (nx,), torch.float64, "one_over_kx"), _check(f, (nx, nv), fdt, "f"), drv,
        _check(e_out, (nx,), torch.float64, "e"), _stream()))
    return e_out
