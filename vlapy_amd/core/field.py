"""Field solve (vlapy/core/field.py) on the GPU."""

import numpy as np
import torch

from .. import ops
from .._lib import Plan
from . import _dev

_plans = {}


def _plan(key, **kw):
    p = _plans.get(key)
    if p is None:
        p = _plans[key] = Plan(**kw)
    return p


def compute_charges(f, dv):
    """n(x) = trapz_v f (field.py:27-36)."""
    is_t = _dev.is_tensor(f)
    dtype = f.dtype if is_t and f.dtype in (torch.float32, torch.float64) else torch.float64
    fd = _dev.to_dev(f, dtype)
    nx, nv = fd.shape
    prec = "fp32" if dtype == torch.float32 else "fp64"
    p = _plan(("charges", nx, nv, float(dv), prec), nx=nx, nv=nv, precision=prec, dv=float(dv))
    out = torch.empty(nx, dtype=torch.float64, device=fd.device)
    ops.compute_charges(p, fd, out)
    return _dev.back(out, f)


def solve_for_field(charge_density, one_over_kx):
    """E = Re ifft(i (1/kx) fft(1 - n)) (field.py:39-63); any nx >= 2."""
    okx = np.asarray(one_over_kx.cpu() if _dev.is_tensor(one_over_kx) else one_over_kx, dtype=np.float64)
    nx = okx.size
    p = _plan(("poisson", nx, okx.tobytes()), nx=nx, nv=4, one_over_kx=okx)
    n = _dev.field_vector(charge_density, nx)
    out = torch.empty(nx, dtype=torch.float64, device=n.device)
    ops.solve_for_field(p, n, out)
    return _dev.back(out, charge_density)


def get_spectral_solver(dv, one_over_kx, precision="fp64"):
    """Closure (driver_field, f) -> driver + E[f] (field.py:66-88)."""
    okx = np.asarray(one_over_kx, dtype=np.float64)
    dtype = _dev.f_dtype(precision)
    plans = {}

    def solve_total_electric_field(driver_field, f):
        fd = _dev.to_dev(f, dtype)
        nx, nv = fd.shape
        if nv not in plans:
            plans[nv] = Plan(nx, nv, precision, one_over_kx=okx, dv=float(dv))
        out = torch.empty(nx, dtype=torch.float64, device=fd.device)
        drv = _dev.field_vector(driver_field, nx)
        ops.field(plans[nv], fd, drv, out)
        return _dev.back(out, f)

    return solve_total_electric_field


def get_field_solver(stuff_for_time_loop, field_solver_implementation="spectral"):
    """Name dispatch (field.py:91-107)."""
    if field_solver_implementation == "spectral":
        return get_spectral_solver(dv=stuff_for_time_loop["dv"], one_over_kx=stuff_for_time_loop["one_over_kx"],
                                   precision=_dev.precision_of(stuff_for_time_loop))
    raise NotImplementedError
