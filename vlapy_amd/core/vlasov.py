"""Advection pushers (vlapy/core/vlasov.py) on the GPU.

get_vdfdx / get_edfdv keep the reference's factory -> closure pattern and
names; the exponential (spectral phase-shift) pushers run as one fused HIP
kernel each.  The semi-Lagrangian and centred-difference alternates are not
implemented on this backend and raise NotImplementedError at factory time,
like any unknown name in the reference.
"""

import numpy as np
import torch

from .. import ops
from .._lib import Plan
from . import _dev


def get_vdfdx_exponential(kx, v, precision="fp64"):
    """x-advection f <- Re ifft_x(exp(-i kx v dt) fft_x f) (vlasov.py:83-110)."""
    kx = np.asarray(kx, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    dtype = _dev.f_dtype(precision)
    plan = Plan(kx.size, v.size, precision, kx=kx, v=v)

    def step_vdfdx_exponential(f, dt):
        fd = _dev.to_dev(f, dtype)
        out = torch.empty_like(fd)
        ops.vdfdx(plan, fd, out, dt)
        return _dev.back(out, f)

    return step_vdfdx_exponential


def get_edfdv_exponential(kv, precision="fp64"):
    """v-advection f <- Re ifft_v(exp(-i kv e dt) fft_v f) (vlasov.py:113-140).

    The plan is made for each nx the closure is called with (the reference
    closure captures kv only).
    """
    kv = np.asarray(kv, dtype=np.float64)
    dtype = _dev.f_dtype(precision)
    plans = {}

    def step_edfdv_exponential(f, e, dt):
        fd = _dev.to_dev(f, dtype)
        nx = fd.shape[0]
        if nx not in plans:
            plans[nx] = Plan(nx, kv.size, precision, kv=kv)
        out = torch.empty_like(fd)
        ops.edfdv(plans[nx], fd, out, _dev.field_vector(e, nx), dt)
        return _dev.back(out, f)

    return step_edfdv_exponential


def get_vdfdx(stuff_for_time_loop, vdfdx_implementation="exponential"):
    """Name dispatch of the x pusher (vlasov.py:213-235)."""
    if vdfdx_implementation == "exponential":
        return get_vdfdx_exponential(kx=stuff_for_time_loop["kx"], v=stuff_for_time_loop["v"],
                                     precision=_dev.precision_of(stuff_for_time_loop))
    raise NotImplementedError(
        "v df/dx: <" + str(vdfdx_implementation) + "> has not yet been implemented on the HIP backend"
    )


def get_edfdv(stuff_for_time_loop, edfdv_implementation="exponential"):
    """Name dispatch of the v pusher (vlasov.py:238-260)."""
    if edfdv_implementation == "exponential":
        return get_edfdv_exponential(kv=stuff_for_time_loop["kv"], precision=_dev.precision_of(stuff_for_time_loop))
    raise NotImplementedError(
        "e df/dv: <" + str(edfdv_implementation) + "> has not yet been implemented on the HIP backend"
    )
