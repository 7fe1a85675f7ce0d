"""Step composition and per-step storage (vlapy/core/step.py) on the GPU."""

import numpy as np
import torch

from .. import ops
from .._lib import NMOMENTS, Plan
from . import _dev, collisions, field, vlasov, vlasov_poisson


def get_vlasov_poisson_step(all_params, stuff_for_time_loop):
    """Pushers + field solver + time integrator from all_params["vlasov-poisson"] (step.py:29-67)."""
    vp = all_params["vlasov-poisson"]
    vdfdx = vlasov.get_vdfdx(stuff_for_time_loop=stuff_for_time_loop, vdfdx_implementation=vp["vdfdx"])
    edfdv = vlasov.get_edfdv(stuff_for_time_loop=stuff_for_time_loop, edfdv_implementation=vp["edfdv"])
    field_solver = field.get_field_solver(stuff_for_time_loop=stuff_for_time_loop,
                                          field_solver_implementation=vp["poisson"])
    return vlasov_poisson.get_time_integrator(time_integrator_name=vp["time"], vdfdx=vdfdx, edfdv=edfdv,
                                              field_solver=field_solver, stuff_for_time_loop=stuff_for_time_loop)


def get_collision_step(stuff_for_time_loop, all_params):
    """Identity for nu == 0, implicit LB/DG step for nu > 0, NotImplementedError
    for nu < 0 (step.py:70-113).  The step runs as one fused HIP kernel
    (moments, diagonals and the tridiagonal solve per x row)."""
    nu = all_params["nu"]
    if nu == 0.0:

        def take_collision_step(f):
            return f

        return take_collision_step
    if not nu > 0.0:
        raise NotImplementedError
    fp = all_params["fokker-planck"]
    if fp["solver"] not in ("batched_tridiagonal", "naive"):
        raise NotImplementedError("Matrix Solver: <" + str(fp["solver"]) + "> has not yet been implemented")
    if fp["type"] not in ("lb", "dg"):
        raise NotImplementedError("Collision Operator: <" + str(fp["type"]) + "> has not yet been implemented")
    precision = _dev.precision_of(stuff_for_time_loop)
    dtype = _dev.f_dtype(precision)
    st = stuff_for_time_loop
    plan = Plan(st["nx"], st["nv"], precision, v=st["v"], dv=st["dv"], nu=st["nu"], dt=st["dt"], op=fp["type"])

    def take_collision_step(f):
        fd = _dev.to_dev(f, dtype)
        out = torch.empty_like(fd)
        ops.collide(plan, fd, out)
        return _dev.back(out, f)

    return take_collision_step


def get_f_update(store_f_rule):
    """f storage rule: the whole f ("all") or its first Fourier modes in x
    (["k0", ...]); step.py:116-140.  As in the reference the number of modes
    kept is len(store_f_rule), the number of keys of the rule."""
    if store_f_rule["space"] == "all":

        def get_f_to_store(f):
            return f

        return get_f_to_store
    if store_f_rule["space"][0] == "k0":
        nm = len(store_f_rule)
        plans = {}

        def get_f_to_store(f):
            fd = _dev.to_dev(f, torch.float64 if not _dev.is_tensor(f) or f.dtype == torch.float64 else torch.float32)
            nx, nv = fd.shape
            key = (nx, nv, fd.dtype)
            if key not in plans:
                plans[key] = Plan(nx, nv, "fp64" if fd.dtype == torch.float64 else "fp32")
            out = torch.empty(nm, nv, dtype=torch.complex128, device=fd.device)
            ops.x_modes(plans[key], fd, nm, out)
            return _dev.back(out, f)

        return get_f_to_store
    raise NotImplementedError


class _Moments:
    """Trapz moments of f on the GPU, one plan per (nx, storage dtype)."""

    def __init__(self, dv, v):
        self.dv, self.v = float(dv), np.asarray(v, dtype=np.float64)
        self.plans = {}

    def __call__(self, f):
        fd = _dev.to_dev(f, f.dtype if _dev.is_tensor(f) and f.dtype == torch.float32 else torch.float64)
        nx, nv = fd.shape
        prec = "fp32" if fd.dtype == torch.float32 else "fp64"
        key = (nx, prec)
        if key not in self.plans:
            self.plans[key] = Plan(nx, nv, prec, v=self.v, dv=self.dv)
        out = torch.empty(NMOMENTS, nx, dtype=torch.float64, device=fd.device)
        ops.moments(self.plans[key], fd, out)
        return out


def _put(arr, i, val):
    """arr[i] = val across numpy/tensor kinds."""
    if _dev.is_tensor(arr):
        arr[i] = val if _dev.is_tensor(val) else torch.as_tensor(np.asarray(val), device=arr.device, dtype=arr.dtype)
    else:
        arr[i] = val.detach().cpu().numpy() if _dev.is_tensor(val) else val


def _scalar(x):
    return float(x.item()) if _dev.is_tensor(x) else float(x)


def _mean_sq(e):
    if _dev.is_tensor(e):
        return float((e.double() ** 2).mean().item())
    return float(np.mean(np.asarray(e) ** 2.0))


_FIELD_ROWS = (("n", 0), ("j", 1), ("T", 2), ("q", 3), ("fv4", 4), ("vN", 5))


def get_fields_update(dv, v):
    """Per-step fields e, driver and the f v^p moments (step.py:143-175)."""
    mom = _Moments(dv, v)

    def update_fields(temp_storage_fields, e, de, f, i, moments=None):
        m = mom(f) if moments is None else moments
        _put(temp_storage_fields["e"], i, e)
        _put(temp_storage_fields["driver"], i, de)
        for key, row in _FIELD_ROWS:
            _put(temp_storage_fields[key], i, m[row])
        return temp_storage_fields

    return update_fields


def get_series_update(dv, v=None):
    """Per-step box averages (step.py:178-228).  mean_f2 and mean_flogf are
    x-averages of trapz(f^2) and trapz(f log f); the latter is NaN once f
    has a negative or zero sample, as in the reference."""
    mom = _Moments(dv, np.zeros(1) if v is None else v)

    def update_series(temp_storage, e, de, f, i, moments=None):
        m = (mom(f) if moments is None else moments).detach().cpu().numpy()
        s = temp_storage["series"]
        fields = temp_storage["fields"]

        def fmean(key):
            val = fields[key][i]
            return float(val.double().mean().item()) if _dev.is_tensor(val) else float(np.mean(val, axis=0))

        s["mean_n"][i] = fmean("n")
        s["mean_j"][i] = fmean("j")
        s["mean_T"][i] = fmean("T")
        s["mean_e2"][i] = _mean_sq(e)
        s["mean_de2"][i] = _mean_sq(de)
        s["mean_f2"][i] = np.mean(m[6])
        s["mean_flogf"][i] = np.mean(m[7])
        return s

    return update_series


def get_storage_step(stuff_for_time_loop):
    """Storage of f, e, fields and series after every step (step.py:231-283)."""
    dv, v = stuff_for_time_loop["dv"], stuff_for_time_loop["v"]
    store_f_function = get_f_update(store_f_rule=stuff_for_time_loop["rules_to_store_f"])
    mom = _Moments(dv, v)
    update_fields = get_fields_update(dv=dv, v=v)
    update_series = get_series_update(dv=dv, v=v)

    def storage_step(temp_storage, e, de, f, i):
        _put(temp_storage["stored_f"], i, store_f_function(f))
        temp_storage["e"] = e
        temp_storage["f"] = f
        m = mom(f)
        temp_storage["fields"] = update_fields(temp_storage_fields=temp_storage["fields"], de=de, e=e, f=f, i=i,
                                               moments=m)
        temp_storage["series"] = update_series(temp_storage=temp_storage, f=f, de=de, e=e, i=i, moments=m)
        return temp_storage

    return storage_step


def get_timestep(all_params, stuff_for_time_loop):
    """VP step, then FP step, then storage (step.py:286-328)."""
    vp_step = get_vlasov_poisson_step(all_params=all_params, stuff_for_time_loop=stuff_for_time_loop)
    fp_step = get_collision_step(all_params=all_params, stuff_for_time_loop=stuff_for_time_loop)
    storage_step = get_storage_step(stuff_for_time_loop=stuff_for_time_loop)

    def timestep(temp_storage, i):
        e = temp_storage["e"]
        f = temp_storage["f"]
        t = temp_storage["time_batch"][i]
        de = temp_storage["driver_array_batch"][i]
        e, f = vp_step(e=e, f=f, t=t)
        f = fp_step(f=f)
        temp_storage = storage_step(temp_storage=temp_storage, e=e, de=de, f=f, i=i)
        return temp_storage, i

    return timestep
