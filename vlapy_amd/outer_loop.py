"""Inner-loop orchestration (vlapy/outer_loop.py) for the HIP backend.

all_params["backend"]["core"] = "hip" selects this backend (the reference
accepts only "numpy", outer_loop.py:49-56).  The inner loop keeps the
reference's temp_storage schema (outer_loop.py:149-215) with host numpy
arrays, so the reference's storage manager and diagnostics can consume it.

With exponential pushers and the spectral field solve, a batch of steps runs
as one call of a fused device loop: vpfp_step_n_store for leapfrog (three
kernels per step, f resident in HBM, the per-step f store or its x-modes
written by the row kernel), vpfp_step_n_scheme for PEFRL and the 6th-order
splitting (their drifts, field solves and kicks as device launches, the
collision, moments and f store fused into the row launch that ends a step).
Any other combination runs the per-operator closures of vlapy_amd.core step
by step.
"""

import numpy as np
import torch

from . import field_driver, initializers, ops
from ._lib import NMOMENTS, Plan
from .core import _dev, step, vlasov_poisson

BACKENDS = ("hip",)


def _check_backend(all_params):
    core = all_params["backend"]["core"]
    if core not in BACKENDS:
        raise NotImplementedError("The backend <" + str(core) + "> has not yet been implemented")


def get_sim_config_and_inner_loop_step(all_params, stuff_for_time_loop, nt_in_loop, store_f_rules):
    """Backend switch, inner-loop stepper and its storage (outer_loop.py:31-64)."""
    _check_backend(all_params)
    do_inner_loop = get_inner_loop_stepper(all_params, stuff_for_time_loop, nt_in_loop)
    sim_config = get_arrays_for_inner_loop(stuff_for_time_loop, nt_in_loop, store_f_rules)
    return sim_config, do_inner_loop


def get_everything_ready_for_outer_loop(diagnostics, all_params, pulse_dictionary, overall_num_steps):
    """Grids, initial f, time axis, driver and e = 0 (outer_loop.py:67-146)."""
    f = initializers.initialize_distribution(nx=all_params["nx"], nv=all_params["nv"], vmax=all_params["vmax"])
    dx, x, kx, one_over_kx = initializers.initialize_spatial_quantities(
        xmin=all_params["xmin"], xmax=all_params["xmax"], nx=all_params["nx"])
    dv, v, kv = initializers.initialize_velocity_quantities(vmax=all_params["vmax"], nv=all_params["nv"])
    t_dummy = np.linspace(0, all_params["tmax"], all_params["nt"])
    dt = t_dummy[1] - t_dummy[0]
    t = dt * np.arange(overall_num_steps)
    driver_function = field_driver.get_driver_function(x=x, pulse_dictionary=pulse_dictionary, np=np)
    driver_array = field_driver.make_driver_array(function=driver_function, x_axis=x, time_axis=t)
    return {
        "e": np.zeros(x.size),
        "f": f,
        "nx": all_params["nx"],
        "kx": kx,
        "x": x,
        "one_over_kx": one_over_kx,
        "v": v,
        "kv": kv,
        "nv": all_params["nv"],
        "dv": dv,
        "driver": driver_array,
        "driver_function": driver_function,
        "dt": dt,
        "nu": all_params["nu"],
        "t": t,
        "rules_to_store_f": diagnostics.rules_to_store_f,
        "vlasov-poisson": all_params["vlasov-poisson"],
        "fokker-planck": all_params["fokker-planck"],
        "precision": all_params["backend"].get("precision", "fp64"),
    }


def get_arrays_for_inner_loop(stuff_for_time_loop, nt_in_loop, store_f_rules, this_np=np):
    """temp_storage for one batch of nt_in_loop steps (outer_loop.py:149-215)."""
    f0 = np.asarray(stuff_for_time_loop["f"], dtype=np.float64)
    e0 = np.asarray(stuff_for_time_loop["e"], dtype=np.float64)
    if store_f_rules["space"] == "all":
        store_f = np.zeros((nt_in_loop,) + f0.shape, dtype=np.float64)
        store_f[0] = f0
    elif isinstance(store_f_rules["space"], list) and store_f_rules["space"][0] == "k0":
        store_f = np.zeros((nt_in_loop, len(store_f_rules["space"]), f0.shape[1]), dtype=np.complex64)
        store_f[0] = np.fft.fft(f0, axis=0)[: len(store_f_rules["space"])]
    else:
        raise NotImplementedError
    shape = (nt_in_loop,) + e0.shape
    return {
        "time_batch": np.zeros(nt_in_loop),
        "e": np.array(e0),
        "f": np.array(f0),
        "stored_f": store_f,
        "mean_cum_de2_previous": 0.0,
        "series": {k: np.zeros(nt_in_loop) for k in
                   ("mean_n", "mean_j", "mean_T", "mean_e2", "mean_de2", "mean_f2", "mean_flogf")},
        "fields": {k: np.zeros(shape) for k in ("e", "driver", "n", "j", "T", "q", "fv4", "vN")},
    }


def post_inner_loop_update(temp_storage, this_np=np):
    """Driver energy bookkeeping after a batch (outer_loop.py:218-245)."""
    s = temp_storage["series"]
    s["mean_cum_de2"] = temp_storage["mean_cum_de2_previous"] + this_np.cumsum(s["mean_de2"])
    s["mean_t_plus_e2_minus_cum_de2"] = s["mean_T"] + (s["mean_e2"] - s["mean_cum_de2"])
    s["mean_t_plus_e2_plus_cum_de2"] = s["mean_T"] + s["mean_e2"] + s["mean_de2"]
    temp_storage["mean_cum_de2_previous"] = s["mean_cum_de2"][-1]
    return temp_storage


def _fused_eligible(all_params):
    vp = all_params["vlasov-poisson"]
    return (vp["time"] in ("leapfrog",) + tuple(ops.SCHEMES) and vp["vdfdx"] == "exponential"
            and vp["edfdv"] == "exponential" and vp["poisson"] == "spectral")


def _scheme_field_times(scheme, time_array, dt):
    """(n, 5) times of the five field solves of each step, as the core
    integrators form them: PEFRL the running sum t + dt1 + dt2 + ...
    (vlasov_poisson.py:115-146), the 6th-order scheme t + a_k dt
    (vlasov_poisson.py:206-226)."""
    t = np.asarray(time_array, dtype=np.float64)
    out = np.empty((t.size, ops.SCHEME_FIELDS))
    if scheme == "pefrl":
        w = (vlasov_poisson._XSI, vlasov_poisson._CHI, 1.0 - 2.0 * (vlasov_poisson._CHI + vlasov_poisson._XSI),
             vlasov_poisson._CHI, vlasov_poisson._XSI)
        tc = t.copy()
        for j in range(ops.SCHEME_FIELDS):
            tc = tc + w[j] * dt
            out[:, j] = tc
    else:
        a1, a2, a3 = vlasov_poisson._H6[:3]
        for j, a in enumerate((a1, a2, a3, a2, a1)):
            out[:, j] = t + a * dt
    return out


class _FusedInnerLoop:
    """Batches of (leapfrog, PEFRL or 6th-order) + collision + storage steps
    on the device.

    One vpfp_step_n_store (leapfrog: three launches per step) or
    vpfp_step_n_scheme call per batch, then one synchronisation: the E
    history, the moments and the f store come back by asynchronous copies
    into page-locked host arrays that this stepper owns and puts into
    temp_storage (reused by every batch, as the reference's storage step
    overwrites its arrays in place)."""

    def __init__(self, all_params, stuff):
        nu = float(all_params["nu"])
        if nu < 0.0:
            raise NotImplementedError
        fp = all_params.get("fokker-planck", {"type": "lb", "solver": "batched_tridiagonal"})
        op = fp["type"] if nu > 0.0 else "lb"
        if nu > 0.0 and fp["solver"] not in ("batched_tridiagonal", "naive"):
            raise NotImplementedError("Matrix Solver: <" + str(fp["solver"]) + "> has not yet been implemented")
        self.precision = stuff.get("precision", all_params["backend"].get("precision", "fp64"))
        self.dt = float(stuff["dt"])
        self.plan = Plan(stuff["nx"], stuff["nv"], self.precision, stuff["kx"], stuff["one_over_kx"], stuff["v"],
                         stuff["kv"], stuff["dv"], nu, self.dt, op)
        self.driver_function = stuff["driver_function"]
        self.rules = stuff["rules_to_store_f"]
        self.nx, self.nv = int(stuff["nx"]), int(stuff["nv"])
        self.scheme = all_params["vlasov-poisson"]["time"]
        self._bufs = {}

    def _driver(self, time_array):
        """driver_function(t + dt) for every step, (n, nx)."""
        return self._driver_at(np.asarray(time_array, dtype=np.float64) + self.dt)

    def _driver_at(self, t):
        """driver_function at each time of t, (t.size, nx): one call on a
        column of times when the function broadcasts over it (the driver of
        field_driver does), checked against single calls at the first and
        last time; otherwise one call per time, as the reference."""
        t = np.asarray(t, dtype=np.float64).reshape(-1)
        n, nx = t.size, self.nx

        def one(ti):
            return np.broadcast_to(np.asarray(self.driver_function(ti), dtype=np.float64), (nx,))

        try:
            d = np.broadcast_to(np.asarray(self.driver_function(t[:, None]), dtype=np.float64), (n, nx))
            for i in sorted({0, n - 1}):
                ref = one(t[i])
                if np.max(np.abs(d[i] - ref)) > 1e-14 * np.max(np.abs(ref)):
                    raise ValueError
            return d
        except (ValueError, TypeError, IndexError):
            return np.stack([one(ti) for ti in t])

    def _buffers(self, n, f_dtype, dev):
        """Device outputs and page-locked host arrays for batches of n steps."""
        key = (n, f_dtype)
        if key not in self._bufs:
            nx, nv = self.nx, self.nv
            space = self.rules["space"]
            nm = 0 if space == "all" else len(space)
            pin = dict(pin_memory=True)
            b = {"nm": nm,
                 "e_hist": torch.empty(n, nx, dtype=torch.float64, device=dev),
                 "mom": torch.empty(n, NMOMENTS, nx, dtype=torch.float64, device=dev),
                 "h_e": torch.empty(n, nx, dtype=torch.float64, **pin),
                 "h_mom": torch.empty(NMOMENTS, n, nx, dtype=torch.float64, **pin)}
            if nm == 0:
                b["store"] = torch.empty(n, nx, nv, dtype=f_dtype, device=dev)
                b["h_store"] = torch.empty(n, nx, nv, dtype=torch.float64, **pin)
            else:
                b["store"] = torch.empty(n, nm, nv, dtype=torch.complex128, device=dev)
                b["h_store"] = torch.empty(n, nm, nv, dtype=torch.complex64, **pin)
            self._bufs[key] = b
        return self._bufs[key]

    def __call__(self, time_array, driver_array, temp_storage):
        dev = _dev.device()
        n = len(time_array)
        temp_storage["time_batch"] = time_array
        temp_storage["driver_array_batch"] = driver_array
        f = _dev.to_dev(temp_storage["f"], _dev.f_dtype(self.precision))
        e = _dev.to_dev(temp_storage["e"], torch.float64)
        if self.scheme == "leapfrog":
            drv = self._driver(time_array)
        else:
            drv = self._driver_at(_scheme_field_times(self.scheme, time_array, self.dt)).reshape(
                n, ops.SCHEME_FIELDS, self.nx)
        drv_d = torch.as_tensor(np.array(drv), device=dev) if np.any(drv != 0.0) else None  # writable copy
        b = self._buffers(n, f.dtype, dev)
        # the whole batch in one fused call; the row kernel writes f (or its
        # copy for the x-modes) after every step's collision into the store
        if self.scheme == "leapfrog":
            ops.step_n_store(self.plan, f, e, n, b["store"], b["nm"], drv_d, b["e_hist"], b["mom"])
        else:
            ops.step_n_scheme(self.plan, self.scheme, f, e, n, drv_d, b["e_hist"], b["mom"], None, b["store"],
                              b["nm"])
        b["h_e"].copy_(b["e_hist"], non_blocking=True)
        b["h_mom"].copy_(b["mom"].transpose(0, 1), non_blocking=True)
        if b["nm"]:
            b["h_store"].copy_(b["store"].to(torch.complex64), non_blocking=True)
        else:
            b["h_store"].copy_(b["store"].double(), non_blocking=True)
        f_host = f.double().cpu().numpy()  # synchronises the stream
        eh, mo = b["h_e"].numpy(), b["h_mom"].numpy()
        flds, ser = temp_storage["fields"], temp_storage["series"]
        flds["e"] = eh
        flds["driver"][:] = np.asarray(driver_array)
        for key, row in (("n", 0), ("j", 1), ("T", 2), ("q", 3), ("fv4", 4), ("vN", 5)):
            flds[key] = mo[row]
        ser["mean_n"][:] = mo[0].mean(axis=1)
        ser["mean_j"][:] = mo[1].mean(axis=1)
        ser["mean_T"][:] = mo[2].mean(axis=1)
        ser["mean_e2"][:] = np.mean(eh ** 2.0, axis=1)
        ser["mean_de2"][:] = np.mean(np.asarray(driver_array) ** 2.0, axis=1)
        ser["mean_f2"][:] = mo[6].mean(axis=1)
        ser["mean_flogf"][:] = mo[7].mean(axis=1)
        hs = b["h_store"].numpy()
        if hs.dtype == temp_storage["stored_f"].dtype and hs.shape == temp_storage["stored_f"].shape:
            temp_storage["stored_f"] = hs
        else:
            temp_storage["stored_f"][:] = hs.astype(temp_storage["stored_f"].dtype, copy=False)
        temp_storage["e"] = eh[-1].copy()
        temp_storage["f"] = f_host
        post_inner_loop_update(temp_storage, np)
        return temp_storage


def get_inner_loop_stepper(all_params, stuff_for_time_loop, steps_in_loop):
    """The inner loop over steps_in_loop steps (outer_loop.py:248-281)."""
    _check_backend(all_params)
    if _fused_eligible(all_params):
        return _FusedInnerLoop(all_params, stuff_for_time_loop)

    one_step = step.get_timestep(all_params=all_params, stuff_for_time_loop=stuff_for_time_loop)

    def inner_loop(time_array, driver_array, temp_storage):
        temp_storage["time_batch"] = time_array
        temp_storage["driver_array_batch"] = driver_array
        for it in range(steps_in_loop):
            temp_storage, _ = one_step(temp_storage, it)
        for key in ("e", "f"):
            if _dev.is_tensor(temp_storage[key]):
                temp_storage[key] = temp_storage[key].double().cpu().numpy()
        post_inner_loop_update(temp_storage, np)
        return temp_storage

    return inner_loop
