"""Batched parameter scans: thousands of independent VPFP instances with
uneven (nx, nv), wavenumber, amplitude and collision frequency advanced
together on one GPU.

Instances are grouped by size class (nx, nv); each class is one batch of the
C ABI (vpfp_batch_*), whose f lives in HBM as [n_inst][nx][nv].  By default
the classes form one group (vpfp_group_*) stepped on one stream: one launch
of the fused kick-collide-kick rows per distinct nv, one drift (with
per-workgroup partial densities) and one field launch per distinct nx, so
the 12 classes of a 4096-instance draw take 10 launches per step.  With
grouped=False each class takes its own three launches (rows, drift, field)
on its own HIP stream.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .core import _dev


def draw_parameters(n, seed, nx_choices=(32, 64, 128), nv_choices=(256, 512, 1024, 2048)):
    """k ~ U[0.25, 0.45], amp = 10^U[-4, -1], nu = 10^U[-5, -2], nx and nv
    uniform over the choices; drawn in that order from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.25, 0.45, n)
    amp = 10.0 ** rng.uniform(-4.0, -1.0, n)
    nu = 10.0 ** rng.uniform(-5.0, -2.0, n)
    nx = rng.choice(np.asarray(nx_choices), n)
    nv = rng.choice(np.asarray(nv_choices), n)
    return {"k": k, "amp": amp, "nu": nu, "nx": nx, "nv": nv}


def instance_grids(nx, k, nv, vmax):
    """x, kx, 1/kx of a box [0, 2 pi/k) and v, kv, dv (initializers.py:56-89)."""
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


class _Class:
    """One size class (nx, nv): a C-ABI batch and its device state."""

    def __init__(self, idx, params, vmax, dt, op, precision, dev, delta=False):
        self.idx = idx
        self.n = idx.size
        self.nx, self.nv = int(params["nx"][idx[0]]), int(params["nv"][idx[0]])
        grids = [instance_grids(self.nx, params["k"][i], self.nv, vmax) for i in idx]
        v, kv, dv = grids[0][3], grids[0][4], grids[0][5]
        self.dv = dv
        kx = np.ascontiguousarray(np.stack([g[1] for g in grids]))
        okx = np.ascontiguousarray(np.stack([g[2] for g in grids]))
        nu = np.ascontiguousarray(params["nu"][idx].astype(np.float64))
        self._keep = (kx, okx, v, kv, nu)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        desc = _lib.BatchDesc(self.n, self.nx, self.nv, _lib.VPFP_F64 if precision == "fp64" else _lib.VPFP_F32,
                              P(kx), P(okx), P(v), P(kv), dv, P(nu), float(dt),
                              _lib.VPFP_OP_DG if op == "dg" else _lib.VPFP_OP_LB)
        h = ctypes.c_void_p()
        _lib.check(_lib.library().vpfp_batch_create(ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h
        # f0 = (1 + amp cos(k x)) exp(-v^2/2)/sqrt(2 pi), formed on the device
        x = torch.as_tensor(np.stack([g[0] for g in grids]), device=dev)
        k = torch.as_tensor(params["k"][idx], device=dev)[:, None]
        amp = torch.as_tensor(params["amp"][idx], device=dev)[:, None]
        dens = 1.0 + amp * torch.cos(k * x)
        vt = torch.as_tensor(v, device=dev)
        maxw = torch.exp(-vt ** 2 / 2.0) / np.sqrt(2.0 * np.pi)
        f0 = dens[:, :, None] * maxw[None, None, :]
        self.delta = bool(delta)
        if self.delta:
            # stored deviation f0 - M, formed in FP64 and rounded once to the
            # storage type (the device's M is the same FP64 Maxwellian)
            _lib.check(_lib.library().vpfp_batch_set_delta(h, 1))
            f0 = f0 - maxw[None, None, :]
        self.f = f0.to(_dev.f_dtype(precision)).contiguous()
        self.e = torch.zeros(self.n, self.nx, dtype=torch.float64, device=dev)
        self.v = torch.as_tensor(v, device=dev)
        self.okx = torch.as_tensor(okx, device=dev)
        self.dx = torch.as_tensor(np.asarray([g[0][1] - g[0][0] for g in grids]), device=dev)

    def full_f(self):
        """f itself (a copy in the storage type), also under delta-f storage."""
        f = self.f.clone()
        if self.delta:
            _lib.check(_lib.library().vpfp_batch_background(
                self.handle, ctypes.c_void_p(f.data_ptr()), 1,
                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return f

    def autotune(self, n_steps, stream):
        t = _lib.Tuning()
        _lib.check(_lib.library().vpfp_batch_autotune(self.handle, int(n_steps), ctypes.c_void_p(stream.cuda_stream),
                                                      ctypes.byref(t)))
        return t.as_dict()

    def tuning(self):
        """Launch shapes in use (vpfp_batch_get_tuning)."""
        t = _lib.Tuning()
        _lib.check(_lib.library().vpfp_batch_get_tuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, t):
        """Launch shapes as returned by autotune()."""
        tu = _lib.Tuning()
        tu.row_q, tu.drift_tile, tu.fuse_field = int(t["row_q"]), int(t["drift_tile"]), 0
        tu.drift_order, tu.drift_tiles = int(t["drift_order"]), int(t["drift_tiles"])
        tu.drift_dft = int(t.get("drift_dft", 0))
        _lib.check(_lib.library().vpfp_batch_set_tuning(self.handle, ctypes.byref(tu)))

    def set_profiling(self, on):
        _lib.check(_lib.library().vpfp_batch_set_profiling(self.handle, 1 if on else 0))

    def kernel_times(self):
        """{kind: (total_ms, launches)} for kinds row, drift, field, other."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_longlong * 4)()
        _lib.check(_lib.library().vpfp_batch_kernel_times(self.handle, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(("row", "drift", "field", "other"))}

    def step(self, n_steps, series, stream):
        _lib.check(_lib.library().vpfp_batch_step_n(
            self.handle, ctypes.c_void_p(self.f.data_ptr()), ctypes.c_void_p(self.e.data_ptr()), int(n_steps),
            ctypes.c_void_p(series.data_ptr()) if series is not None else None,
            ctypes.c_void_p(stream.cuda_stream)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.vpfp_batch_destroy(h)


class Scan:
    """A ragged parameter scan.

    params: per-instance arrays k, amp, nu, nx, nv (see draw_parameters).
    Each instance: x in [0, 2 pi/k), v in [-vmax, vmax], f0 = (1 + amp cos kx)
    times a unit Maxwellian, e0 = 0, then every step the leapfrog split and
    the implicit collision step (op "dg" or "lb").  delta=True stores f - M
    (vpfp_batch_set_delta): same algorithm, FP32 rounding relative to the
    deviation from the Maxwellian instead of to f.
    """

    def __init__(self, params, vmax=6.4, dt=0.1, op="dg", precision="fp32", delta=False, grouped=True):
        self.params = {key: np.asarray(val) for key, val in params.items()}
        self.n = self.params["k"].size
        self.vmax, self.dt, self.precision = float(vmax), float(dt), precision
        dev = _dev.device()
        keys = sorted(set(zip(self.params["nx"].tolist(), self.params["nv"].tolist())))
        self.classes = []
        for nx, nv in keys:
            idx = np.nonzero((self.params["nx"] == nx) & (self.params["nv"] == nv))[0]
            self.classes.append(_Class(idx, self.params, vmax, dt, op, precision, dev, delta))
        self.streams = [torch.cuda.Stream() for _ in self.classes]
        self.cells = int(np.sum(self.params["nx"].astype(np.int64) * self.params["nv"].astype(np.int64)))
        # grouped: every class in one step of 10 launches at most (one row
        # launch per nv, one drift and one field launch per nx; vpfp_group_*)
        self.group = None
        if grouped:
            nb = len(self.classes)
            arr = (ctypes.c_void_p * nb)(*[c.handle.value for c in self.classes])
            h = ctypes.c_void_p()
            _lib.check(_lib.library().vpfp_group_create(arr, nb, ctypes.byref(h)))
            self.group = h
        self.nv_order = list(dict.fromkeys(c.nv for c in self.classes))
        self.nx_order = list(dict.fromkeys(c.nx for c in self.classes))

    def _ptrs(self, tensors):
        return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() if t is not None else None for t in tensors])

    def _group_step(self, n_steps, series, stream):
        lib = _lib.library()
        ser = self._ptrs(series) if series is not None else None
        _lib.check(lib.vpfp_group_step_n(self.group, self._ptrs([c.f for c in self.classes]),
                                         self._ptrs([c.e for c in self.classes]), int(n_steps), ser,
                                         ctypes.c_void_p(stream.cuda_stream)))

    def shapes(self):
        """Grouped scans: the group's launch shapes per class {(nx, nv): tuning}."""
        nb = len(self.classes)
        q, C, per, dft = ((ctypes.c_int * nb)() for _ in range(4))
        _lib.check(_lib.library().vpfp_group_get_shapes(self.group, q, C, per, dft))
        return {(c.nx, c.nv): {"row_q": q[i], "drift_tile": C[i], "fuse_field": 0, "drift_order": 0,
                               "drift_tiles": per[i], "us_per_step": 0.0, "drift_dft": dft[i]}
                for i, c in enumerate(self.classes)}

    def autotune(self, n_steps=5):
        """Pick the launch shapes by timing them (scratch state; the instances
        are not advanced): per nv and per nx for a grouped scan, per size
        class otherwise.  Returns {(nx, nv): tuning}."""
        if self.group is not None:
            us = ctypes.c_double()
            _lib.check(_lib.library().vpfp_group_autotune(self.group, int(n_steps),
                                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                                          ctypes.byref(us)))
            out = self.shapes()
            for t in out.values():
                t["us_per_step"] = us.value
            return out
        out = {}
        for c, s in zip(self.classes, self.streams):
            out[(c.nx, c.nv)] = c.autotune(n_steps, s)
        return out

    def set_tuning(self, tuning):
        """Launch shapes as returned by autotune(): {(nx, nv): tuning}."""
        if self.group is not None:
            nb = len(self.classes)
            t = [tuning[(c.nx, c.nv)] for c in self.classes]
            q = (ctypes.c_int * nb)(*[int(x["row_q"]) for x in t])
            C = (ctypes.c_int * nb)(*[int(x["drift_tile"]) for x in t])
            per = (ctypes.c_int * nb)(*[max(1, int(x["drift_tiles"])) for x in t])
            dft = (ctypes.c_int * nb)(*[int(x.get("drift_dft", 0)) for x in t])
            _lib.check(_lib.library().vpfp_group_set_shapes(self.group, q, C, per, dft))
            return
        for c in self.classes:
            c.set_tuning(tuning[(c.nx, c.nv)])

    def capture(self, n_steps):
        """Record n_steps of every size class (each class's launches on its own
        stream, forked from and joined back into the capture stream) into one
        HIP graph.  replay() then advances every instance n_steps with one
        graph launch; results equal run(n_steps, record=False) bit for bit.
        Call after autotune(): the graph keeps the launch shapes it recorded."""
        self._graph = None
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream()
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=cap, capture_error_mode="relaxed"):
            if self.group is not None:
                self._group_step(n_steps, None, cap)
            else:
                for c, s in zip(self.classes, self.streams):
                    s.wait_stream(cap)
                    c.step(n_steps, None, s)
                for s in self.streams:
                    cap.wait_stream(s)
        torch.cuda.current_stream().wait_stream(cap)
        self._graph = (g, int(n_steps))
        return self._graph[1]

    def replay(self, times=1):
        """Launch the captured graph `times` times on the current stream."""
        if getattr(self, "_graph", None) is None:
            raise RuntimeError("no graph captured (call capture(n_steps) first)")
        for _ in range(int(times)):
            self._graph[0].replay()

    def run(self, n_steps, record=True):
        """Advance every instance n_steps.  Returns <E^2>(t) as [n_steps, n_inst]
        (host) when record is true.  Work is enqueued on the classes' streams,
        ordered after torch's current stream and joined back into it."""
        dev = _dev.device()
        series = [torch.zeros(n_steps, c.n, 2, dtype=torch.float64, device=dev) if record else None
                  for c in self.classes]
        cur = torch.cuda.current_stream()
        if self.group is not None:
            self._group_step(n_steps, series if record else None, cur)
        else:
            for c, s, ser in zip(self.classes, self.streams, series):
                s.wait_stream(cur)
                c.step(n_steps, ser, s)
            for s in self.streams:
                cur.wait_stream(s)
        if not record:
            return None
        out = np.zeros((n_steps, self.n))
        for c, ser in zip(self.classes, series):
            out[:, c.idx] = ser[:, :, 0].cpu().numpy()
        return out

    def set_profiling(self, on):
        """Per-kernel HIP event timing of every launch (off by default)."""
        if self.group is not None:
            _lib.check(_lib.library().vpfp_group_set_profiling(self.group, 1 if on else 0))
            return
        for c in self.classes:
            c.set_profiling(on)

    def kernel_times(self):
        """{launch: {kind: (total_ms, launches)}} since profiling was switched on.
        Per size class (nx, nv) for a scan of separate classes; for a grouped
        scan per row launch ("nv", nv) and per drift / field launch ("nx", nx)."""
        if self.group is None:
            return {(c.nx, c.nv): c.kernel_times() for c in self.classes}
        lib = _lib.library()
        out = {}
        for kind, name, order, key in ((0, "row", self.nv_order, "nv"), (1, "drift", self.nx_order, "nx"),
                                       (2, "field", self.nx_order, "nx")):
            for j, size in enumerate(order):
                ms, n = ctypes.c_double(), ctypes.c_longlong()
                _lib.check(lib.vpfp_group_launch_times(self.group, kind, j, ctypes.byref(ms), ctypes.byref(n)))
                out.setdefault((key, size), {})[name] = (ms.value, n.value)
        return out

    def launch_cells(self):
        """Cells each launch of kernel_times() covers: {launch: cells}."""
        if self.group is None:
            return {(c.nx, c.nv): c.n * c.nx * c.nv for c in self.classes}
        out = {}
        for c in self.classes:
            for key in (("nv", c.nv), ("nx", c.nx)):
                out[key] = out.get(key, 0) + c.n * c.nx * c.nv
        return out

    def __del__(self):
        g = getattr(self, "group", None)
        if g is not None and g.value and _lib is not None and _lib._lib is not None:
            _lib._lib.vpfp_group_destroy(g)
            self.group = None

    def energy(self, self_consistent=False):
        """Total energy per instance (host): kinetic dx sum_x trapz_v(f v^2 / 2)
        plus field dx sum_x E^2 / 2.  E is the field of the last step, or with
        self_consistent the Poisson field of the present f (E = Re ifft(i/kx
        fft(1 - n)), as at the start of a run, where the stored e is zero)."""
        out = np.zeros(self.n)
        for c in self.classes:
            f = c.full_f().double()
            w = f * (0.5 * c.v ** 2)[None, None, :]
            kin = c.dv * (w.sum(-1) - 0.5 * (w[..., 0] + w[..., -1]))
            if self_consistent:
                n = c.dv * (f.sum(-1) - 0.5 * (f[..., 0] + f[..., -1]))
                e = torch.fft.ifft(1j * c.okx * torch.fft.fft(1.0 - n, dim=-1), dim=-1).real
            else:
                e = c.e
            tot = c.dx * (kin.sum(-1) + 0.5 * (e ** 2).sum(-1))
            out[c.idx] = tot.cpu().numpy()
        return out

    def density(self):
        """Box-averaged trapz density of every instance (host)."""
        out = np.zeros(self.n)
        for c in self.classes:
            f = c.full_f().double()
            n = c.dv * (f.sum(-1) - 0.5 * (f[..., 0] + f[..., -1]))
            out[c.idx] = n.mean(-1).cpu().numpy()
        return out
