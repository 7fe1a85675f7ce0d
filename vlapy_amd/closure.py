"""Learned coarse-velocity closure (no counterpart in the reference).

A 4-layer 1-D convolutional network over v (64 channels, kernel 5) that
corrects the plain solver on a coarse velocity grid (nv = 256) after every
step:  f(x, .) <- f(x, .) + out_scale * net(in_scale * f(x, .)).
Training, inference and the Adam update run in the BF16/MFMA kernels of
libvpfp.so (include/vpfc.h); this module holds the host side: parameter
initialisation and files, training data from the method's own fine-grid runs,
the training loop and the rollouts that evaluate the closure against the fine
grid.
"""

import ctypes

import numpy as np
import torch

from . import _lib

NV, CH, KS = 256, 64, 5
NPARAMS = 41793
_LAYOUT = (("W1", CH * KS, KS), ("b1", CH, None), ("W2", CH * CH * KS, CH * KS), ("b2", CH, None),
           ("W3", CH * CH * KS, CH * KS), ("b3", CH, None), ("W4", CH * KS, CH * KS), ("b4", 1, None))


def init_params(seed, out_gain=0.1):
    """Uniform Glorot initialisation for tanh layers (limit sqrt(6/(fan_in+fan_out))),
    zero biases; the output layer is scaled by out_gain so that an untrained
    closure starts close to the identity map of f."""
    rng = np.random.default_rng(seed)
    parts = []
    fan_out = {"W1": CH * KS, "W2": CH * KS, "W3": CH * KS, "W4": KS}
    for name, n, fan_in in _LAYOUT:
        if fan_in is None:
            parts.append(np.zeros(n))
            continue
        lim = np.sqrt(6.0 / (fan_in + fan_out[name]))
        w = rng.uniform(-lim, lim, n)
        if name == "W4":
            w *= out_gain
        parts.append(w)
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAMS
    return p


class Closure:
    """One network on the GPU (parameters, Adam moments, training workspace
    for up to max_rows rows)."""

    def __init__(self, params=None, in_scale=1.0, out_scale=1.0, max_rows=4096, seed=0, bg=None):
        lib = _lib.library()
        h = ctypes.c_void_p()
        _lib.check(lib.vpfc_create(int(max_rows), ctypes.byref(h)), closure=True)
        self.handle = h
        self.max_rows = int(max_rows)
        self.set_params(init_params(seed) if params is None else params)
        self.set_scales(in_scale, out_scale)
        self.set_background(bg)
        self.cons = None

    # parameters ---------------------------------------------------------
    def set_params(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32))
        if p.shape != (NPARAMS,):
            raise ValueError("expected %d parameters, got %s" % (NPARAMS, p.shape))
        _lib.check(_lib.library().vpfc_set_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)), closure=True)

    def params(self):
        p = np.zeros(NPARAMS, dtype=np.float32)
        _lib.check(_lib.library().vpfc_get_params(self.handle, p.ctypes.data_as(ctypes.c_void_p)), closure=True)
        return p

    def set_scales(self, in_scale, out_scale):
        self.in_scale, self.out_scale = float(in_scale), float(out_scale)
        _lib.check(_lib.library().vpfc_set_scales(self.handle, self.in_scale, self.out_scale), closure=True)

    def set_background(self, bg):
        """Background profile [256] (see include/vpfc.h) or None."""
        if bg is None:
            self.bg = None
            _lib.check(_lib.library().vpfc_set_background(self.handle, None), closure=True)
            return
        self.bg = np.ascontiguousarray(np.asarray(bg, dtype=np.float32))
        if self.bg.shape != (NV,):
            raise ValueError("background must hold %d values" % NV)
        _lib.check(_lib.library().vpfc_set_background(self.handle, self.bg.ctypes.data_as(ctypes.c_void_p)),
                   closure=True)

    def set_conservation(self, v, bg):
        """Project every correction so that it adds no density, momentum or
        kinetic energy (include/vpfc.h); v None switches it off."""
        if v is None:
            self.cons = None
            _lib.check(_lib.library().vpfc_set_conservation(self.handle, None, None), closure=True)
            return
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        bg = np.ascontiguousarray(np.asarray(bg, dtype=np.float64))
        if v.shape != (NV,) or bg.shape != (NV,):
            raise ValueError("v and bg must hold %d values" % NV)
        self.cons = (v, bg)
        _lib.check(_lib.library().vpfc_set_conservation(self.handle, v.ctypes.data_as(ctypes.c_void_p),
                                                        bg.ctypes.data_as(ctypes.c_void_p)), closure=True)

    def save(self, path, **meta):
        extra = {} if self.bg is None else {"bg": self.bg}
        if self.cons is not None:
            extra.update(cons_v=self.cons[0], cons_bg=self.cons[1])
        np.savez(path, params=self.params(), in_scale=self.in_scale, out_scale=self.out_scale, **extra,
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, max_rows=4096):
        with np.load(path, allow_pickle=False) as z:
            bg = z["bg"] if "bg" in z.files else None
            net = cls(z["params"], float(z["in_scale"]), float(z["out_scale"]), max_rows=max_rows, bg=bg)
            if "cons_v" in z.files:
                net.set_conservation(z["cons_v"], z["cons_bg"])
            return net

    # compute ------------------------------------------------------------
    @staticmethod
    def _rows(a, name):
        if not (isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float32 and a.is_contiguous()):
            raise ValueError("%s must be a contiguous float32 CUDA tensor" % name)
        if a.shape[-1] != NV:
            raise ValueError("%s rows must hold %d values" % (name, NV))
        return a.numel() // NV

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, x):
        """y = net(in_scale * x) for rows x (..., 256)."""
        rows = self._rows(x, "x")
        y = torch.empty_like(x)
        _lib.check(_lib.library().vpfc_forward(self.handle, ctypes.c_void_p(x.data_ptr()), rows,
                                               ctypes.c_void_p(y.data_ptr()), self._stream()), closure=True)
        return y

    def apply(self, f):
        """f <- f + out_scale * net(in_scale * f), in place."""
        rows = self._rows(f, "f")
        _lib.check(_lib.library().vpfc_apply(self.handle, ctypes.c_void_p(f.data_ptr()), rows, self._stream()),
                   closure=True)
        return f

    def gradients(self, x, target):
        """(grad [NPARAMS] device tensor, loss device scalar) of the mean
        squared error of net(in_scale * x) against target."""
        rows = self._rows(x, "x")
        if self._rows(target, "target") != rows:
            raise ValueError("x and target hold different row counts")
        g = torch.empty(NPARAMS, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        _lib.check(_lib.library().vpfc_gradients(self.handle, ctypes.c_void_p(x.data_ptr()),
                                                 ctypes.c_void_p(target.data_ptr()), rows,
                                                 ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                                 self._stream()), closure=True)
        return g, loss

    def train_step(self, x, target, lr, beta1=0.9, beta2=0.999, eps=1e-8, loss=None):
        rows = self._rows(x, "x")
        if self._rows(target, "target") != rows:
            raise ValueError("x and target hold different row counts")
        lp = ctypes.c_void_p(loss.data_ptr()) if loss is not None else None
        _lib.check(_lib.library().vpfc_train_step(self.handle, ctypes.c_void_p(x.data_ptr()),
                                                  ctypes.c_void_p(target.data_ptr()), rows, float(lr), float(beta1),
                                                  float(beta2), float(eps), lp, self._stream()), closure=True)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.vpfc_destroy(h)
            self.handle = None


# ---------------------------------------------------------------------------
# data, training and evaluation
# ---------------------------------------------------------------------------
NX, DT, VMAX, NV_FINE = 64, 0.05, 6.4, 4096


def draw_instances(n, seed):
    """Parameters of n runs at nx = 64, dt = 0.05: k ~ U[0.25, 0.45],
    amplitude 10^U[-4, -1], nu 10^U[-5, -2] (Dougherty), drawn in that order
    from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.25, 0.45, n)
    amp = 10.0 ** rng.uniform(-4.0, -1.0, n)
    nu = 10.0 ** rng.uniform(-5.0, -2.0, n)
    return {"k": k, "amp": amp, "nu": nu}


def _scan(inst, nv, precision, delta=False):
    from .scan import Scan

    n = inst["k"].size
    p = dict(inst, nx=np.full(n, NX), nv=np.full(n, nv))
    # one size class: per-class batches (the closure hook attaches to a batch)
    return Scan(p, vmax=VMAX, dt=DT, op="dg", precision=precision, delta=delta, grouped=False)


def background(dtype=np.float32):
    """The unit Maxwellian of the fine grid averaged onto the coarse grid
    ([256], FP32 by default): the coarse representation of the equilibrium."""
    dvf = 2 * VMAX / NV_FINE
    vf = np.linspace(-VMAX + dvf / 2.0, VMAX - dvf / 2.0, NV_FINE)
    m = np.exp(-vf ** 2 / 2.0) / np.sqrt(2.0 * np.pi)
    return m.reshape(NV, NV_FINE // NV).mean(-1).astype(dtype)


def coarse_v():
    """Cell centres of the coarse velocity grid on [-VMAX, VMAX]."""
    dv = 2 * VMAX / NV
    return np.linspace(-VMAX + dv / 2.0, VMAX - dv / 2.0, NV)


def conservation_projector(v, bg):
    """P [256, 256] (float64): P y = y - bg (c0 + c1 v + c2 v^2) with the
    trapz moments of order 0, 1, 2 of P y zero - the map the kernel applies
    to every correction (vpfc_set_conservation)."""
    w = np.ones(NV)
    w[0] = w[-1] = 0.5
    B = np.stack([bg, v * bg, v * v * bg], 1)
    Vm = np.stack([w, w * v, w * v * v], 0)
    return np.eye(NV) - B @ np.linalg.solve(Vm @ B, Vm)


def downsample(f, factor):
    """Cell average of `factor` consecutive velocity cells: f [n, nx, nv] ->
    [n, nx, nv / factor] (conserves sum f dv)."""
    n, nx, nv = f.shape
    return f.reshape(n, nx, nv // factor, factor).mean(-1)


def training_pairs(inst, n_steps, every, span=1, fine_precision="fp64", log=None):
    """(input, target) rows from fine runs (nv = 4096) of the instances.

    Every `every` steps the fine state is averaged onto the coarse grid,
    c = D f(t).  The input is one plain coarse FP32 step s = S(c); the target
    is the per-step correction averaged over `span` steps,
    (D f(t + span dt) - S^span(c)) / span, which keeps the secular part of the
    coarse-grid error and averages out its phase-dependent part."""
    fine = _scan(inst, NV_FINE, fine_precision)
    coarse = _scan(inst, NV, "fp32")
    fc, cc = fine.classes[0], coarse.classes[0]
    factor = NV_FINE // NV
    xs, ts = [], []
    done = 0
    while done + max(every, span) <= n_steps:
        if every > span:
            fine.run(every - span, record=False)
        cc.f.copy_(downsample(fc.f, factor).float())
        cc.e.copy_(fc.e)
        coarse.run(1, record=False)
        s = cc.f.reshape(-1, NV).clone()
        if span > 1:
            coarse.run(span - 1, record=False)
        fine.run(span, record=False)
        done += max(every, span)
        t = (downsample(fc.f, factor).float().reshape(-1, NV) - cc.f.reshape(-1, NV)) / span
        xs.append(s)
        ts.append(t)
        if log is not None:
            log(done)
    return torch.cat(xs), torch.cat(ts)


def train(x, t, steps, batch=4096, lr=1e-3, seed=0, net=None, log=None, bg=None, conserve=False):
    """Adam on minibatches of rows drawn without replacement per epoch.
    With a background bg the network works on d / s (d = x - bg, s = rms_v d
    per row) and predicts t / s.  Scales: in_scale = 1 / max |d / s|,
    out_scale = rms(t / s), so that inputs and targets are of order 1."""
    if conserve:
        # targets projected like the corrections the kernel will apply
        P = torch.as_tensor(conservation_projector(coarse_v(), np.asarray(background(), np.float64)),
                            dtype=torch.float32, device=t.device)
        t = (t @ P.T).contiguous()
    if bg is not None:
        d = x - torch.as_tensor(bg, device=x.device)[None, :]
        srow = (d.pow(2).mean(-1, keepdim=True) + 1e-30).sqrt()
        in_scale = 1.0 / float((d / srow).abs().max())
        out_scale = float((t / srow).pow(2).mean().sqrt())
        del d, srow
    else:
        in_scale = 1.0 / float(x.abs().max())
        out_scale = float(t.pow(2).mean().sqrt())
    tn = (t / out_scale).contiguous()
    if net is None:
        net = Closure(in_scale=in_scale, out_scale=out_scale, max_rows=batch, seed=seed, bg=bg)
    else:
        net.set_scales(in_scale, out_scale)
        net.set_background(bg)
    if conserve:
        net.set_conservation(coarse_v(), np.asarray(background(), np.float64))
    gen = torch.Generator(device=x.device)
    gen.manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=x.device)
    perm = torch.randperm(x.shape[0], generator=gen, device=x.device)
    pos = 0
    hist = []
    for it in range(steps):
        if pos + batch > x.shape[0]:
            perm = torch.randperm(x.shape[0], generator=gen, device=x.device)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        # cosine decay of the step size
        lr_i = lr * 0.5 * (1.0 + np.cos(np.pi * it / steps))
        net.train_step(x[idx].contiguous(), tn[idx].contiguous(), lr_i, loss=loss)
        if log is not None and (it % 200 == 0 or it == steps - 1):
            hist.append((it, float(loss.item())))
            log(it, hist[-1][1])
    return net, hist


def rollout(inst, nv, precision, n_steps, net=None, delta=False):
    """<E^2>(t) [n_steps, n] of plain runs on an nv grid, or of FP32 coarse
    runs with the closure applied after every step (net given).  Coarse runs
    start from the fine initial state averaged onto their grid,
    (1 + amp cos kx) background(), the projection the closure is trained on.
    delta: store f - M (vpfp_batch_set_delta); the coarse initial state is
    then formed from the FP64 cell average."""
    sc = _scan(inst, nv, precision, delta)
    if nv == NV:
        from .scan import instance_grids

        c = sc.classes[0]
        x = np.stack([instance_grids(NX, k, NV, VMAX)[0] for k in inst["k"]])
        dens = 1.0 + inst["amp"][:, None] * np.cos(inst["k"][:, None] * x)
        if delta:
            v = coarse_v()
            f0 = dens[:, :, None] * background(np.float64)[None, None, :]
            f0 = f0 - (np.exp(-v ** 2 / 2.0) / np.sqrt(2.0 * np.pi))[None, None, :]
        else:
            f0 = dens[:, :, None] * background().astype(np.float64)[None, None, :]
        c.f.copy_(torch.as_tensor(f0, device=c.f.device).to(c.f.dtype))
    if net is not None:
        _lib.check(_lib.library().vpfp_batch_set_closure(sc.classes[0].handle, net.handle))
    n0 = sc.density()
    e2 = sc.run(n_steps)
    if net is not None:
        _lib.check(_lib.library().vpfp_batch_set_closure(sc.classes[0].handle, None))
    rollout.density_drift = np.abs(sc.density() / n0 - 1.0)
    return e2


def damping_errors(inst, e2_ref, e2, t):
    """|gamma / gamma_ref - 1| per instance from the standing-wave estimator on
    a window after one period, before half the trapping time (pi / sqrt(amp)), before
    coarse-grid recurrence (pi / (k dv)) and before <E^2> falls 1e8 below its
    start; NaN where the window holds fewer than three maxima."""
    from .diagnostics.low_level_helpers import get_standing_wave_rate_and_frequency as rate

    dv = 2 * VMAX / NV
    out = np.full(inst["k"].size, np.nan)
    for i in range(out.size):
        y = e2_ref[:, i]
        floor = np.nonzero(y < 1e-8 * y[:20].max())[0]
        hi = min(np.pi / np.sqrt(inst["amp"][i]), np.pi / (inst["k"][i] * dv), t[floor[0]] if floor.size else t[-1])
        try:
            g_ref, _ = rate(y, t, 6.0, hi)
            g, _ = rate(e2[:, i], t, 6.0, hi)
        except ValueError:
            continue
        out[i] = abs(g / g_ref - 1.0)
    return out


def history_errors(e2_ref, e2):
    """Relative L2 error of log10 <E^2>(t) per instance."""
    a, b = np.log10(np.maximum(e2_ref, 1e-300)), np.log10(np.maximum(e2, 1e-300))
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(a, axis=0)
