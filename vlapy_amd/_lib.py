"""ctypes binding of libvpfp.so (the C ABI declared in include/vpfp.h).

The library is built in-tree (vlapy_amd/libvpfp.so, see vlapy_amd/build.py).
There is no fallback: if the library is missing or no GPU is visible, every
entry point raises.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvpfp.so")

VPFP_OK, VPFP_EBADARG, VPFP_EUNSUPPORTED, VPFP_EHIP = 0, 1, 2, 3
VPFP_F64, VPFP_F32 = 0, 1
VPFP_OP_LB, VPFP_OP_DG = 0, 1
NMOMENTS = 8
NOALLOC_MAX_NX = 4096  # VPFP_NOALLOC_MAX_NX
MOMENT_NAMES = ("n", "j", "T", "q", "fv4", "vN", "f2", "flogf")

_c_p = ctypes.c_void_p
_c_d = ctypes.c_double
_c_i = ctypes.c_int


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("nx", _c_i),
        ("nv", _c_i),
        ("precision", _c_i),
        ("kx", ctypes.POINTER(_c_d)),
        ("one_over_kx", ctypes.POINTER(_c_d)),
        ("v", ctypes.POINTER(_c_d)),
        ("kv", ctypes.POINTER(_c_d)),
        ("dv", _c_d),
        ("nu", _c_d),
        ("dt", _c_d),
        ("op", _c_i),
    ]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_inst", _c_i),
        ("nx", _c_i),
        ("nv", _c_i),
        ("precision", _c_i),
        ("kx", ctypes.POINTER(_c_d)),
        ("one_over_kx", ctypes.POINTER(_c_d)),
        ("v", ctypes.POINTER(_c_d)),
        ("kv", ctypes.POINTER(_c_d)),
        ("dv", _c_d),
        ("nu", ctypes.POINTER(_c_d)),
        ("dt", _c_d),
        ("op", _c_i),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("row_q", _c_i),
        ("drift_tile", _c_i),
        ("fuse_field", _c_i),
        ("drift_order", _c_i),
        ("drift_tiles", _c_i),
        ("us_per_step", _c_d),
        ("drift_dft", _c_i),
    ]

    def as_dict(self):
        return {"row_q": self.row_q, "drift_tile": self.drift_tile, "fuse_field": self.fuse_field,
                "drift_order": self.drift_order, "drift_tiles": self.drift_tiles, "us_per_step": self.us_per_step,
                "drift_dft": self.drift_dft}


# name -> (argtypes); every function returns int status unless listed in _OTHER
SIGNATURES = {
    "vpfp_plan_create": [ctypes.POINTER(GridDesc), ctypes.POINTER(_c_p)],
    "vpfp_plan_destroy": [_c_p],
    "vpfp_vdfdx": [_c_p, _c_p, _c_p, _c_d, _c_p],
    "vpfp_edfdv": [_c_p, _c_p, _c_p, _c_p, _c_d, _c_p],
    "vpfp_compute_charges": [_c_p, _c_p, _c_p, _c_p],
    "vpfp_solve_for_field": [_c_p, _c_p, _c_p, _c_p],
    "vpfp_field": [_c_p, _c_p, _c_p, _c_p, _c_p],
    "vpfp_field_noalloc": [_c_i, _c_i, _c_i, _c_d, _c_p, _c_p, _c_p, _c_p, _c_p],
    "vpfp_collision_diagonals": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p],
    "vpfp_tridiag_solve": [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p],
    "vpfp_collide": [_c_p, _c_p, _c_p, _c_p],
    "vpfp_step_n": [_c_p, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p],
    "vpfp_step_n_store": [_c_p, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p],
    "vpfp_step_n_scheme": [_c_p, _c_i, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p],
    "vpfp_moments": [_c_p, _c_p, _c_p, _c_p],
    "vpfp_x_modes": [_c_p, _c_p, _c_i, _c_p, _c_p],
    "vpfp_batch_create": [ctypes.POINTER(BatchDesc), ctypes.POINTER(_c_p)],
    "vpfp_batch_step_n": [_c_p, _c_p, _c_p, _c_i, _c_p, _c_p],
    "vpfp_batch_destroy": [_c_p],
    "vpfp_batch_set_delta": [_c_p, _c_i],
    "vpfp_plan_set_delta": [_c_p, _c_i],
    "vpfp_plan_background": [_c_p, _c_p, _c_i, _c_p],
    "vpfp_batch_background": [_c_p, _c_p, _c_i, _c_p],
    "vpfp_plan_set_profiling": [_c_p, _c_i],
    "vpfp_plan_kernel_times": [_c_p, ctypes.POINTER(_c_d), ctypes.POINTER(ctypes.c_longlong)],
    "vpfp_batch_set_profiling": [_c_p, _c_i],
    "vpfp_group_create": [ctypes.POINTER(_c_p), _c_i, ctypes.POINTER(_c_p)],
    "vpfp_group_step_n": [_c_p, ctypes.POINTER(_c_p), ctypes.POINTER(_c_p), _c_i, ctypes.POINTER(_c_p), _c_p],
    "vpfp_group_autotune": [_c_p, _c_i, _c_p, ctypes.POINTER(_c_d)],
    "vpfp_group_get_shapes": [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_i),
                              ctypes.POINTER(_c_i)],
    "vpfp_group_set_shapes": [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_i),
                              ctypes.POINTER(_c_i)],
    "vpfp_group_launch_times": [_c_p, _c_i, _c_i, ctypes.POINTER(_c_d), ctypes.POINTER(ctypes.c_longlong)],
    "vpfp_group_set_profiling": [_c_p, _c_i],
    "vpfp_group_kernel_times": [_c_p, ctypes.POINTER(_c_d), ctypes.POINTER(ctypes.c_longlong)],
    "vpfp_group_destroy": [_c_p],
    "vpfp_batch_kernel_times": [_c_p, ctypes.POINTER(_c_d), ctypes.POINTER(ctypes.c_longlong)],
    "vpfp_plan_get_tuning": [_c_p, ctypes.POINTER(Tuning)],
    "vpfp_plan_set_tuning": [_c_p, ctypes.POINTER(Tuning)],
    "vpfp_plan_autotune": [_c_p, _c_i, _c_p, ctypes.POINTER(Tuning)],
    "vpfp_batch_get_tuning": [_c_p, ctypes.POINTER(Tuning)],
    "vpfp_batch_set_tuning": [_c_p, ctypes.POINTER(Tuning)],
    "vpfp_batch_autotune": [_c_p, _c_i, _c_p, ctypes.POINTER(Tuning)],
    "vpfp_last_error": [],
    "vpfp_version": [],
    # learned closure (include/vpfc.h)
    "vpfc_create": [_c_i, ctypes.POINTER(_c_p)],
    "vpfc_destroy": [_c_p],
    "vpfc_set_params": [_c_p, _c_p],
    "vpfc_get_params": [_c_p, _c_p],
    "vpfc_set_scales": [_c_p, ctypes.c_float, ctypes.c_float],
    "vpfc_set_background": [_c_p, _c_p],
    "vpfc_set_conservation": [_c_p, _c_p, _c_p],
    "vpfc_forward": [_c_p, _c_p, _c_i, _c_p, _c_p],
    "vpfc_apply": [_c_p, _c_p, _c_i, _c_p],
    "vpfc_gradients": [_c_p, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p],
    "vpfc_train_step": [_c_p, _c_p, _c_p, _c_i, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                        _c_p, _c_p],
    "vpfp_batch_set_closure": [_c_p, _c_p],
    "vpfc_last_error": [],
}
_OTHER = {"vpfp_last_error": ctypes.c_char_p, "vpfp_version": _c_i, "vpfc_last_error": ctypes.c_char_p}

_lib = None


def library():
    """Load libvpfp.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libvpfp.so is not built (expected %s); run `python -m vlapy_amd.build`" % LIB_PATH
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _OTHER.get(name, _c_i)
        _lib = lib
    return _lib


def check(status, closure=False):
    """Map a C ABI status to the reference's Python exceptions (closure:
    the message is vpfc_last_error's)."""
    if status == VPFP_OK:
        return
    err = library().vpfc_last_error if closure else library().vpfp_last_error
    msg = err().decode(errors="replace")
    if status == VPFP_EBADARG:
        raise ValueError(msg)
    if status == VPFP_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(_c_d))


class Plan:
    """Owns one vpfp plan (device grid copies, twiddle tables, scratch)."""

    def __init__(self, nx, nv, precision="fp64", kx=None, one_over_kx=None, v=None, kv=None, dv=1.0,
                 nu=0.0, dt=0.0, op="lb"):
        lib = library()
        self.nx, self.nv = int(nx), int(nv)
        self.precision = precision
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be fp64 or fp32")
        if op not in ("lb", "dg"):
            raise NotImplementedError("Collision Operator: <" + str(op) + "> has not yet been implemented")

        def arr(a, n):
            if a is None:
                return np.zeros(n)
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            if a.shape != (n,):
                raise ValueError("grid array has shape %s, expected (%d,)" % (a.shape, n))
            return a

        # keep the host copies alive for the duration of the call
        self._kx, self._okx = arr(kx, self.nx), arr(one_over_kx, self.nx)
        self._v, self._kv = arr(v, self.nv), arr(kv, self.nv)
        desc = GridDesc(self.nx, self.nv, VPFP_F64 if precision == "fp64" else VPFP_F32,
                        _dptr(self._kx), _dptr(self._okx), _dptr(self._v), _dptr(self._kv),
                        float(dv), float(nu), float(dt), VPFP_OP_LB if op == "lb" else VPFP_OP_DG)
        handle = _c_p()
        check(lib.vpfp_plan_create(ctypes.byref(desc), ctypes.byref(handle)))
        self.handle = handle
        self.dv, self.nu, self.dt, self.op = float(dv), float(nu), float(dt), op

    def set_profiling(self, on):
        check(library().vpfp_plan_set_profiling(self.handle, 1 if on else 0))

    def set_delta(self, on):
        """Delta-f storage (vpfp_plan_set_delta): f arrays hold f - M(v)."""
        check(library().vpfp_plan_set_delta(self.handle, 1 if on else 0))

    def background(self, f, sign, stream=None):
        """f += sign * M(v) in place on a device tensor of the plan's storage type
        (sign -1: f to the stored deviation, +1 back)."""
        import torch

        from . import ops

        # (nx, nv), the storage dtype, contiguous, on the device: the kernel
        # writes nx * nv elements of that type
        ptr = ops._f(self, f)
        if int(sign) not in (1, -1):
            raise ValueError("sign must be 1 or -1")
        s = stream if stream is not None else torch.cuda.current_stream()
        check(library().vpfp_plan_background(self.handle, _c_p(ptr), int(sign), _c_p(s.cuda_stream)))

    def kernel_times(self):
        """{kind: (total_ms, launches)} for kinds row, drift, field, other."""
        ms = (_c_d * 4)()
        n = (ctypes.c_longlong * 4)()
        check(library().vpfp_plan_kernel_times(self.handle, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(("row", "drift", "field", "other"))}

    def tuning(self):
        """Launch shapes in use: row_q, drift_tile, fuse_field, us_per_step."""
        t = Tuning()
        check(library().vpfp_plan_get_tuning(self.handle, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, row_q=None, drift_tile=None, fuse_field=None, drift_order=None, drift_tiles=None,
                   drift_dft=None):
        t = Tuning()
        check(library().vpfp_plan_get_tuning(self.handle, ctypes.byref(t)))
        if row_q is not None:
            t.row_q = int(row_q)
        if drift_tile is not None:
            t.drift_tile = int(drift_tile)
        if fuse_field is not None:
            t.fuse_field = int(fuse_field)
        if drift_order is not None:
            t.drift_order = int(drift_order)
        if drift_tiles is not None:
            t.drift_tiles = int(drift_tiles)
        if drift_dft is not None:
            t.drift_dft = int(drift_dft)
        check(library().vpfp_plan_set_tuning(self.handle, ctypes.byref(t)))

    def autotune(self, n_steps=20, stream=None):
        """Time every valid launch shape of the step (on scratch state) and
        keep the fastest; returns the chosen tuning."""
        t = Tuning()
        check(library().vpfp_plan_autotune(self.handle, int(n_steps), stream, ctypes.byref(t)))
        return t.as_dict()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.vpfp_plan_destroy(h)
            self.handle = None
