"""Array plumbing for the vlapy.core-compatible closures.

A closure called with numpy arrays copies them to the GPU, runs the HIP
kernels and returns numpy float64 arrays (drop-in mode, synchronous).  Called
with torch device tensors it stays on the device and returns tensors
(resident mode, asynchronous on torch's current stream).
"""

import numpy as np
import torch


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("the HIP backend needs a GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def is_tensor(a):
    return isinstance(a, torch.Tensor)


def to_dev(a, dtype):
    """Contiguous device tensor of `dtype` holding `a` (numpy, scalar or tensor)."""
    if is_tensor(a):
        t = a if a.is_cuda else a.to(device())
        if t.dtype != dtype:
            t = t.to(dtype)
        return t.contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)), device=device()).to(dtype)


def field_vector(e, nx):
    """Device float64 (nx,) tensor from an array, tensor or scalar (e.g. a zero driver)."""
    if is_tensor(e) or np.ndim(e) > 0:
        t = to_dev(e, torch.float64)
        if tuple(t.shape) != (nx,):
            raise ValueError("field has shape %s, expected (%d,)" % (tuple(t.shape), nx))
        return t
    return torch.full((nx,), float(e), dtype=torch.float64, device=device())


def back(t, like):
    """Return `t` in the kind of `like`: numpy float64 for numpy input, tensor otherwise."""
    if is_tensor(like):
        return t
    return t.detach().to(torch.float64).cpu().numpy() if not t.is_complex() else t.cpu().numpy()


def f_dtype(precision):
    if precision == "fp64":
        return torch.float64
    if precision == "fp32":
        return torch.float32
    raise ValueError("precision must be 'fp64' or 'fp32', got %r" % (precision,))


def precision_of(stuff):
    return stuff.get("precision", "fp64") if isinstance(stuff, dict) else "fp64"
