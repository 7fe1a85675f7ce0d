"""CPU oracle of the learned coarse-velocity closure (test infrastructure only).

Only tests/, the smoke entry point and the benchmark's CPU leg use this
module; the product path runs the HIP kernels of vlapy_amd/csrc/vpfc.hip.

The closure has no counterpart in the reference (VlaPy has no learned
component); this file restates the network and its arithmetic as declared
in include/vpfc.h:

    net: conv1d over v (kernel 5, zero padding 2), 1 -> 64 -> 64 -> 64 -> 1,
         tanh after the first three layers, linear output
    input in_scale * d / s with d = f - bg, s = rms_v d per row (with a
    background bg; d = f, s = 1 without one); target compared as target / s
    out[co][v] = b[co] + sum_{k, ci} W[co][k*64 + ci] in[ci][v + k - 2]

With bf16=True every factor of every product is rounded to BF16 (round to
nearest even) exactly as the GPU rounds it, post-tanh activations and the
pre-activation gradients dz are stored rounded, and sums run in float64 (the
GPU sums in FP32).  With bf16=False the same formulas run in float64 with no
rounding: that mode is pinned by central finite differences of the loss
(tests/test_closure_oracle.py), which checks every term of the backward pass.

Flat parameter layout (41793 floats):
    W1 [64][5] | b1 [64] | W2 [64][320] | b2 [64] | W3 [64][320] | b3 [64] |
    W4 [320] | b4 [1]
"""

import numpy as np

NV, CH, KS = 256, 64, 5
K = CH * KS
OFF = {}
_o = 0
for _name, _n in (("W1", CH * KS), ("b1", CH), ("W2", CH * K), ("b2", CH), ("W3", CH * K), ("b3", CH),
                  ("W4", K), ("b4", 1)):
    OFF[_name] = (_o, _o + _n)
    _o += _n
NPARAMS = _o


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even); returned as
    float32 (NaN is not handled: the network never produces it)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    """Views of the flat vector: W1 (64,5), b1, W2 (64,5,64) [co, k, ci], b2,
    W3, b3, W4 (5,64) [k, ci], b4."""
    p = np.asarray(p)
    g = lambda n: p[OFF[n][0]:OFF[n][1]]  # noqa: E731
    return {"W1": g("W1").reshape(CH, KS), "b1": g("b1"),
            "W2": g("W2").reshape(CH, KS, CH), "b2": g("b2"),
            "W3": g("W3").reshape(CH, KS, CH), "b3": g("b3"),
            "W4": g("W4").reshape(KS, CH), "b4": g("b4")}


def _pad(a):
    """zero padding of 2 on the last axis"""
    w = [(0, 0)] * (a.ndim - 1) + [(2, 2)]
    return np.pad(a, w)


def _conv(Wck, x):
    """out[r, co, v] = sum_{k, ci} W[co, k, ci] x[r, ci, v + k - 2]  (float64)"""
    xp = _pad(x)
    out = np.zeros((x.shape[0], Wck.shape[0], NV))
    for k in range(KS):
        out += np.einsum("oc,rcv->rov", Wck[:, k, :], xp[:, :, k:k + NV], optimize=True)
    return out


def _conv_t(Wck, dz):
    """input gradient of _conv: dx[r, ci, u] = sum_{k, co} W[co, k, ci] dz[r, co, u - k + 2]"""
    dp = _pad(dz)
    out = np.zeros((dz.shape[0], Wck.shape[2], NV))
    for k in range(KS):
        # u - k + 2 = v  ->  slice of the padded dz starting at 4 - k
        out += np.einsum("oc,rov->rcv", Wck[:, k, :], dp[:, :, 4 - k:4 - k + NV], optimize=True)
    return out


def _wgrad(dz, x):
    """dW[co, k, ci] = sum_{r, v} dz[r, co, v] x[r, ci, v + k - 2]"""
    xp = _pad(x)
    return np.stack([np.einsum("rov,rcv->oc", dz, xp[:, :, k:k + NV], optimize=True) for k in range(KS)], axis=1)


def row_scale(f, bg):
    """d = f - bg and s = sqrt(mean_v d^2 + 1e-30) per row (s = 1 without bg)."""
    f = np.asarray(f, dtype=np.float64)
    if bg is None:
        return f, np.ones((f.shape[0], 1))
    d = f - np.asarray(bg, dtype=np.float64)[None, :]
    return d, np.sqrt(np.mean(d * d, axis=1, keepdims=True) + 1e-30)


def forward(p, f, in_scale, bf16=True, bg=None):
    """y = net(in_scale * d / s) for rows f [R, 256] (d, s from row_scale);
    returns y and the cache."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    q = unpack(p)
    d, s = row_scale(f, bg)
    if bf16:
        # FP32 as the GPU forms it: (in_scale / s) * d
        x = r(np.asarray((np.float32(in_scale) / s.astype(np.float32)) * d.astype(np.float32), dtype=np.float32))
    else:
        x = in_scale * d / s
    x = np.asarray(x, dtype=np.float64)
    W1 = np.asarray(r(q["W1"]), np.float64)
    z1 = np.zeros((x.shape[0], CH, NV))
    xp = _pad(x)
    for k in range(KS):
        z1 += W1[None, :, k, None] * xp[:, None, k:k + NV]
    a1 = np.asarray(r(np.tanh(z1 + q["b1"][None, :, None])), np.float64)
    a2 = np.asarray(r(np.tanh(_conv(np.asarray(r(q["W2"]), np.float64), a1) + q["b2"][None, :, None])), np.float64)
    a3 = np.asarray(r(np.tanh(_conv(np.asarray(r(q["W3"]), np.float64), a2) + q["b3"][None, :, None])), np.float64)
    W4 = np.asarray(r(q["W4"]), np.float64)
    a3p = _pad(a3)
    y = np.full((x.shape[0], NV), float(q["b4"][0]))
    for k in range(KS):
        y += np.einsum("c,rcv->rv", W4[k], a3p[:, :, k:k + NV], optimize=True)
    return y, {"x": x, "a1": a1, "a2": a2, "a3": a3, "s": s}


def loss_and_grad(p, f, target, in_scale, bf16=True, bg=None):
    """L = mean (y - target / s)^2 and dL/dp (flat, float64); s = 1 without bg."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    rr = (lambda a: np.asarray(bf16_round(a), np.float64)) if bf16 else (lambda a: np.asarray(a, np.float64))
    q = unpack(p)
    y, c = forward(p, f, in_scale, bf16, bg)
    t = np.asarray(target, dtype=np.float64) / c["s"]
    count = y.size
    d = y - t
    loss = float(np.sum(d * d) / count)
    if bf16:
        dy = np.asarray(np.float32(2.0 / count) * d.astype(np.float32), np.float64)
    else:
        dy = 2.0 * d / count
    g = np.zeros(NPARAMS)

    def put(name, val):
        g[OFF[name][0]:OFF[name][1]] = np.asarray(val).reshape(-1)

    dyr = rr(dy)
    a3 = c["a3"]
    W4 = np.asarray(r(q["W4"]), np.float64)
    # output layer
    a3p = _pad(a3)
    put("W4", np.stack([np.einsum("rv,rcv->c", dyr, a3p[:, :, k:k + NV]) for k in range(KS)]))
    put("b4", [np.sum(dy)])
    dyp = _pad(dyr)
    s3 = np.zeros_like(a3)
    for k in range(KS):
        s3 += W4[k][None, :, None] * dyp[:, None, 4 - k:4 - k + NV]
    dz3 = rr(s3 * (1.0 - a3 * a3))
    # layer 3
    put("W3", _wgrad(dz3, c["a2"]))
    put("b3", dz3.sum(axis=(0, 2)))
    dz2 = rr(_conv_t(np.asarray(r(q["W3"]), np.float64), dz3) * (1.0 - c["a2"] ** 2))
    # layer 2
    put("W2", _wgrad(dz2, c["a1"]))
    put("b2", dz2.sum(axis=(0, 2)))
    dz1 = rr(_conv_t(np.asarray(r(q["W2"]), np.float64), dz2) * (1.0 - c["a1"] ** 2))
    # layer 1
    xp = _pad(c["x"])
    put("W1", np.stack([np.einsum("rcv,rv->c", dz1, xp[:, k:k + NV]) for k in range(KS)], axis=1))
    put("b1", dz1.sum(axis=(0, 2)))
    return loss, g


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1); t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
