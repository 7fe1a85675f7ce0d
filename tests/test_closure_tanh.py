"""The closure kernel's tanh (vpfc.hip tanh_act): the rational x P(x^2)/Q(x^2)
whose coefficients are the literals in the source, evaluated in FP32 the way
the kernel does (clamp, fma Horner, reciprocal, product), against float64
tanh.  The bound is what keeps GPU and oracle BF16 activations agreeing; the
test reads the literals from the source so that an edit of any of them is
checked.  Host only."""

import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "vlapy_amd", "csrc", "vpfc.hip")


def _coefficients():
    src = open(SRC).read()
    body = src[src.index("float tanh_act(float x) {"):]
    body = body[:body.index("\n}\n")]
    p_line = re.search(r"const float p =\s*(.*?);", body, re.S).group(1)
    q_line = re.search(r"const float q =\s*(.*?);", body, re.S).group(1)
    lit = lambda s: [np.float32(v) for v in re.findall(r"([0-9.]+e[-+][0-9]+)f", s)]
    # Horner literals appear highest degree first, then the next ones as additions
    p, q = lit(p_line), lit(q_line)
    assert len(p) == 4 and len(q) == 4 and p_line.rstrip().endswith("y, 1.f)") and q_line.rstrip().endswith("y, 1.f)")
    clamp = re.search(r"fmed3f\(x, -([0-9.]+)f, ([0-9.]+)f\)", body)
    assert clamp and clamp.group(1) == clamp.group(2)
    return [np.float32(1)] + p[::-1], [np.float32(1)] + q[::-1], np.float32(clamp.group(1))


def _eval(x, p, q, xmax):
    c = np.clip(np.asarray(x, np.float32), -xmax, xmax)
    y = (c * c).astype(np.float32)
    pp = np.full_like(y, p[-1])
    qq = np.full_like(y, q[-1])
    for k in range(len(p) - 2, -1, -1):
        pp = (pp * y + p[k]).astype(np.float32)
        qq = (qq * y + q[k]).astype(np.float32)
    return ((c * pp).astype(np.float32) * (np.float32(1) / qq).astype(np.float32)).astype(np.float32)


def _bf16(v):
    """Round FP32 to BF16 (nearest even) and back."""
    u = v.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_kernel_tanh_relative_error_and_shape():
    p, q, xmax = _coefficients()
    assert xmax == 9.0
    x = np.concatenate([np.linspace(0.0, 12.0, 1200001), np.logspace(-10, 0, 100001)]).astype(np.float32)
    x = np.concatenate([x, -x])
    t = _eval(x, p, q, xmax).astype(np.float64)
    ref = np.tanh(x.astype(np.float64))
    rel = np.abs(t - ref) / np.maximum(np.abs(ref), 1e-300)
    ok = x != 0
    assert rel[ok].max() <= 4.5e-7, rel[ok].max()
    assert np.array_equal(t[: t.size // 2], -t[t.size // 2:])  # odd
    # at most 1 + 4e-7 in magnitude, which rounds to a BF16 activation of exactly +-1
    assert np.all(np.abs(t) <= 1 + 4.5e-7) and np.all(np.abs(t[np.abs(x) >= 9]) >= 1 - 5e-7)
    assert np.all(np.abs(_bf16(t.astype(np.float32))) <= 1.0)
    # the BF16 activations are monotone in x (the rational's wiggles stay below
    # a BF16 step)
    xs = np.sort(x)
    assert np.all(np.diff(_bf16(_eval(xs, p, q, xmax)).astype(np.float64)) >= 0)
