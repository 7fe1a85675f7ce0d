"""Pins of the closure oracle (CPU).

* BF16 rounding equals torch's float32 -> bfloat16 conversion (round to
  nearest even) on random values, ties and subnormals.
* The backward pass of the oracle (float64, no rounding) equals central
  finite differences of the loss for parameters of every layer, with and
  without the background-relative row scaling.
* Adam restates Kingma & Ba (2015): the first step moves every parameter by
  lr * g / (|g| + eps) and repeated steps on a quadratic converge.
* The BF16 mode differs from the exact mode only by BF16 rounding.
"""

import numpy as np
import pytest
import torch

from oracle import closure_oracle as C


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 1e-40, -3e-39, 0.0],
                                 dtype=np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(C.bf16_round(x), ref)


def _problem(rows=2, seed=1):
    from vlapy_amd.closure import init_params

    rng = np.random.default_rng(seed)
    p = init_params(seed, out_gain=1.0).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.size)  # non-zero biases too
    v = np.linspace(-6.4 + 0.025, 6.4 - 0.025, 256)
    f = np.exp(-v ** 2 / 2)[None, :] / np.sqrt(2 * np.pi) * (1 + 0.3 * rng.standard_normal((rows, 256)))
    t = 0.1 * rng.standard_normal((rows, 256))
    return p, f, t


@pytest.mark.parametrize("with_bg", [False, True])
def test_backward_equals_finite_differences(with_bg):
    p, f, t = _problem()
    v = np.linspace(-6.4 + 0.025, 6.4 - 0.025, 256)
    bg = np.exp(-v ** 2 / 2) / np.sqrt(2 * np.pi) if with_bg else None
    loss, g = C.loss_and_grad(p, f, t, 2.5, bf16=False, bg=bg)
    rng = np.random.default_rng(7)
    picks = []
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        lo, hi = C.OFF[name]
        picks += list(rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False))
    for i in picks:
        h = 1e-5 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = C.loss_and_grad(pp, f, t, 2.5, bf16=False, bg=bg)
        lm, _ = C.loss_and_grad(pm, f, t, 2.5, bf16=False, bg=bg)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_bf16_mode_is_the_rounded_exact_mode():
    p, f, t = _problem(rows=3, seed=2)
    p32 = p.astype(np.float32)
    y_bf, _ = C.forward(p32, f, 2.5, bf16=True)
    y_ex, _ = C.forward(p32.astype(np.float64), f, 2.5, bf16=False)
    # BF16 keeps 8 significant bits: the output moves by a few 2^-8 relative
    err = np.abs(y_bf - y_ex).max() / np.abs(y_ex).max()
    assert 1e-5 < err < 5e-2
    l_bf, g_bf = C.loss_and_grad(p32, f, t, 2.5, bf16=True)
    l_ex, g_ex = C.loss_and_grad(p32.astype(np.float64), f, t, 2.5, bf16=False)
    assert abs(l_bf / l_ex - 1) < 5e-2
    assert np.linalg.norm(g_bf - g_ex) / np.linalg.norm(g_ex) < 5e-2


def test_adam_first_step_and_quadratic():
    g = np.array([3.0, -0.5, 1e-3])
    p, m, v = C.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, lr=0.1, eps=1e-12)
    np.testing.assert_allclose(p, -0.1 * g / (np.abs(g) + 1e-12), rtol=1e-12)
    # minimise 0.5 * a * p^2
    a = np.array([1.0, 10.0, 100.0])
    p, m, v = np.ones(3), np.zeros(3), np.zeros(3)
    for t in range(1, 3001):
        p, m, v = C.adam(p, m, v, a * p, t, lr=1e-2)
    assert np.abs(p).max() < 1e-2


@pytest.mark.parametrize("with_bg", [False, True])
def test_forward_equals_torch_conv1d(with_bg):
    """The declared network, evaluated by torch.nn.functional.conv1d (zero
    padding 2, cross-correlation) in float64, equals the oracle's exact-mode
    forward pass: an implementation that shares no code with the oracle."""
    import torch.nn.functional as F

    p, f, _ = _problem(rows=3, seed=5)
    v = np.linspace(-6.4 + 0.025, 6.4 - 0.025, 256)
    bg = np.exp(-v ** 2 / 2) / np.sqrt(2 * np.pi) if with_bg else None
    y, _ = C.forward(p, f, 2.5, bf16=False, bg=bg)

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))  # noqa: E731
    o = C.OFF
    W1 = t(p[o["W1"][0]:o["W1"][1]].reshape(64, 1, 5))
    # oracle layout [co][k][ci] -> torch [co][ci][k]
    W2 = t(p[o["W2"][0]:o["W2"][1]].reshape(64, 5, 64).transpose(0, 2, 1))
    W3 = t(p[o["W3"][0]:o["W3"][1]].reshape(64, 5, 64).transpose(0, 2, 1))
    W4 = t(p[o["W4"][0]:o["W4"][1]].reshape(5, 64).T[None, :, :])
    b = {n: t(p[o[n][0]:o[n][1]]) for n in ("b1", "b2", "b3", "b4")}
    ff = np.asarray(f, dtype=np.float64)
    if with_bg:
        d = ff - bg[None, :]
        s = np.sqrt(np.mean(d * d, axis=1, keepdims=True) + 1e-30)
        x = 2.5 * d / s
    else:
        x = 2.5 * ff
    h = t(x)[:, None, :]
    h = torch.tanh(F.conv1d(h, W1, b["b1"], padding=2))
    h = torch.tanh(F.conv1d(h, W2, b["b2"], padding=2))
    h = torch.tanh(F.conv1d(h, W3, b["b3"], padding=2))
    ref = F.conv1d(h, W4, b["b4"], padding=2)[:, 0, :].numpy()
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err <= 1e-13, err
