"""GPU parity of the learned closure against its oracle.

BF16 operands make exact agreement impossible only where the GPU's FP32 sums
and the oracle's float64 sums straddle a BF16 rounding boundary: such an
activation moves by one BF16 ulp (2^-8 relative) and the outputs it feeds by
up to a few 1e-3 of max |y|.  The bars are set above that and far below any
layout or indexing mistake (which produces O(1) relative errors):
  forward output   ||gpu - oracle|| / ||oracle|| <= 2e-3, at least 95 % of
                   outputs within 1e-4 max |y| (one flip reaches the ~9
                   outputs within two kernel widths), none beyond 1e-2 max |y|
  loss             relative <= 1e-4
  gradient         ||gpu - oracle|| / ||oracle|| <= 5e-3 for each block
"""

import numpy as np
import pytest

from oracle import closure_oracle as C

pytestmark = pytest.mark.gpu


def _setup(rows, seed):
    import torch
    from vlapy_amd.closure import Closure, init_params

    rng = np.random.default_rng(seed)
    p = init_params(seed, out_gain=1.0)
    p += (0.05 * rng.standard_normal(p.size)).astype(np.float32)
    v = np.linspace(-6.4 + 0.025, 6.4 - 0.025, 256)
    f = (np.exp(-v ** 2 / 2)[None, :] / np.sqrt(2 * np.pi) * (1 + 0.3 * rng.standard_normal((rows, 256))))
    f = f.astype(np.float32)
    t = (0.1 * rng.standard_normal((rows, 256))).astype(np.float32)
    net = Closure(p, in_scale=2.5, out_scale=0.01, max_rows=max(rows, 8))
    dev = torch.device("cuda:0")
    return net, p, f, t, torch.as_tensor(f, device=dev), torch.as_tensor(t, device=dev)


@pytest.mark.parametrize("rows", [1, 7, 300])
def test_forward_matches_oracle(rows):
    net, p, f, t, fd, td = _setup(rows, 3)
    print("closure forward rows", rows, flush=True)
    y = net.forward(fd).cpu().numpy()
    y_ref, _ = C.forward(p, f, 2.5, bf16=True)
    scale = np.abs(y_ref).max()
    err = np.abs(y - y_ref)
    l2 = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    frac = np.mean(err > 1e-4 * scale)
    print("forward rel l2", l2, "max", err.max() / scale, "fraction > 1e-4", frac, flush=True)
    assert l2 <= 2e-3 and frac <= 0.05 and err.max() <= 1e-2 * scale
    # apply: f + out_scale * y, in place
    g = fd.clone()
    net.apply(g)
    np.testing.assert_allclose(g.cpu().numpy(), f + np.float32(0.01) * y, rtol=0, atol=1e-6)


@pytest.mark.parametrize("rows", [2, 33])
def test_gradients_match_oracle(rows):
    net, p, f, t, fd, td = _setup(rows, 4)
    print("closure gradients rows", rows, flush=True)
    g, loss = net.gradients(fd, td)
    g = g.cpu().numpy().astype(np.float64)
    l_ref, g_ref = C.loss_and_grad(p, f, t, 2.5, bf16=True)
    assert abs(loss.item() / l_ref - 1) <= 1e-4
    for name in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4"):
        lo, hi = C.OFF[name]
        e = np.linalg.norm(g[lo:hi] - g_ref[lo:hi]) / np.linalg.norm(g_ref[lo:hi])
        print(name, "grad rel err", e, flush=True)
        assert e <= 5e-3, name


def test_adam_step_matches_oracle_and_training_lowers_loss():
    net, p, f, t, fd, td = _setup(16, 5)
    g, _ = net.gradients(fd, td)
    g = g.cpu().numpy().astype(np.float64)
    net.train_step(fd, td, lr=1e-3)
    p1 = net.params().astype(np.float64)
    p_ref, _, _ = C.adam(p.astype(np.float64), 0.0, 0.0, g, 1, lr=1e-3)
    np.testing.assert_allclose(p1, p_ref, rtol=0, atol=1e-6)
    # a few hundred steps fit a smooth target
    import torch

    v = torch.linspace(-1, 1, 256, device=fd.device)
    tgt = (torch.sin(3 * v)[None, :] * torch.ones(16, 1, device=fd.device)).contiguous()
    _, l0 = net.gradients(fd, tgt)
    loss = torch.zeros(1, device=fd.device)
    for _ in range(300):
        net.train_step(fd, tgt, lr=2e-3, loss=loss)
    _, l1 = net.gradients(fd, tgt)
    print("loss", l0.item(), "->", l1.item(), flush=True)
    assert l1.item() < 0.25 * l0.item()
