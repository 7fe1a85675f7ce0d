"""Coefficients of the closure kernel's tanh (k_fwd in vpfc.hip).

tanh(x) ~ x P(x^2) / Q(x^2) on |x| <= 9 (clamped beyond: tanh(9) = 1 - 3e-8),
P and Q of degree 4 in x^2 with P(0) = Q(0) = 1.  The coefficients minimise
the maximum relative error by iteratively reweighted linear least squares
(the linearised residual x P - tanh(x) Q, weights raised where the error is
largest, Lawson's scheme).  Run it to regenerate the literals and print the
error of their FP32 evaluation (fma Horner, reciprocal, product):

    python vlapy_amd/csrc/tanh_fit.py
"""

import numpy as np

X_MAX = 9.0
DEG = 5  # coefficients of P and of Q


def fit(iters=120):
    x = np.unique(np.concatenate([np.linspace(1e-4, X_MAX, 200000), np.logspace(-6, np.log10(X_MAX), 20000)]))
    t = np.tanh(x)
    y = x * x
    w = np.ones_like(x)
    best = None
    for _ in range(iters):
        a = np.hstack([x[:, None] * y[:, None] ** np.arange(DEG)[None, :],
                       -(t[:, None] * y[:, None] ** np.arange(1, DEG)[None, :])])
        sw = (w / t)[:, None]
        c, *_ = np.linalg.lstsq(a * sw, t * sw[:, 0], rcond=None)
        p, q = c[:DEG], np.concatenate([[1.0], c[DEG:]])
        err = np.abs(x * np.polyval(p[::-1], y) / np.polyval(q[::-1], y) / t - 1)
        if best is None or err.max() < best[0]:
            best = (err.max(), p, q)
        w = w * np.sqrt(err / err.mean())
        w /= w.mean()
    return best[1].astype(np.float32), best[2].astype(np.float32)


def eval_f32(x, p, q):
    """The kernel's arithmetic in FP32: clamp, Horner in x^2, x P * (1 / Q)."""
    c = np.clip(np.asarray(x, np.float32), np.float32(-X_MAX), np.float32(X_MAX))
    y = (c * c).astype(np.float32)
    pp = np.full_like(y, p[-1])
    qq = np.full_like(y, q[-1])
    for k in range(DEG - 2, -1, -1):
        pp = (pp * y + p[k]).astype(np.float32)
        qq = (qq * y + q[k]).astype(np.float32)
    return ((c * pp).astype(np.float32) * (np.float32(1) / qq).astype(np.float32)).astype(np.float32)


if __name__ == "__main__":
    p, q = fit()
    print("P:", ", ".join("%.9ef" % v for v in p))
    print("Q:", ", ".join("%.9ef" % v for v in q))
    x = np.concatenate([np.linspace(0, X_MAX, 2000001), np.logspace(-8, 0, 100001)]).astype(np.float32)
    ref = np.tanh(x.astype(np.float64))
    rel = np.abs(eval_f32(x, p, q) - ref) / np.maximum(ref, 1e-30)
    for lo, hi in ((0, 0.2), (0.2, 1), (1, 3.5), (3.5, X_MAX)):
        s = (x >= lo) & (x < hi)
        print("|x| in [%g, %g): max relative error %.2e" % (lo, hi, rel[s].max()))
