"""The FP64 log of the f log f moment (vpfp_device.h log_f64): the atanh
series coefficients and the ln 2 split are read from the source and the
algorithm is evaluated in float64 against a 40-digit decimal log, over normal
and subnormal arguments, with the special values of log.  Host only."""

import math
import os
import re
from decimal import Decimal, getcontext

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "vlapy_amd", "csrc", "vpfp_device.h")


def _constants():
    src = open(SRC).read()
    body = src[src.index("VP_HD double log_f64(double x) {"):]
    body = body[:body.index("\n}\n")]
    num = r"(-?[0-9]+\.[0-9]+(?:e[-+][0-9]+)?)"
    r0 = float(re.search(r"double r = " + num + ";", body).group(1))
    coef = [r0] + [float(v) for v in re.findall(r"r = fma\(z, r, " + num + r"\);", body)]
    hi, lo = re.search(r"fma\(de, " + num + r", fma\(de, " + num + r", l1p\)\)", body).groups()
    return coef, float(hi), float(lo)


def _log(x, coef, hi, lo):
    if not x > 0:
        return -math.inf if x == 0 else math.nan
    if x == math.inf:
        return x
    m, e = math.frexp(x)
    if m < 0.70710678118654752:
        m, e = 2.0 * m, e - 1
    f = m - 1.0
    s = f * (1.0 / (2.0 + f))
    z = s * s
    r = coef[0]
    for c in coef[1:]:
        r = r * z + c
    r *= z
    return e * hi + (e * lo + (f - s * (f - r)))


def test_log_coefficients_and_ln2_split():
    coef, hi, lo = _constants()
    assert len(coef) == 11
    for k, v in enumerate(coef[::-1], start=1):      # 2 / (2k + 1) for z^k
        assert v == 2.0 / (2 * k + 1)
    getcontext().prec = 40
    ln2 = Decimal(2).ln()
    assert hi == float(ln2) and abs(Decimal(hi) + Decimal(lo) - ln2) < Decimal("1e-33")


def test_log_accuracy_and_special_values():
    coef, hi, lo = _constants()
    getcontext().prec = 40
    rng = np.random.default_rng(5)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2.0, 3000),
                         1.0 + rng.uniform(-1e-7, 1e-7, 500), [1.0, 2.0, 0.5, math.sqrt(0.5), math.sqrt(2.0),
                                                                 1e-310, 5e-324, 1.7e308]])
    worst = 0.0
    for x in xs:
        x = float(x)
        ref = Decimal(x).ln()
        got = Decimal(_log(x, coef, hi, lo))
        err = abs(got) if ref == 0 else abs((got - ref) / ref)
        worst = max(worst, float(err))
    assert worst <= 3 * 2.0 ** -52, worst
    assert _log(0.0, coef, hi, lo) == -math.inf
    assert math.isnan(_log(-1.0, coef, hi, lo)) and math.isnan(_log(math.nan, coef, hi, lo))
    assert _log(math.inf, coef, hi, lo) == math.inf
