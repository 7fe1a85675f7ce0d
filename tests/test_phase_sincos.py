"""The FP32 phase factors' sin/cos (vpfp_device.h sincos_half): Taylor series
in y^2 whose coefficients are the literals in the source, evaluated in FP32
with one rounding per fma the way the kernel does, against float64 sin/cos on
|y| <= pi/2.  The literals are read from the source, so an edit of any of them
is checked.  Host only."""

import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "vlapy_amd", "csrc", "vpfp_device.h")


def _literals():
    """Coefficients highest degree first, as they appear in the nested fmaf calls."""
    src = open(SRC).read()
    body = src[src.index("VP_HD void sincos_half(float y, float& s, float& c) {"):]
    body = body[:body.index("\n}\n")]
    lit = lambda t: [float(v) for v in re.findall(r"(-?[0-9]+\.[0-9]*(?:e[-+][0-9]+)?)f", t)]
    ps = lit(re.search(r"const float ps =(.*?);", body, re.S).group(1))
    pc = lit(re.search(r"const float pc =(.*?);", body, re.S).group(1))
    assert "s = fmaf(y * y2, ps, y);" in body and "c = fmaf(y2, pc, 1.f);" in body
    return ps, pc


def _fma(a, b, c):
    # one rounding of a*b + c: the FP32 product is exact in float64, and the
    # float64 sum's own rounding is far below an FP32 ulp
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _eval(y, ps, pc):
    y = np.asarray(y, np.float32)
    y2 = (y * y).astype(np.float32)
    p = np.full_like(y, ps[0])
    for cf in ps[1:]:
        p = _fma(y2, p, np.full_like(y, cf))
    s = _fma((y * y2).astype(np.float32), p, y)
    q = np.full_like(y, pc[0])
    for cf in pc[1:]:
        q = _fma(y2, q, np.full_like(y, cf))
    c = _fma(y2, q, np.ones_like(y))
    return s, c


def test_sincos_half_literals_are_the_taylor_coefficients():
    ps, pc = _literals()
    assert len(ps) == 6 and len(pc) == 7
    for k, v in enumerate(ps[::-1], start=1):   # y^(2k+1) / (2k+1)!
        assert abs(v - (-1) ** k / math.factorial(2 * k + 1)) <= 2e-9 * abs(v)
    for k, v in enumerate(pc[::-1], start=1):   # y^(2k) / (2k)!
        assert abs(v - (-1) ** k / math.factorial(2 * k)) <= 2e-9 * abs(v)


def test_sincos_half_fp32_error():
    ps, pc = _literals()
    y = np.concatenate([np.linspace(-math.pi / 2 - 1e-6, math.pi / 2 + 1e-6, 400001),
                        np.logspace(-12, 0, 20001)]).astype(np.float32)
    s, c = _eval(y, ps, pc)
    yd = y.astype(np.float64)
    es = np.abs(s - np.sin(yd))
    ec = np.abs(c - np.cos(yd))
    ulp1 = 2.0 ** -23
    assert es.max() <= 1.5 * ulp1 and ec.max() <= 1.5 * ulp1, (es.max(), ec.max())
    # sin keeps relative precision for small arguments (the seeds of exp(-i th) - 1)
    small = np.abs(yd) < 1e-3
    assert np.all(np.abs(s[small] - np.sin(yd[small])) <= 1.0 * np.abs(yd[small]) * 2.0 ** -23)


def _literals64():
    src = open(SRC).read()
    body = src[src.index("VP_HD void sincos_half(double y, double& s, double& c) {"):]
    body = body[:body.index("\n}\n")]
    num = r"(-?[0-9]+\.[0-9]*(?:e[-+][0-9]+)?)"
    p0 = float(re.search(r"double p = " + num + ";", body).group(1))
    q0 = float(re.search(r"double q = " + num + ";", body).group(1))
    ps = [p0] + [float(v) for v in re.findall(r"p = fma\(y2, p, " + num + r"\);", body)]
    pc = [q0] + [float(v) for v in re.findall(r"q = fma\(y2, q, " + num + r"\);", body)]
    assert "s = fma(y * y2, p, y);" in body and "c = fma(y2, q, 1.0);" in body
    return ps, pc


def test_sincos_half_fp64_literals_and_error():
    ps, pc = _literals64()
    assert len(ps) == 11 and len(pc) == 12
    for k, v in enumerate(ps[::-1], start=1):
        assert abs(v - (-1) ** k / math.factorial(2 * k + 1)) <= 4e-16 * abs(v)
    for k, v in enumerate(pc[::-1], start=1):
        assert abs(v - (-1) ** k / math.factorial(2 * k)) <= 4e-16 * abs(v)
    # float64 Horner (separate rounding of product and sum: at most the kernel's fma error)
    y = np.linspace(-math.pi / 2, math.pi / 2, 20001)
    y2 = y * y
    p = np.full_like(y, ps[0])
    for cf in ps[1:]:
        p = p * y2 + cf
    s = y + (y * y2) * p
    q = np.full_like(y, pc[0])
    for cf in pc[1:]:
        q = q * y2 + cf
    c = 1.0 + y2 * q
    # reference: the float64 library (correctly rounded to within an ulp)
    assert np.abs(s - np.sin(y)).max() <= 4 * 2.0 ** -52 and np.abs(c - np.cos(y)).max() <= 4 * 2.0 ** -52


def test_reduce_2pi_constants():
    src = open(SRC).read()
    body = src[src.index("VP_HD double reduce_2pi(double th) {"):]
    body = body[:body.index("\n}\n")]
    c2, c1 = [float(v) for v in re.findall(r"fma\(-n, ([0-9.e-]+),", body)]
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    two_pi = Decimal("6.2831853071795864769252867665590057683943387987502")
    assert c1 == float(two_pi)                                   # C1: 2 pi rounded to double
    assert abs(Decimal(c1) + Decimal(c2) - two_pi) < Decimal("1e-31")  # C1 + C2 = 2 pi to ~1e-31
