"""The FP32 form's sine/cosine and atan2 (sincos_fit, atan2_fit in
pystamps_amd/csrc/ps2_device.hpp): their coefficients are read from the
header and evaluated here in float32 arithmetic (fused multiply-adds through
float64, then rounded), against float64 numpy.  Pins the error bounds the
header states.  CPU."""
import re
from pathlib import Path

import numpy as np
import pytest

HDR = Path(__file__).resolve().parents[1] / "pystamps_amd" / "csrc" / "ps2_device.hpp"
F32 = np.float32
HEX = re.compile(r"-?0x1\.[0-9a-f]+p[+-]?\d+")


def _body(name):
    text = HDR.read_text()
    start = text.index(f"__device__ inline {name}")
    return text[start:text.index("\n}\n", start)]


def _consts(name):
    return [F32(float.fromhex(h)) for h in HEX.findall(_body(name))]


def _fma(a, b, c):
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(F32)


def sincos_fit(x):
    two_over_pi, p1, p2, p3, s3, s2, s1, c4, c3, c2 = _consts("void sincos_fit")
    x = x.astype(F32)
    q = np.rint((x * two_over_pi).astype(F32)).astype(F32)
    r = _fma(-q, p1, x)
    r = _fma(-q, p2, r)
    r = _fma(-q, p3, r)
    s = (r * r).astype(F32)
    ps = _fma(_fma(s3, s, s2), s, s1)
    sn = _fma((r * s).astype(F32), ps, r)
    pc = _fma(_fma(_fma(c4, s, c3), s, c2), s, F32(-0.5))
    cn = _fma(s, pc, F32(1.0))
    qi = q.astype(np.int64)
    odd = (qi & 1) != 0
    S = np.where(odd, cn, sn)
    C = np.where(odd, sn, cn)
    S = np.where((qi & 2) != 0, -S, S)
    C = np.where(((qi + 1) & 2) != 0, -C, C)
    return S, C


def atan2_fit(y, x, rcp_ulps=0.0):
    c = _consts("float atan2_fit")
    coef, half_pi, pi = c[:8], c[8], c[9]
    x = x.astype(F32)
    y = y.astype(F32)
    ax, ay = np.abs(x), np.abs(y)
    mx = np.maximum(np.maximum(ax, ay), F32(1e-30))
    mn = np.minimum(ax, ay)
    rc = (1.0 / mx.astype(np.float64)).astype(F32)
    rc = (rc * (1.0 + rcp_ulps * 2.0 ** -23)).astype(F32)   # the hardware reciprocal's 1 ulp
    t = (mn * rc).astype(F32)
    s = (t * t).astype(F32)
    p = np.full_like(s, coef[0])
    for k in coef[1:]:
        p = _fma(p, s, k)
    a = _fma((t * s).astype(F32), p, t)
    a = np.where(ay > ax, (half_pi - a).astype(F32), a)
    a = np.where(x < 0, (pi - a).astype(F32), a)
    xy = (x + y).astype(F32)
    return np.where(np.isnan(xy), xy, np.copysign(a, y))


def test_constants_are_pi_halves_and_two_over_pi():
    two_over_pi, p1, p2, p3 = _consts("void sincos_fit")[:4]
    assert two_over_pi == F32(2 / np.pi)
    # p1 + p2 + p3 = pi/2 far beyond double precision (each part is the
    # rounding of what the earlier ones leave)
    assert p1 == F32(np.pi / 2)
    assert abs((float(p1) - np.pi / 2) + float(p2)) < 2.0 ** -48
    assert abs(float(p3)) < 2.0 ** -48


def _ulp_err(v, ref):
    ulp = np.spacing(np.abs(ref).astype(F32)).astype(np.float64)
    return np.abs(v.astype(np.float64) - ref) / np.maximum(ulp, 2.0 ** -149)


def test_sincos_fit_within_1p5_ulp():
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-8192, 8192, 1_000_000), rng.uniform(-60, 60, 1_000_000),
                        rng.uniform(-1, 1, 200_000),
                        # near multiples of pi/2, where the reduction cancels
                        (np.arange(-5000, 5000) * (np.pi / 2)) + rng.uniform(-1e-3, 1e-3, 10000)]).astype(F32)
    S, C = sincos_fit(x)
    xd = x.astype(np.float64)
    assert _ulp_err(S, np.sin(xd)).max() <= 1.5
    assert _ulp_err(C, np.cos(xd)).max() <= 1.5
    assert np.abs(S.astype(np.float64) - np.sin(xd)).max() < 7.5e-8
    assert np.abs(C.astype(np.float64) - np.cos(xd)).max() < 7.5e-8


def test_sincos_fit_special_values():
    S, C = sincos_fit(np.array([0.0, -0.0, np.pi / 2, np.pi, -np.pi / 2], F32))
    assert S[0] == 0 and C[0] == 1 and not np.signbit(S[0])
    assert S[1] == 0 and C[1] == 1   # sin(-0) is +0 here (stated in the header)
    np.testing.assert_allclose(S[2:], [1.0, 0.0, -1.0], atol=1e-7)
    np.testing.assert_allclose(C[2:], [0.0, -1.0, 0.0], atol=1e-7)


@pytest.mark.parametrize("rcp_ulps", [0.0, 1.0, -1.0])
def test_atan2_fit_within_3e7(rcp_ulps):
    rng = np.random.default_rng(11)
    n = 2_000_000
    ang = rng.uniform(-np.pi, np.pi, n)
    mag = np.exp(rng.uniform(-20, 5, n))
    x = (mag * np.cos(ang)).astype(F32)
    y = (mag * np.sin(ang)).astype(F32)
    a = atan2_fit(y, x, rcp_ulps)
    err = np.abs(a.astype(np.float64) - np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() < 3.5e-7


def test_atan2_fit_axes_zeros_and_nan():
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, 1e-30, np.nan, 1.0], F32)
    x = np.array([0.0, 0.0, -1.0, -1.0, 0.0, 0.0, 1.0, -1.0, 1.0, np.nan], F32)
    a = atan2_fit(y, x)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(a[:8], ref[:8], rtol=0, atol=3e-7)
    assert np.signbit(a[1]) and not np.signbit(a[0])
    assert np.isnan(a[8]) and np.isnan(a[9])


def _fma_f32(a, b, c):
    # the product of two float32 values is exact in long double (64-bit
    # significand); one rounding to long double then float32 for the sum
    L = np.longdouble
    return (L(a) * L(b) + L(c)).astype(F32)


def _div_from_rcp(e, d):
    # the 8-trial scan's quotient (div_from_rcp in ps2_kernels.hpp): per-pixel
    # reciprocal, two corrections written negated to keep a zero's sign
    rcp = (F32(1) / d).astype(F32)
    q = (e * rcp).astype(F32)
    q = -_fma_f32(-_fma_f32(-q, d, e), rcp, -q)
    q = -_fma_f32(-_fma_f32(-q, d, e), rcp, -q)
    return q


@pytest.mark.skipif(np.finfo(np.longdouble).nmant < 63, reason="needs an 80-bit long double for exact FMA emulation")
def test_scan_quotient_has_ieee_bits():
    """bperp / range from the reciprocal with two FMA corrections equals the
    IEEE float32 quotient bit for bit (Markstein's theorem), so the scan that
    computes its unit steps from bperp sees the step table's values."""
    rng = np.random.default_rng(5)
    n = 2_000_000
    cases = [
        (rng.uniform(-3000, 3000, n), rng.uniform(7e5, 1.2e6, n)),          # bperp, slant range
        (rng.uniform(1, 2, n) * 2.0 ** rng.integers(-100, 100, n) * rng.choice([-1, 1], n),
         rng.uniform(1, 2, n) * 2.0 ** rng.integers(-24, 24, n)),          # any normal quotient
        (np.concatenate([np.zeros(500), -np.zeros(500)]), rng.uniform(7e5, 1.2e6, 1000)),
    ]
    for e, d in cases:
        e, d = e.astype(F32), d.astype(F32)
        q = _div_from_rcp(e, d)
        np.testing.assert_array_equal(q.view(np.uint32), (e / d).view(np.uint32))


def test_f64_baseline_split_is_exact():
    """The FP32 form stores its f64 baselines as h = float32(v) plus an int32
    residual lo = (v - h) 2^(180 - be) (be the biased exponent bits of h), and
    reads them back as h + lo 2^(be - 180).  Every double whose float32 image
    is normal splits exactly, with |lo| <= 2^29: random doubles over the f32
    normal range, doubles next to powers of two (where h rounds up into the
    next binade) and every representable residual sign."""
    rng = np.random.default_rng(7)
    v = np.concatenate([
        rng.uniform(-600.0, 600.0, 200_000),
        np.exp(rng.uniform(np.log(1.2e-38), np.log(3.0e38), 100_000)) * rng.choice([-1.0, 1.0], 100_000),
        np.nextafter(2.0 ** np.arange(-120, 120, dtype=np.float64), 0.0),
        np.nextafter(2.0 ** np.arange(-120, 120, dtype=np.float64), np.inf),
        2.0 ** np.arange(-120, 120, dtype=np.float64) * (1.0 - 2.0 ** -25),
    ])
    h = v.astype(np.float32)
    assert np.all(np.isfinite(h)) and np.all(np.abs(h) >= np.finfo(np.float32).tiny)
    be = ((h.view(np.uint32) >> 23) & 0xFF).astype(np.int64)
    r = (v - h.astype(np.float64)) * np.ldexp(1.0, 180 - be)
    assert np.all(r == np.rint(r)) and np.all(np.abs(r) <= 2.0 ** 29)
    lo = r.astype(np.int32)
    back = h.astype(np.float64) + np.ldexp(lo.astype(np.float64), be - 180)
    assert np.array_equal(back, v)
    # a wrong scale (one bit short) loses residual bits somewhere
    r_bad = (v - h.astype(np.float64)) * np.ldexp(1.0, 179 - be)
    assert not np.all(r_bad == np.rint(r_bad))
