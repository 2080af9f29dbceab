"""Closed-form checks of the rewrites the MC kernel's pair loop makes of the
reference's pair terms (src/kernel_cuda_l2_calcenergy.cu:105-184). Each test
states the reference's form and the kernel's form and checks that they agree
(polynomial identities exactly, bounds on a dense grid in FP64)."""
import math

import numpy as np
from numpy.polynomial import polynomial as P

LN2 = math.log(2.0)
LOG2E = 1.0 / LN2


def test_hpc_switch_factors_exactly():
    """1 - (3.5x - 4.5x^2 + 2.5x^3 - 0.5x^4) (calcenergy.cu:116-121) equals
    (1 - x)^3 (1 - x/2), i.e. u^3 (1 + u) / 2 with u = 1 - x, and is zero at
    the cut-off x = 1 (d = 9 A) with zero slope."""
    ref = np.array([1.0, -3.5, 4.5, -2.5, 0.5])
    fac = P.polymul(P.polypow([1.0, -1.0], 3), [1.0, -0.5])
    np.testing.assert_array_equal(ref, fac)
    x = np.linspace(0.0, 1.0, 1001)
    u = 1.0 - x
    np.testing.assert_allclose(P.polyval(x, ref), u ** 3 * (1.0 + u) / 2.0, rtol=0, atol=1e-15)
    assert P.polyval(1.0, ref) == 0.0
    assert P.polyval(1.0, P.polyder(ref)) == 0.0
    # beyond the cut-off the reference adds nothing; max(u, 0) gives 0 there
    assert max(1.0 - 1.5, 0.0) ** 3 * (1.0 + max(1.0 - 1.5, 0.0)) == 0.0


def test_lj_denominator_rewrite():
    """p1 lj0 (1 + lj1 d^2) + 1 with p1 = p1a / d^9 (calcenergy.cu:130-135),
    multiplied through by d^9, is p1a lj0 + p1a lj0 lj1 d^2 + d^6 d^3: the
    kernel's fma(d^6, d^3, a1 d^2 + a0) with a0 = p1a lj0, a1 = p1a lj0 lj1."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        p1a, p2a = rng.uniform(1e2, 1e7), rng.uniform(1e1, 1e5)
        lj0, lj1 = rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)
        d = rng.uniform(0.5, 30.0)
        p1, p2 = p1a / d ** 9, p2a / d ** 6
        vdw_ref = (p1 - p2) / (p1 * lj0 * (1.0 + lj1 * d * d) + 1.0)
        a0, a1 = p1a * lj0, p1a * lj0 * lj1
        d3 = d * d * d
        vdw_ker = (p1a - p2a * d3) / ((d3 * d3) * d3 + (a1 * d * d + a0))
        assert abs(vdw_ker - vdw_ref) <= 1e-12 * max(1.0, abs(vdw_ref))


def test_pmf_exponent_peak():
    """(6 - d/2)(d - pmf0) (calcenergy.cu:157-160) peaks at d = 6 + pmf0/2
    with value (pmf0 - 12)^2 / 8; its log2 form A d^2 + B d + C uses
    A = -log2(e)/2, B = log2(e)(6 + pmf0/2), C = -6 log2(e) pmf0."""
    d = np.linspace(0.0, 60.0, 600001)
    for pmf0 in (-20.0, -6.0, 2.69, 4.0, 7.38, 30.0, 45.0):
        w = (6.0 - d / 2.0) * (d - pmf0)
        peak = (pmf0 - 12.0) ** 2 / 8.0
        if 6.0 + pmf0 / 2.0 >= 0.0:
            assert abs(w.max() - peak) <= 1e-6 * max(1.0, peak)
        assert w.max() <= peak + 1e-9
        A, B, C = -LOG2E / 2.0, LOG2E * (6.0 + pmf0 / 2.0), -6.0 * LOG2E * pmf0
        np.testing.assert_allclose(A * d * d + B * d + C, w * LOG2E, rtol=1e-12, atol=1e-9)
    # the parameter file's range keeps the exponent far below exp2's range:
    # the kernel drops its clamp where the peak is at most 60 (in log2 units)
    assert LOG2E * (2.69 - 12.0) ** 2 / 8.0 < 16.0
    assert LOG2E * (7.38 - 12.0) ** 2 / 8.0 < 16.0


def test_hbond_gaussian_in_base_2():
    """The kernel's H-bond Gaussian HC exp2(-(d HA + HB)^2), HA = hdb1
    sqrt(log2(e) / 2), HB = -hdb0 HA, is the reference's
    exp(-((d - hdb0) hdb1)^2 / 2) (src/kernel_cuda_l2_calcenergy.cu:179-184)
    at every distance."""
    rng = np.random.default_rng(1)
    d = np.linspace(0.0, 25.0, 5001)
    for _ in range(100):
        hdb0, sig = rng.uniform(3.29, 4.55), rng.uniform(0.6, 1.48)
        hdb1 = 1.0 / sig
        ha = hdb1 * math.sqrt(LOG2E / 2.0)
        hb = -hdb0 * ha
        g_ref = np.exp(-((d - hdb0) * hdb1) ** 2 / 2.0)
        g_ker = np.exp2(-(d * ha + hb) ** 2)
        np.testing.assert_allclose(g_ker, g_ref, rtol=1e-9, atol=1e-300)


def _fma32(a, b, c):
    """FP32 fused multiply-add: the exact FP64 product of two FP32 values plus
    an FP32 addend, rounded once to FP32 (the FP64 sum's own rounding can
    only matter for addends far below the product's last bit)."""
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def test_shared_reciprocal_with_fused_products():
    """vdW (num/den) and the PMF factor 1/(1 + e) share one reciprocal
    r = 1/(den (1 + e)): vdW = num (1 + e) r and 1/(1 + e) = den r. The kernel
    forms den (1 + e) and num (1 + e) as fma(den, e, den) and fma(num, e, num),
    one rounding each. Checked in FP32 against FP64 over the range of d
    (0.5 to 60 A: den up to 60^9 = 1e16) and of the exponent (e up to 2^64,
    the clamp), with no overflow."""
    rng = np.random.default_rng(2)
    f32 = np.float32
    worst_v, worst_p = 0.0, 0.0
    for _ in range(20000):
        d = rng.uniform(0.5, 60.0)
        p1a, p2a = rng.uniform(1e2, 1e7), rng.uniform(1e1, 1e5)
        a0, a1 = p1a * rng.uniform(0.5, 1.0), p1a * rng.uniform(0.25, 1.0)
        e = 2.0 ** rng.uniform(-150.0, 64.0)
        num = f32(p1a - p2a * d ** 3)
        den = f32(a0 + a1 * d * d + d ** 9)
        e32 = f32(e)
        dene = _fma32(den, e32, den)
        nume = _fma32(num, e32, num)
        assert np.isfinite(dene) and np.isfinite(nume)
        r = f32(1.0) / dene
        vdw = np.float64(nume) * np.float64(r)
        pmf = np.float64(den) * np.float64(r)
        ref_v = np.float64(num) / np.float64(den)
        ref_p = 1.0 / (1.0 + np.float64(e32))
        worst_v = max(worst_v, abs(vdw - ref_v) / abs(ref_v))
        worst_p = max(worst_p, abs(pmf - ref_p) / ref_p)
    # two FP32 roundings in each form (the fma, the reciprocal): a few ulp
    assert worst_v < 4 * 2.0 ** -23 and worst_p < 4 * 2.0 ** -23


def test_hpc_clamp_is_the_fma_output_clamp():
    """u = 1 - d^2/81 never exceeds 1 (d^2 >= 0), so clamping it to [0, 1],
    as the fma's output modifier does, is the switch's max(u, 0)."""
    q2 = np.concatenate([np.linspace(0.0, 200.0, 20001), [0.0, 81.0, 1e-30, 1e30]])
    u = 1.0 - q2 / 81.0
    assert np.all(u <= 1.0)
    np.testing.assert_array_equal(np.clip(u, 0.0, 1.0), np.maximum(u, 0.0))
