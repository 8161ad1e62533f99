"""The sign of the reference global phase next to odd multiples of pi (CPU).

mentpy's amplitude-sum partial trace (`mentpy/calculator/state_ops.py:68-73`)
multiplies each step's state by sgn(cos(theta/2)) e^{i theta/2}.  The GPU
kernels keep this as exp(i (sum theta/2 - pi sum m)) with the integer
m = rint(theta / 2 pi), so (-1)^m must equal sgn(cos(theta/2)) for every
float32 angle.  It changes sign at theta = (2k+1) pi, and float32(pi) lies
8.7e-8 above pi, so that angle needs m = 1 (phase -i).

The kernel computes m as fma(double(theta), fl64(1/2pi), 1.5 * 2^52), which
rounds the exact product theta * fl64(1/2pi) to the nearest integer (ties to
even).  This file emulates that arithmetic exactly with Python integers and
compares it with the true rounding of theta / 2pi (pi to 60 digits) for every
float32 within 64 ulps of +-(2k+1)pi, over |theta| < 2^12.  Farther angles
lie more than 64 ulps (>= 1e-5 relative) from a sign change, far outside the
product's relative error of 1.1e-16, so these cases are the only ones that
can fail.  The float32 product the kernel used before (rint(theta *
fl32(1/2pi)) in float32) is emulated the same way and must fail, at
float32(pi) among others.
"""
from fractions import Fraction

import numpy as np

from oracle.closed_forms import step_phase

PI_60 = Fraction("3.141592653589793238462643383279502884197169399375105820974944")
RANGE = 2.0 ** 12
ULPS = 64


def _rint_even(p, q):
    """round-half-even of p / q for integers (q > 0)."""
    f, r = divmod(p, q)
    if 2 * r > q or (2 * r == q and f & 1):
        f += 1
    return f


def _angles_near_odd_pi():
    out = []
    k = 0
    while (2 * k + 1) * np.pi < RANGE:
        for sgn in (1.0, -1.0):
            c = np.float32(sgn * (2 * k + 1) * np.pi)
            lo = c
            hi = c
            for _ in range(ULPS):
                lo = np.nextafter(lo, np.float32(-np.inf))
                hi = np.nextafter(hi, np.float32(np.inf))
            bits_lo, bits_hi = sorted((int(lo.view(np.int32)), int(hi.view(np.int32))))
            vals = np.arange(bits_lo, bits_hi + 1, dtype=np.int64).astype(np.int32).view(np.float32)
            out.append(vals)
        k += 1
    return np.concatenate(out)


ANGLES = _angles_near_odd_pi()


def _m_true(theta):
    n, d = float(theta).as_integer_ratio()
    return _rint_even(n * PI_60.denominator, 2 * d * PI_60.numerator)


def _m_kernel(theta):
    """fma(double(theta), fl64(1/2pi), 1.5 * 2^52): the exact product rounded
    to an integer with ties to even (the FMA rounds once, at unit spacing)."""
    n, d = float(theta).as_integer_ratio()
    cn, cd = (0.5 / np.pi).as_integer_ratio()
    return _rint_even(n * cn, d * cd)


def _m_float32_product(theta):
    """fmaf(theta, fl32(1/2pi), 1.5 * 2^23): the previous float32 rounding."""
    n, d = float(theta).as_integer_ratio()
    cn, cd = float(np.float32(0.5 / np.pi)).as_integer_ratio()
    return _rint_even(n * cn, d * cd)


def test_angle_set_covers_float32_pi():
    assert np.float32(np.pi) in ANGLES
    assert np.float32(-np.pi) in ANGLES
    assert np.float32(3 * np.pi) in ANGLES
    assert ANGLES.size > 100_000


def test_kernel_rounding_matches_true_sign_near_odd_multiples_of_pi():
    bad = [float(t) for t in ANGLES if _m_kernel(t) != _m_true(t)]
    assert not bad, f"{len(bad)} mismatches, first {bad[:5]}"


def test_no_float32_angle_within_the_rounding_error_of_a_sign_change():
    """The smallest distance of theta / 2pi from a half-integer over the set
    (which holds the nearest float32 to every (2k+1)pi) is far above the
    product's worst error |theta / 2pi| * 1.1e-16 < 7.2e-14."""
    cn, cd = (0.5 / np.pi).as_integer_ratio()
    rel = abs(Fraction(cn, cd) * 2 * PI_60 - 1)
    assert rel < 2.0 ** -53
    dmin = 1.0
    for t in ANGLES:
        n, d = float(t).as_integer_ratio()
        p, q = n * PI_60.denominator, 2 * d * PI_60.numerator
        r = p % q
        dmin = min(dmin, abs(2 * r - q) / (2 * q))
    print(f"closest float32 theta/2pi to a half-integer: {dmin:.3e}")
    assert dmin > 1e-9
    assert dmin > 100 * (RANGE / (2 * np.pi)) * float(rel)


def test_oracle_step_phase_sign_matches_true_rounding():
    """The oracle's sgn(cos(theta/2)) (double cos of the float32 angle) has the
    sign of (-1)^m on every angle of the set."""
    s = np.real(step_phase(ANGLES.astype(np.float64)) * np.exp(-0.5j * ANGLES.astype(np.float64)))
    m = np.array([_m_true(t) for t in ANGLES])
    assert np.array_equal(np.sign(s), np.where(m % 2 == 0, 1.0, -1.0))


def test_float32_product_fails_at_float32_pi():
    """The float32 rounding (a 4e-8 too small constant) is caught: it gives
    m = 0 (phase +i) at float32(pi), where the reference has -i."""
    t = np.float32(np.pi)
    assert _m_true(t) == 1 and _m_kernel(t) == 1
    assert _m_float32_product(t) == 0
    bad = sum(_m_float32_product(x) != _m_true(x) for x in ANGLES)
    assert bad > 0
