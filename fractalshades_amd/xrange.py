"""Host-side conversions between mpmath numbers and extended-range pairs.

An extended-range ("Xrange") value is ``mantissa * 2**exp`` with a float64
(or complex128) mantissa and an int32 exponent.  These helpers produce the
(mantissa, exponent) pairs the engine consumes; they follow the value
semantics of the reference's ``mpf_to_Xrange`` / ``mpc_to_Xrange``
(src/fractalshades/numpy_utils/xrange.py:8-50) and of MPFR's
``mpfr_get_d_2exp`` used for the orbit register (mpmath_utils/FP_loop.pyx:
423-446).
"""
import math

import mpmath


def mpf_to_xr(x):
    """(m, e) with 0.5 <= |m| < 1 (or m == 0) and m * 2**e == x rounded to 53 bits."""
    x = mpmath.mpf(x)
    if x == 0:
        return 0.0, 0
    m, e = mpmath.frexp(x)
    fm = float(m)          # round to nearest: |m| in [0.5, 1) is a normal number
    if abs(fm) == 1.0:     # rounding carried into the next binade
        fm = math.copysign(0.5, fm)
        e += 1
    return fm, int(e)


def mpc_to_xr(z):
    """Complex (re_m, im_m, e): both parts share the larger exponent."""
    z = mpmath.mpc(z)
    rm, re_ = mpf_to_xr(z.real)
    im, ie = mpf_to_xr(z.imag)
    if rm == 0.0 and im == 0.0:
        return 0.0, 0.0, 0
    if rm == 0.0:
        return 0.0, im, ie
    if im == 0.0:
        return rm, 0.0, re_
    e = max(re_, ie)
    return math.ldexp(rm, re_ - e), math.ldexp(im, ie - e), e


def xr_to_float(m, e):
    """Standard float of an Xrange pair (0 below the float64 range)."""
    return math.ldexp(m, e) if m != 0 else 0.0


def xr_to_mpf(m, e):
    return mpmath.ldexp(mpmath.mpf(m), int(e))


def mpf_to_double(x):
    """float64 nearest to the mpf x, correctly rounded down to subnormals;
    a value that rounds to zero keeps its sign (mpfr_get_d, MPFR_RNDN, as
    the reference's orbit loop reads its iterates, FP_loop.pyx:376-377)."""
    x = mpmath.mpf(x)
    if x == 0:
        return 0.0
    m, e = mpmath.frexp(x)
    if e - 1 >= -1022:
        return float(x)
    # subnormal range: round once, on the 2**-1074 grid
    q = int(mpmath.nint(mpmath.ldexp(abs(x), 1074)))
    return math.copysign(math.ldexp(q, -1074), -1. if x < 0 else 1.)
