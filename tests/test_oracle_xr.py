"""Pins the oracle's extended-range arithmetic against float64 closed forms,
on the reference's seeded test vectors (tests/test_numba_xr.py:23-46:
generate_random_xr, seeds 100 / 800, |exp| < 200 / 800): add and sub are
exact, products within 2 ulp, to_standard and normalize exact."""
import math

import numpy as np
import pytest

import oracle as orc


def gen_xr(nvec, max_bin_exp=200, seed=100, complex_=True):
    rg = np.random.default_rng(seed)
    if complex_:
        m = (rg.random([nvec]) * 2. - 1.) + 1j * (2. * rg.random([nvec]) - 1.)
    else:
        m = rg.random([nvec]) * 2. - 1.
    exp = rg.integers(low=-max_bin_exp, high=max_bin_exp, size=[nvec]).astype(np.int32)
    std = m * (2. ** exp)
    return m, exp, std


def to_std_c(om, oe):
    return orc.xrc_to_std(om, oe).view(np.complex128)


@pytest.mark.parametrize("op", ["add", "sub"])
def test_add_sub_exact(op):
    ma, ea, sa = gen_xr(10000)
    mb, eb, sb = gen_xr(10000, seed=800)
    code = 0 if op == "add" else 1
    om, oe = orc.xrc_binop(code, ma.view(np.float64).copy(), ea, mb.view(np.float64).copy(), eb)
    res = to_std_c(om, oe)
    expected = sa + sb if op == "add" else sa - sb
    np.testing.assert_array_equal(res, expected)


def test_real_add_sub_exact():
    ma, ea, sa = gen_xr(10000, complex_=False)
    mb, eb, sb = gen_xr(10000, seed=800, complex_=False)
    for code, exp_ in ((0, sa + sb), (1, sa - sb)):
        om, oe = orc.xrf_binop(code, ma, ea, mb, eb)
        np.testing.assert_array_equal(np.ldexp(om, oe), exp_)


def test_mul_2ulp():
    ma, ea, sa = gen_xr(100000, max_bin_exp=200)
    mb, eb, sb = gen_xr(100000, max_bin_exp=200, seed=800)
    om, oe = orc.xrc_binop(2, ma.view(np.float64).copy(), ea, mb.view(np.float64).copy(), eb)
    res = to_std_c(om, oe)
    expected = sa * sb
    np.testing.assert_allclose(res, expected, rtol=2 * np.finfo(np.float64).eps)


def test_mul_large_exponents_stays_normalised():
    # products far outside the float64 range keep a normalised mantissa
    ma, ea, _ = gen_xr(1000, max_bin_exp=800)
    mb, eb, _ = gen_xr(1000, max_bin_exp=800, seed=800)
    ea = ea * 2
    eb = eb * 2
    om, oe = orc.xrc_binop(2, ma.view(np.float64).copy(), ea, mb.view(np.float64).copy(), eb)
    m = om.view(np.complex128)
    mant_exp = np.frexp(np.maximum(abs(m.real), abs(m.imag)))[1]
    assert np.all(abs(mant_exp) <= 102)
    # value: compare log2 magnitudes in float (mantissa products are exact to 2 ulp)
    lhs = np.log2(abs(m)) + oe
    rhs = np.log2(abs(ma)) + ea + np.log2(abs(mb)) + eb
    np.testing.assert_allclose(lhs, rhs, atol=1e-9)


def test_to_standard_exact():
    m, e, std = gen_xr(5000, max_bin_exp=800)
    res = to_std_c(m.view(np.float64).copy(), e)
    np.testing.assert_array_equal(res, std)


def test_to_standard_signed_zeros():
    """to_standard of a complex (numba_xr.py:825-828) evaluates
    (co_nm_real + 1j * co_nm_imag) * ldexp(1., co_nexp) with the float
    operands promoted to complex, so the sign of a zero component (an exact
    zero, or a product that underflows) follows complex arithmetic, not a
    per-component product.  CPython 3.10 promotes mixed float / complex
    operands the same way, so the expression evaluated by Python is the
    check.  Inputs are normalised (larger component in [1, 2), the other
    within 2^-60 of it, or a zero of either sign), where the co-exponent
    step leaves the mantissas as they are."""
    import sys
    assert sys.version_info < (3, 14)    # 3.14 adds mixed-mode complex arithmetic
    rg = np.random.default_rng(11)
    n = 4000
    big = (1. + rg.random(n)) * rg.choice([-1., 1.], n)
    small = ldexp_arr((1. + rg.random(n)) * rg.choice([-1., 1.], n), -rg.integers(0, 60, n))
    small[rg.random(n) < 0.3] = 0.
    small[rg.random(n) < 0.15] = -0.
    swap = rg.random(n) < 0.5
    re, im = np.where(swap, small, big), np.where(swap, big, small)
    # half of the exponents where the products underflow to zero
    e = np.where(rg.random(n) < 0.5, rg.integers(-1200, -1070, n), rg.integers(-1000, 60, n)).astype(np.int32)
    got = orc.xrc_to_std(np.stack([re, im], 1).ravel().copy(), e).view(np.complex128)
    want = np.array([(float(x) + 1j * float(y)) * math.ldexp(1., int(k)) for x, y, k in zip(re, im, e)])
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    # the cases where the promoted product differs from the per-component one occur
    naive = np.array([complex(float(x) * math.ldexp(1., int(k)), float(y) * math.ldexp(1., int(k)))
                      for x, y, k in zip(re, im, e)])
    assert np.count_nonzero(naive.view(np.uint64) != want.view(np.uint64)) > 50


def ldexp_arr(m, k):
    return np.array([math.ldexp(float(a), int(b)) for a, b in zip(m, k)])


def test_normalize_exact():
    m, e, std = gen_xr(5000, max_bin_exp=200)
    om, oe = orc.xrc_normalize(m.view(np.float64).copy(), e)
    got = om.view(np.complex128) * (2. ** oe.astype(np.float64))
    np.testing.assert_array_equal(got, std)
    mm = om.view(np.complex128)
    big = np.maximum(abs(mm.real), abs(mm.imag))
    assert np.all((big >= 1.) & (big < 2.))


def test_hypot_within_one_ulp():
    rg = np.random.default_rng(3)
    x = rg.standard_normal(20000) * 10. ** rg.integers(-300, 300, 20000)
    y = rg.standard_normal(20000) * 10. ** rg.integers(-300, 300, 20000)
    for a, b in zip(x, y):
        h = orc.hypot(a, b)
        ref = math.hypot(a, b)
        assert abs(h - ref) <= math.ulp(ref), (a, b, h, ref)
    assert orc.hypot(0., 0.) == 0.
    assert orc.hypot(3., 4.) == 5.
    assert math.isinf(orc.hypot(math.inf, 1.))
