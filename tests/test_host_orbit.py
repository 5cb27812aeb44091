"""Reference orbits on the host: the multi-limb fixed-point loop
(fsg_ref_orbit, fractalshades_amd.fixed_orbit) against

* the reference's own known answers (tests/test_FP_loop.py, fixtures in
  tests/golden/fp_loop_kats.json): the Xrange register of a 10^7-step orbit at
  3520 digits, and the ball-method period 4252;
* mpmath at the same precision (mpmath_loop): bit-identical complex128 orbit,
  escape index and Xrange register wherever the orbit is determined at that
  precision, for z^2 + c, z^d + c (d = 3..8) and the five burning-ship
  flavours;
* mpmath at a higher precision where it is not (chaotic orbits, iterates at
  the noise level of a deep nucleus): the fixed-point orbit, which carries 128
  more bits, follows the higher-precision one at least as far.
"""
import json
import os

import mpmath
import numpy as np
import pytest

from fractalshades_amd import fixed_orbit as fo
from fractalshades_amd import mpmath_loop as ml
from fractalshades_amd import settings
from fractalshades_amd._native import FsgError

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "fp_loop_kats.json")))
CENTRES = json.load(open(os.path.join(os.path.dirname(HERE), "fractalshades_amd", "data",
                                      "serve_centres.json")))


def _c(x, y, prec):
    with mpmath.workprec(prec):
        return mpmath.mpc(mpmath.mpf(x), mpmath.mpf(y))


def _assert_same(a, b):
    assert a[0] == b[0]
    np.testing.assert_array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
    assert a[2] == b[2]


def _first_diff(a, b):
    w = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    return int(w[0]) // 2 if len(w) else len(a)


@pytest.mark.parametrize("prec", [120, 1100, 3500])
@pytest.mark.parametrize("c", [("-0.12", "0.75"), ("-1.3", "0.02"), ("0.3", "0.5"), ("-0.7454294", "0.113089")])
def test_m2_orbit_equals_mpmath(prec, c):
    cc = _c(*c, prec)
    _assert_same(fo.holo_orbit(cc, 2, 1500, 2000., True, prec),
                 ml.holo_orbit(cc, 2, 1500, 2000., True, prec))


@pytest.mark.parametrize("d", [3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("c", [("0.1", "0.2"), ("-0.5", "0.55"), ("0.45", "0.45")])
def test_mn_orbit_equals_mpmath(d, c):
    cc = _c(*c, 300)
    _assert_same(fo.holo_orbit(cc, d, 800, 2000., True, 300),
                 ml.holo_orbit(cc, d, 800, 2000., True, 300))


@pytest.mark.parametrize("flavor", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("c", [("-0.3", "0.1"), ("-0.6", "-0.25"), ("0.4", "0.6")])
def test_bs_orbit_equals_mpmath(flavor, c):
    cc = _c(*c, 300)
    _assert_same(fo.bs_orbit(cc, flavor, 800, 2000., True, 300),
                 ml.bs_orbit(cc, flavor, 800, 2000., True, 300))


def test_bounded_cases_are_bounded():
    """The comparisons above include orbits that never escape."""
    assert fo.holo_orbit(_c("-0.12", "0.75", 120), 2, 1500, 2000., False, 120)[0] == 1500
    assert fo.holo_orbit(_c("0.1", "0.2", 300), 5, 800, 2000., False, 300)[0] == 800
    assert fo.bs_orbit(_c("-0.3", "0.1", 300), 1, 800, 2000., False, 300)[0] == 800


@pytest.mark.parametrize("flavor", [1, 5])
def test_chaotic_orbit_follows_higher_precision(flavor):
    """A chaotic burning-ship orbit loses its binary64 digits once the
    precision is spent; until then all agree, and the fixed-point orbit keeps
    following the 256-bit-deeper one at least as long as mpmath does."""
    prec = 300
    cc = _c("-1.75", "0.02", prec)
    hi = ml.bs_orbit(cc, flavor, 3000, 2000., False, prec + 256)[1]
    lo = ml.bs_orbit(cc, flavor, 3000, 2000., False, prec)[1]
    nat = fo.bs_orbit(cc, flavor, 3000, 2000., False, prec)[1]
    k_lo = _first_diff(lo, hi)
    assert _first_diff(nat, lo) >= k_lo
    assert _first_diff(nat, hi) >= k_lo


def test_deep_nucleus_register():
    """A period-76 nucleus at 2100 digits: the iterates at multiples of the
    period are ~1e-2083, at the rounding noise of a 7003-bit orbit.  The
    binary64 orbit and the register's indices agree with mpmath's; the
    register's values agree with a 512-bit-deeper orbit to its noise, and the
    fixed-point ones are at least as close as mpmath's."""
    cen = CENTRES["centres"][0]
    prec = int(CENTRES["digits"] * 3.33) + 10
    cc = _c(cen["x"], cen["y"], prec)
    n = 400
    nat = fo.holo_orbit(cc, 2, n, 2000., True, prec)
    mp = ml.holo_orbit(cc, 2, n, 2000., True, prec)
    hi = ml.holo_orbit(cc, 2, n, 2000., True, prec + 512)
    assert nat[0] == mp[0] == n
    np.testing.assert_array_equal(nat[1].view(np.uint64), mp[1].view(np.uint64))
    assert sorted(nat[2]) == sorted(mp[2]) == sorted(hi[2]) == [76 * k for k in range(1, 6)]

    def val(t):
        return complex(t[0], t[1]) * 2.0 ** (t[2] + 6900)
    for i in nat[2]:
        h = val(hi[2][i])
        e_nat, e_mp = abs(val(nat[2][i]) - h), abs(val(mp[2][i]) - h)
        assert e_nat <= max(e_mp, 1e-12 * abs(h))
        assert e_mp <= 1e-4 * abs(h)


@pytest.mark.parametrize("threads", [2, 3, 4])
def test_threads_give_the_same_orbit(threads):
    cen = CENTRES["centres"][1]
    prec = int(CENTRES["digits"] * 3.33) + 10
    cc = _c(cen["x"], cen["y"], prec)
    _assert_same(fo.holo_orbit(cc, 2, 300, 2000., True, prec, n_threads=threads),
                 fo.holo_orbit(cc, 2, 300, 2000., True, prec, n_threads=1))
    _assert_same(fo.bs_orbit(cc, 3, 300, 2000., True, prec, n_threads=threads),
                 fo.bs_orbit(cc, 3, 300, 2000., True, prec, n_threads=1))


def test_edge_cases():
    # c = 0: every iterate is 0, every one goes to the register as (0, 0, 0)
    _assert_same(fo.holo_orbit(0, 2, 50, 2000., True, 100), ml.holo_orbit(0, 2, 50, 2000., True, 100))
    assert fo.holo_orbit(0, 2, 50, 2000., True, 100)[2][49] == (0.0, 0.0, 0)
    # a register larger than the first buffer (regrown)
    _assert_same(fo.holo_orbit(0, 2, 9000, 2000., True, 100), ml.holo_orbit(0, 2, 9000, 2000., True, 100))
    # no iterate; escape at the first step
    _assert_same(fo.holo_orbit(1, 2, 1, 2000., True, 100), ml.holo_orbit(1, 2, 1, 2000., True, 100))
    _assert_same(fo.holo_orbit(3, 2, 10, 2., True, 100), ml.holo_orbit(3, 2, 10, 2., True, 100))
    _assert_same(fo.bs_orbit(3j, 2, 10, 2., True, 100), ml.bs_orbit(3j, 2, 10, 2., True, 100))
    # the last index: no escape up to max_iter
    _assert_same(fo.holo_orbit(-1, 2, 7, 2000., True, 100), ml.holo_orbit(-1, 2, 7, 2000., True, 100))


def test_to_limbs_rounds_to_nearest_even():
    with mpmath.workprec(200):
        v = mpmath.mpf(3) * mpmath.mpf(2) ** -66          # 0.75 units of 2^-64
        assert fo.to_limbs(v, 64, 2)[1][0] == 1
        v = mpmath.mpf(2) ** -65                          # half a unit: ties to even (0)
        assert fo.to_limbs(v, 64, 2)[1][0] == 0
        v = mpmath.mpf(3) * mpmath.mpf(2) ** -65          # 1.5 units: to 2
        assert fo.to_limbs(v, 64, 2)[1][0] == 2
        neg, limbs = fo.to_limbs(-mpmath.mpf("1.5"), 64, 2)
        assert neg == 1 and limbs[1] == 1 and limbs[0] == 1 << 63
        assert fo.to_limbs(mpmath.mpf(0), 64, 2)[0] == 0
    with pytest.raises(ValueError):
        fo.to_limbs(mpmath.mpf(2) ** 70, 64, 2)


def test_escape_radius_too_large_for_layout():
    import ctypes
    from fractalshades_amd import _native
    lib = _native.load()
    d = _native.OrbitDesc()
    limbs = np.zeros(2, dtype=np.uint64)
    d.kind, d.degree, d.flavor, d.need_xr = 1, 8, 1, 0
    d.n_limbs, d.frac_limbs = 2, 1
    d.c_re = d.c_im = limbs.ctypes.data
    d.max_iter, d.M = 5, 1e6                  # 8 log2(1e6) > 64 integer bits
    Zn = np.zeros(6, dtype=np.complex128)
    o = _native.OrbitOut()
    o.Zn = Zn.ctypes.data
    with pytest.raises(FsgError):
        _native.check(lib.fsg_ref_orbit(ctypes.byref(d), ctypes.byref(o)))


def test_ball_method_kat():
    """tests/test_FP_loop.py:166-193: period 4252."""
    k = KATS["ball_method"]
    prec = k["precision_bits"]
    order = ml.holo_ball_method(_c(k["x"], k["y"], prec), mpmath.mpf(k["radius"]), 2,
                                k["maxiter"], k["M_divergence"], prec)
    assert order == k["order"]


def _ref_path_frame(max_iter):
    import fractalshades_amd.models as fsm
    k = KATS["ref_path"]
    old = settings.no_newton
    settings.no_newton = True          # test_FP_loop.py:94
    try:
        f = fsm.Perturbation_mandelbrot(None)
        f.zoom(precision=k["precision"], x=k["x"], y=k["y"], dx=k["dx"], nx=k["nx"], xy_ratio=1.0,
               theta_deg=0.)
        f.calc_std_div(calc_name="test", subset=None, max_iter=max_iter,
                       M_divergence=k["M_divergence"], epsilon_stationnary=1.e-3,
                       interior_detect=False, BLA_eps=1.e-6, calc_dzndc=False)
        f.get_FP_orbit()
    finally:
        settings.no_newton = old
    return f


def _check_ref_path(f, max_iter):
    k = KATS["ref_path"]
    assert settings.orbit_backend == "native"
    xr = f.FP_params["xr"]
    want = [i for i in k["xr_indices"] if i <= max_iter]
    assert sorted(xr) == want
    for i, is_xr in k["membership"]:
        if i <= max_iter:
            assert (i in xr) == is_xr
    assert f.FP_params["div_iter"] == max_iter + 1


def test_ref_path_kat_first_two_hits():
    """The reference's 3520-digit orbit (test_FP_loop.py:88-158) up to
    1,351,600 steps: the first two Xrange iterations, 675728 and 1351568."""
    max_iter = 1_351_600
    _check_ref_path(_ref_path_frame(max_iter), max_iter)


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("FSG_SLOW"), reason="10^7-step orbit (minutes); FSG_SLOW=1 runs it")
def test_ref_path_kat_full():
    """The whole 10^7-step orbit: the 14 Xrange iterations of test_numba_path
    and the membership checks of numba_path_loop (test_FP_loop.py:66-80)."""
    max_iter = KATS["ref_path"]["max_iter"]
    _check_ref_path(_ref_path_frame(max_iter), max_iter)


# ---- period and nucleus ------------------------------------------------------
def test_ball_method_kat_native():
    """tests/test_FP_loop.py:166-193 with the fixed-point loop: period 4252."""
    k = KATS["ball_method"]
    prec = k["precision_bits"]
    order = fo.holo_ball_method(_c(k["x"], k["y"], prec), mpmath.mpf(k["radius"]), 2,
                                k["maxiter"], k["M_divergence"], prec)
    assert order == k["order"]


@pytest.mark.parametrize("degree", [2, 3, 5])
@pytest.mark.parametrize("seed", range(4))
def test_holo_ball_method_equals_mpmath(degree, seed):
    rng = np.random.default_rng(100 * degree + seed)
    prec = 160
    c = _c(repr(rng.uniform(-1.6, 0.4)), repr(rng.uniform(0., 1.0)), prec)
    for px in ("1e-3", "1e-9", "1e-25"):
        a = fo.holo_ball_method(c, mpmath.mpf(px), degree, 4000, 1000., prec)
        b = ml.holo_ball_method(c, mpmath.mpf(px), degree, 4000, 1000., prec)
        assert a == b


@pytest.mark.parametrize("flavor", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("seed", range(3))
def test_bs_ball_method_equals_mpmath(flavor, seed):
    rng = np.random.default_rng(10 * flavor + seed)
    prec = 160
    c = _c(repr(rng.uniform(-1.9, 0.4)), repr(rng.uniform(-1.0, 0.2)), prec)
    for px in ("1e-3", "1e-9", "1e-25"):
        a = fo.bs_ball_method(c, mpmath.mpf(px), flavor, 3000, 1000., prec)
        b = ml.bs_ball_method(c, mpmath.mpf(px), flavor, 3000, 1000., prec)
        assert a == b


def test_newton_values_equal_mpmath():
    """fsg_newton_eval's orbit point and derivative at the order and at its
    divisors agree with a 200-bit-deeper mpmath orbit to the working
    precision (the fixed-point loop carries 128 guard bits)."""
    from fractalshades_amd import _native
    prec = 300
    c = _c("-1.7548776662466927600495", "0.0003", prec)
    order = 12
    vals = fo._newton_values(_native.FSG_KIND_M2, c, 2, 1, order, prec, True)
    with mpmath.workprec(prec + 200):
        z, dz, want = mpmath.mpc(0), mpmath.mpc(0), []
        for i in range(1, order + 1):
            dz = 2 * dz * z + 1
            z = z * z + c
            if order % i == 0:
                want += [z.real, z.imag, dz.real, dz.imag]
        assert len(vals) == len(want) == 4 * 6
        for v, w in zip(vals, want):
            assert abs(v - w) <= max(abs(w), 1) * mpmath.ldexp(1, -prec)


@pytest.mark.parametrize("name", ["M2_divref5", "M2_int32", "BS_ex14", "mn3", "mn5"])
def test_frame_reference_equals_mpmath(name):
    """The whole reference-point search of a test frame (ball-method period,
    Newton nucleus, orbit) with either backend: same period, nucleus within
    the working precision, identical complex128 orbit and Xrange register."""
    import frames
    out = {}
    old = settings.orbit_backend
    try:
        for be in ("mpmath", "native"):
            settings.orbit_backend = be
            f = frames.mn_frame(int(name[2:])) if name.startswith("mn") else frames.make(name)
            f.get_FP_orbit()
            out[be] = (f.FP_params, f.Zn_path, f.prec)
    finally:
        settings.orbit_backend = old
    (pm, Zm, prec), (pn, Zn, _) = out["mpmath"], out["native"]
    assert pm["order"] == pn["order"] is not None
    assert pm["div_iter"] == pn["div_iter"]
    with mpmath.workprec(prec):
        assert abs(pm["ref_point"] - pn["ref_point"]) <= abs(pm["ref_point"]) * mpmath.ldexp(1, -prec + 4)
    np.testing.assert_array_equal(Zm.view(np.uint64), Zn.view(np.uint64))
    assert pm["xr"] == pn["xr"]


# ---- the remaining known answers of tests/test_FP_loop.py ----------------------
def test_newton_nucleus_kat():
    """tests/test_FP_loop.py:231-272: the divisor-removal Newton of z^2 + c at
    order 4252 and 11622 bits from the reference's start point converges to
    its ~3500-digit nucleus (mpmath.almosteq at precision - 10 bits, as the
    reference compares).  The orbit and derivative come from the fixed-point
    loop (fsg_newton_eval), the step is formed in mpmath."""
    k = KATS["newton"]
    prec = k["precision_bits"]
    eps = mpmath.mpf(2) ** -prec
    eps_pix = mpmath.mpf(2) ** -k["pix_bits"]
    ok, val = fo.find_nucleus_m2(_c(k["x_start"], k["y_start"], prec), k["order"], eps_pix, prec,
                                 max_newton=k["max_newton"], eps_cv=eps)
    assert ok
    with mpmath.workprec(prec - 10):
        assert mpmath.almosteq(mpmath.mpf(k["x_nucleus"]), val.real)
        assert mpmath.almosteq(mpmath.mpf(k["y_nucleus"]), val.imag)
    # the start point itself is off by ~1e-82: the test is not vacuous
    with mpmath.workprec(prec):
        assert abs(mpmath.mpf(k["x_start"]) - mpmath.mpf(k["x_nucleus"])) > mpmath.mpf("1e-90")


@pytest.mark.parametrize("backend", ["native", "mpmath"])
def test_bs_path_kat(backend):
    """tests/test_FP_loop.py:278-309: 10 burning-ship steps from c = 0.1 + 0.1i
    equal the binary64 recurrence x' = x^2 - y^2 + a, y' = 2|x y| - b.  The
    reference compares bit for bit against MPFR at 53 bits (one rounding per
    operation); the orbit loops here round each exact iterate once, so the
    two may differ by an ulp of an iterate: compared within 4 ulp."""
    k = KATS["bs_path"]
    a, b = k["x"], k["y"]
    loop = fo if backend == "native" else ml
    # npts entries, iterates 1 .. npts - 1 (FP_loop's max_iter = npts - 1)
    n, Zn, _ = loop.bs_orbit(_c(repr(a), repr(b), 300), 1, k["npts"], k["M"], False, 300)
    assert n == k["npts"]
    z = 0j
    for i in range(k["npts"]):
        assert abs(Zn[i] - z) <= 4 * np.finfo(float).eps * abs(z), (i, Zn[i], z)
        x, y = z.real, z.imag
        z = complex(x * x - y * y + a, 2 * abs(x * y) - b)


@pytest.mark.parametrize("backend", ["native", "mpmath"])
def test_bs_ball_order_kat(backend):
    """tests/test_FP_loop.py:311-341: burning-ship ball-method period 3 at
    two points (px 0.01, 100 steps, M 1000, 53 bits)."""
    k = KATS["bs_order"]
    loop = fo if backend == "native" else ml
    prec = k["precision_bits"]
    for (x, y, px), want in zip(k["points"], k["orders"]):
        order = loop.bs_ball_method(_c(x, y, prec), mpmath.mpf(px), 1, k["maxiter"],
                                    k["M_divergence"], prec)
        assert order == want


@pytest.mark.parametrize("backend", ["native", "mpmath"])
def test_bs_newton_kat(backend):
    """tests/test_FP_loop.py:343-382: the burning ship's 2x2 Newton at order 3
    from the two start points.  The reference only checks that it runs; here
    the result must also be a period-3 nucleus (x_3, y_3 = 0 to 1e-12 by
    direct iteration) within the seed radius of the start point."""
    from fractalshades_amd.mpmath_loop import bs_iterate
    k = KATS["bs_newton"]
    loop = fo if backend == "native" else ml
    for case in k["cases"]:
        prec = case["prec"]
        start = _c(case["x"], case["y"], prec)
        ok, val = loop.bs_find_any_nucleus(start, case["order"], 1, mpmath.mpf(case["px"]), prec,
                                           max_newton=k["max_newton"],
                                           eps_cv=mpmath.mpf(2) ** -prec)
        assert ok, case
        with mpmath.workprec(200):
            a, b = mpmath.mpf(val.real), mpmath.mpf(val.imag)
            x = y = mpmath.mpf(0)
            for _ in range(case["order"]):
                x, y = bs_iterate(1, x, y, a, b)
            assert mpmath.hypot(x, y) < 1e-12
            assert abs(mpmath.mpc(a, b) - start) < 10 * mpmath.mpf(case["px"])
