"""Reference orbits through the library's multi-limb fixed-point loop.

``holo_orbit`` and ``bs_orbit`` have the signature and results of their
mpmath counterparts in ``mpmath_loop`` (and of the reference's MPFR loops,
mpmath_utils/FP_loop.pyx:274-421 and 1828-1979) but iterate in C++
(``fsg_ref_orbit``, csrc/fsg_fixed.cpp): the reference point is rounded to
``prec`` bits as mpmath rounds it, then handed over exactly as a fixed-point
number with at least ``prec + 128`` fraction bits.
"""
import ctypes
import math

import mpmath
import numpy as np

from . import _native

GUARD_BITS = 128


def layout(prec, M, degree):
    """(n_limbs, frac_limbs) for values of |.| <= M^degree at prec bits."""
    nf = -(-(int(prec) + GUARD_BITS) // 64)
    int_bits = degree * math.log2(M + 2.) + 3.
    ni = max(1, math.ceil(int_bits / 64.))
    return nf + ni, nf


def to_limbs(v, frac_bits, n_limbs):
    """(neg, uint64[n_limbs]) of round(v * 2^frac_bits), ties to even."""
    sign, man, exp, _ = mpmath.mpf(v)._mpf_      # value = (-1)^sign man 2^exp
    if not man and exp:
        raise ValueError("infinite or nan value")
    neg = bool(sign)
    q = int(man)
    sh = int(exp) + frac_bits
    if sh >= 0:
        q <<= sh
    else:
        r = q & ((1 << -sh) - 1)
        q >>= -sh
        half = 1 << (-sh - 1)
        if r > half or (r == half and q & 1):
            q += 1
    if q.bit_length() > 64 * n_limbs - 1:
        raise ValueError("value too large for the fixed-point layout")
    limbs = np.frombuffer(q.to_bytes(8 * n_limbs, "little"), dtype=np.uint64).copy()
    return int(neg and q != 0), limbs


def _desc(kind, c, degree, flavor, max_iter, M, prec, n_threads=0, extra_bits=0., need_xr=False):
    """(OrbitDesc, keep-alive arrays, frac_limbs) for c rounded to prec bits."""
    with mpmath.workprec(int(prec)):
        c = mpmath.mpc(c)
        n, nf = layout(prec, M * 2. ** (extra_bits / max(degree, 2)),
                       degree if kind != _native.FSG_KIND_BS else 2)
        re_neg, re_l = to_limbs(c.real, 64 * nf, n)
        im_neg, im_l = to_limbs(c.imag, 64 * nf, n)
    d = _native.OrbitDesc()
    d.kind, d.degree, d.flavor, d.need_xr = kind, degree, flavor, int(bool(need_xr))
    d.n_limbs, d.frac_limbs = n, nf
    d.c_re_neg, d.c_im_neg = re_neg, im_neg
    d.c_re, d.c_im = re_l.ctypes.data, im_l.ctypes.data
    d.max_iter = int(max_iter)
    d.M = float(M)
    d.n_threads = int(n_threads)
    return d, (re_l, im_l), nf


def _run(kind, c, degree, flavor, ref_len, M, need_xr, prec, n_threads):
    lib = _native.load()
    d, keep, _ = _desc(kind, c, degree, flavor, ref_len - 1, M, prec, n_threads, need_xr=need_xr)
    Zn = np.zeros(ref_len, dtype=np.complex128)
    cap = min(ref_len, max(4096, ref_len // 16)) if need_xr else 0
    while True:
        idx = np.zeros(cap, dtype=np.int64)
        mant = np.zeros((cap, 2), dtype=np.float64)
        ex = np.zeros((cap, 2), dtype=np.int32)
        o = _native.OrbitOut()
        o.Zn = Zn.ctypes.data
        o.xr_cap = cap
        o.xr_index, o.xr_mant, o.xr_exp = (a.ctypes.data if cap else None for a in (idx, mant, ex))
        rc = lib.fsg_ref_orbit(ctypes.byref(d), ctypes.byref(o))
        if rc == _native.FSG_E_RANGE and o.n_xr > cap:
            cap = int(o.n_xr)
            continue
        _native.check(rc)
        break
    k = int(o.n_xr)
    return int(o.div_iter), Zn, idx[:k], mant[:k], ex[:k]


def holo_orbit(c, degree, ref_len, M, need_xr, prec, n_threads=0):
    """Orbit of z -> z^d + c from 0 (mpmath_loop.holo_orbit's results):
    (div_iter, Zn_path[ref_len], {i: (re_m, im_m, exp)}).  n_threads: host
    threads sharing each product (0: the library's choice)."""
    kind = _native.FSG_KIND_M2 if degree == 2 else _native.FSG_KIND_MN
    div, Zn, idx, mant, ex = _run(kind, c, degree, 1, ref_len, M, need_xr, prec, n_threads)
    xr = {int(i): (float(m[0]), float(m[1]), int(e[0])) for i, m, e in zip(idx, mant, ex)}
    return div, Zn, xr


def bs_orbit(c, flavor, ref_len, M, need_xr, prec, n_threads=0):
    """Orbit of the burning-ship family from (0, 0) (mpmath_loop.bs_orbit's
    results): (div_iter, Zn_path, {i: ((x_m, x_e), (y_m, y_e))})."""
    div, Zn, idx, mant, ex = _run(_native.FSG_KIND_BS, c, 2, flavor, ref_len, M, need_xr, prec, n_threads)
    xr = {int(i): ((float(m[0]), int(e[0])), (float(m[1]), int(e[1]))) for i, m, e in zip(idx, mant, ex)}
    return div, Zn, xr


# ---------------------------------------------------------------------------
# period (ball method) and Newton search for the nucleus
def _kind(degree):
    return _native.FSG_KIND_M2 if degree == 2 else _native.FSG_KIND_MN


def _ball(kind, c, px, degree, flavor, maxiter, M_divergence, prec):
    lib = _native.load()
    with mpmath.workprec(int(prec)):
        m, e = mpmath.frexp(mpmath.mpf(px))
    d, keep, _ = _desc(kind, c, degree, flavor, maxiter, M_divergence, prec, extra_bits=8.)
    order = ctypes.c_int64()
    _native.check(lib.fsg_ball_period(ctypes.byref(d), float(m), int(e), ctypes.byref(order)))
    return None if order.value < 0 else int(order.value)


def holo_ball_method(c, px, degree, maxiter, M_divergence, prec):
    """First n with |z_n / (dz_n/dc)| < px, None if the orbit escapes
    (mpmath_loop.holo_ball_method; FP_loop.pyx:631-758)."""
    return _ball(_kind(degree), c, px, degree, 1, maxiter, M_divergence, prec)


def bs_ball_method(c, px, flavor, maxiter, M_divergence, prec):
    """Ball-method period of the burning-ship family with the 2x2 Jacobian
    (mpmath_loop.bs_ball_method; FP_loop.pyx:2357-2462)."""
    return _ball(_native.FSG_KIND_BS, c, px, 2, flavor, maxiter, M_divergence, prec)


NEWTON_M = 2. ** 16       # the fixed-point range of the Newton orbits


def _newton_values(kind, c, degree, flavor, order, prec, divisors):
    """The mpf values of fsg_newton_eval at c, or None when the orbit left
    the fixed-point range."""
    lib = _native.load()
    d, keep, nf = _desc(kind, c, degree, flavor, order, NEWTON_M, prec, extra_bits=8.)
    n = d.n_limbs
    per = 6 if kind == _native.FSG_KIND_BS else 4
    n_rec = (sum(1 for i in range(1, order) if order % i == 0) + 1) if divisors else 1
    cap = per * n_rec
    limbs = np.zeros((cap, n), dtype=np.uint64)
    neg = np.zeros(cap, dtype=np.int32)
    ex = np.zeros(cap, dtype=np.int64)
    v = _native.MpValues()
    v.cap, v.limbs, v.neg, v.exp = cap, limbs.ctypes.data, neg.ctypes.data, ex.ctypes.data
    rc = lib.fsg_newton_eval(ctypes.byref(d), int(order), int(bool(divisors)), ctypes.byref(v))
    if rc == _native.FSG_E_RANGE:
        return None
    _native.check(rc)
    out = []
    with mpmath.workprec(int(prec)):
        for k in range(int(v.n)):
            man = int.from_bytes(limbs[k].tobytes(), "little")
            val = mpmath.ldexp(mpmath.mpf(man), int(ex[k]) - 64 * nf)
            out.append(-val if neg[k] else val)
    return out


def _nan():
    return mpmath.mpc("nan", "nan")


def find_nucleus_m2(c, order, eps_pixel, prec, max_newton=80, eps_cv=None):
    """Newton on z_order(c) / prod_{i | order, i < order} z_i(c) for z^2 + c
    (mpmath_loop.find_nucleus_m2; FP_loop.pyx:900-1116): the orbit and its
    derivative come from fsg_newton_eval, the step is formed at prec bits."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            vals = _newton_values(_native.FSG_KIND_M2, c, 2, 1, order, prec, True)
            if vals is None:
                return False, _nan()
            recs = [(mpmath.mpc(vals[k], vals[k + 1]), mpmath.mpc(vals[k + 2], vals[k + 3]))
                    for k in range(0, len(vals), 4)]
            h = mpmath.mpc(1)
            dh = mpmath.mpc(0)
            for zi, dzi in recs[:-1]:
                h = h * zi
                dh = dzi / zi + dh
            zr, dzrdc = recs[-1]
            f = zr / h
            df = (dzrdc - zr * dh) / h
            step = f / df
            c = c - step
            if abs(step) <= eps:
                return abs(zr) <= eps_pixel, c
    return False, _nan()


def find_any_nucleus(c, order, degree, eps_pixel, prec, max_newton=80, eps_cv=None):
    """Newton on z_order(c) = 0 for z^d + c (mpmath_loop.find_any_nucleus;
    FP_loop.pyx:1159-1329)."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            vals = _newton_values(_kind(degree), c, degree, 1, order, prec, False)
            if vals is None:
                return False, _nan()
            zr = mpmath.mpc(vals[0], vals[1])
            dzrdc = mpmath.mpc(vals[2], vals[3])
            step = zr / dzrdc
            c = c - step
            if abs(step) <= eps:
                return abs(zr) <= eps_pixel, c
    return False, _nan()


def bs_find_any_nucleus(c, order, flavor, eps_pixel, prec, max_newton=80, eps_cv=None):
    """2x2 Newton on (x_order, y_order)(a, b) = 0 for the burning-ship family
    (mpmath_loop.bs_find_any_nucleus; FP_loop.pyx:2564-2752)."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        a, b = c.real, c.imag
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            vals = _newton_values(_native.FSG_KIND_BS, mpmath.mpc(a, b), 2, flavor, order, prec, False)
            if vals is None:
                return False, _nan()
            x, y, D0, D1, D2, D3 = vals
            inv = 1 / (D0 * D3 - D2 * D1)
            da = (D3 * x - D1 * y) * inv
            db = (D0 * y - D2 * x) * inv
            a = a - da
            b = b - db
            if mpmath.hypot(da, db) <= eps:
                return mpmath.hypot(x, y) <= eps_pixel, mpmath.mpc(a, b)
    return False, _nan()
