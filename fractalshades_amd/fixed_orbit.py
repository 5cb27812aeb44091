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


def _run(kind, c, degree, flavor, ref_len, M, need_xr, prec, n_threads):
    lib = _native.load()
    with mpmath.workprec(int(prec)):
        c = mpmath.mpc(c)
        n, nf = layout(prec, M, degree if kind != _native.FSG_KIND_BS else 2)
        re_neg, re_l = to_limbs(c.real, 64 * nf, n)
        im_neg, im_l = to_limbs(c.imag, 64 * nf, n)
    d = _native.OrbitDesc()
    d.kind, d.degree, d.flavor, d.need_xr = kind, degree, flavor, int(bool(need_xr))
    d.n_limbs, d.frac_limbs = n, nf
    d.c_re_neg, d.c_im_neg = re_neg, im_neg
    d.c_re, d.c_im = re_l.ctypes.data, im_l.ctypes.data
    d.max_iter = ref_len - 1
    d.M = float(M)
    d.n_threads = int(n_threads)
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
