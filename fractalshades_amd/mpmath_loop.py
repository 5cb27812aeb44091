"""Arbitrary-precision host routines: reference orbit, period and nucleus search.

These stay on the host, in arbitrary precision, as in the reference (which
runs them in MPFR/MPC from a Cython module).  This module uses mpmath; every
routine works at the caller's precision (``prec`` in bits) inside
``mpmath.workprec`` so that concurrent calculations do not share the global
mpmath context.

Reference counterparts (src/fractalshades/mpmath_utils/FP_loop.pyx):
  * orbit: perturbation_mandelbrotN_select_FP_loop 274-421,
    perturbation_nonholomorphic_FP_loop 1828-1979
  * ball-method period: perturbation_mandelbrotN_select_ball_method 631-758,
    perturbation_nonholomorphic_ball_method 2357-2462
  * Newton: perturbation_mandelbrot_find_nucleus 900-1116 (divisor removal),
    perturbation_mandelbrotN_select_find_any_nucleus 1159-1329,
    perturbation_nonholomorphic_find_any_nucleus 2564-2752
  * size estimate: perturbation_mandelbrotN_select_nucleus_size_estimate 487-602
"""
import math

import mpmath
import numpy as np

from .xrange import mpf_to_double, mpf_to_xr, mpc_to_xr

XR_TSHOLD = 1.e-300


# ---------------------------------------------------------------------------
# elementary iterations
def _zpow(z, d):
    if d == 2:
        return z * z
    return z ** d


def bs_iterate(flavor, x, y, a, b):
    """One step of the burning-ship family (burning_ship.py:81-123)."""
    if flavor == 1:
        return x * x - y * y + a, 2 * abs(x * y) - b
    if flavor == 2:
        return x * x - y * y + a, 2 * x * abs(y) - b
    if flavor == 3:
        return x * x - y * abs(y) + a, 2 * x * y - b
    if flavor == 4:
        return abs(x * x - y * y) + a, 2 * x * y - b
    if flavor == 5:
        return abs(x * x - y * y) + a, 2 * abs(x * y) - b
    raise ValueError(flavor)


def _sgn(x):
    return -1 if x < 0 else 1


def bs_jacobian(flavor, x, y):
    """Jacobian [[dfx/dx, dfx/dy], [dfy/dx, dfy/dy]] (burning_ship.py:440-532)."""
    if flavor == 1:
        return 2 * x, -2 * y, 2 * _sgn(x) * abs(y), 2 * _sgn(y) * abs(x)
    if flavor == 2:
        return 2 * x, -2 * y, 2 * abs(y), 2 * _sgn(y) * x
    if flavor == 3:
        return 2 * x, -2 * abs(y), 2 * y, 2 * x
    q = _sgn(x * x - y * y)
    if flavor == 4:
        return 2 * q * x, -2 * q * y, 2 * y, 2 * x
    if flavor == 5:
        return 2 * q * x, -2 * q * y, 2 * _sgn(x) * abs(y), 2 * _sgn(y) * abs(x)
    raise ValueError(flavor)


def _hypot_f(x, y):
    return math.hypot(mpf_to_double(x), mpf_to_double(y))


# ---------------------------------------------------------------------------
# reference orbits
def holo_orbit(c, degree, ref_len, M, need_xr, prec):
    """Orbit of z -> z^d + c from 0, rounded to complex128.

    Returns (div_iter, Zn_path[ref_len], xr) where xr maps an index i with
    |Z_i| < 1e-300 to its (re_m, im_m, exp) extended-range value and div_iter
    is the first invalid index (escape |Z| > M, or ref_len when the orbit
    stays bounded).  Entries past an escape are left at 0.
    """
    Zn = np.zeros(ref_len, dtype=np.complex128)
    xr = {}
    i = 0
    abs_i = 0.
    max_iter = ref_len - 1
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        z = mpmath.mpc(0)
        for i in range(1, max_iter + 1):
            z = _zpow(z, degree) + c
            x = mpf_to_double(z.real)
            y = mpf_to_double(z.imag)
            Zn[i] = complex(x, y)
            abs_i = math.hypot(x, y)
            if abs_i > M:
                break
            if need_xr and abs_i < XR_TSHOLD:
                xr[i] = mpc_to_xr(z)
    if i == max_iter and abs_i <= M:
        i += 1
    return i, Zn, xr


def bs_orbit(c, flavor, ref_len, M, need_xr, prec):
    """Orbit of the burning-ship family from (0, 0); xr holds indices where
    either coordinate is below 1e-300, as ((xm, xe), (ym, ye))."""
    Zn = np.zeros(ref_len, dtype=np.complex128)
    xr = {}
    i = 0
    abs_i = 0.
    max_iter = ref_len - 1
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        a, b = c.real, c.imag
        x = mpmath.mpf(0)
        y = mpmath.mpf(0)
        for i in range(1, max_iter + 1):
            x, y = bs_iterate(flavor, x, y, a, b)
            fx = mpf_to_double(x)
            fy = mpf_to_double(y)
            Zn[i] = complex(fx, fy)
            abs_i = math.hypot(fx, fy)
            if abs_i > M:
                break
            if need_xr and (abs(fx) < XR_TSHOLD or abs(fy) < XR_TSHOLD):
                xr[i] = (mpf_to_xr(x), mpf_to_xr(y))
    if i == max_iter and abs_i <= M:
        i += 1
    return i, Zn, xr


# ---------------------------------------------------------------------------
# period detection
def holo_ball_method(c, px, degree, maxiter, M_divergence, prec):
    """First n with |z_n / (dz_n/dc)| < px, None if the orbit escapes."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        inv_px = 1 / mpmath.mpf(px)
        z = mpmath.mpc(0)
        dz = mpmath.mpc(0)
        for i in range(1, maxiter + 1):
            if degree == 2:
                dz = 2 * (z * dz) + 1
            else:
                dz = degree * (dz * z ** (degree - 1)) + 1
            z = _zpow(z, degree) + c
            r = (z / dz) * inv_px
            if _hypot_f(z.real, z.imag) > M_divergence:
                return None
            if abs(r) < 1:
                return i
    return None


def bs_ball_method(c, px, flavor, maxiter, M_divergence, prec):
    """Ball-method period of the burning-ship family with the 2x2 Jacobian."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        a, b = c.real, c.imag
        inv_px = 1 / mpmath.mpf(px)
        x = y = mpmath.mpf(0)
        D = [mpmath.mpf(0)] * 4
        for i in range(1, maxiter + 1):
            J = bs_jacobian(flavor, x, y)
            D = [J[0] * D[0] + J[1] * D[2] + 1, J[0] * D[1] + J[1] * D[3],
                 J[2] * D[0] + J[3] * D[2], J[2] * D[1] + J[3] * D[3] - 1]
            x, y = bs_iterate(flavor, x, y, a, b)
            det = D[0] * D[3] - D[2] * D[1]
            rx = (D[3] * x - D[1] * y) / det * inv_px
            ry = (D[0] * y - D[2] * x) / det * inv_px
            if _hypot_f(x, y) > M_divergence:
                return None
            if abs(rx) < 1 and abs(ry) < 1:
                if math.hypot(mpf_to_double(rx), mpf_to_double(ry)) < 1.:
                    return i
    return None


# ---------------------------------------------------------------------------
# Newton search for the nucleus
def find_nucleus_m2(c, order, eps_pixel, prec, max_newton=80, eps_cv=None):
    """Newton on z_order(c) / prod_{i | order, i < order} z_i(c) for z^2 + c
    (the divisor-removal variant); returns (converged, nucleus)."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            zr = mpmath.mpc(0)
            dzrdc = mpmath.mpc(0)
            h = mpmath.mpc(1)
            dh = mpmath.mpc(0)
            for i in range(1, order + 1):
                dzrdc = 2 * dzrdc * zr + 1
                zr = zr * zr + c
                if i < order and order % i == 0:
                    h = h * zr
                    dh = dzrdc / zr + dh
            f = zr / h
            df = (dzrdc - zr * dh) / h
            step = f / df
            c = c - step
            if abs(step) <= eps:
                return abs(zr) <= eps_pixel, c
    return False, mpmath.mpc("nan", "nan")


def find_any_nucleus(c, order, degree, eps_pixel, prec, max_newton=80, eps_cv=None):
    """Newton on z_order(c) = 0 for z^d + c, divisors of the order allowed."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            zr = mpmath.mpc(0)
            dzrdc = mpmath.mpc(0)
            for _i in range(1, order + 1):
                if degree == 2:
                    dzrdc = 2 * (zr * dzrdc) + 1
                else:
                    dzrdc = degree * (dzrdc * zr ** (degree - 1)) + 1
                zr = _zpow(zr, degree) + c
            step = zr / dzrdc
            c = c - step
            if abs(step) <= eps:
                return abs(zr) <= eps_pixel, c
    return False, mpmath.mpc("nan", "nan")


def bs_find_any_nucleus(c, order, flavor, eps_pixel, prec, max_newton=80, eps_cv=None):
    """2x2 Newton on (x_order, y_order)(a, b) = 0 for the burning-ship family."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        a, b = c.real, c.imag
        eps = (mpmath.mpf(eps_cv) if eps_cv is not None else mpmath.ldexp(1, -prec)) * 64
        eps_pixel = mpmath.mpf(eps_pixel)
        for _ in range(max_newton):
            x = y = mpmath.mpf(0)
            D = [mpmath.mpf(0)] * 4
            for _i in range(1, order + 1):
                J = bs_jacobian(flavor, x, y)
                D = [J[0] * D[0] + J[1] * D[2] + 1, J[0] * D[1] + J[1] * D[3],
                     J[2] * D[0] + J[3] * D[2], J[2] * D[1] + J[3] * D[3] - 1]
                x, y = bs_iterate(flavor, x, y, a, b)
            inv = 1 / (D[0] * D[3] - D[2] * D[1])
            da = (D[3] * x - D[1] * y) * inv
            db = (D[0] * y - D[2] * x) * inv
            a = a - da
            b = b - db
            if mpmath.hypot(da, db) <= eps:
                return mpmath.hypot(x, y) <= eps_pixel, mpmath.mpc(a, b)
    return False, mpmath.mpc("nan", "nan")


def nucleus_size_estimate(c, order, degree, prec):
    """Hyperbolic component size estimate 1 / (b l^(d/(d-1)))."""
    with mpmath.workprec(prec):
        c = mpmath.mpc(c)
        z = mpmath.mpc(0)
        l = mpmath.mpc(1)
        b = mpmath.mpc(1)
        for _ in range(1, order):
            z = _zpow(z, degree) + c
            if degree == 2:
                l = 2 * (z * l)
            else:
                l = degree * (l * z ** (degree - 1))
            b = b + 1 / l
        if degree == 2:
            ll = l * l
        else:
            ll = l ** (mpmath.mpf(degree) / (degree - 1))
        return 1 / (b * ll)
