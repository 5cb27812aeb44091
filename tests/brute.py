"""Brute-force ground truth: direct arbitrary-precision iteration of a pixel,
without perturbation (mpmath)."""
import mpmath
import numpy as np


def pixel_c(f, pix):
    """centre + dx * (lin_mat . f(pix)): rotation and skew included
    (lin_proj_impl, perturbation.py:2665-2672); f the projection, rounded to
    complex128 as the reference's loop has it (the escape count of a pixel
    near the boundary follows that rounding)."""
    m = f.lin_mat
    proj = getattr(f, "projection", None)
    if proj is not None and type(proj).__name__ != "Cartesian":
        pix = complex(proj.f(np.array([pix]))[0])
    px, py = float(pix.real), float(pix.imag)
    with mpmath.workprec(f.prec + 64):
        ox = mpmath.mpf(float(m[0, 0])) * px + mpmath.mpf(float(m[0, 1])) * py
        oy = mpmath.mpf(float(m[1, 0])) * px + mpmath.mpf(float(m[1, 1])) * py
        return mpmath.mpc(f.x + ox * f.dx, f.y + oy * f.dx)


def holo_escape(f, pix, degree, max_iter, M, prec=600):
    """(n, z_n, dz_n/dc * dx) at the first n with |z_n| > M, or at max_iter."""
    c = pixel_c(f, pix)
    M2 = M * M
    with mpmath.workprec(prec):
        z = mpmath.mpc(0)
        dz = mpmath.mpc(0)
        for n in range(1, max_iter + 1):
            dz = degree * dz * (z ** (degree - 1) if degree > 2 else z) + 1
            z = z ** degree + c if degree > 2 else z * z + c
            if z.real ** 2 + z.imag ** 2 > M2:
                break
        return n, complex(z), complex(dz * f.dx)


def bs_escape(f, pix, flavor, max_iter, M, prec=600):
    from fractalshades_amd.mpmath_loop import bs_iterate
    c = pixel_c(f, pix)
    M2 = M * M
    with mpmath.workprec(prec):
        a, b = c.real, c.imag
        x = y = mpmath.mpf(0)
        for n in range(1, max_iter + 1):
            x, y = bs_iterate(flavor, x, y, a, b)
            if x * x + y * y > M2:
                break
        return n, complex(x, y)


def holo_point(f, pix, degree, k, prec=600):
    """z_k of the pixel's own orbit (z_0 = 0)."""
    c = pixel_c(f, pix)
    with mpmath.workprec(prec):
        z = mpmath.mpc(0)
        for _ in range(max(k, 0)):
            z = z ** degree + c if degree > 2 else z * z + c
        return complex(z)


def bs_point(f, pix, flavor, k, prec=600):
    from fractalshades_amd.mpmath_loop import bs_iterate
    c = pixel_c(f, pix)
    with mpmath.workprec(prec):
        a, b = c.real, c.imag
        x = y = mpmath.mpf(0)
        for _ in range(max(k, 0)):
            x, y = bs_iterate(flavor, x, y, a, b)
        return complex(x, y)
