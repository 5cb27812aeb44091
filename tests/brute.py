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


def _pow(z, k):
    """z^k for an integer k >= 1 by repeated products (mpmath's generic
    complex power is much slower at thousands of bits)."""
    r = z
    for _ in range(k - 1):
        r = r * z
    return r


def holo_escape(f, pix, degree, max_iter, M, prec=600):
    """(n, z_n, dz_n/dc * dx) at the first n with |z_n| > M, or at max_iter."""
    c = pixel_c(f, pix)
    M2 = M * M
    with mpmath.workprec(prec):
        z = mpmath.mpc(0)
        dz = mpmath.mpc(0)
        for n in range(1, max_iter + 1):
            zp = _pow(z, degree - 1)
            dz = degree * dz * zp + 1
            z = zp * z + c
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
            z = _pow(z, degree) + c
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


def holo_escape_dz(f, pix, degree, max_iter, M, prec=600):
    """(n, z_n, dz_n/dc * dx, dz_n/dz_1) of the pixel's own orbit at the first
    n with |z_n| > M (or at max_iter): dz/dz is the derivative with respect
    to the first iterate, 1 at n = 1 and f'(z_n) times itself after."""
    c = pixel_c(f, pix)
    M2 = M * M
    with mpmath.workprec(prec):
        z = mpmath.mpc(0)
        dc = mpmath.mpc(0)
        dz = mpmath.mpc(0)
        for n in range(1, max_iter + 1):
            zp = _pow(z, degree - 1)
            fp = degree * zp
            dc = fp * dc + 1
            dz = fp * dz if n > 1 else mpmath.mpc(1)
            z = zp * z + c
            if z.real ** 2 + z.imag ** 2 > M2:
                break
        return n, complex(z), complex(dc * f.dx), complex(dz)


class Dual:
    """A real value with its partial derivatives with respect to (a, b):
    forward-mode differentiation of the burning-ship step, independent of the
    hand-written Jacobians of the model."""
    __slots__ = ("v", "da", "db")

    def __init__(self, v, da=0, db=0):
        self.v, self.da, self.db = v, da, db

    @staticmethod
    def _of(o):
        return o if isinstance(o, Dual) else Dual(o)

    def __add__(self, o):
        o = Dual._of(o)
        return Dual(self.v + o.v, self.da + o.da, self.db + o.db)
    __radd__ = __add__

    def __sub__(self, o):
        o = Dual._of(o)
        return Dual(self.v - o.v, self.da - o.da, self.db - o.db)

    def __rsub__(self, o):
        return Dual._of(o) - self

    def __neg__(self):
        return Dual(-self.v, -self.da, -self.db)

    def __mul__(self, o):
        o = Dual._of(o)
        return Dual(self.v * o.v, self.da * o.v + self.v * o.da, self.db * o.v + self.v * o.db)
    __rmul__ = __mul__

    def __abs__(self):
        return -self if self.v < 0 else self


def bs_escape_jac(f, pix, flavor, max_iter, M, prec=600):
    """(n, x_n + i y_n, J * dx) with J = d(x_n, y_n)/d(a, b) as the four values
    (dx/da, dx/db, dy/da, dy/db), by forward differentiation of the direct
    iteration in mpmath."""
    from fractalshades_amd.mpmath_loop import bs_iterate
    c = pixel_c(f, pix)
    M2 = M * M
    with mpmath.workprec(prec):
        A = Dual(c.real, mpmath.mpf(1), mpmath.mpf(0))
        B = Dual(c.imag, mpmath.mpf(0), mpmath.mpf(1))
        x = y = Dual(mpmath.mpf(0), mpmath.mpf(0), mpmath.mpf(0))
        for n in range(1, max_iter + 1):
            x, y = bs_iterate(flavor, x, y, A, B)
            if x.v * x.v + y.v * y.v > M2:
                break
        dx = f.dx
        J = tuple(float(v * dx) for v in (x.da, x.db, y.da, y.db))
        return n, complex(x.v, y.v), J
