"""Seeded workload generation: the burning-ship descent towards the set.

The descent's distance estimate d = |z|^2 ln|z| / |J^T z| rests on the 2x2
Jacobian J = d(x_n, y_n)/d(a, b) that it carries per flavour; here every
entry of J is checked against central finite differences of the orbit in
mpmath (same n), and d against G / |grad G| with G = ln|z_n| / 2^n."""
import mpmath
import numpy as np
import pytest

from fractalshades_amd import configs
from fractalshades_amd.models import get_flavor_int

FLAVORS = ["Burning ship", "Perpendicular burning ship", "Shark fin", "Celtic", "Buffalo"]


def bs_step(flavor, x, y, a, b):
    """One step of the burning-ship family (burning_ship.py:63-76)."""
    if flavor == 3:
        X, Y = x * x - y * abs(y), 2 * x * y
    else:
        X = x * x - y * y
        if flavor in (4, 5):
            X = abs(X)
        Y = 2 * x * abs(y) if flavor == 2 else 2 * x * y
        if flavor in (1, 5):
            Y = abs(Y)
    return X + a, Y - b


def bs_orbit_z(flavor, a, b, n):
    """(x_n, y_n) from 0."""
    x = y = mpmath.mpf(0)
    for _ in range(n):
        x, y = bs_step(flavor, x, y, a, b)
    return x, y


def escape_n(flavor, a, b, nmax=20000):
    """First n with |z_n|^2 > 1e10 (the descent's escape radius)."""
    x = y = mpmath.mpf(0)
    for n in range(1, nmax + 1):
        x, y = bs_step(flavor, x, y, a, b)
        if x * x + y * y > 1e10:
            return n
    return None


@pytest.mark.parametrize("flavor", FLAVORS)
def test_bs_descent_distance_estimate(flavor):
    fl = get_flavor_int(flavor)
    rng = np.random.default_rng(fl)
    prec = 200
    with mpmath.workprec(prec):
        for _ in range(40):
            c0 = complex(rng.uniform(-2., 0.5), rng.uniform(0., 1.2))
            res = configs.descend_to_boundary_bs(c0, -30., prec, 4000, fl)
            if res is not None:
                break
        assert res is not None
        c, d_log2 = res
        assert d_log2 < -30.
        a, b = c.real, c.imag
        n = escape_n(fl, a, b, 4000)
        assert n is not None                    # an exterior point

        def G(a, b):
            x, y = bs_orbit_z(fl, a, b, n)
            return mpmath.log(x * x + y * y) / 2 / mpmath.mpf(2) ** n

        # J of the descent against central differences of the orbit
        p = prec
        esc = configs.bs_escape_jacobian(int(mpmath.ldexp(a, p)), int(mpmath.ldexp(b, p)), p, 4000, fl)
        assert esc is not None and esc[0] == n
        J = np.array([[esc[3], esc[4]], [esc[5], esc[6]]], dtype=object)
        h = mpmath.ldexp(1, int(d_log2) - 24)
        for col, (da, db) in enumerate([(h, 0), (0, h)]):
            xp, yp = bs_orbit_z(fl, a + da, b + db, n)
            xm, ym = bs_orbit_z(fl, a - da, b - db, n)
            fd = [(xp - xm) / (2 * h), (yp - ym) / (2 * h)]
            for row in range(2):
                got = mpmath.ldexp(J[row, col], -p)
                assert abs(got - fd[row]) <= 1e-6 * abs(fd[row]) + 1e-12 * abs(got), (row, col)

        ga = (G(a + h, b) - G(a - h, b)) / (2 * h)
        gb = (G(a, b + h) - G(a, b - h)) / (2 * h)
        d_fd = G(a, b) / mpmath.hypot(ga, gb)
        assert abs(float(mpmath.log(d_fd, 2)) - d_log2) < 0.05


def _check_jacobian(fl, a, b, prec, nmax=4000):
    """J of bs_escape_jacobian at c = (a, b) against central differences."""
    esc = configs.bs_escape_jacobian(int(mpmath.ldexp(a, prec)), int(mpmath.ldexp(b, prec)), prec, nmax, fl)
    if esc is None:
        return False
    n, J = esc[0], ((esc[3], esc[4]), (esc[5], esc[6]))
    x, y = mpmath.ldexp(esc[1], -prec), mpmath.ldexp(esc[2], -prec)
    jn = max(abs(mpmath.ldexp(v, -prec)) for row in J for v in row)
    h = mpmath.sqrt(x * x + y * y) / jn * mpmath.mpf(2) ** -30
    for col, (da, db) in enumerate([(h, 0), (0, h)]):
        xp, yp = bs_orbit_z(fl, a + da, b + db, n)
        xm, ym = bs_orbit_z(fl, a - da, b - db, n)
        fd = [(xp - xm) / (2 * h), (yp - ym) / (2 * h)]
        for row in range(2):
            got = mpmath.ldexp(J[row][col], -prec)
            assert abs(got - fd[row]) <= 1e-6 * jn, (fl, row, col)
    return True


@pytest.mark.parametrize("flavor", FLAVORS)
def test_bs_jacobian_scattered_points(flavor):
    """The Jacobian at exterior points spread over both half-planes, where
    the folds take every sign combination."""
    fl = get_flavor_int(flavor)
    rng = np.random.default_rng(100 + fl)
    checked = 0
    with mpmath.workprec(200):
        while checked < 6:
            a, b = mpmath.mpf(rng.uniform(-2., 1.)), mpmath.mpf(rng.uniform(-1.5, 1.5))
            checked += _check_jacobian(fl, a, b, 200)
