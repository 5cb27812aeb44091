"""Pins of the post-processing formulas (fractalshades_amd.postproc:
continuous iteration, DEM, normal map; the reference's postproc.py:352-406,
684-731, 572-628, formula 1001-1009) against the potential of the set,
computed by brute force from each pixel's own orbit in mpmath.

Outside the set, G(c) = lim log|z_n| / d^n (the Green function: |z|^d + c
grows like |z|^d, and so do the burning-ship flavours 1, 2 and 5, whose step
has modulus |z|^2 exactly).  Three consequences, each independent of the
formulas under test:
* continuous iteration: nu = n - log_d(log|z_n| / log M) tends to
  log_d(log M / G(c)), the same for every escape radius M and step n;
* DEM: |z_n| ln|z_n| / |dz_n/dc| tends to G / |grad G|;
* normal map: z_n / (dz_n/dc) points along grad G, and so does J^T z_n for
  the burning ship.
G is iterated here to |z| > 1e80 and grad G taken by central differences in
c, at a step 1e-12 of the pixel spacing (the burning ship's G folds along
the preimages of the axes, which near the set lie closer together than 1e-6
of a pixel).  At the loop's escape (|z_n| >
M = 1e3) the limits are reached to ~1e-7 (the next step changes log|z| by
|c| / |z|^d), hence the tolerances.  Pixel values (z_n, dz_n/dc times dx,
stop_iter) come from the oracle with BLA off (with BLA on, z_n carries the
approximation's own error, 1e-3 relative at the escape of these frames); the
GPU layers are compared with these formulas in the GPU tests."""
import mpmath
import numpy as np
import pytest

import brute
import frames
import oracle as orc
from fractalshades_amd.mpmath_loop import bs_iterate
from fractalshades_amd.postproc import continuous_iter, dem, normal_map

R_FAR = mpmath.mpf(10) ** 80


def green(c, degree=2, flavor=None, prec=600, cap=100000):
    """G(c) from the orbit of 0, iterated until |z| > 1e80."""
    with mpmath.workprec(prec):
        if flavor is None:
            z = mpmath.mpc(0)
            for n in range(1, cap):
                z = brute._pow(z, degree) + c
                if abs(z) > R_FAR:
                    return mpmath.log(abs(z)) / mpmath.mpf(degree) ** n
        else:
            x = y = mpmath.mpf(0)
            for n in range(1, cap):
                x, y = bs_iterate(flavor, x, y, c.real, c.imag)
                if mpmath.hypot(x, y) > R_FAR:
                    return mpmath.log(mpmath.hypot(x, y)) / mpmath.mpf(2) ** n
    raise AssertionError("no escape")


def green_grad(f, pix, degree=2, flavor=None, prec=600):
    """The limits at the pixel's c = a + i b, from G and grad G = dG/da +
    i dG/db (mpmath values: G is below binary64's range past ~1100 steps):
    (log_d(log M / G) for M = 1e3, G / |grad G|, grad G / |grad G|)."""
    c = brute.pixel_c(f, pix)
    with mpmath.workprec(prec):
        h = mpmath.mpf(f.dx) * mpmath.mpf("1e-12")
        g = green(c, degree, flavor, prec)
        ga = (green(c + h, degree, flavor, prec) - green(c - h, degree, flavor, prec)) / (2 * h)
        gb = (green(c + 1j * h, degree, flavor, prec) - green(c - 1j * h, degree, flavor, prec)) / (2 * h)
        ag = mpmath.hypot(ga, gb)
        nu = mpmath.log(mpmath.log(1000) / g) / mpmath.log(degree)
        return float(nu), float(g / ag), complex(float(ga / ag), float(gb / ag))


def escaped_sample(res, n_take, seed):
    sr = res[2].reshape(-1)
    idx = np.flatnonzero(sr == 1)
    assert idx.size >= n_take
    return np.random.default_rng(seed).choice(idx, n_take, replace=False)


@pytest.mark.parametrize("name,degree", [("M2_E20", 2), ("M2_divref5", 2), ("MN3", 3)])
def test_holomorphic_nu_dem_normal_vs_potential(name, degree):
    if name == "MN3":
        f = frames.mn_frame(3, nx=24, calc_dzndc=True, interior_detect=False)
    else:
        f = frames.make(name, nx=24, calc_dzndc=True, interior_detect=False)
    args, c_pix = frames.frame_inputs(f)
    res = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    M = args.fields["M_divergence"]
    assert M == 1000.
    dx = float(f.dx)
    for i in escaped_sample(res, 6, seed=1):
        zn, dzndc, n = res[0][0, i], res[0][-1, i], res[3][0, i]
        nu_lim, de_lim, grad_dir = green_grad(f, c_pix[i], degree, prec=f.prec + 160)
        nu = continuous_iter(np.array([zn]), np.array([n]), d=degree, M=M)[0]
        assert abs(nu - nu_lim) < 2e-6, (i, nu, nu_lim)
        de = dem(np.array([zn]), np.array([dzndc]))[0] * dx
        assert abs(de / de_lim - 1.) < 1e-5, (i, de, de_lim)
        nrm = normal_map(np.array([zn]), np.array([dzndc]))[0]
        assert abs(nrm - grad_dir) < 1e-5, (i, nrm, grad_dir)


@pytest.mark.parametrize("name", ["BS_ex14", "PBS_ex18"])
def test_burning_ship_nu_dem_normal_vs_potential(name):
    """Flavours 1 and 2 (|step| = |z|^2): nu and the normal J^T z_n as for
    the holomorphic maps; the DEM uses the Jacobian's larger singular value
    (checked against numpy's SVD), so it bounds G / |grad G| from below
    (|J^T z| <= sigma_max |z|) instead of reaching it."""
    f = frames.make(name, nx=24, max_iter=20000)
    args, c_pix = frames.frame_inputs(f)
    flavor = args.fields["flavor"]
    assert flavor in (1, 2)
    res = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    M = args.fields["M_divergence"]
    assert M == 1000.
    dx = float(f.dx)
    for i in escaped_sample(res, 5, seed=2):
        zn = complex(res[0][0, i], res[0][1, i])
        jac = tuple(np.array([res[0][k, i]]) for k in range(2, 6))
        n = res[3][0, i]
        nu_lim, de_lim, grad_dir = green_grad(f, c_pix[i], flavor=flavor, prec=f.prec + 160)
        nu = continuous_iter(np.array([zn]), np.array([n]), d=2, M=M)[0]
        assert abs(nu - nu_lim) < 2e-6, (i, nu, nu_lim)
        J = np.array([[jac[0][0], jac[1][0]], [jac[2][0], jac[3][0]]])
        smax = np.linalg.svd(J, compute_uv=False)[0]
        de = dem(np.array([zn]), jac=jac)[0]
        assert abs(de - abs(zn) * np.log(abs(zn)) / smax) <= 1e-12 * de
        assert de * dx <= de_lim * (1. + 1e-6)
        nrm = normal_map(np.array([zn]), jac=jac)[0]
        assert abs(nrm - grad_dir) < 1e-5, (i, nrm, grad_dir)


def test_expmap_screen_derivative_vs_brute_force():
    """Exponential map by bands (set_exp_zoom_step): the derivative the
    layers use -- the loop's dz/dc (scaled by dx exp(h_band)), times the
    band's dzndc_modifier exp(Re t + h_moy - h_band), times df = exp(i Im t)
    (screen_derivatives), t = pix_to_ht pix -- is the derivative of z_n with
    respect to the screen position divided by pix_to_ht: c = centre +
    dx exp(h_moy + t) gives dz_n/dpix = dz_n/dc dx pix_to_ht exp(h_moy + t).
    Checked on sampled pixels of every band against central differences of
    the brute-force z_n (n fixed) in pix, within 1e-6 (the step moves z_n by
    ~1e-5, where z_n is linear in pix to ~1e-8)."""
    from fractalshades_amd.postproc import screen_derivatives
    from fractalshades_amd import settings
    f = frames.expmap_frame("M2_E20", nx=400, hmin=0., hmax=10., dx="1e-24")
    p = f.projection
    assert f.nx // settings.chunk_size >= 2
    rng = np.random.default_rng(5)
    checked = n_bands = 0
    for fields, arrays, chunks in f.expmap_bands("t", exp_zoom_step=settings.chunk_size):
        n_bands += 1
        assert p.use_step and p.df_kind == 3
        for ix, ixx, iy, iyy in chunks[:4]:
            pix = np.ravel(f.chunk_pixel_pos((ix, ixx, iy, iyy)))
            pix = pix[rng.choice(pix.size, min(8, pix.size), replace=False)]
            off = np.ascontiguousarray(p.f(pix))
            Zo, Uo, sro, sio = orc.OracleCtx(dict(fields, bla_on=0), arrays).cycles(off)
            dz = Zo[-1] * p.dzndc_modifier(pix)
            scr = screen_derivatives(p, pix, dzndc=dz)
            for q in range(pix.size):
                n = int(sio[0, q])
                if sro[0, q] != 1:
                    continue
                with mpmath.workprec(f.prec + 300):
                    # a step that moves z_n by ~1e-5 (|z_n| ~ 1e3): in the linear regime
                    h = mpmath.mpf(1e-8) * 1e3 / max(abs(scr[q]), 1.)
                    zs = []
                    for sgn in (1, -1, 0):
                        t = (mpmath.mpc(complex(pix[q])) + sgn * h) * mpmath.mpf(p.pix_to_ht)
                        c = mpmath.mpc(f.x, f.y) + f.dx * mpmath.exp(mpmath.mpf(p.hmoy) + t)
                        z = mpmath.mpc(0)
                        for _ in range(n):
                            z = z * z + c
                        zs.append(z)
                    fd = complex((zs[0] - zs[1]) / (2 * h) / mpmath.mpf(p.pix_to_ht))
                # the binary64 loop tracks the exact orbit (else rounding decides)
                if abs(complex(zs[2]) - (Zo[0, q])) > 1e-6 * abs(complex(zs[2])):
                    continue
                assert abs(scr[q] - fd) <= 1e-6 * abs(fd), (n, scr[q], fd)
                checked += 1
    assert n_bands >= 2 and checked >= 8
