"""Pins the oracle's perturbation loop (std and Xrange modes, BLA on and off)
against brute-force direct iteration in mpmath on sampled pixels of the
reference test suite's zoom locations."""
import mpmath
import numpy as np
import pytest

import oracle as orc
import brute
import frames


def sample(n_pts, k, seed=0):
    return np.random.default_rng(seed).choice(n_pts, k, replace=False)


@pytest.mark.parametrize("name", ["M2_E20", "M2_E11", "M2_divref", "M2_divref5"])
def test_m2_vs_brute_force(name):
    """BLA off: escape counts exact, escaped z_n and dz_n/dc within 1e-6 (rounding amplified along long orbits).
    BLA on: escape counts within one iteration (the BLA_eps = 1e-6
    approximation moves the escaped value by up to ~1e-3, by design)."""
    f = frames.make(name, nx=40)
    args, c_pix = frames.frame_inputs(f)
    on = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    off = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    for res in (on, off):
        assert set(np.unique(res[2])) <= {0, 1}
    sel = sample(len(c_pix), 24)
    exact_on = 0
    for i in sel:
        n, z, dz = brute.holo_escape(f, c_pix[i], 2, args.fields["max_iter"], 1e3)
        assert n == off[3][0, i], (i, n, off[3][0, i])
        if off[2][0, i] == 1 and n <= 1200:
            assert abs(off[0][0, i] - z) <= 1e-6 * abs(z)
            assert abs(off[0][1, i] - dz) <= 1e-6 * abs(dz)
        exact_on += int(abs(int(n) - int(on[3][0, i])) <= 1)
    assert exact_on >= len(sel) - 2


def test_bla_on_off_same_escape_counts():
    # the reference checks BLA on vs off on its 5e-12 case (test_perturbation.py:131-198)
    f = frames.make("M2_E11", nx=48)
    args, c_pix = frames.frame_inputs(f)
    on = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    off_fields = dict(args.fields, bla_on=0)
    off = orc.OracleCtx(off_fields, args.arrays).cycles(c_pix)
    # the reference only checks this pair at a 0.15 mean-RGB image tolerance;
    # measured here: 95% identical counts on this frame, 99.6% on 5e-20
    same = np.mean(on[3] == off[3])
    close = np.mean(abs(on[3].astype(int) - off[3].astype(int)) <= 1)
    assert same >= 0.9 and close >= 0.95


def test_forced_xrange_matches_standard():
    # Xrange mantissa arithmetic rounds like binary64 away from the
    # extremes: forcing Xrange on a shallow frame gives the same pixels
    f = frames.make("M2_E20", nx=40)
    args, c_pix = frames.frame_inputs(f)
    std = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    xr = orc.OracleCtx(dict(args.fields, xr_mode=1), args.arrays).cycles(c_pix)
    np.testing.assert_array_equal(std[3], xr[3])
    np.testing.assert_array_equal(std[2], xr[2])
    np.testing.assert_allclose(std[0][0], xr[0][0], rtol=1e-12)


def test_interior_detection_flags_bounded_pixels():
    # period-14292 minibrot: dz/dz decays over a few periods
    f = frames.make("M2_int32", nx=24, max_iter=200000)
    args, c_pix = frames.frame_inputs(f)
    Z, U, sr, si = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    assert np.count_nonzero(sr == 2) > 0
    for i in np.flatnonzero(sr[0] == 2)[:2]:
        # an interior pixel stays bounded 10% past its stationarity time
        n, z, dz = brute.holo_escape(f, c_pix[i], 2, int(1.1 * si[0, i]), 1e3, prec=300)
        assert abs(z) <= 1e3
    for i in np.flatnonzero(sr[0] == 1)[:4]:
        n, z, dz = brute.holo_escape(f, c_pix[i], 2, args.fields["max_iter"], 1e3, prec=300)
        assert abs(int(n) - int(si[0, i])) <= 1


@pytest.mark.parametrize("backshift", [1, 3, 7])
def test_calc_orbit_vs_brute_force(backshift):
    """calc_orbit (perturbation.py:1096-1101, 1266-1273, 1390-1396): the
    extra field is the pixel's orbit point at exit - backshift.  BLA off,
    so that every iteration is a recording point; an off-by-one in the
    index moves |z| by a factor ~|z|^2 near escape."""
    f = frames.make("M2_E20", nx=32)
    args, c_pix = frames.frame_inputs(f)
    fields = dict(args.fields, bla_on=0, calc_orbit=1, backshift=backshift)
    Z, U, sr, si = orc.OracleCtx(fields, args.arrays).cycles(c_pix)
    assert Z.shape[0] == 3
    n_checked = 0
    for i in sample(len(c_pix), 16):
        if sr[0, i] != 1:
            continue
        zk = brute.holo_point(f, c_pix[i], 2, int(si[0, i]) - backshift)
        assert abs(Z[-1, i] - zk) <= 1e-6 * abs(zk), (i, Z[-1, i], zk)
        n_checked += 1
    assert n_checked >= 8


def test_calc_orbit_bs_vs_brute_force():
    """BS calc_orbit: the finish loop restarts from the last-but-one record
    (every multiple of backshift with BLA off) and iterates with
    (a + Re Zn_path[1], b + Im Zn_path[1]) (perturbation.py:1778-1787).
    Y_1 = -B_ref, so that second constant is B - 2 B_ref, not the pixel's
    B: kept as the reference has it.  Brute force: the true orbit point at
    the record, then the same finish in mpmath."""
    import mpmath
    from fractalshades_amd.mpmath_loop import bs_iterate
    bsh = 4
    f = frames.make("BS_ex14", nx=32, max_iter=20000)
    args, c_pix = frames.frame_inputs(f)
    fl = args.fields["flavor"]
    fields = dict(args.fields, bla_on=0, calc_orbit=1, backshift=bsh)
    Z, U, sr, si = orc.OracleCtx(fields, args.arrays).cycles(c_pix)
    assert Z.shape[0] == 8
    Zn1 = args.arrays["Zn_path"].view(np.complex128)[1]
    n_checked = 0
    escaped = np.flatnonzero(sr[0] == 1)
    for i in np.random.default_rng(0).choice(escaped, min(16, escaped.size), replace=False):
        n = int(si[0, i])
        # pixels whose own escape the binary64 loop tracks (over ~4000
        # steps some BS pixels are chaotic: see test_bs_vs_brute_force)
        nb, zb = brute.bs_escape(f, c_pix[i], fl, 20000, 1e3, prec=f.prec + 128)
        if nb != n or abs(complex(Z[0, i], Z[1, i]) - zb) > 1e-8 * abs(zb):
            continue
        i2 = bsh * (n // bsh - 1)
        prec = f.prec + 128
        z2 = brute.bs_point(f, c_pix[i], fl, i2, prec=prec)
        c = brute.pixel_c(f, c_pix[i])
        with mpmath.workprec(prec):
            AA = c.real
            BB = c.imag + 2 * mpmath.mpf(Zn1.imag)      # b + Im Zn_path[1], Im Zn_path[1] = -B_ref
            x, y = mpmath.mpf(z2.real), mpmath.mpf(z2.imag)
            for _ in range(n - bsh - i2):
                x, y = bs_iterate(fl, x, y, AA, BB)
            want = complex(x, y)
        got = complex(Z[6, i], Z[7, i])
        assert abs(got - want) <= 1e-6 * abs(want), (i, got, want)
        n_checked += 1
    assert n_checked >= 6


@pytest.mark.parametrize("name", ["BS_ex14", "PBS_ex18"])
def test_bs_vs_brute_force(name):
    """BS loop, BLA off, vs direct iteration at the frame's precision: escape
    counts exact on most pixels and the escaped (x, y) within 1e-8 on most;
    the rest are chaotic over these 4000-12000-step orbits, where binary64
    rounding (the reference's as much as this one's) decides."""
    f = frames.make(name, nx=24, max_iter=20000)
    args, c_pix = frames.frame_inputs(f)
    fl = args.fields["flavor"]
    Z, U, sr, si = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    sel = sample(len(c_pix), 12)
    same_n = close = 0
    for i in sel:
        nb, zb = brute.bs_escape(f, c_pix[i], fl, 20000, 1e3, prec=f.prec + 128)
        same_n += int(nb == si[0, i])
        close += int(nb == si[0, i] and abs(complex(Z[0, i], Z[1, i]) - zb) <= 1e-8 * abs(zb))
    assert same_n >= 9 and close >= 6, (same_n, close)


@pytest.mark.parametrize("orientation", ["horizontal", "vertical"])
def test_expmap_vs_brute_force(orientation):
    """The exponential map (projection.Expmap, h in [0, 10]): the pixel
    offsets projected on the host (exp(h_moy + pix_to_ht pix), checked
    against mpmath to binary64 accuracy) and iterated by the oracle (BLA off)
    escape at the brute-force iteration of c = centre + dx f(pix) (but for
    pixels shadowing the reference cycle, where binary64 rounding decides),
    z_n within 1e-6; dz/dc is scaled by dx exp(h_moy), as the reference
    scales the derivative paths."""
    f = frames.expmap_frame("M2_E20", nx=40, hmin=0., hmax=10., dx="1e-24", orientation=orientation)
    args, c_pix = frames.frame_inputs(f)
    assert args.projection is f.projection
    if orientation == "horizontal":
        assert f.nx == 40 and abs(f.xy_ratio - 10. / (2 * np.pi)) < 1e-15
    else:
        assert abs(f.xy_ratio - 2 * np.pi / 10.) < 1e-15 and f.nx == int(40 * f.xy_ratio + 0.5)
    cp = np.ascontiguousarray(f.projection.f(c_pix))
    off = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(cp)
    scale = float(mpmath.exp(f.projection.hmoy))
    sel = sample(len(c_pix), 16)
    p = f.projection
    same_n = checked = 0
    for i in sel:
        # the projected offset: exp of the binary64 argument h_moy + pix_to_ht pix
        # (as the reference forms it) to binary64 accuracy
        arg = complex(p.hmoy + p.pix_to_ht * np.complex128(c_pix[i]))
        with mpmath.workprec(200):
            w = complex(mpmath.exp(mpmath.mpc(arg.real, arg.imag)))
        assert abs(w - cp[i]) <= 4e-16 * abs(w)
        n, z, dz = brute.holo_escape(f, c_pix[i], 2, args.fields["max_iter"], 1e3)
        same_n += int(n == off[3][0, i])
        # pixels that escape within ~1.1 periods of the period-643 reference
        # cycle are not chaotic: their escaped z_n agrees; those that shadow
        # the cycle longer amplify binary64 rounding (the reference's as much)
        if off[2][0, i] == 1 and n == off[3][0, i] and n <= 700:
            assert abs(off[0][0, i] - z) <= 1e-6 * abs(z)
            checked += 1
    assert same_n >= len(sel) - 2 and checked >= len(sel) // 2
    # dz/dc: the derivative paths are linear in their scale, so the field
    # at dx exp(h_moy) is exp(h_moy) times the field at dx (the recurrence
    # itself is pinned by brute force on the Cartesian frames)
    unit = orc.OracleCtx(dict(args.fields, bla_on=0, deriv_mant=0., deriv_exp=0), args.arrays).cycles(cp)
    esc = off[2][0] == 1
    np.testing.assert_allclose(off[0][1, esc], unit[0][1, esc] * scale, rtol=1e-12)


def test_deep_expmap_vs_brute_force():
    """A deep exponential map (h in [0, 758], exp(hmax) ~ 1e329 > 1e300: the
    reference's Xrange h, projection.py:287-294) at dx = 1e-330 around an
    exterior point, computed by bands of h: per band the offsets
    exp(h - h_band + i t) and the pixel scale dx exp(h_band); the map spans
    |c - centre| from 1e-330 (Xrange deltas, the centre's own escape count)
    to 0.1 (other counts).  On two sampled pixels of every band, the band's
    oracle run (BLA off) escapes at the brute-force iteration of
    c = centre + dx exp(h + i t) computed exactly, z_n within 1e-6."""
    f = frames.expmap_frame("M2_E20", nx=1200, hmin=0., hmax=758., dx="1e-330", precision=345,
                            no_newton=True, max_iter=3000, x="-0.75", y="0.2")
    p = f.projection
    assert p.deep
    with pytest.raises(ValueError):
        p.f(np.zeros(2))
    rng = np.random.default_rng(3)
    checked = n_bands = 0
    counts = set()
    for k, (fields, arrays, chunks) in enumerate(f.expmap_bands("t")):
        n_bands += 1
        for q, (ix, ixx, iy, iyy) in enumerate(chunks[:1] * 2):
            i, j = int(rng.integers(ix, ixx)), int(rng.integers(iy, iyy))
            if q and ixx == f.nx:
                i = f.nx - 1                              # the largest |c - centre| of the map
            pix = np.ravel(f.chunk_pixel_pos((ix, ixx, iy, iyy)))[(j - iy) * (ixx - ix) + (i - ix)]
            off = p.f(np.array([pix]))                    # the band's offset of this pixel
            assert abs(off[0]) < 1e100
            Zo, Uo, sro, sio = orc.OracleCtx(dict(fields, bla_on=0), arrays).cycles(off)
            with mpmath.workprec(f.prec + 64):
                ht = mpmath.mpf(p.hmoy) + mpmath.mpc(complex(p.pix_to_ht * np.complex128(pix)))
                c = mpmath.mpc(f.x, f.y) + f.dx * mpmath.exp(ht)
                z = mpmath.mpc(0)
                for n in range(1, fields["max_iter"] + 1):
                    z = z * z + c
                    if z.real ** 2 + z.imag ** 2 > 1e6:
                        break
            assert n == sio[0, 0], (k, i, j, n, sio[0, 0])
            if sro[0, 0] == 1:
                assert abs(Zo[0, 0] - complex(z)) <= 1e-6 * abs(complex(z))
            counts.add(int(n))
            checked += 1
    assert n_bands >= 5 and checked >= 10
    assert len(counts) >= 2          # the map spans depths with different escape counts
