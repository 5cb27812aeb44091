"""Independent pins of the oracle's loops that the z^2 + c brute-force tests
do not reach: z^d + c, the burning ship's BLA and Jacobian, dz/dz values, and
the Xrange loop at the depths where its register and rebase confirmation run
(1e-300, 1e-800, 1e-2000).  Each check compares the oracle with direct
arbitrary-precision iteration of the pixel (tests/brute.py), so that a wrong
binomial coefficient, a dropped Jacobian term, a wrong path index or a
missing rebase correction fails.

Reference loops pinned: numba_iterate (perturbation.py:1065-1400) with the
Mn kernels (mandelbrot_Mn.py:646-729), numba_iterate_BS and its BLA
(perturbation.py:1468-1811, 1884-2105), the dzndz block
(perturbation.py:1189-1246) and the Xrange rebase (1317-1357).
"""
import numpy as np
import pytest

import oracle as orc
import brute
import frames
from fractalshades_amd import configs


def sample(n_pts, k, seed=0):
    return np.random.default_rng(seed).choice(n_pts, k, replace=False)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ---- z^d + c ---------------------------------------------------------------
@pytest.mark.parametrize("xr", [0, 1])
@pytest.mark.parametrize("d", [3, 4, 6])
def test_mn_vs_brute_force(d, xr):
    """z^d + c (mandelbrot_Mn.py:659-729), binary64 and forced-Xrange deltas:
    with BLA off the escape count is exact and z_n, dz_n/dc agree within
    1e-6; with BLA on the count is within one iteration."""
    f = frames.mn_frame(d, nx=24)
    args, c_pix = frames.frame_inputs(f)
    fields = dict(args.fields, xr_mode=xr)
    off = orc.OracleCtx(dict(fields, bla_on=0), args.arrays).cycles(c_pix)
    on = orc.OracleCtx(fields, args.arrays).cycles(c_pix)
    mi = fields["max_iter"]
    dc_row = 2 if fields["interior_detect"] else 1
    n_esc = n_on = 0
    for i in sample(len(c_pix), 16, seed=d):
        n, z, dc = brute.holo_escape(f, c_pix[i], d, mi, 1e3, prec=f.prec + 128)
        assert n == off[3][0, i], (i, n, off[3][0, i])
        if off[2][0, i] == 1:
            assert rel(off[0][0, i], z) <= 1e-6, (i, off[0][0, i], z)
            assert rel(off[0][dc_row, i], dc) <= 1e-6, (i, off[0][dc_row, i], dc)
            n_esc += 1
        n_on += int(abs(int(n) - int(on[3][0, i])) <= 1)
    assert n_esc >= 8
    assert n_on >= 15


# ---- burning ship: BLA and Jacobian ------------------------------------------
@pytest.mark.parametrize("name", ["BS_ex14", "PBS_ex18"])
def test_bs_bla_and_jacobian_vs_brute_force(name):
    """The burning-ship loop with its BLA (perturbation.py:1793-1811,
    1884-2105): on the pixels whose binary64 orbit tracks the exact one with
    BLA off (escape count equal to brute force), BLA on stays within one
    iteration; and the four Jacobian rows dx/da, dx/db, dy/da, dy/db (times
    dx) of the BLA-off loop agree with forward differentiation of the direct
    iteration within 1e-6 (of the largest entry)."""
    f = frames.make(name, nx=24, max_iter=20000)
    args, c_pix = frames.frame_inputs(f)
    fl = args.fields["flavor"]
    assert args.fields["calc_dzndc"]
    off = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    on = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    tracked = on_ok = jac_ok = 0
    for i in sample(len(c_pix), 12, seed=3):
        n, z, J = brute.bs_escape_jac(f, c_pix[i], fl, 20000, 1e3, prec=f.prec + 128)
        if n != off[3][0, i] or off[2][0, i] != 1:
            continue
        if abs(complex(off[0][0, i], off[0][1, i]) - z) > 1e-8 * abs(z):
            continue
        tracked += 1
        on_ok += int(abs(int(n) - int(on[3][0, i])) <= 1)
        jn = max(abs(v) for v in J)
        got = off[0][2:6, i]
        jac_ok += int(all(abs(g - w) <= 1e-6 * jn for g, w in zip(got, J)))
    assert tracked >= 5, tracked
    assert on_ok >= tracked - 1
    assert jac_ok == tracked


# ---- dz/dz ----------------------------------------------------------------------
@pytest.mark.parametrize("name", ["M2_E20", "M2_divref5"])
def test_dzndz_values_vs_brute_force(name):
    """Interior detection's dz/dz (perturbation.py:1189-1246): the derivative
    of z_n with respect to z_1 along the pixel's own orbit.  The loop writes
    the delta; before the first rebase the orbit's dZndz path value at U is to
    be added, after it (nullify_dZndz) the delta is the whole derivative,
    because dZndz_path[0] = 0 (perturbation.py:2501-2512).  Reference orbit
    from the image centre (no cycle), so that a pixel has rebased exactly when
    U != stop_iter.  BLA off, binary64 deltas.  The second frame's reference
    orbit escapes at 2781, before its pixels do (2773-2805), so that they also
    rebase at the end of the reference (w >= div_iter - 1,
    perturbation.py:1283-1313)."""
    f = frames.make(name, nx=32, no_newton=True, interior_detect=True)
    args, c_pix = frames.frame_inputs(f)
    assert args.fields["ref_order"] >= 1 << 62
    ctx = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays)
    Z, U, sr, si = ctx.cycles(c_pix)
    m, _ = ctx.path(1)
    dz_path = m.view(np.complex128)
    # pixels whose last rebase was at n = div_iter - 1: among them those that
    # reach the end of the reference orbit without a dynamic rebase, where the
    # dZndz path is added for the first time at that end
    at_end = np.flatnonzero(si[0] - U[0] == args.fields["ref_div_iter"] - 1)[:40]
    n_reb = n_plain = 0
    for i in np.concatenate([sample(len(c_pix), 20, seed=5), at_end]):
        if sr[0, i] != 1:
            continue
        n, z, dc, dzz = brute.holo_escape_dz(f, c_pix[i], 2, args.fields["max_iter"], 1e3)
        assert n == si[0, i]
        rebased = U[0, i] != si[0, i]
        total = Z[1, i] + (0 if rebased else dz_path[U[0, i]])
        assert rel(total, dzz) <= 1e-6, (i, rebased, total, dzz)
        n_reb += int(rebased)
        n_plain += int(not rebased)
    assert n_reb >= 3
    if name == "M2_divref5":
        assert at_end.size >= 3


def test_forced_xrange_interior_frame_matches_standard():
    """The Xrange loop's dz/dz (written as 0, perturbation.py:1377-1380) only
    shows in the stationarity test: forcing Xrange on the period-14292
    interior frame stops every pixel at the same iteration, for the same
    reason, as the binary64 loop, which the test above pins by brute force."""
    f = frames.make("M2_int32", nx=24, max_iter=200000)
    args, c_pix = frames.frame_inputs(f)
    std = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    xr = orc.OracleCtx(dict(args.fields, xr_mode=1), args.arrays).cycles(c_pix)
    # 6 stationary, 35 at max_iter, the rest escaped
    assert np.count_nonzero(std[2] == 2) >= 5 and np.count_nonzero(std[2] == 0) >= 20
    np.testing.assert_array_equal(std[2], xr[2])
    np.testing.assert_array_equal(std[3], xr[3])
    np.testing.assert_array_equal(std[1], xr[1])


# ---- deep frames: the Xrange loop where it matters ----------------------------
DEEP = {"C2": 16, "C2_2000": 6, "C3_d3": 8, "C3_d4": 6, "C3_d6": 6, "C3_bs": 4}


@pytest.mark.parametrize("name", list(DEEP))
def test_deep_frame_vs_brute_force(name):
    """The seeded benchmark frames at 1e-300 (2048^2 recipe), 1e-2000 and
    1e-800 (z^3 + c, z^6 + c, burning ship), resized to 48x48 over the same
    region: every delta is below binary64's range, so the Xrange register,
    the Xrange rebase confirmation and the extended derivative paths run.
    BLA off: exact escape counts, z_n and dz_n/dc (burning ship: x_n, y_n and
    the Jacobian) within 1e-6; BLA on: within one iteration."""
    w = configs.WORKLOADS[name]
    f = configs.build(w, nx=48, ny=48)
    args, c_pix = frames.frame_inputs(f, "bench")
    assert args.fields["xr_mode"] == 1
    off = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    on = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    bs = w.model == "BS"
    dc_row = 2 if args.fields["interior_detect"] else 1
    rebases = 0
    for i in sample(len(c_pix), DEEP[name], seed=11):
        if bs:
            n, z, J = brute.bs_escape_jac(f, c_pix[i], args.fields["flavor"],
                                          args.fields["max_iter"], 1e3, prec=f.prec + 128)
            got_z = complex(off[0][0, i], off[0][1, i])
        else:
            n, z, dc = brute.holo_escape(f, c_pix[i], w.degree, args.fields["max_iter"], 1e3,
                                         prec=f.prec + 128)
            got_z = off[0][0, i]
        assert n == off[3][0, i], (i, n, off[3][0, i])
        assert off[2][0, i] == 1
        assert rel(got_z, z) <= 1e-6, (i, got_z, z)
        if bs:
            jn = max(abs(v) for v in J)
            for g, v in zip(off[0][2:6, i], J):
                assert abs(g - v) <= 1e-6 * jn, (i, off[0][2:6, i], J)
        else:
            assert rel(off[0][dc_row, i], dc) <= 1e-6, (i, off[0][dc_row, i], dc)
        assert abs(int(n) - int(on[3][0, i])) <= 1
        rebases += int(off[1][0, i] != off[3][0, i])
    # the sample crosses at least one rebase of the reference orbit
    assert rebases >= 1


# ---- the reference's edge-case frames -----------------------------------------
EDGE = {"M2_E213": (350000, 3), "M2_flake": (50000, 4), "M2_ultradeep": (2_000_000, 1)}
_SLOW = pytest.mark.skipif(not __import__("os").environ.get("FSG_SLOW"),
                           reason="brute force over 231,000 iterations at 1950 bits (~40 s); FSG_SLOW=1 runs it")


@pytest.mark.parametrize("name", ["M2_E213", "M2_flake",
                                  pytest.param("M2_ultradeep", marks=[pytest.mark.slow, _SLOW])])
def test_edge_frame_vs_brute_force(name):
    """Deep zoom at 3.2e-213 (~134,000 iterations per pixel), the flake
    (dynamic glitches, ~31,000) and the Xrange interior frame at 5.07e-433
    (escapes from 231,000 iterations): on the earliest-escaping pixels of a
    40-pixel-wide version, the escape count with BLA off is the brute-force
    one, with BLA on within one iteration; z_n within 1e-6."""
    max_iter, k = EDGE[name]
    f = frames.make(name, nx=40, max_iter=max_iter)
    args, c_pix = frames.frame_inputs(f)
    off = orc.OracleCtx(dict(args.fields, bla_on=0), args.arrays).cycles(c_pix)
    on = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    esc = np.flatnonzero(off[2][0] == 1)
    for i in esc[np.argsort(off[3][0, esc], kind="stable")][:k]:
        n, z, _ = brute.holo_escape(f, c_pix[i], 2, max_iter, 1e3, prec=f.prec + 128)
        assert n == off[3][0, i], (i, n, off[3][0, i])
        assert abs(int(n) - int(on[3][0, i])) <= 1
        assert rel(off[0][0, i], z) <= 1e-6, (i, off[0][0, i], z)
    if name == "M2_ultradeep":
        assert np.count_nonzero(on[2] == 2) > 0 and args.fields["xr_mode"] == 1


@pytest.mark.parametrize("jitter", [0, 1])
def test_supersampling_pixel_offsets(jitter):
    """chunk_pixel_pos with 3x3 supersampling (core.py:1767-1830, the
    reference's test_supersampling, test_perturbation.py:376-420): each
    pixel's 3x3 sub-pixels are centred on it, a third of a pixel apart, rows
    flipped; jitter moves each by at most half its rescaled amplitude,
    reproducibly (seeded)."""
    f = frames.make("M2_supersampling", nx=12)
    nx, ny, s = f.nx, f.ny, 3
    plain = f.chunk_pixel_pos((0, nx, 0, ny))
    ss = f.chunk_pixel_pos((0, nx, 0, ny), 0, s)
    assert ss.shape == (ny * s, nx * s)
    blocks = ss.reshape(ny, s, nx, s).mean(axis=(1, 3))
    np.testing.assert_allclose(blocks, plain, atol=1e-15)
    step_x = 1. / (nx - 1) / s
    np.testing.assert_allclose(np.diff(ss.real, axis=1), step_x, rtol=1e-9)
    np.testing.assert_allclose(np.diff(ss.imag, axis=0), -1. / (ny - 1) / s / f.xy_ratio, rtol=1e-9)
    if jitter:
        jt = f.chunk_pixel_pos((0, nx, 0, ny), 1, s)
        d = jt - ss
        amp = 0.5 * 0.7071067811865476 / (nx - 1) / s
        assert np.all(np.abs(d.real) <= amp) and np.all(np.abs(d.imag) <= amp)
        assert np.std(d.real) > amp / 4
        np.testing.assert_array_equal(jt, f.chunk_pixel_pos((0, nx, 0, ny), 1, s))
