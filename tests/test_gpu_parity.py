"""GPU engine vs the oracle, through the C ABI: bit-exact on every output
(escape iteration, stop reason, reference position, z_n, dz_n/dc, dz_n/dz),
for binary64 and Xrange deltas, BLA on and off, on the reference test
suite's zoom locations and the seeded benchmark workloads."""
import numpy as np
import pytest

import oracle as orc
import frames
from gpu_util import run_gpu, assert_same
from fractalshades_amd import configs
from fractalshades_amd.perturbation import CycleIndepArgs

pytestmark = pytest.mark.gpu


def with_fields(args, **over):
    return CycleIndepArgs(dict(args.fields, **over), args.arrays, None, projection=args.projection)


@pytest.mark.parametrize("name", ["M2_E20", "M2_E11", "M2_divref", "M2_divref5", "M2_int32"])
@pytest.mark.parametrize("xr", [0, 1])
def test_m2_bitexact(name, xr):
    f = frames.make(name, nx=64)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case {name} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)
    assert stats["plain_steps"] > 0


@pytest.mark.parametrize("dc,dz", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_m2_xr_derivative_combinations_bitexact(dc, dz):
    # the extended-range z^2 + c kernel is compiled once per combination of
    # derivatives: each one against the oracle, on a strided subsample of the
    # interior-dominated workload (2048x2048 at 5e-601, max_iter 2e5), where
    # pixels stop on all three tests
    if not _cached("C2_int"):
        pytest.skip("no cached seeded centre for C2_int")
    f = configs.build(configs.WORKLOADS["C2_int"])
    args = with_fields(f._calc_data["bench"]["cycle_indep_args"], calc_dzndc=dc, interior_detect=dz)
    c_pix = np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny)))
    c_pix = np.ascontiguousarray(c_pix[np.arange(0, c_pix.size, 2039)])
    print(f"case C2_int dc={dc} dz={dz} npts={c_pix.size}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)
    assert stats["bla_steps"] > 0 and stats["rebases"] > 0
    hist = np.bincount(ref[2].ravel().astype(np.int64), minlength=3)
    assert hist[0] > 0 and hist[1] > 0 and (hist[2] > 0) == bool(dz)


@pytest.mark.parametrize("dc,dz", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_m2_std_derivative_combinations_bitexact(dc, dz):
    # the binary64 z^2 + c kernel with its tables in LDS is compiled once per
    # combination of derivatives too: each one against the oracle on the
    # reference's E11 frame in binary64 deltas with BLA on, where pixels stop
    # on all three tests (max_iter, escape, interior)
    f = frames.make("M2_E11", nx=64)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=0, calc_dzndc=dc, interior_detect=dz)
    print(f"case M2_E11 std dc={dc} dz={dz} npts={c_pix.size}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)
    assert stats["plain_steps"] > 0 and stats["bla_steps"] > 0
    hist = np.bincount(ref[2].ravel().astype(np.int64), minlength=3)
    assert hist[0] > 0 and hist[1] > 0 and (hist[2] > 0) == bool(dz)


@pytest.mark.parametrize("bla", [0, 1])
def test_m2_bla_switch_bitexact(bla):
    f = frames.make("M2_E20", nx=64)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, bla_on=bla)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)
    assert (stats["bla_steps"] > 0) == bool(bla)


def test_bla_tree_and_paths_bitexact():
    f = frames.make("M2_int32", nx=32)
    args, _ = frames.frame_inputs(f)
    o = orc.OracleCtx(args.fields, args.arrays)
    M, r, stages = o.bla()
    bl, st = args.ctx.bla_shape()
    assert bl == r.size and st == stages
    np.testing.assert_array_equal(args.ctx.read_table(0, M.shape, np.float64).view(np.uint64), M.view(np.uint64))
    np.testing.assert_array_equal(args.ctx.read_table(1, r.shape, np.float64).view(np.uint64), r.view(np.uint64))
    L = args.fields["ref_len"]
    m, e = o.path(0)
    np.testing.assert_array_equal(args.ctx.read_table(2, (2 * L,), np.float64), m)
    m, e = o.path(1)
    np.testing.assert_array_equal(args.ctx.read_table(4, (2 * (L + 1),), np.float64), m)


def _cached(name):
    try:
        configs.center_for(name, compute=False)
        return True
    except KeyError:
        return False


@pytest.mark.parametrize("name", ["C1", "C2", "C2_2000", "C3_d3", "C3_d4", "C3_d6", "C3_bs"])
def test_workload_subsample_bitexact(name):
    # full-size seeded benchmark frame on the GPU (C3: z^d + c and the
    # burning ship at 1e-800, 1920x1080, max_iter 1e6), checked on a strided
    # subsample against the oracle
    if not _cached(name):
        pytest.skip(f"no cached seeded centre for {name}")
    w = configs.WORKLOADS[name]
    f = configs.build(w)
    args = f._calc_data["bench"]["cycle_indep_args"]
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
    print(f"case {name} npts={c_pix.size}", flush=True)
    got, stats = run_gpu(args, c_pix)
    sel = np.arange(0, c_pix.size, 997)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix[sel])
    sub = tuple(np.ascontiguousarray(a[..., sel]) for a in got)
    assert_same(sub, ref)
    assert np.all(got[2] >= 0)
    # the whole frame again in 8x8 pixel blocks (the product path's order):
    # the same results at the same indices
    print(f"case {name} grid_w={f.nx}", flush=True)
    got_grid, stats_grid = run_gpu(args, c_pix, grid_w=f.nx)
    assert_same(got_grid, got)
    assert stats_grid["plain_steps"] == stats["plain_steps"]


@pytest.mark.parametrize("frame", ["M2_E20-1", "Mn3-0", "BS_ex14-Celtic-1"])
@pytest.mark.parametrize("precision", [0, 1])
def test_grid_order_bitexact(frame, precision):
    """Pixels taken in 8x8 blocks of rows of grid_w: the oracle's results at
    the same indices, for row widths that divide the frame, that leave a
    ragged right strip and a short last row, narrower than a block, of one
    pixel, and wider than the whole input."""
    parts = frame.split("-")
    if parts[0] == "Mn3":
        f = frames.mn_frame(3, nx=45)
    elif parts[0].startswith("M2"):
        f = frames.make(parts[0], nx=50)
    else:
        f = frames.make(parts[0], nx=44, model_kw={"flavor": parts[1]}, max_iter=6000)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=int(parts[-1]), precision=precision)
    if precision == 0:
        ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    else:
        ref, _ = run_gpu(args, c_pix)      # binary32 path: index order
    n = c_pix.size
    for w in (f.nx, 37, 8, 5, 1, n + 3):
        print(f"case grid {frame} precision={precision} grid_w={w} n={n}", flush=True)
        got, _ = run_gpu(args, c_pix, grid_w=w)
        assert_same(got, ref)


@pytest.mark.parametrize("d", [3, 4, 6])
@pytest.mark.parametrize("xr", [0, 1])
def test_mn_bitexact(d, xr):
    f = frames.mn_frame(d, nx=64)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case Mn d={d} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)


@pytest.mark.parametrize("name,flavor,xr", [
    ("BS_ex14", None, 0), ("BS_ex14", None, 1), ("PBS_ex18", None, 0),
    ("SF_ex22", None, 1), ("BS_ex14", "Celtic", 0), ("BS_ex14", "Buffalo", 1),
])
def test_bs_bitexact(name, flavor, xr):
    f = frames.make(name, nx=48, model_kw={"flavor": flavor} if flavor else None,
                    max_iter=6000)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case {name} {flavor} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)


@pytest.mark.parametrize("case", [
    ("M2_E20", 0, 3, 1), ("M2_E20", 1, 3, 1), ("M2_divref5", 0, 1, 1), ("M2_E11", 1, 7, 0),
])
def test_calc_orbit_bitexact(case):
    """calc_orbit / backshift: the extra zn_orbit field, BLA on and off,
    binary64 and Xrange deltas."""
    name, xr, backshift, bla = case
    f = frames.make(name, nx=48, calc_orbit=True, backshift=backshift)
    args, c_pix = frames.frame_inputs(f)
    assert args.fields["calc_orbit"] == 1 and args.fields["backshift"] == backshift
    args = with_fields(args, xr_mode=xr, bla_on=bla)
    print(f"case orbit {case}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, _ = run_gpu(args, c_pix)
    assert got[0].shape[0] == f._calc_data["t"]["state"].codes[0].__len__()
    assert_same(got, ref)


@pytest.mark.parametrize("d,xr", [(3, 0), (4, 1)])
def test_calc_orbit_mn_bitexact(d, xr):
    f = frames.mn_frame(d, nx=40, calc_orbit=True, backshift=2)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case orbit Mn d={d} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, _ = run_gpu(args, c_pix)
    assert_same(got, ref)


@pytest.mark.parametrize("flavor,xr", [(None, 0), ("Celtic", 1)])
def test_calc_orbit_bs_bitexact(flavor, xr):
    f = frames.make("BS_ex14", nx=40, model_kw={"flavor": flavor} if flavor else None,
                    max_iter=6000, calc_orbit=True, backshift=4)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case orbit BS {flavor} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, _ = run_gpu(args, c_pix)
    assert got[0].shape[0] == 8
    assert_same(got, ref)


@pytest.mark.parametrize("d", [2, 5, 7, 8])
@pytest.mark.parametrize("xr", [0, 1])
def test_mn_other_degrees_bitexact(d, xr):
    f = frames.mn_frame(d, nx=40)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    print(f"case Mn d={d} xr={xr}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, _ = run_gpu(args, c_pix)
    assert_same(got, ref)


@pytest.mark.parametrize("case", ["M2_E20-1", "M2_divref5-0", "Mn3-1", "BS_ex14-Celtic-1", "PBS_ex18-None-0"])
def test_tables_global_and_lds_bitexact(case, monkeypatch):
    """The tables staged in LDS (default when they fit) and read from global
    memory (FSG_TABLES=global) give the oracle's results bit for bit."""
    parts = case.split("-")
    if parts[0] == "Mn3":
        f = frames.mn_frame(3, nx=40)
    elif parts[0].startswith("M2"):
        f = frames.make(parts[0], nx=48)
    else:
        fl = None if parts[1] == "None" else parts[1]
        f = frames.make(parts[0], nx=40, model_kw={"flavor": fl} if fl else None, max_iter=6000)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=int(parts[-1]))
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    # all tables in LDS, in global memory, and partial plans (the orbit and
    # whatever else fits a capped budget in LDS, the rest global)
    for mode in ("lds", "global", "6000", "12000", "24000"):
        print(f"case tables {case} {mode}", flush=True)
        monkeypatch.delenv("FSG_TABLES", raising=False)
        monkeypatch.delenv("FSG_LDS_BYTES", raising=False)
        if mode == "global":
            monkeypatch.setenv("FSG_TABLES", "global")
        elif mode != "lds":
            monkeypatch.setenv("FSG_LDS_BYTES", mode)
        got, _ = run_gpu(with_fields(args), c_pix)
        assert_same(got, ref)


@pytest.mark.parametrize("name,orientation,xr", [("M2_E20", "horizontal", 0), ("M2_E20", "vertical", 1),
                                                 ("M2_E11", "horizontal", 0), ("BS_ex14", "horizontal", 0),
                                                 ("BS_ex14", "vertical", 1)])
def test_expmap_bitexact(name, orientation, xr):
    """The exponential map through the product path (numba_cycle_call:
    pixel offsets projected on the host, the GPU loop, dz/dc times the
    projection's modifier) against the oracle on the same projected offsets
    with the same modifier."""
    kw = {"max_iter": 6000} if name.startswith("BS") else {}
    f = frames.expmap_frame(name, nx=48, hmin=0., hmax=10., orientation=orientation, **kw)
    args, c_pix = frames.frame_inputs(f)
    args = with_fields(args, xr_mode=xr)
    assert args.projection is f.projection
    print(f"case expmap {name} {orientation} xr={xr}", flush=True)
    state = f._calc_data["t"]["state"]
    got = f._result_arrays(state, c_pix)
    assert f.numba_cycle_call(got, args) == 0
    cp = np.ascontiguousarray(f.projection.f(c_pix))
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(cp)
    mod = f.projection.dzndc_modifier(c_pix)
    for row in args.dzndc_rows():
        ref[0][row] *= mod
    assert_same(got[1:], ref)


@pytest.mark.parametrize("orientation", ["horizontal", "vertical"])
def test_expmap_steps_bitexact(orientation, monkeypatch):
    """The exponential map by bands of h (calc_raw_by_steps: per band, the
    BLA radii and derivative scale of its |c| bound, on the shared device
    orbit): every band's tiles equal the oracle run with that band's
    descriptor and modifier."""
    from fractalshades_amd import settings
    monkeypatch.setattr(settings, "chunk_size", 16)
    f = frames.expmap_frame("M2_E11", nx=48, hmin=0., hmax=10., orientation=orientation)
    print(f"case expmap steps {orientation}", flush=True)
    res = f.calc_raw_by_steps("t", exp_zoom_step=16, keep_bands=True)
    assert len(res["bands"]) >= 3
    nx, ny = f.nx, f.ny
    proj = f.projection
    kcs = set()
    for r, (args, chunks) in enumerate(res["bands"]):
        kcs.add((args.fields["kc_mant"], args.fields["kc_exp"]))
        for ix, ixx, iy, iyy in chunks:
            c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((ix, ixx, iy, iyy))))
            ref = orc.OracleCtx(args.fields, args.arrays).cycles(np.ascontiguousarray(proj.f(c_pix)))
            # the band's modifier (the projection is back out of step mode)
            hshift = proj.hmoy - 0.5 * (args.step_h[0] + args.step_h[1])
            mod = np.exp(np.real(proj.pix_to_ht * c_pix) + hshift)
            for row in args.dzndc_rows():
                ref[0][row] *= mod
            sel = (slice(None), slice(iy, iyy), slice(ix, ixx))
            got = tuple(np.ascontiguousarray(a.reshape(a.shape[0], ny, nx)[sel]).reshape(a.shape[0], -1)
                        for a in (res["Z"], res["U"], res["stop_reason"], res["stop_iter"]))
            assert_same(got, ref)
    assert len(kcs) >= 3           # each band has its own BLA bound


def test_deep_expmap_bitexact():
    """A deep exponential map (h in [0, 758] at dx = 1e-330, exp(hmax) above
    1e300: the reference's Xrange h) through the product path: calc_raw runs
    it by bands of h, each band's offsets relative to exp(h_band) and its
    pixel scale dx exp(h_band); every band's pixels equal the oracle run with
    that band's descriptor, offsets and modifier."""
    f = frames.expmap_frame("M2_E20", nx=1200, hmin=0., hmax=758., dx="1e-330", precision=345,
                            no_newton=True, max_iter=3000, x="-0.75", y="0.2")
    assert f.projection.deep
    print("case deep expmap", flush=True)
    res = f.calc_raw("t")
    nx, ny = f.nx, f.ny
    proj = f.projection
    n_bands = 0
    for fields, arrays, chunks in f.expmap_bands("t"):
        n_bands += 1
        for ix, ixx, iy, iyy in chunks:
            c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((ix, ixx, iy, iyy))))
            ref = orc.OracleCtx(fields, arrays).cycles(np.ascontiguousarray(proj.f(c_pix)))
            mod = proj.dzndc_modifier(c_pix)
            for row in (1,):
                ref[0][row] *= mod
            sel = (slice(None), slice(iy, iyy), slice(ix, ixx))
            got = tuple(np.ascontiguousarray(a.reshape(a.shape[0], ny, nx)[sel]).reshape(a.shape[0], -1)
                        for a in (res["Z"], res["U"], res["stop_reason"], res["stop_iter"]))
            assert_same(got, ref)
    assert n_bands >= 5
    assert len(np.unique(res["stop_iter"])) >= 2


@pytest.mark.parametrize("model", ["M2", "Mn3"])
@pytest.mark.parametrize("interior", [0, 1])
def test_critical_orbit_bitexact(model, interior):
    """Shallow frames (width above 1e-5) run around the critical orbit with
    BLA off, as the reference does (perturbation.py:669-672): the orbit is
    constant 0, so every step rebases; with interior detection the dz/dz
    path's wrapped value sits at [L] with L = order + 1."""
    import fractalshades_amd as fsa
    f = fsa.Perturbation_mandelbrot(None) if model == "M2" else fsa.Perturbation_mandelbrot_N(None, 3)
    f.zoom(precision=20, x="-0.6", y="0.1", dx="2.5", nx=48,
           xy_ratio=1.0, theta_deg=0.)
    f.calc_std_div(calc_name="t", subset=None, max_iter=3000, M_divergence=1.e3,
                   epsilon_stationnary=1.e-3, BLA_eps=1.e-6, interior_detect=bool(interior))
    args, c_pix = frames.frame_inputs(f)
    assert args.fields["ref_order"] == 100 and args.fields["bla_on"] == 0
    print(f"case critical {model} interior={interior}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, _ = run_gpu(args, c_pix)
    assert_same(got, ref)


# the reference's edge cases (test_perturbation.py:203-297, 376-420, 503-576,
# 651-716) at reduced size: deep zoom with calc_orbit, Dinkydau's flake (the
# dynamic-glitch case), an Xrange frame at 5.07e-433 with interior
# detection, and 3x3 supersampling (with and without jitter)
EDGE_RUNS = {"M2_E213": dict(nx=48), "M2_flake": dict(nx=64),
             "M2_ultradeep": dict(nx=40, max_iter=2_000_000), "M2_supersampling": dict(nx=24)}


@pytest.mark.parametrize("jitter", [0, 1])
@pytest.mark.parametrize("name", list(EDGE_RUNS))
def test_edge_frames_bitexact(name, jitter):
    if jitter and name != "M2_supersampling":
        pytest.skip("jitter is run on the supersampled frame")
    f = frames.make(name, **EDGE_RUNS[name])
    args = f._calc_data["t"]["cycle_indep_args"]
    ssg = frames.EDGE[name]["supersampling"]
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny), jitter, ssg)))
    print(f"case edge {name} jitter={jitter} npts={c_pix.size}", flush=True)
    ref = orc.OracleCtx(args.fields, args.arrays).cycles(c_pix)
    got, stats = run_gpu(args, c_pix)
    assert_same(got, ref)
    got_grid, _ = run_gpu(args, c_pix, grid_w=f.nx * (ssg or 1))
    assert_same(got_grid, ref)
    # the product path (calc_raw) gives the same pixels
    res = f.calc_raw("t", jitter=jitter, supersampling=ssg)
    np.testing.assert_array_equal(res["stop_iter"], ref[3])
    np.testing.assert_array_equal(res["Z"].view(np.uint64), ref[0].view(np.uint64))
