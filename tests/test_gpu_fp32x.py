"""The binary32-mantissa + int32-exponent form against the binary64 drop-in
on the same frames (the drop-in is bit-exact with the oracle, see
test_gpu_parity.py): escape counts, stop reasons, continuous iteration and
DEM, with the errors printed and bounded."""
import numpy as np
import pytest

import frames
from gpu_util import run_gpu
from fractalshades_amd import configs
from fractalshades_amd.postproc import error_report

pytestmark = pytest.mark.gpu

FP32X = 1


def compare(args, c_pix, d=2):
    ref, _ = run_gpu(args, c_pix)
    got, stats = run_gpu(args.with_fields(precision=FP32X), c_pix)
    rep = error_report(got, ref, holomorphic=args.fields["kind"] != 2, d=d)
    print(rep, flush=True)
    return rep


@pytest.mark.parametrize("name", ["M2_E20", "M2_E11", "M2_divref5", "M2_int32"])
def test_fp32x_frames(name):
    f = frames.make(name, nx=128)
    args, c_pix = frames.frame_inputs(f)
    rep = compare(args, c_pix)
    assert rep["unfinished"] == 0 and rep["non_finite_zn"] == 0
    assert rep["escape_count_mismatch_frac"] < 0.05


@pytest.mark.parametrize("d", [3, 4, 6])
def test_fp32x_mn(d):
    f = frames.mn_frame(d, nx=128)
    args, c_pix = frames.frame_inputs(f)
    rep = compare(args, c_pix, d=d)
    assert rep["unfinished"] == 0 and rep["escape_count_mismatch_frac"] < 0.05


@pytest.mark.parametrize("name", ["C1", "C2"])
def test_fp32x_workloads(name):
    w = configs.WORKLOADS[name]
    f = configs.build(w, nx=512, ny=512)
    args = f._calc_data["bench"]["cycle_indep_args"]
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
    rep = compare(args, c_pix)
    assert rep["unfinished"] == 0 and rep["non_finite_zn"] == 0
    assert rep["escape_count_mismatch_frac"] <= 1e-4


@pytest.mark.parametrize("name", ["C3_d4", "C3_bs"])
def test_fp32x_deep_workloads(name):
    """The 1e-800 frames (reduced to 480x270 pixels): BLA coefficients and
    derivatives far outside the binary32 range."""
    try:
        configs.center_for(name, compute=False)
    except KeyError:
        pytest.skip(f"no cached seeded centre for {name}")
    w = configs.WORKLOADS[name]
    f = configs.build(w, nx=480, ny=270)
    args = f._calc_data["bench"]["cycle_indep_args"]
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
    print(f"case fp32x {name} npts={c_pix.size}", flush=True)
    rep = compare(args, c_pix, d=w.degree if w.model == "MN" else 2)
    assert rep["unfinished"] == 0 and rep["non_finite_zn"] == 0
    assert rep["escape_count_mismatch_frac"] < 0.01
    assert rep["stop_reason_mismatch"] == 0


@pytest.mark.parametrize("name,flavor", [("BS_ex14", "Burning ship"), ("PBS_ex18", "Perpendicular burning ship")])
@pytest.mark.parametrize("xr", [0, 1])
def test_fp32x_burning_ship_deep(name, flavor, xr):
    """The burning-ship family in the binary32 form (the binary64 kernels'
    step and Jacobian templates on a binary32 mantissa with an int32
    exponent) on the reference examples' deep frames.  These frames are
    chaotic at binary64 precision already: switching BLA off moves 9 % (BS)
    and 0.2 % (PBS) of the escape counts of the drop-in, and binary32 moves
    16 % and 2 %; the continuous iteration of the escaped pixels agrees to
    ~1e-8 (median)."""
    f = frames.make(name, nx=96, model_kw={"flavor": flavor}, max_iter=6000)
    args, c_pix = frames.frame_inputs(f)
    args = args.with_fields(xr_mode=xr)
    print(f"case fp32x {name} {flavor} xr={xr}", flush=True)
    rep = compare(args, c_pix)
    assert rep["unfinished"] == 0 and rep["non_finite_zn"] == 0
    assert rep["escape_count_mismatch_frac"] < 0.25
    assert rep["stop_reason_mismatch"] <= 0.01 * rep["pixels"]
    assert rep["nu_rel_err_median"] < 1e-5 and rep["nu_rel_err_p99"] < 1e-3


@pytest.mark.parametrize("flavor", ["Burning ship", "Perpendicular burning ship", "Shark fin", "Celtic", "Buffalo"])
def test_fp32x_burning_ship_flavours(flavor):
    """Every flavour in the binary32 form on a whole-set view (perturbation
    around the critical orbit, as the reference does above 1e-5)."""
    import fractalshades_amd as fsa
    f = fsa.Perturbation_burning_ship("/tmp/fsa_test", flavor=flavor)
    f.zoom(precision=20, x="-0.5", y="-0.5", dx="3.0", nx=96, xy_ratio=1.0, theta_deg=0.)
    f.calc_std_div(calc_name="t", subset=None, max_iter=2000, M_divergence=1.e3, BLA_eps=1.e-6,
                   calc_hessian=True)
    args, c_pix = frames.frame_inputs(f)
    print(f"case fp32x flavour {flavor}", flush=True)
    rep = compare(args, c_pix)
    assert rep["unfinished"] == 0 and rep["non_finite_zn"] == 0
    assert rep["escape_count_mismatch_frac"] < 0.05
    assert rep["stop_reason_mismatch"] <= 0.01 * rep["pixels"]
    assert rep["nu_rel_err_median"] < 1e-5 and rep["nu_rel_err_p99"] < 1e-3


@pytest.mark.parametrize("case", ["M2_E20", "M2_divref5", "Mn3", "C2", "C2_2000", "BS_ex14", "C3_bs"])
def test_fp32x_lds_matches_global(case, monkeypatch):
    """The binary32 loop with its tables staged in LDS (k_tile_lds_x, the
    default when they fit; at 1e-2000 only the xs orbit and the BLA tables
    fit and the rest stays in global memory; k_tile_lds_bsx for the burning
    ship) gives the same bits as with the tables in global memory
    (FSG_TABLES=global): only where the tables are read from changes."""
    from gpu_util import assert_same
    if case == "Mn3":
        f = frames.mn_frame(3, nx=96)
        args, c_pix = frames.frame_inputs(f)
    elif case in ("C2", "C2_2000", "C3_bs"):
        try:
            configs.center_for(case, compute=False)
        except KeyError:
            pytest.skip(f"no cached seeded centre for {case}")
        f = configs.build(configs.WORKLOADS[case], nx=512, ny=512)
        args = f._calc_data["bench"]["cycle_indep_args"]
        c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))))
    elif case == "BS_ex14":
        f = frames.make(case, nx=96, max_iter=6000)
        args, c_pix = frames.frame_inputs(f)
    else:
        f = frames.make(case, nx=96)
        args, c_pix = frames.frame_inputs(f)
    a32 = args.with_fields(precision=FP32X)
    print(f"case fp32x tables {case}", flush=True)
    monkeypatch.delenv("FSG_TABLES", raising=False)
    monkeypatch.delenv("FSG_LDS_BYTES", raising=False)
    got, _ = run_gpu(a32, c_pix, grid_w=f.nx)
    monkeypatch.setenv("FSG_TABLES", "global")
    ref, _ = run_gpu(a32, c_pix, grid_w=f.nx)
    assert_same(got, ref)
    assert np.all(got[2] >= 0)
