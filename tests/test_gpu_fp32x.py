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
