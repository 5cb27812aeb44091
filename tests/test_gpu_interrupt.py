"""The interrupt flag during a launch (the reference polls _interrupted after
every pixel, perturbation.py:1042-1043): set from the host while a long tile
runs, the kernel takes no new pixel, the call returns FSG_INTERRUPTED well
before the tile would have finished, the pixels never taken keep stop_reason
-1, and every pixel that was computed equals the uninterrupted run's."""
import threading
import time

import numpy as np
import pytest

import frames
from gpu_util import run_gpu
from fractalshades_amd import _native

pytestmark = pytest.mark.gpu


def test_interrupt_mid_launch():
    # 3.2e-213 zoom, BLA off: ~134,000 plain steps per pixel
    f = frames.make("M2_E213", nx=1152, max_iter=350000, calc_orbit=False, backshift=0)
    args, c_pix = frames.frame_inputs(f)
    args = args.with_fields(bla_on=0)
    print(f"case interrupt npts={c_pix.size}", flush=True)
    t0 = time.perf_counter()
    full, st_full = run_gpu(args, c_pix)
    t_full = time.perf_counter() - t0
    print(f"uninterrupted {t_full:.3f} s", flush=True)
    assert t_full > 0.3

    n = c_pix.size
    nz = _native.n_z_fields(args.fields)
    Z = np.zeros((nz, n), dtype=np.complex128)
    U = np.zeros((1, n), dtype=np.int32)
    sr = np.zeros((1, n), dtype=np.int8)
    si = np.zeros((1, n), dtype=np.int32)
    flag = np.zeros(1, dtype=np.bool_)
    out = {}

    def run():
        t = time.perf_counter()
        out["rc"], _ = args.ctx.tile(c_pix, Z, U, sr, si, interrupted=flag)
        out["t"] = time.perf_counter() - t

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.25 * t_full)
    t_set = time.perf_counter()
    flag[0] = True
    th.join(timeout=60)
    assert not th.is_alive()
    after = time.perf_counter() - t_set
    print(f"interrupted: rc {out['rc']}, returned {after:.3f} s after the flag, "
          f"{np.count_nonzero(sr == -1)} of {n} pixels not taken", flush=True)
    assert out["rc"] == _native.FSG_INTERRUPTED
    assert np.count_nonzero(sr == -1) > n // 10
    assert after < 0.6 * t_full
    done = sr[0] >= 0
    assert np.count_nonzero(done) > 0
    np.testing.assert_array_equal(si[0, done], full[3][0, done])
    np.testing.assert_array_equal(sr[0, done], full[2][0, done])
    np.testing.assert_array_equal(np.ascontiguousarray(Z[:, done]).view(np.uint64),
                                  np.ascontiguousarray(full[0][:, done]).view(np.uint64))
    # the context is usable again, uninterrupted
    again, _ = run_gpu(args, c_pix[:4096])
    np.testing.assert_array_equal(again[3], full[3][:, :4096])
