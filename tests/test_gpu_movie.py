"""The zoom-movie workload on the GPU: one seeded 2100-digit period-76
nucleus, its orbit computed unwrapped to 5,000,001 points (the config's
~5e6-entry orbit, Xrange register every period) and kept resident; frames of
the zoom sequence at max_iter 5e6 in its three regimes -- a shallow frame
around the critical orbit with BLA off (width > 1e-5, perturbation.py:669-672),
a binary64 frame and an Xrange frame on the resident orbit -- run through the
serving loop, each bit-identical to the same frame run alone and, on a strided
subsample, to the oracle."""
import numpy as np
import pytest

import oracle as orc
from gpu_util import run_gpu, assert_same
from fractalshades_amd import serving
from fractalshades_amd.serving import TileServer, PinnedTile, n_fields

pytestmark = pytest.mark.gpu

MAX_ITER = 5_000_000


@pytest.fixture(scope="module")
def centre():
    return serving.movie_centre(MAX_ITER)


def test_movie_frames_three_regimes(centre):
    Zn_path, div_iter, xr = centre.orbits[True]
    assert Zn_path.size == MAX_ITER + 1 and len(xr) > 60000
    widths = ["2.5e-3", "1e-150", "3.2e-1187", "1e-2000"]
    specs = [serving.RequestSpec(0, 480, 270, w, MAX_ITER, 0.0, 0.0) for w in widths]
    bases, reqs = {}, []
    for sp in specs:
        f, cia, c_pix = serving.build_request(sp, [centre], "fp64")
        if cia.resident_key not in bases:
            bases[cia.resident_key] = cia.ctx
        else:
            cia = cia.shared_with(bases[cia.resident_key])
        reqs.append((sp, cia, c_pix))
    regimes = [(cia.resident_key[0], cia.fields["xr_mode"], cia.fields["bla_on"]) for _, cia, _ in reqs]
    assert regimes[0] == ("critical", 0, 0)
    assert regimes[1] == (0, 0, 1) and regimes[2] == (0, 1, 1) and regimes[3] == (0, 1, 1)
    srv = TileServer(lanes=3, chunk_px=65536, hbm_budget=2 << 30)
    try:
        tiles = []
        for sp, cia, c_pix in reqs:
            t = PinnedTile(c_pix.size, n_fields(cia.fields), True)
            t.c_pix[:] = c_pix
            t.stop_reason[:] = -1
            tiles.append((t, srv.submit(cia.ctx, t)))
        infos = [srv.wait(tk, 600000) for _, tk in tiles]
    finally:
        srv.close()
    for (sp, cia, c_pix), (t, _), info in zip(reqs, tiles, infos):
        hist = np.bincount(t.stop_reason.reshape(-1).astype(np.int64) + 1, minlength=4)
        print(f"movie frame {sp.width}: stop reasons (-1, max_iter, escaped, interior) {hist.tolist()}, "
              f"max stop_iter {int(t.stop_iter.max())}", flush=True)
        assert info["status"] == 0 and hist[0] == 0
        alone, _ = run_gpu(cia, c_pix)
        assert_same((t.Z, t.U, t.stop_reason, t.stop_iter), alone)
        sel = np.arange(0, c_pix.size, 997)
        ref = orc.OracleCtx(cia.fields, cia.arrays).cycles(c_pix[sel])
        assert_same(tuple(np.ascontiguousarray(a[..., sel]) for a in (t.Z, t.U, t.stop_reason, t.stop_iter)), ref)
        t.close()
