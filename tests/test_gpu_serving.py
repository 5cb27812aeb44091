"""Serving loop (fsg_server_*): every request's results are bit-identical
whether it runs alone (fsg_tile) or in a mix of concurrently running
requests of other sizes, models, depths and precisions; admission by
device-memory footprint; request bookkeeping."""
import numpy as np
import pytest

import frames
from gpu_util import run_gpu, assert_same
from fractalshades_amd._native import FsgError
from fractalshades_amd.serving import TileServer, PinnedTile, n_fields

pytestmark = pytest.mark.gpu


def with_fields(args, **over):
    return args.with_fields(**over) if over else args


def _mix():
    """(label, CycleIndepArgs, c_pix, row width) for a mix of models, sizes,
    depths and precisions."""
    out = []
    f = frames.make("M2_E20", nx=96)
    a, c = frames.frame_inputs(f)
    out += [("M2_E20 fp64", a, c, 96), ("M2_E20 xr", with_fields(a, xr_mode=1), c, 96),
            ("M2_E20 fp32x", with_fields(a, precision=1), c, 96)]
    f = frames.make("M2_divref5", nx=64)
    a, c = frames.frame_inputs(f)
    out += [("M2_divref5", a, c, 64)]
    f = frames.make("M2_E11", nx=40, interior_detect=True)
    a, c = frames.frame_inputs(f)
    out += [("M2_E11 interior xr", with_fields(a, xr_mode=1), c, 40)]
    f = frames.mn_frame(3, nx=48)
    a, c = frames.frame_inputs(f)
    out += [("Mn3", a, c, 48)]
    f = frames.make("BS_ex14", nx=56, max_iter=4000)
    a, c = frames.frame_inputs(f)
    out += [("BS_ex14", a, c, 56), ("BS_ex14 xr", with_fields(a, xr_mode=1), c, 56),
            ("BS_ex14 fp32x", with_fields(a, precision=1), c, 56)]
    return out


def _tile_of(args, c_pix):
    t = PinnedTile(c_pix.size, n_fields(args.fields), args.fields["kind"] != 2)
    t.c_pix[:] = c_pix
    t.stop_reason[:] = -1
    return t


def test_server_mix_equals_alone():
    mix = _mix()
    alone = []
    for label, a, c, _ in mix:
        print(f"alone {label} n={c.size}", flush=True)
        alone.append(run_gpu(a, c)[0])
    srv = TileServer(lanes=3, chunk_px=1024, hbm_budget=256 << 20)
    try:
        tiles, tickets = [], []
        # every request twice, interleaved: in index order, then in 8x8
        # blocks of its rows (chunks of 1024 pixels start inside rows)
        for rep in range(2):
            for label, a, c, w in mix:
                t = _tile_of(a, c)
                tiles.append((label, t))
                tickets.append(srv.submit(a.ctx, t, grid_w=w if rep else 0))
        infos = [srv.wait(tk, 120000) for tk in tickets]
    finally:
        srv.close()
    for k, ((label, t), info) in enumerate(zip(tiles, infos)):
        print(f"mix {label} chunks={info['chunks']}", flush=True)
        assert info is not None and info["status"] == 0
        assert info["chunks"] == (t.n + 1023) // 1024
        assert info["t_submit_ms"] <= info["t_admit_ms"] <= info["t_first_ms"] <= info["t_done_ms"]
        assert_same((t.Z, t.U, t.stop_reason, t.stop_iter), alone[k % len(mix)])


def test_server_admission_by_footprint():
    f = frames.make("M2_E20", nx=128)
    a, c = frames.frame_inputs(f)
    ref = run_gpu(a, c)[0]
    probe = TileServer(lanes=2, chunk_px=4096, hbm_budget=64 << 20)
    foot = probe.footprint(a.ctx, c.size)
    probe.close()
    # room for one request at a time: the others wait for the arena
    budget = max(1 << 20, int(foot * 1.5))
    srv = TileServer(lanes=2, chunk_px=4096, hbm_budget=budget)
    try:
        tiles = [_tile_of(a, c) for _ in range(4)]
        tickets = [srv.submit(a.ctx, t) for t in tiles]
        infos = [srv.wait(tk, 120000) for tk in tickets]
        big = _tile_of(a, np.ascontiguousarray(np.tile(c, 4)))
        with pytest.raises(FsgError):
            srv.submit(a.ctx, big)      # alone larger than the budget
    finally:
        srv.close()
    for t, info in zip(tiles, infos):
        assert info["footprint"] == foot
        assert_same((t.Z, t.U, t.stop_reason, t.stop_iter), ref)
    # one at a time: each admission comes after the previous request is done
    for prev, nxt in zip(infos, infos[1:]):
        assert nxt["t_admit_ms"] >= prev["t_done_ms"] - 1e-3


def test_server_empty_request():
    f = frames.make("M2_E20", nx=16)
    a, c = frames.frame_inputs(f)
    srv = TileServer(lanes=1)
    try:
        t = _tile_of(a, c[:0])
        info = srv.wait(srv.submit(a.ctx, t), 10000)
        assert info["status"] == 0 and info["chunks"] == 0
    finally:
        srv.close()


@pytest.mark.parametrize("xr", [0, 1])
def test_shared_context_equals_own(xr):
    """fsg_ctx_create_shared: same tables and results as a context of its own
    for a different width / kc / max_iter on the same orbit."""
    f = frames.make("M2_E20", nx=64, interior_detect=True)
    a, c = frames.frame_inputs(f)
    fields = dict(a.fields, xr_mode=xr, dx_mant=a.fields["dx_mant"] * 1.37,
                  kc_mant=a.fields["kc_mant"] * 1.37, max_iter=a.fields["max_iter"] // 2)
    base = a.ctx
    own = a.__class__(fields, a.arrays, a.interrupted)
    shared = a.__class__(fields, a.arrays, a.interrupted).shared_with(base)
    print(f"case shared xr={xr}", flush=True)
    bl, _ = own.ctx.bla_shape()
    for which, shape, dt in ((0, (bl, 4), np.float64), (1, (bl,), np.float64),
                             (2, (2 * fields["ref_len"],), np.float64),
                             (4, (2 * (fields["ref_len"] + 1),), np.float64)):
        np.testing.assert_array_equal(own.ctx.read_table(which, shape, dt).view(np.uint64),
                                      shared.ctx.read_table(which, shape, dt).view(np.uint64))
    assert_same(run_gpu(shared, c)[0], run_gpu(own, c)[0])


def test_shared_context_rejects_other_orbit():
    f = frames.make("M2_E20", nx=16)
    a, _ = frames.frame_inputs(f)
    g = frames.make("M2_E11", nx=16)
    b, _ = frames.frame_inputs(g)
    with pytest.raises(FsgError):
        _ = b.shared_with(a.ctx).ctx


@pytest.mark.parametrize("name,model_kw", [("M2_E20", None), ("BS_ex14", {"flavor": "Burning ship"})])
def test_server_layers_match_host_postproc(name, model_kw):
    """fsg_server_submit_layers: the continuous iteration count, the DEM and
    the normal map computed on the device from the raw results equal the host
    formulas of postproc.py (Continuous_iter_pp / DEM_pp / Attr_normal_pp
    restated in numpy, the burning-ship frame with its skew) on the raw
    results of fsg_tile, to float32 resolution; NaN where the pixel did not
    diverge."""
    from fractalshades_amd.postproc import continuous_iter, dem, normal_map, post_desc
    from fractalshades_amd.serving import LayerTile
    kw = {"max_iter": 6000} if name.startswith("BS") else {}
    f = frames.make(name, nx=64, model_kw=model_kw, **kw)
    a, c = frames.frame_inputs(f)
    (Z, U, sr, si), _ = run_gpu(a, c)
    esc = sr[0] == 1
    assert esc.sum() > 100
    holo = a.fields["kind"] != 2
    zn = Z[0] if holo else Z[0] + 1j * Z[1]
    d = f.potential_d
    nu_ref = continuous_iter(zn[esc], si[0][esc], d=d, a_d=f.potential_a_d, M=1e3)
    dem_ref = dem(zn[esc], dzndc=Z[-1][esc]) if holo else dem(zn[esc], jac=Z[2:6, esc])
    nrm_ref = (normal_map(zn[esc], dzndc=Z[-1][esc], skew=f._skew) if holo
               else normal_map(zn[esc], jac=Z[2:6, esc], skew=f._skew))
    assert holo or f._skew is not None
    pp = post_desc(f, "t")
    t = LayerTile(c.size, normal=True)
    t.c_pix[:] = c
    srv = TileServer(lanes=2, chunk_px=1024)
    try:
        info = srv.wait(srv.submit_layers(a.ctx, t, pp), 60000)
    finally:
        srv.close()
    assert info["status"] == 0
    np.testing.assert_array_equal(t.stop_iter, si)
    np.testing.assert_array_equal(t.stop_reason, sr)
    assert np.all(np.isnan(t.nu[~esc])) and np.all(np.isnan(t.dem[~esc]))
    np.testing.assert_allclose(t.nu[esc], nu_ref.astype(np.float32), rtol=2e-7, atol=0)
    np.testing.assert_allclose(t.dem[esc], dem_ref.astype(np.float32), rtol=2e-7, atol=0)
    assert np.all(np.isnan(t.normal[~esc]))
    np.testing.assert_allclose(t.normal[esc, 0], nrm_ref.real.astype(np.float32), rtol=0, atol=1.2e-7)
    np.testing.assert_allclose(t.normal[esc, 1], nrm_ref.imag.astype(np.float32), rtol=0, atol=1.2e-7)
    t.close()


@pytest.mark.parametrize("name", ["M2_E20", "BS_ex14"])
@pytest.mark.parametrize("banded", [False, True])
def test_server_layers_expmap(name, banded):
    """The device layers of an exponential-map frame: the derivatives take
    the loop's dzndc_modifier and the projection's df (over the whole map
    exp(t); by bands exp(i Im t); the burning ship's 2x2 matrices), formed on
    the device from each pixel's projected offset, equal the host formulas
    (postproc.screen_derivatives on the screen pixels) to float32
    resolution."""
    from fractalshades_amd.postproc import continuous_iter, dem, normal_map, post_desc, screen_derivatives
    from fractalshades_amd.serving import LayerTile
    kw = {"max_iter": 6000} if name.startswith("BS") else {}
    f = frames.expmap_frame(name, nx=256, hmin=0., hmax=8., **kw)
    p = f.projection
    a, _ = frames.frame_inputs(f)

    def check(a, pix, pp):
        w = np.ascontiguousarray(p.f(pix))
        (Z, U, sr, si), _ = run_gpu(a, w)
        esc = sr[0] == 1
        assert esc.sum() > 50
        holo = a.fields["kind"] != 2
        zn = Z[0] if holo else Z[0] + 1j * Z[1]
        mod = p.dzndc_modifier(pix)
        if holo:
            dz = screen_derivatives(p, pix[esc], dzndc=Z[-1][esc] * mod[esc])
            dem_ref, nrm_ref = dem(zn[esc], dzndc=dz), normal_map(zn[esc], dzndc=dz, skew=f._skew)
        else:
            jac = screen_derivatives(p, pix[esc], jac=tuple(Z[k][esc] * mod[esc] for k in range(2, 6)))
            dem_ref, nrm_ref = dem(zn[esc], jac=jac), normal_map(zn[esc], jac=jac, skew=f._skew)
        nu_ref = continuous_iter(zn[esc], si[0][esc], d=f.potential_d, a_d=f.potential_a_d, M=1e3)
        t = LayerTile(w.size, normal=True)
        t.c_pix[:] = w
        srv = TileServer(lanes=2, chunk_px=1024)
        try:
            info = srv.wait(srv.submit_layers(a.ctx, t, pp), 60000)
        finally:
            srv.close()
        assert info["status"] == 0
        np.testing.assert_array_equal(t.stop_iter, si)
        np.testing.assert_allclose(t.nu[esc], nu_ref.astype(np.float32), rtol=2e-7, atol=0)
        np.testing.assert_allclose(t.dem[esc], dem_ref.astype(np.float32), rtol=5e-7, atol=0)
        np.testing.assert_allclose(t.normal[esc, 0], nrm_ref.real.astype(np.float32), rtol=0, atol=5e-7)
        np.testing.assert_allclose(t.normal[esc, 1], nrm_ref.imag.astype(np.float32), rtol=0, atol=5e-7)
        t.close()

    if not banded:
        pp = post_desc(f, "t")
        assert pp.proj == 1 and pp.mod == 1
        check(a, np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))), pp)
        return
    from fractalshades_amd.perturbation import CycleIndepArgs
    n_bands = 0
    for fields, arrays, chunks in f.expmap_bands("t", exp_zoom_step=200):
        pp = post_desc(f, "t")
        assert pp.proj == 3 and pp.mod == 1
        ab = CycleIndepArgs(fields, arrays, None, projection=p).shared_with(a.ctx)
        pix = np.concatenate([np.ravel(f.chunk_pixel_pos(ch)) for ch in chunks])
        print(f"case layers expmap band {n_bands} of {name}", flush=True)
        check(ab, pix, pp)
        ab.close()
        n_bands += 1
    assert n_bands >= 2


def test_server_wait_from_several_threads():
    """Two threads waiting on one ticket both get the request's status and
    times; the ticket is released after them, so a later wait fails cleanly
    (no access to a freed request)."""
    import threading
    f = frames.make("M2_E20", nx=96)
    a, c = frames.frame_inputs(f)
    ref = run_gpu(a, c)[0]
    srv = TileServer(lanes=2, chunk_px=1024, hbm_budget=64 << 20)
    try:
        t = _tile_of(a, c)
        tk = srv.submit(a.ctx, t)
        got = [None, None]

        def waiter(k):
            got[k] = srv.wait(tk, 60000)

        ths = [threading.Thread(target=waiter, args=(k,)) for k in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join(60)
        assert all(g is not None and g["status"] == 0 for g in got)
        assert got[0]["t_done_ms"] == got[1]["t_done_ms"]
        with pytest.raises(FsgError):
            srv.wait(tk, 1000)
        assert_same((t.Z, t.U, t.stop_reason, t.stop_iter), ref)
    finally:
        srv.close()


def test_serving_stream_subset():
    """The seeded request stream of the serving benchmark (512 requests on
    eight resident 2100-digit nuclei: sizes 256^2, 512^2, 1024^2, 4096x256;
    widths log-uniform in [1e-2000, 1e-20]; max_iter log-uniform in
    [1e3, 1e7]), a seeded subset of six requests of each size, plus every
    other one in the binary32 form: all submitted at once to one server
    (continuous batching across requests of different depth, size and
    arithmetic), each result bit-identical to the same request run alone
    (fsg_tile), and a strided subsample of each binary64 result bit-exact
    with the oracle."""
    import oracle as orc
    from fractalshades_amd import serving
    centres = serving.serve_centres()
    specs = serving.serve_requests(512, len(centres), seed=0)
    rng = np.random.default_rng(7)
    pick = []
    for size in serving.SIZES:
        idx = [k for k, sp in enumerate(specs) if (sp.nx, sp.ny) == size]
        pick += list(rng.choice(idx, 6, replace=False))
    chosen = [specs[k] for k in pick]
    assert min(int(sp.width.split("e")[1]) for sp in chosen) < -1500
    assert max(sp.max_iter for sp in chosen) > 1e6
    bases, reqs = {}, []
    for k, sp in enumerate(chosen):
        precision = "fp32x" if k % 2 else "fp64"
        f, cia, c_pix = serving.build_request(sp, centres, precision)
        key = (cia.resident_key, precision)
        if key not in bases:
            bases[key] = cia.ctx
        else:
            cia = cia.shared_with(bases[key])
        reqs.append((sp, precision, cia, c_pix))
    srv = TileServer(lanes=3, chunk_px=65536, hbm_budget=4 << 30)
    try:
        tiles, tickets = [], []
        for sp, precision, cia, c_pix in reqs:
            t = _tile_of(cia, c_pix)
            tiles.append(t)
            tickets.append(srv.submit(cia.ctx, t))
        infos = [srv.wait(tk, 300000) for tk in tickets]
    finally:
        srv.close()
    for (sp, precision, cia, c_pix), t, info in zip(reqs, tiles, infos):
        print(f"stream {sp.nx}x{sp.ny} width {sp.width} max_iter {sp.max_iter} {precision} "
              f"chunks {info['chunks']}", flush=True)
        assert info["status"] == 0
        alone, _ = run_gpu(cia, c_pix)
        assert_same((t.Z, t.U, t.stop_reason, t.stop_iter), alone)
        if precision == "fp64":
            sel = np.arange(0, c_pix.size, 4099)
            ref = orc.OracleCtx(cia.fields, cia.arrays).cycles(c_pix[sel])
            assert_same(tuple(np.ascontiguousarray(a[..., sel]) for a in (t.Z, t.U, t.stop_reason, t.stop_iter)),
                        ref)
        t.close()
    for _, _, cia, _ in reqs:
        cia.close()
