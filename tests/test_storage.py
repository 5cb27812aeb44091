"""On-disk raw results in the reference's layout (fractalshades_amd.storage,
src/fractalshades/core.py:2125-2512): file names, memmap shapes and dtypes,
the tile report, the fingerprint test; on the GPU, stored tiles equal the
tiles computed one by one, and an interrupted calculation resumes with only
its missing tiles."""
import os

import numpy as np
import pytest

import frames
from fractalshades_amd import settings, storage


def _frame(tmp_path, name="M2_E20", nx=250, **kw):
    f = frames.make(name, nx=nx, **kw)
    f.directory = str(tmp_path)
    return f


def test_layout_and_fingerprint(tmp_path):
    f = _frame(tmp_path)
    assert f.chunks_count == 4 and settings.chunk_size == 200      # 200 + 50 per axis
    storage.init_files(f, "t")
    paths = storage.data_path(f, "t")
    n_pts = f.nx * f.ny
    state = f._calc_data["t"]["state"]
    n_Z, n_U = len(state.codes[0]), len(state.codes[1])
    for key, shape, dt in (("Z", (n_Z, n_pts), np.complex128), ("U", (n_U, n_pts), np.int32),
                           ("stop_reason", (1, n_pts), np.int8), ("stop_iter", (1, n_pts), np.int32)):
        assert paths[key].endswith(f"data{os.sep}t_{key}.arr")
        a = np.load(paths[key], mmap_mode="r")
        assert a.shape == shape and a.dtype == dt
    rep = np.load(storage.report_path(f, "t"))
    assert rep.dtype == np.int32 and rep.shape == (4, 4)
    pts = [(ixx - ix) * (iyy - iy) for ix, ixx, iy, iyy in f.chunk_slices()]
    assert list(rep[:, 2]) == pts and rep[0, 0] == 0 and list(rep[1:, 0]) == list(rep[:-1, 1])
    assert rep[-1, 1] == n_pts and not rep[:, 3].any()
    assert storage.res_available(f, "t")
    assert not storage.res_available(f, "t", next(iter(f.chunk_slices())))
    fp = storage.fingerprint(f, "t")
    assert fp["calc_class"] == "Perturbation_mandelbrot" and fp["calc_callable"] == "calc_std_div"
    assert fp["calc_kwargs"]["max_iter"] == 5000
    # another max_iter: the stored results no longer match
    g = _frame(tmp_path, max_iter=4000)
    assert not storage.res_available(g, "t")


def _add_subset_calc(f, name="t2", func="x == 1", key="stop_reason"):
    """A second calculation on the frame, over the pixels selected by a
    Fractal_array of the stored calculation "t"."""
    sub = storage.Fractal_array(f, "t", key, func=func)
    kw = dict(f._calc_data["t"]["calc_kwargs"])
    kw.update(calc_name=name, subset=sub)
    getattr(f, f._calc_data["t"]["calc_callable"])(**kw)
    return sub


def test_subset_layout(tmp_path):
    """A calculation over a Fractal_array subset stores only the selected
    points: compressed report and arrays, the mask in <calc>_subset.arr,
    tile after tile (core.py:1745-1762, 2404-2424)."""
    f = _frame(tmp_path)
    storage.init_files(f, "t")
    # a stored stop_reason pattern for "t", as if computed
    sr = np.lib.format.open_memmap(storage.data_path(f, "t")["stop_reason"], mode="r+")
    rng = np.random.default_rng(0)
    pattern = rng.integers(0, 3, sr.shape[1]).astype(np.int8)
    sr[0, :] = pattern
    del sr
    sub = _add_subset_calc(f)
    rep_t = np.load(storage.report_path(f, "t"))
    masks = []
    for r, cs in enumerate(f.chunk_slices()):
        m = sub[cs]
        want = pattern[rep_t[r, 0]:rep_t[r, 1]] == 1
        np.testing.assert_array_equal(m, want)
        masks.append(want)
    np.testing.assert_array_equal(sub[None], pattern == 1)
    np.testing.assert_array_equal((~sub)[next(iter(f.chunk_slices()))], ~masks[0])
    storage.init_files(f, "t2")
    rep = np.load(storage.report_path(f, "t2"))
    counts = [int(m.sum()) for m in masks]
    assert list(rep[:, 2]) == counts and rep[-1, 1] == sum(counts)
    z = np.load(storage.data_path(f, "t2")["Z"], mmap_mode="r")
    assert z.shape[1] == sum(counts)
    mask_file = np.load(storage.data_path(f, "t2")["subset"])
    assert mask_file.dtype == bool and mask_file.shape == (f.nx * f.ny,)
    np.testing.assert_array_equal(mask_file, np.concatenate(masks))
    # reload: the tile's mask and its selected pixel offsets
    cs = list(f.chunk_slices())[2]
    m, c_pix, Z, U, sr2, si = storage.reload_data(f, cs, "t2")
    np.testing.assert_array_equal(m, masks[2])
    np.testing.assert_array_equal(c_pix, np.ravel(f.chunk_pixel_pos(cs))[masks[2]])
    assert Z.shape[1] == counts[2]
    # the fingerprint keeps the subset and matches after a reload from disk
    assert storage.fingerprint_matching(f, "t2")
    g_sub = storage.Fractal_array(f, "t", "stop_reason", func="x == 2")
    assert g_sub != sub and storage.Fractal_array(f, "t", "stop_reason", func="x == 1") == sub


def _tiles_equal_one_by_one(f):
    for cs in f.chunk_slices():
        _, c_pix, Z, U, sr, si = storage.reload_data(f, cs, "t")
        ret, (c1, Z1, U1, sr1, si1) = f.compute_tile("t", cs)
        assert ret == 0
        np.testing.assert_array_equal(c_pix, c1)
        for a, b in ((Z, Z1), (U, U1), (sr, sr1), (si, si1)):
            np.testing.assert_array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["M2_E20", "BS_ex14"])
def test_stored_tiles_equal_tiles_one_by_one(tmp_path, name):
    f = _frame(tmp_path, name, nx=250, **({"max_iter": 4000} if name.startswith("BS") else {}))
    print(f"case storage {name}", flush=True)
    assert storage.calc_to_disk(f, "t") == 4
    assert all(storage.res_available(f, "t", cs) for cs in f.chunk_slices())
    _tiles_equal_one_by_one(f)
    assert storage.calc_to_disk(f, "t") == 0          # nothing left to do


@pytest.mark.gpu
def test_resume_computes_only_missing_tiles(tmp_path):
    f = _frame(tmp_path)
    print("case storage resume", flush=True)
    storage.calc_to_disk(f, "t")
    # forget two tiles, as after an interruption
    rep = np.lib.format.open_memmap(storage.report_path(f, "t"), mode="r+")
    rep[1, 3] = 0
    rep[3, 3] = 0
    beg, end = int(rep[1, 0]), int(rep[1, 1])
    del rep
    z = np.lib.format.open_memmap(storage.data_path(f, "t")["Z"], mode="r+")
    z[:, beg:end] = 0
    del z
    assert storage.calc_to_disk(f, "t") == 2
    _tiles_equal_one_by_one(f)


@pytest.mark.gpu
def test_subset_calc_to_disk(tmp_path):
    """A calculation over the pixels of a stored one with an even escape
    count: its stored tiles equal the tiles computed one by one with the
    same subset, and the stored calculation's results at the selected
    pixels."""
    f = _frame(tmp_path)
    print("case storage subset", flush=True)
    storage.calc_to_disk(f, "t")
    _add_subset_calc(f, key="stop_iter", func="x % 2 == 0")
    assert storage.calc_to_disk(f, "t2") == 4
    n_sel = 0
    for cs in f.chunk_slices():
        m, c_pix, Z, U, sr, si = storage.reload_data(f, cs, "t2")
        ret, (c1, Z1, U1, sr1, si1) = f.compute_tile("t2", cs)
        assert ret == 0
        np.testing.assert_array_equal(c_pix, c1)
        for a, b in ((Z, Z1), (U, U1), (sr, sr1), (si, si1)):
            np.testing.assert_array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
        _, _, Zt, Ut, srt, sit = storage.reload_data(f, cs, "t")
        sel = np.ascontiguousarray(np.asarray(Zt)[:, m])
        np.testing.assert_array_equal(np.ascontiguousarray(Z).view(np.uint8), sel.view(np.uint8))
        for a, b in ((U, Ut), (sr, srt), (si, sit)):
            np.testing.assert_array_equal(a, np.asarray(b)[:, m])
        assert np.all(np.asarray(si) % 2 == 0)
        n_sel += int(m.sum())
    assert 0 < n_sel < f.nx * f.ny
    assert storage.calc_to_disk(f, "t2") == 0
    # the whole frame at once with the same subset (calc_raw assembles the
    # mask from the tiles): the full frame's results at the selected pixels
    full = f.calc_raw("t")
    part = f.calc_raw("t2")
    mask = f._image_mask(f._calc_data["t2"]["state"].subset)
    assert part["stop_iter"].shape[1] == n_sel == int(mask.sum())
    np.testing.assert_array_equal(part["stop_iter"], full["stop_iter"][:, mask])
    np.testing.assert_array_equal(np.ascontiguousarray(part["Z"]).view(np.uint8),
                                  np.ascontiguousarray(full["Z"][:, mask]).view(np.uint8))


def test_image_mask_from_tiles(tmp_path):
    """A tile-indexed subset assembled into a whole-frame mask: each tile's
    pixels (chunk_pixel_pos, row-major (y, x)) sit at [iy:iyy, ix:ixx] of
    the frame's own grid."""
    f = _frame(tmp_path)
    full = np.ravel(f.chunk_pixel_pos((0, f.nx, 0, f.ny))).reshape(f.ny, f.nx)
    pattern = np.random.default_rng(1).random((f.ny, f.nx)) < 0.3

    class TileSubset:
        def __getitem__(self, cs):
            ix, ixx, iy, iyy = cs
            return np.ravel(pattern[iy:iyy, ix:ixx])

    for cs in f.chunk_slices():
        # (a tile's own linspace may differ from the frame's in the last bit)
        ix, ixx, iy, iyy = cs
        np.testing.assert_allclose(np.ravel(f.chunk_pixel_pos(cs)), np.ravel(full[iy:iyy, ix:ixx]),
                                   rtol=0, atol=1e-15)
    np.testing.assert_array_equal(f._image_mask(TileSubset()), np.ravel(pattern))
    np.testing.assert_array_equal(f._image_mask(pattern.ravel()), np.ravel(pattern))


def test_reference_orbit_is_stored_and_reused(tmp_path, monkeypatch):
    """The reference orbit is stored in <directory>/data/ref_pt.{json,npy}
    (the reference's ref_pt.dat, perturbation.py:256-290, without pickle) and
    reused by a later calculation at the same location with the same or a
    smaller max_iter and precision (ref_point_matching, perturbation.py:211-253):
    the same orbit, register and parameters, and no orbit is computed.  A
    larger max_iter computes a new one."""
    import fractalshades_amd as fsa
    from fractalshades_amd import models

    def frame(max_iter):
        f = fsa.Perturbation_mandelbrot(str(tmp_path))
        f.zoom(precision=360, x="-1.74920463345912691", y="-0.00028684660237361114", dx="1e-310",
               nx=32, xy_ratio=1.0, theta_deg=0.)
        f.no_newton = True
        f.calc_std_div(calc_name="t", subset=None, max_iter=max_iter, M_divergence=1.e3,
                       epsilon_stationnary=1.e-3, BLA_eps=1.e-6, interior_detect=True)
        return f

    a = frame(3000)
    a.get_FP_orbit()
    fp, Zn = a.FP_params, np.array(a.Zn_path)
    assert os.path.exists(os.path.join(str(tmp_path), "data", "ref_pt.json"))
    calls = []
    orig = models.Perturbation_mandelbrot.FP_loop
    monkeypatch.setattr(models.Perturbation_mandelbrot, "FP_loop",
                        lambda self, *args: calls.append(args) or orig(self, *args))
    b = frame(2000)
    assert b.ref_point_matching()
    b.get_FP_orbit()
    assert calls == []
    np.testing.assert_array_equal(b.Zn_path.view(np.uint64), Zn.view(np.uint64))
    assert b.FP_params["xr"] == fp["xr"] and b.FP_params["div_iter"] == fp["div_iter"]
    assert b.FP_params["ref_point"] == fp["ref_point"]
    assert b.get_path_data()["xr_index"].tolist() == a.get_path_data()["xr_index"].tolist()
    c = frame(4000)          # its calculation computes and stores a new orbit
    assert len(calls) == 1
    assert c.FP_params["max_iter"] == 4000 and c.Zn_path.size == 4001
