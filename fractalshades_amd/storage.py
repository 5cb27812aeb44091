"""Raw results on disk in the reference's layout, so that a reference-style
plotter can reload them and an interrupted calculation can resume.

The reference keeps, per calculation, under ``<directory>/data/``
(src/fractalshades/core.py:2125-2512):

* ``<calc>_Z.arr``, ``<calc>_U.arr``, ``<calc>_stop_reason.arr``,
  ``<calc>_stop_iter.arr``: ``.npy`` memory maps of shape (fields, points),
  dtypes complex128 (float64 for the burning ship) / int32 / int8 / int32,
  the points of each tile (``chunk_slices`` order, each tile row-major as
  ``chunk_pixel_pos`` lays it out) one after the other;
* ``<calc>.report``: a ``.npy`` int32 memory map (tiles, 4) of
  ``chunk1d_begin, chunk1d_end, chunk_pts, done`` (core.py:1266-1271,
  2171-2211);
* ``<calc>.fingerprint``: the pickled dict of ``calc_class``,
  ``calc_callable``, ``calc_kwargs``, ``zoom_kwargs``, ``init_kwargs``
  (core.py:1962-1971, 2130-2134), compared key by key before results are
  reused (core.py:1851-1881).

``calc_to_disk`` computes the tiles whose ``done`` flag is 0 -- all of them
in one launch, with each tile's own pixel grid as the reference computes it
-- and writes them in place.  Calculations with a ``subset`` are not
supported (the reference stores the subset mask per tile as well).  The
reference orbit file ``ref_pt.dat`` is not written: its Xrange register holds
objects of the reference's own Xrange class.
"""
import os
import pickle

import numpy as np
from numpy.lib.format import open_memmap

REPORT_ITEMS = ["chunk1d_begin", "chunk1d_end", "chunk_pts", "done"]   # core.py:1266-1271
SAVE_ARRS = ["Z", "U", "stop_reason", "stop_iter"]                     # core.py:1273-1278


def data_path(f, calc_name):
    """{key: path} of the raw arrays (core.py:2353-2358)."""
    return {key: os.path.join(f.directory, "data", f"{calc_name}_{key}.arr")
            for key in ["subset"] + SAVE_ARRS}


def report_path(f, calc_name):
    return os.path.join(f.directory, "data", calc_name + ".report")


def fingerprint_path(f, calc_name):
    return os.path.join(f.directory, "data", calc_name + ".fingerprint")


def fingerprint(f, calc_name):
    """The calculation's parameters as the reference fingerprints them."""
    cd = f._calc_data[calc_name]
    return {
        "calc_class": cd["calc_class"],
        "calc_callable": cd["calc_callable"],
        "calc_kwargs": dict(cd["calc_kwargs"]),
        "zoom_kwargs": dict(cd["zoom_kwargs"]),
        "init_kwargs": dict(f.init_kwargs),
    }


def _flatten(d, prefix=""):
    out = {}
    for k, v in d.items():
        key = f"{prefix}@{k}" if prefix else str(k)
        if isinstance(v, dict):
            out.update(_flatten(v, key))
        else:
            out[key] = v
    return out


def fingerprint_matching(f, calc_name):
    """True when the stored fingerprint matches the calculation's."""
    try:
        with open(fingerprint_path(f, calc_name), "rb") as fh:
            stored = pickle.load(fh)          # a file this module wrote
    except (OSError, EOFError, pickle.UnpicklingError):
        return False
    got, want = _flatten(stored), _flatten(fingerprint(f, calc_name))
    return all(k in got and _same(got[k], v) for k, v in want.items())


def _same(a, b):
    try:
        return bool(a == b)
    except Exception:
        return False


def _state(f, calc_name):
    state = f._calc_data[calc_name]["state"]
    if getattr(state, "subset", None) is not None:
        raise NotImplementedError("on-disk results of a calculation with a subset")
    return state


def _chunk_pts(f):
    return np.array([(ixx - ix) * (iyy - iy) for ix, ixx, iy, iyy in f.chunk_slices()], dtype=np.int64)


def init_files(f, calc_name):
    """Create the fingerprint, the report (no tile done) and the raw arrays
    (core.py:2171-2211, 2361-2404)."""
    state = _state(f, calc_name)
    os.makedirs(os.path.join(f.directory, "data"), exist_ok=True)
    with open(fingerprint_path(f, calc_name), "wb") as fh:
        pickle.dump(fingerprint(f, calc_name), fh, pickle.HIGHEST_PROTOCOL)
    pts = _chunk_pts(f)
    cum = np.zeros(len(pts) + 1, dtype=np.int64)
    np.cumsum(pts, out=cum[1:])
    rep = open_memmap(report_path(f, calc_name), mode="w+", dtype=np.int32, shape=(len(pts), len(REPORT_ITEMS)))
    rep[:, REPORT_ITEMS.index("chunk1d_begin")] = cum[:-1]
    rep[:, REPORT_ITEMS.index("chunk1d_end")] = cum[1:]
    rep[:, REPORT_ITEMS.index("chunk_pts")] = pts
    rep[:, REPORT_ITEMS.index("done")] = 0
    del rep
    n_pts = int(cum[-1])
    n_Z, n_U = len(state.codes[0]), len(state.codes[1])
    dims = {"Z": ((n_Z, n_pts), state.complex_type), "U": ((n_U, n_pts), f.int_type),
            "stop_reason": ((1, n_pts), f.termination_type), "stop_iter": ((1, n_pts), f.int_type)}
    paths = data_path(f, calc_name)
    for key in SAVE_ARRS:
        shape, dt = dims[key]
        mm = open_memmap(paths[key], mode="w+", dtype=np.dtype(dt), shape=shape)
        del mm


def res_available(f, calc_name, chunk_slice=None):
    """The reference's test (core.py:1883-1915): the fingerprint matches and,
    for a tile, its results are done."""
    if not fingerprint_matching(f, calc_name):
        return False
    try:
        rep = np.load(report_path(f, calc_name), mmap_mode="r")
    except OSError:
        return False
    if chunk_slice is None:
        return True
    return bool(rep[f.chunk_rank(chunk_slice), REPORT_ITEMS.index("done")] > 0)


def calc_to_disk(f, calc_name):
    """Compute the tiles not yet done and store them in the reference's
    layout.  Returns the number of tiles computed."""
    _state(f, calc_name)
    if not res_available(f, calc_name):
        init_files(f, calc_name)
    rep = open_memmap(report_path(f, calc_name), mode="r+")
    todo = [(r, cs) for r, cs in enumerate(f.chunk_slices())
            if rep[r, REPORT_ITEMS.index("done")] == 0]
    if not todo:
        return 0
    # every tile's own pixel grid, in the stored order, in one launch
    c_pix = np.concatenate([np.ravel(f.chunk_pixel_pos(cs)) for _, cs in todo])
    state = f._calc_data[calc_name]["state"]
    args = f._result_arrays(state, c_pix)
    ret = f.numba_cycle_call(args, f._calc_data[calc_name]["cycle_indep_args"])
    if ret != 0:
        raise RuntimeError(f"calculation interrupted ({ret})")
    _, Z, U, stop_reason, stop_iter = args
    arrs = {"Z": Z, "U": U, "stop_reason": stop_reason, "stop_iter": stop_iter}
    paths = data_path(f, calc_name)
    mms = {key: open_memmap(paths[key], mode="r+") for key in SAVE_ARRS}
    o = 0
    for r, _ in todo:
        beg = int(rep[r, REPORT_ITEMS.index("chunk1d_begin")])
        end = int(rep[r, REPORT_ITEMS.index("chunk1d_end")])
        for key in SAVE_ARRS:
            mms[key][:, beg:end] = arrs[key][:, o:o + end - beg]
        o += end - beg
    for mm in mms.values():
        mm.flush()
    del mms
    # tiles are marked done only once their results are on disk
    rep[[r for r, _ in todo], REPORT_ITEMS.index("done")] = 1
    rep.flush()
    del rep
    return len(todo)


def reload_data(f, chunk_slice, calc_name):
    """(subset, c_pix, Z, U, stop_reason, stop_iter) of one stored tile
    (core.py:2474-2512)."""
    _state(f, calc_name)
    rep = np.load(report_path(f, calc_name), mmap_mode="r")
    r = f.chunk_rank(chunk_slice)
    beg = int(rep[r, REPORT_ITEMS.index("chunk1d_begin")])
    end = int(rep[r, REPORT_ITEMS.index("chunk1d_end")])
    paths = data_path(f, calc_name)
    arr = {key: np.load(paths[key], mmap_mode="r")[:, beg:end] for key in SAVE_ARRS}
    c_pix = np.ravel(f.chunk_pixel_pos(chunk_slice))
    return None, c_pix, arr["Z"], arr["U"], arr["stop_reason"], arr["stop_iter"]
