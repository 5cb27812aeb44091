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
-- and writes them in place.

A calculation with a ``subset`` (usually a ``Fractal_array`` over another
calculation's stored results, postproc.py:829-960) stores only the pixels
of the mask: the points of a tile and the arrays' second dimension count
the selected pixels, and ``<calc>_subset.arr`` holds the mask itself, a
bool memmap of nx * ny entries, tile after tile (core.py:1745-1762,
2404-2424, 2490-2507).

The reference orbit file ``ref_pt.dat`` is not written: its Xrange register
holds objects of the reference's own Xrange class.
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
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return isinstance(a, np.ndarray) and isinstance(b, np.ndarray) and np.array_equal(a, b)
    try:
        return bool(a == b)
    except Exception:
        return False


class Fractal_array:
    """One stored field of a calculation, tile by tile (postproc.py:829-960):
    ``arr[chunk_slice]`` is the field over the tile's stored points (all of
    them with ``chunk_slice`` None), with ``func`` applied (a callable or an
    expression in ``x``, e.g. ``"x != 1"``) and ``~`` inverting a boolean
    result.  Passed as the ``subset`` of another calculation, it selects
    the pixels that calculation computes."""

    def __init__(self, fractal, calc_name, key, func=None, inv=False):
        self.fractal = fractal
        self.calc_name = calc_name
        self.key = key
        self._func = func
        self.func = self.parse_func(func)
        self.inv = inv

    @staticmethod
    def parse_func(func):
        if isinstance(func, str):
            code = compile(func, "<Fractal_array func>", "eval")
            return lambda x: eval(code, {"np": np, "__builtins__": {}}, {"x": x})
        return func

    def __reduce__(self):
        # the fractal is not part of the pickled state (it holds device
        # contexts); an unpickled array compares equal and is rebound with
        # bind_fractal
        return (self.__class__, (None, self.calc_name, self.key, self._func, self.inv))

    def bind_fractal(self, fractal):
        self.fractal = fractal

    def __eq__(self, other):
        return (isinstance(other, Fractal_array) and self._func == other._func
                and self.calc_name == other.calc_name and self.key == other.key
                and self.inv == other.inv)

    def __invert__(self):
        return Fractal_array(self.fractal, self.calc_name, self.key, self._func, not self.inv)

    def _field(self):
        codes = self.fractal._calc_data[self.calc_name]["state"].codes
        complex_codes, int_codes, _ = codes
        if self.key in complex_codes:
            return "Z", complex_codes.index(self.key)
        if self.key in int_codes:
            return "U", int_codes.index(self.key)
        if self.key in ("stop_reason", "stop_iter"):
            return self.key, 0
        raise ValueError(f"unknown key {self.key!r} for {self.calc_name}")

    def __getitem__(self, chunk_slice):
        f = self.fractal
        if f is None:
            raise ValueError("Fractal_array without a fractal: use bind_fractal")
        arr_key, row = self._field()
        mm = np.load(data_path(f, self.calc_name)[arr_key], mmap_mode="r")
        if chunk_slice is None:
            arr = np.array(mm[row, :])
        else:
            rep = np.load(report_path(f, self.calc_name), mmap_mode="r")
            r = f.chunk_rank(chunk_slice)
            beg = int(rep[r, REPORT_ITEMS.index("chunk1d_begin")])
            end = int(rep[r, REPORT_ITEMS.index("chunk1d_end")])
            arr = np.array(mm[row, beg:end])
        if self.func is not None:
            arr = self.func(arr)
        if self.inv:
            arr = ~arr
        return arr


def _state(f, calc_name):
    return f._calc_data[calc_name]["state"]


def _mask(subset, cs):
    return np.ravel(np.asarray(subset[cs], dtype=bool))


def _full_pts(f):
    """Pixels of each tile, in chunk order (the uncompressed layout)."""
    return np.array([(ixx - ix) * (iyy - iy) for ix, ixx, iy, iyy in f.chunk_slices()], dtype=np.int64)


def _chunk_pts(f, state=None):
    """Stored points of each tile: the selected pixels under a subset
    (core.py:1745-1762)."""
    subset = getattr(state, "subset", None)
    if subset is None:
        return _full_pts(f)
    return np.array([int(np.count_nonzero(_mask(subset, cs))) for cs in f.chunk_slices()], dtype=np.int64)


def init_files(f, calc_name):
    """Create the fingerprint, the report (no tile done) and the raw arrays
    (core.py:2171-2211, 2361-2404)."""
    state = _state(f, calc_name)
    os.makedirs(os.path.join(f.directory, "data"), exist_ok=True)
    with open(fingerprint_path(f, calc_name), "wb") as fh:
        pickle.dump(fingerprint(f, calc_name), fh, pickle.HIGHEST_PROTOCOL)
    pts = _chunk_pts(f, state)
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
    subset = getattr(state, "subset", None)
    if subset is not None:
        # the mask, fixed size whatever it selects (core.py:2404-2424)
        full = _full_pts(f)
        ucum = np.concatenate([[0], np.cumsum(full)])
        mm = open_memmap(paths["subset"], mode="w+", dtype=bool, shape=(int(ucum[-1]),))
        for r, cs in enumerate(f.chunk_slices()):
            mm[ucum[r]:ucum[r + 1]] = _mask(subset, cs)
        mm.flush()
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
    # every tile's own pixel grid (its selected pixels under a subset), in
    # the stored order, in one launch
    state = f._calc_data[calc_name]["state"]
    subset = getattr(state, "subset", None)
    parts = []
    for _, cs in todo:
        c = np.ravel(f.chunk_pixel_pos(cs))
        parts.append(c if subset is None else c[_mask(subset, cs)])
    c_pix = np.concatenate(parts)
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
    (core.py:2474-2512); subset is the tile's mask, or None."""
    state = _state(f, calc_name)
    rep = np.load(report_path(f, calc_name), mmap_mode="r")
    r = f.chunk_rank(chunk_slice)
    beg = int(rep[r, REPORT_ITEMS.index("chunk1d_begin")])
    end = int(rep[r, REPORT_ITEMS.index("chunk1d_end")])
    paths = data_path(f, calc_name)
    arr = {key: np.load(paths[key], mmap_mode="r")[:, beg:end] for key in SAVE_ARRS}
    c_pix = np.ravel(f.chunk_pixel_pos(chunk_slice))
    mask = None
    if getattr(state, "subset", None) is not None:
        full = _full_pts(f)
        ubeg = int(full[:r].sum())
        mask = np.array(np.load(paths["subset"], mmap_mode="r")[ubeg:ubeg + int(full[r])])
        c_pix = c_pix[mask]
    return mask, c_pix, arr["Z"], arr["U"], arr["stop_reason"], arr["stop_iter"]
