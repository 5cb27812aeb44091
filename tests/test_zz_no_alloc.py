"""fsg_tile_grid_noalloc (static per-device staging, no allocation) against
fsg_tile_grid, through the C ABI: bit-exact on every output, on two calls in
a row (the second on a smaller tile, so that stale staging from the first
would show), in both arithmetic modes, and an oversize tile rejected with
FSG_E_RANGE before anything runs."""
import numpy as np
import pytest

import frames
from fractalshades_amd import _native
from fractalshades_amd.perturbation import CycleIndepArgs

pytestmark = pytest.mark.gpu


def _bufs(args, n):
    nz = _native.n_z_fields(args.fields)
    zdt = np.float64 if args.fields["kind"] == 2 else np.complex128
    return (np.zeros((nz, n), dtype=zdt), np.zeros((1, n), dtype=np.int32),
            -np.ones((1, n), dtype=np.int8), np.zeros((1, n), dtype=np.int32))


def _run(args, c_pix, grid_w, noalloc):
    Z, U, sr, si = _bufs(args, c_pix.size)
    rc, st = args.ctx.tile(np.ascontiguousarray(c_pix), Z, U, sr, si, grid_w=grid_w, noalloc=noalloc)
    assert rc == 0
    return (Z, U, sr, si), st


def _same(a, b):
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


@pytest.mark.parametrize("case", [("M2_E20", 0), ("M2_int32", 1), ("BS_ex14", 0)])
def test_noalloc_equals_tile_grid(case):
    name, precision = case
    f = frames.make(name, nx=96, max_iter=20000)
    args, c_pix = frames.frame_inputs(f)
    if precision:
        args = CycleIndepArgs(dict(args.fields, precision=precision), args.arrays, None,
                              projection=args.projection)
    print(f"case noalloc {name} precision={precision} n={c_pix.size}", flush=True)
    big = (c_pix, f.nx)
    small = (c_pix[: 37 * 41], 37)
    for cp, w in (big, small):
        want, st_w = _run(args, cp, w, False)
        got, st_g = _run(args, cp, w, True)
        _same(got, want)
        assert st_g["plain_steps"] == st_w["plain_steps"] > 0
    # and once more in index order
    want, _ = _run(args, c_pix[:1000], 0, False)
    got, _ = _run(args, c_pix[:1000], 0, True)
    _same(got, want)


def test_noalloc_rejects_oversize_tile():
    f = frames.make("M2_E20", nx=16)
    args, c_pix = frames.frame_inputs(f)
    # at least complex128 input + one complex field + U, stop_iter, stop_reason: 41 B/px
    n = _native.FSG_NOALLOC_SCRATCH_BYTES // 41 + 1000
    cp = np.zeros(n, dtype=np.complex128)
    Z, U, sr, si = _bufs(args, n)
    with pytest.raises(_native.FsgError) as e:
        args.ctx.tile(cp, Z, U, sr, si, noalloc=True)
    assert "noalloc" in str(e.value)
    assert np.all(sr == -1)
