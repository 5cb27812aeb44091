"""Helpers for GPU tests: run a frame on the engine with the reference's
buffer layout."""
import numpy as np

from fractalshades_amd._native import n_z_fields


def run_gpu(args, c_pix, grid_w=0):
    kind = args.fields["kind"]
    n = c_pix.shape[0]
    nz = n_z_fields(args.fields)
    Z = np.zeros((nz, n), dtype=np.float64 if kind == 2 else np.complex128)
    U = np.zeros((1, n), dtype=np.int32)
    sr = -np.ones((1, n), dtype=np.int8)
    si = np.zeros((1, n), dtype=np.int32)
    rc, stats = args.ctx.tile(np.ascontiguousarray(c_pix), Z, U, sr, si, grid_w=grid_w)
    assert rc == 0
    return (Z, U, sr, si), stats


def assert_same(got, ref):
    Zg, Ug, srg, sig = got
    Zr, Ur, srr, sir = ref
    np.testing.assert_array_equal(sig, sir)
    np.testing.assert_array_equal(srg, srr)
    np.testing.assert_array_equal(Ug, Ur)
    # bit-exact (NaN-aware) on every output field
    bad = Zg.view(np.uint64) != Zr.view(np.uint64)
    if np.any(bad):
        rows, cols = np.nonzero(bad.reshape(Zg.view(np.uint64).shape[0], -1))
        k = cols[:6]
        flat_g = Zg.view(np.float64).reshape(Zg.shape[0], -1)
        flat_r = Zr.view(np.float64).reshape(Zr.shape[0], -1)
        detail = [(int(r), int(c), float(flat_g[r, c]), float(flat_r[r, c]))
                  for r, c in zip(rows[:6], cols[:6])]
        raise AssertionError(f"{int(bad.sum())} of {bad.size} values differ; rows {sorted(set(rows.tolist()))}; "
                             f"(row, float index, got, want): {detail}; stop_iter {sig.reshape(-1)[k // 2 if np.iscomplexobj(Zg) else k].tolist()}")
