"""ctypes binding of libfsgpu.so (declarations in include/fsgpu.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded, every entry point raises ``FsgError`` -- there is no CPU
fallback.
"""
import ctypes
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfsgpu.so")

FSG_KIND_M2, FSG_KIND_MN, FSG_KIND_BS = 0, 1, 2
FSG_PREC_FP64, FSG_PREC_FP32X = 0, 1
FSG_INTERRUPTED = 1

# entry points declared in include/fsgpu.h
EXPORTS = (
    "fsg_abi_version", "fsg_ctx_create", "fsg_ctx_create_shared", "fsg_ctx_destroy", "fsg_tile",
    "fsg_tile_device", "fsg_tile_grid", "fsg_tile_grid_device", "fsg_work_bytes", "fsg_work_stats",
    "fsg_ctx_read_table", "fsg_ctx_bla_shape", "fsg_last_error",
    "fsg_server_create", "fsg_server_destroy", "fsg_server_footprint", "fsg_server_clock",
    "fsg_server_submit", "fsg_server_wait", "fsg_host_alloc", "fsg_host_free",
    "fsg_postproc_device", "fsg_server_submit_layers", "fsg_ref_orbit",
    "fsg_ball_period", "fsg_newton_eval", "fsg_tile_grid_noalloc",
)
# device staging of fsg_tile_grid_noalloc (FSG_NOALLOC_SCRATCH_BYTES)
FSG_NOALLOC_SCRATCH_BYTES = 32 << 20
FSG_E_INVALID, FSG_E_HIP, FSG_E_NOMEM, FSG_E_UNSUPPORTED, FSG_E_RANGE = -1, -2, -3, -4, -5
FSG_PENDING = 2


def n_z_fields(fields):
    """Rows of the Z result array of a descriptor (the reference's
    complex_codes): holomorphic zn, [dzndz], [dzndc], [zn_orbit]; burning
    ship xn, yn, [4 Jacobian], [xn_orbit, yn_orbit]."""
    orb = int(bool(fields.get("calc_orbit", 0)))
    if fields["kind"] == FSG_KIND_BS:
        return 2 + (4 if fields["calc_dzndc"] else 0) + 2 * orb
    return 1 + int(bool(fields["calc_dzndc"])) + int(bool(fields["interior_detect"])) + orb


class FsgError(RuntimeError):
    """Error raised by the native engine (or its absence)."""


class RefDesc(ctypes.Structure):
    """fsg_ref_desc (include/fsgpu.h)."""
    _fields_ = [
        ("kind", ctypes.c_int32), ("degree", ctypes.c_int32),
        ("flavor", ctypes.c_int32), ("xr_mode", ctypes.c_int32),
        ("ref_len", ctypes.c_int64), ("ref_div_iter", ctypes.c_int64),
        ("ref_order", ctypes.c_int64),
        ("Zn_path", ctypes.c_void_p),
        ("n_xr", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("xr_index", ctypes.c_void_p), ("xr_mant", ctypes.c_void_p),
        ("xr_exp", ctypes.c_void_p),
        ("dx_mant", ctypes.c_double), ("dx_exp", ctypes.c_int32), ("pad1", ctypes.c_int32),
        ("drift_mant", ctypes.c_double * 2), ("drift_exp", ctypes.c_int32 * 2),
        ("lin_mat", ctypes.c_double * 4),
        ("kc_mant", ctypes.c_double), ("kc_exp", ctypes.c_int32), ("bla_on", ctypes.c_int32),
        ("bla_eps", ctypes.c_double),
        ("max_iter", ctypes.c_int64),
        ("M_divergence", ctypes.c_double), ("epsilon_stationnary", ctypes.c_double),
        ("calc_dzndc", ctypes.c_int32), ("interior_detect", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("calc_orbit", ctypes.c_int32),
        ("backshift", ctypes.c_int32), ("pad2", ctypes.c_int32),
        ("deriv_mant", ctypes.c_double), ("deriv_exp", ctypes.c_int32), ("pad3", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    """fsg_stats (include/fsgpu.h)."""
    _fields_ = [
        ("plain_steps", ctypes.c_uint64), ("bla_steps", ctypes.c_uint64),
        ("bla_iters", ctypes.c_uint64), ("rebases", ctypes.c_uint64),
        ("bla_lookups", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
        ("redo_pixels", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PostDesc(ctypes.Structure):
    """fsg_post_desc (include/fsgpu.h)."""
    _fields_ = [
        ("holomorphic", ctypes.c_int32), ("dzndc_row", ctypes.c_int32),
        ("d", ctypes.c_double), ("a_d", ctypes.c_double), ("M", ctypes.c_double),
        ("floor_iter", ctypes.c_double), ("px_snap", ctypes.c_double),
        ("skew", ctypes.c_double * 4),
        ("proj", ctypes.c_int32), ("mod", ctypes.c_int32), ("w_scale", ctypes.c_double),
        ("mod_scale", ctypes.c_double),
    ]


class ServerOpts(ctypes.Structure):
    """fsg_server_opts (include/fsgpu.h)."""
    _fields_ = [("lanes", ctypes.c_int32), ("chunk_px", ctypes.c_int32), ("hbm_budget", ctypes.c_int64)]


class RequestInfo(ctypes.Structure):
    """fsg_request_info (include/fsgpu.h)."""
    _fields_ = [
        ("status", ctypes.c_int32), ("chunks", ctypes.c_int32), ("footprint", ctypes.c_int64),
        ("t_submit_ms", ctypes.c_double), ("t_admit_ms", ctypes.c_double),
        ("t_first_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_lock = threading.Lock()


class OrbitDesc(ctypes.Structure):
    """fsg_orbit_desc (include/fsgpu.h)."""
    _fields_ = [
        ("kind", ctypes.c_int32), ("degree", ctypes.c_int32),
        ("flavor", ctypes.c_int32), ("need_xr", ctypes.c_int32),
        ("n_limbs", ctypes.c_int32), ("frac_limbs", ctypes.c_int32),
        ("c_re_neg", ctypes.c_int32), ("c_im_neg", ctypes.c_int32),
        ("c_re", ctypes.c_void_p), ("c_im", ctypes.c_void_p),
        ("max_iter", ctypes.c_int64), ("M", ctypes.c_double),
        ("n_threads", ctypes.c_int32), ("pad0", ctypes.c_int32),
    ]


class OrbitOut(ctypes.Structure):
    """fsg_orbit_out (include/fsgpu.h)."""
    _fields_ = [
        ("Zn", ctypes.c_void_p), ("xr_cap", ctypes.c_int64),
        ("xr_index", ctypes.c_void_p), ("xr_mant", ctypes.c_void_p),
        ("xr_exp", ctypes.c_void_p),
        ("n_xr", ctypes.c_int64), ("div_iter", ctypes.c_int64),
    ]


class MpValues(ctypes.Structure):
    """fsg_mp_values (include/fsgpu.h)."""
    _fields_ = [
        ("cap", ctypes.c_int64), ("limbs", ctypes.c_void_p), ("neg", ctypes.c_void_p),
        ("exp", ctypes.c_void_p), ("n", ctypes.c_int64),
    ]


def load(path=None):
    """Load libfsgpu.so once; raise FsgError when it is not built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise FsgError(
                f"{p} is missing: build it with `python -m fractalshades_amd.build`"
            )
        try:
            lib = ctypes.CDLL(p)
        except OSError as exc:
            raise FsgError(f"cannot load {p}: {exc}") from exc
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        sig = {
            "fsg_abi_version": ([], ctypes.c_int),
            "fsg_ctx_create": ([ctypes.POINTER(RefDesc), ctypes.POINTER(vp)], ctypes.c_int),
            "fsg_ctx_create_shared": ([ctypes.POINTER(RefDesc), vp, ctypes.POINTER(vp)], ctypes.c_int),
            "fsg_ctx_destroy": ([vp], ctypes.c_int),
            "fsg_tile": ([vp, vp, i64, vp, vp, vp, vp, vp, ctypes.POINTER(Stats)], ctypes.c_int),
            "fsg_tile_device": ([vp, vp, i64, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "fsg_tile_grid": ([vp, vp, i64, i64, vp, vp, vp, vp, vp, ctypes.POINTER(Stats)], ctypes.c_int),
            "fsg_tile_grid_device": ([vp, vp, i64, i64, vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "fsg_tile_grid_noalloc": ([vp, vp, i64, i64, vp, vp, vp, vp, vp, ctypes.POINTER(Stats)],
                                      ctypes.c_int),
            "fsg_work_bytes": ([], ctypes.c_size_t),
            "fsg_work_stats": ([vp, ctypes.POINTER(Stats)], ctypes.c_int),
            "fsg_ctx_read_table": ([vp, ctypes.c_int, vp, ctypes.c_size_t], ctypes.c_int),
            "fsg_ctx_bla_shape": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i32)], ctypes.c_int),
            "fsg_last_error": ([], ctypes.c_char_p),
            "fsg_server_create": ([ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)], ctypes.c_int),
            "fsg_server_destroy": ([vp], ctypes.c_int),
            "fsg_server_footprint": ([vp, i64, ctypes.POINTER(i64)], ctypes.c_int),
            "fsg_server_clock": ([vp, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            "fsg_server_submit": ([vp, vp, vp, i64, i64, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)],
                                  ctypes.c_int),
            "fsg_server_wait": ([vp, ctypes.c_uint64, i32, ctypes.POINTER(RequestInfo)], ctypes.c_int),
            "fsg_postproc_device": ([ctypes.POINTER(PostDesc), vp, i64, i64, vp, vp, vp, vp, vp, vp, vp],
                                    ctypes.c_int),
            "fsg_server_submit_layers": ([vp, vp, vp, i64, i64, ctypes.POINTER(PostDesc), vp, vp, vp, vp, vp,
                                          ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "fsg_host_alloc": ([ctypes.c_size_t], vp),
            "fsg_host_free": ([vp], ctypes.c_int),
            "fsg_ref_orbit": ([ctypes.POINTER(OrbitDesc), ctypes.POINTER(OrbitOut)], ctypes.c_int),
            "fsg_ball_period": ([ctypes.POINTER(OrbitDesc), ctypes.c_double, ctypes.c_int64,
                                 ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
            "fsg_newton_eval": ([ctypes.POINTER(OrbitDesc), ctypes.c_int64, ctypes.c_int32,
                                 ctypes.POINTER(MpValues)], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if lib.fsg_abi_version() != 6:
            raise FsgError("libfsgpu.so ABI version mismatch")
        if path is None:
            _lib = lib
        return lib


def check(rc):
    """Map a negative return code to FsgError."""
    if rc < 0:
        msg = load().fsg_last_error()
        raise FsgError(f"fsgpu error {rc}: {msg.decode() if msg else ''}")
    return rc


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class GpuContext:
    """A device-resident reference (fsg_ctx): orbit, derivative paths, BLA."""

    def __init__(self, desc_fields, arrays, share=None):
        """share: a GpuContext on the same orbit whose device orbit and BLA
        coefficients this one reuses (fsg_ctx_create_shared)."""
        lib = load()
        # keep the input arrays alive until fsg_ctx_create has copied them
        self._arrays = arrays
        desc = RefDesc()
        for k, v in desc_fields.items():
            if k in ("drift_mant", "drift_exp", "lin_mat"):
                arr = getattr(desc, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(desc, k, v)
        desc.Zn_path = _ptr(arrays["Zn_path"])
        desc.xr_index = _ptr(arrays["xr_index"])
        desc.xr_mant = _ptr(arrays["xr_mant"])
        desc.xr_exp = _ptr(arrays["xr_exp"])
        desc.n_xr = int(arrays["xr_index"].size)
        self.desc = desc
        handle = ctypes.c_void_p()
        if share is None:
            check(lib.fsg_ctx_create(ctypes.byref(desc), ctypes.byref(handle)))
        else:
            check(lib.fsg_ctx_create_shared(ctypes.byref(desc), share.handle, ctypes.byref(handle)))
        self._base = share            # the base outlives this context
        self._h = handle
        self.kind = desc.kind

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load().fsg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tile(self, c_pix, Z, U, stop_reason, stop_iter, interrupted=None, grid_w=0, noalloc=False):
        """Run one tile in place; returns (rc, stats dict).  grid_w > 0: c_pix
        holds rows of grid_w pixels, computed in 8x8 blocks (same results).
        noalloc: through fsg_tile_grid_noalloc (static device staging, no
        allocation; at most FSG_NOALLOC_SCRATCH_BYTES of staging)."""
        for name, arr in (("c_pix", c_pix), ("Z", Z), ("U", U),
                          ("stop_reason", stop_reason), ("stop_iter", stop_iter)):
            if not arr.flags.c_contiguous:
                raise ValueError(f"{name} must be C-contiguous")
        n = c_pix.shape[0]
        if c_pix.dtype != np.complex128:
            raise TypeError("c_pix must be complex128")
        zdt = np.float64 if self.kind == FSG_KIND_BS else np.complex128
        if Z.dtype != zdt or Z.shape[-1] != n:
            raise TypeError("Z has the wrong dtype or shape")
        if U.dtype != np.int32 or stop_iter.dtype != np.int32 or stop_reason.dtype != np.int8:
            raise TypeError("U / stop_iter must be int32, stop_reason int8")
        st = Stats()
        intr = None
        if interrupted is not None:
            intr = interrupted.ctypes.data
        entry = load().fsg_tile_grid_noalloc if noalloc else load().fsg_tile_grid
        rc = check(entry(
            self._h, _ptr(c_pix), n, int(grid_w), _ptr(Z), _ptr(U), _ptr(stop_reason),
            _ptr(stop_iter), intr, ctypes.byref(st)))
        return rc, st.as_dict()

    def tile_device(self, c_pix_ptr, n, Z_ptr, U_ptr, sr_ptr, si_ptr, work_ptr, stream_ptr, grid_w=0):
        """Enqueue one tile on device buffers (raw pointers); asynchronous."""
        return check(load().fsg_tile_grid_device(
            self._h, c_pix_ptr, n, int(grid_w), Z_ptr, U_ptr, sr_ptr, si_ptr, work_ptr, stream_ptr))

    def bla_shape(self):
        bl = ctypes.c_int64()
        st = ctypes.c_int32()
        check(load().fsg_ctx_bla_shape(self._h, ctypes.byref(bl), ctypes.byref(st)))
        return bl.value, st.value

    def read_table(self, which, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        check(load().fsg_ctx_read_table(self._h, which, out.ctypes.data, out.nbytes))
        return out


def work_bytes():
    return load().fsg_work_bytes()


def work_stats(work_ptr):
    st = Stats()
    check(load().fsg_work_stats(work_ptr, ctypes.byref(st)))
    return st.as_dict()
