"""ctypes wrapper of the C oracle (oracle/fs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The
product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libfs_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


class _Desc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("degree", ctypes.c_int32),
        ("flavor", ctypes.c_int32), ("xr", ctypes.c_int32),
        ("L", ctypes.c_int64), ("div_iter", ctypes.c_int64), ("order", ctypes.c_int64),
        ("Zn", ctypes.c_void_p),
        ("n_xr", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("xr_idx", ctypes.c_void_p), ("xr_m", ctypes.c_void_p), ("xr_e", ctypes.c_void_p),
        ("dx_m", ctypes.c_double), ("dx_e", ctypes.c_int32), ("pad1", ctypes.c_int32),
        ("drift_m", ctypes.c_double * 2), ("drift_e", ctypes.c_int32 * 2),
        ("lin_mat", ctypes.c_double * 4),
        ("kc_m", ctypes.c_double), ("kc_e", ctypes.c_int32), ("bla_on", ctypes.c_int32),
        ("bla_eps", ctypes.c_double),
        ("max_iter", ctypes.c_int64),
        ("M2", ctypes.c_double), ("eps2", ctypes.c_double),
        ("calc_dzndc", ctypes.c_int32), ("calc_dzndz", ctypes.c_int32),
        ("calc_orbit", ctypes.c_int32), ("backshift", ctypes.c_int32),
        ("sd_m", ctypes.c_double), ("sd_e", ctypes.c_int32), ("pad3", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        L.or_ctx_create.argtypes = [ctypes.POINTER(_Desc)]
        L.or_ctx_create.restype = vp
        L.or_ctx_destroy.argtypes = [vp]
        L.or_bla_info.argtypes = [vp, ctypes.POINTER(i32)]
        L.or_bla_info.restype = i64
        L.or_bla_copy.argtypes = [vp, vp, vp]
        L.or_path_copy.argtypes = [vp, ctypes.c_int, vp, vp]
        L.or_cycles.argtypes = [vp, vp, i64, vp, vp, vp, vp, ctypes.c_int]
        L.or_cycles.restype = ctypes.c_int
        for name in ("or_xrc_binop", "or_xrf_binop"):
            getattr(L, name).argtypes = [ctypes.c_int, i64, vp, vp, vp, vp, vp, vp]
        L.or_xrc_to_std.argtypes = [i64, vp, vp, vp]
        L.or_xrc_normalize.argtypes = [i64, vp, vp, vp, vp]
        L.or_hypot.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_hypot.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class OracleCtx:
    """The oracle's counterpart of a GPU context, built from the same
    descriptor fields and arrays (PerturbationFractal.get_ref_desc)."""

    def __init__(self, fields, arrays):
        d = _Desc()
        d.kind, d.degree, d.flavor = fields["kind"], fields["degree"], fields["flavor"]
        d.xr = fields["xr_mode"]
        d.L, d.div_iter, d.order = fields["ref_len"], fields["ref_div_iter"], fields["ref_order"]
        self._arrays = arrays
        d.Zn = _p(arrays["Zn_path"])
        d.n_xr = int(arrays["xr_index"].size)
        d.xr_idx, d.xr_m, d.xr_e = _p(arrays["xr_index"]), _p(arrays["xr_mant"]), _p(arrays["xr_exp"])
        d.dx_m, d.dx_e = fields["dx_mant"], fields["dx_exp"]
        for i in range(2):
            d.drift_m[i] = fields["drift_mant"][i]
            d.drift_e[i] = fields["drift_exp"][i]
        for i in range(4):
            d.lin_mat[i] = fields["lin_mat"][i]
        d.kc_m, d.kc_e = fields["kc_mant"], fields["kc_exp"]
        d.bla_on, d.bla_eps = fields["bla_on"], fields["bla_eps"]
        d.max_iter = fields["max_iter"]
        d.M2 = fields["M_divergence"] * fields["M_divergence"]
        d.eps2 = fields["epsilon_stationnary"] * fields["epsilon_stationnary"]
        d.calc_dzndc, d.calc_dzndz = fields["calc_dzndc"], fields["interior_detect"]
        d.calc_orbit = int(fields.get("calc_orbit", 0))
        d.backshift = int(fields.get("backshift", 0))
        d.sd_m, d.sd_e = float(fields.get("deriv_mant", 0.)), int(fields.get("deriv_exp", 0))
        self.fields = fields
        self.kind = fields["kind"]
        self._h = lib().or_ctx_create(ctypes.byref(d))
        if not self._h:
            raise ValueError("oracle rejected the descriptor")

    def __del__(self):
        try:
            if self._h:
                lib().or_ctx_destroy(self._h)
        except Exception:
            pass

    def n_z(self):
        f = self.fields
        orb = bool(f.get("calc_orbit", 0))
        if self.kind == 2:
            return 2 + (4 if f["calc_dzndc"] else 0) + (2 if orb else 0)
        return 1 + int(bool(f["calc_dzndc"])) + int(bool(f["interior_detect"])) + int(orb)

    def cycles(self, c_pix, nthreads=0):
        c_pix = np.ascontiguousarray(c_pix, dtype=np.complex128)
        n = c_pix.shape[0]
        zdt = np.float64 if self.kind == 2 else np.complex128
        Z = np.zeros((self.n_z(), n), dtype=zdt)
        U = np.zeros((1, n), dtype=np.int32)
        sr = -np.ones((1, n), dtype=np.int8)
        si = np.zeros((1, n), dtype=np.int32)
        lib().or_cycles(self._h, _p(c_pix), n, _p(Z), _p(U), _p(sr), _p(si), int(nthreads))
        return Z, U, sr, si

    def bla(self):
        st = ctypes.c_int32()
        n = lib().or_bla_info(self._h, ctypes.byref(st))
        dim = 8 if self.kind == 2 else 4
        M = np.zeros(n * dim, dtype=np.float64)
        r = np.zeros(n, dtype=np.float64)
        lib().or_bla_copy(self._h, _p(M), _p(r))
        return M, r, st.value

    def path(self, which):
        L = self.fields["ref_len"]
        if which == 0:
            m, e = np.zeros(2 * L), np.zeros(L, dtype=np.int32)
        elif which == 1:
            m, e = np.zeros(2 * (L + 1)), np.zeros(L + 1, dtype=np.int32)
        else:
            m, e = np.zeros(L), np.zeros(L, dtype=np.int32)
        lib().or_path_copy(self._h, which, _p(m), _p(e))
        return m, e


def xrc_binop(op, am, ae, bm, be):
    n = ae.size
    om, oe = np.zeros(2 * n), np.zeros(n, dtype=np.int32)
    lib().or_xrc_binop(op, n, _p(am), _p(ae), _p(bm), _p(be), _p(om), _p(oe))
    return om, oe


def xrf_binop(op, am, ae, bm, be):
    n = ae.size
    om, oe = np.zeros(n), np.zeros(n, dtype=np.int32)
    lib().or_xrf_binop(op, n, _p(am), _p(ae), _p(bm), _p(be), _p(om), _p(oe))
    return om, oe


def xrc_to_std(am, ae):
    n = ae.size
    out = np.zeros(2 * n)
    lib().or_xrc_to_std(n, _p(am), _p(ae), _p(out))
    return out


def xrc_normalize(am, ae):
    n = ae.size
    om, oe = np.zeros(2 * n), np.zeros(n, dtype=np.int32)
    lib().or_xrc_normalize(n, _p(am), _p(ae), _p(om), _p(oe))
    return om, oe


def hypot(x, y):
    return lib().or_hypot(float(x), float(y))
