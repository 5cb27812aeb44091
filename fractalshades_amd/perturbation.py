"""PerturbationFractal: host side of the perturbation path.

Mirrors the reference class (src/fractalshades/perturbation.py:22-977):
``zoom`` (43-135), reference-orbit search and caching (get_FP_orbit 651-768,
compute_FP_orbit 808-854, ref_point_matching 211-253), path data
(get_path_data 317-375), ``ref_point_kc`` (166-194) and the per-calculation
preparation ``get_cycle_indep_args`` (431-562).  The per-tile loop
(``numba_cycle_call``, 414-428) runs on the GPU through libfsgpu.so.
"""
import json
import logging
import os

import mpmath
import numpy as np

from . import projection as fsproj
from . import settings
from . import mpmath_loop as fsFP
from ._native import GpuContext, FSG_KIND_BS, FSG_PREC_FP64, FSG_PREC_FP32X
from .core import Fractal
from .utils import zoom_options
from .xrange import mpf_to_xr, mpc_to_xr, xr_to_float

logger = logging.getLogger(__name__)

REF_ORDER_NONE = 1 << 62   # perturbation.py:455-456
PRECISIONS = {"fp64": FSG_PREC_FP64, "fp32x": FSG_PREC_FP32X}


def _without_dir(kw):
    """Model parameters other than the directory (the reference compares
    those only, perturbation.py:219-220)."""
    return {k: v for k, v in kw.items() if k != "directory"}


class CycleIndepArgs:
    """What the per-tile call needs, computed once per calculation: the
    engine descriptor (fields + arrays) and, created on first use, the
    device-resident context built from it."""

    def __init__(self, fields, arrays, interrupted, share=None, projection=None):
        self.fields = fields
        self.arrays = arrays
        self.interrupted = interrupted
        self.last_stats = None
        self.share = share      # GpuContext whose orbit / BLA coefficients to reuse
        self.projection = projection   # non-identity screen projection, applied on the host
        self.resident_key = None       # identity of the orbit, for sharing its device copy
        self._ctx = None

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = GpuContext(self.fields, self.arrays, share=self.share)
        return self._ctx

    def close(self):
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None

    def with_fields(self, **over):
        """A copy with some descriptor fields replaced (its own context)."""
        c = CycleIndepArgs(dict(self.fields, **over), self.arrays, self.interrupted, self.share,
                           self.projection)
        c.resident_key = self.resident_key
        return c

    def shared_with(self, base_ctx):
        """A copy whose context reuses base_ctx's device orbit and BLA
        coefficients (same orbit required)."""
        c = CycleIndepArgs(self.fields, self.arrays, self.interrupted, base_ctx, self.projection)
        c.resident_key = self.resident_key
        return c

    def dzndc_rows(self):
        """Rows of the raw Z array that hold dz/dc (the burning ship's four
        Jacobian rows): zn, [dzndz], dzndc, [zn_orbit] / xn, yn, 4 x J."""
        f = self.fields
        if not f["calc_dzndc"]:
            return []
        if f["kind"] == FSG_KIND_BS:
            return [2, 3, 4, 5]
        return [1 + int(bool(f["interior_detect"]))]


class PerturbationFractal(Fractal):
    """Escape-time fractal z_{n+1} = f(z_n) + c computed by perturbation
    around an arbitrary-precision reference orbit."""

    def __init__(self, directory):
        super().__init__(directory)
        self._FP_params = None
        self._Zn_path = None

    # ---- zoom ----------------------------------------------------------------
    @zoom_options
    def zoom(self, *, precision: int, x, y, dx, nx: int, xy_ratio: float,
             theta_deg: float, projection="cartesian", has_skew: bool = False,
             skew_00: float = 1., skew_01: float = 0., skew_10: float = 0.,
             skew_11: float = 1.):
        """Frame definition; precision in decimal digits (perturbation.py:43-135)."""
        self.prec = mpmath.libmp.dps_to_prec(precision)
        with mpmath.workprec(self.prec):
            self.x = mpmath.mpf(x)
            self.y = mpmath.mpf(y)
            self.dx = dx = mpmath.mpf(dx)
        # a str selects the Cartesian projection, as the reference's
        # backward-compatible path does (perturbation.py:95-104)
        if projection is None or isinstance(projection, str):
            projection = fsproj.Cartesian()
        if not isinstance(projection, fsproj.Projection):
            raise TypeError("projection must be a fractalshades_amd.projection.Projection")
        self.projection = projection
        self._skew = None
        if has_skew:
            self._skew = np.array(((skew_00, skew_01), (skew_10, skew_11)), dtype=np.float64)
        self.dx_std = float(dx)
        self.dx_xr = mpf_to_xr(dx)
        self.lin_mat = self.get_lin_mat()
        projection.adjust_to_zoom(self)
        nx = self.nx
        # precision check (perturbation.py:123-135)
        with mpmath.workdps(6):
            pix = projection.min_local_scale * dx / nx
            required_dps = int(-mpmath.log10(pix / nx) + 1)
        if required_dps > precision:
            raise ValueError(
                "Precision is too low for min. pixel size and shall be increased "
                f"to {required_dps} (current setting: {precision})."
            )

    @property
    def xr_detect_activated(self):
        """Extended range needed below dx = 1e-300 (perturbation.py:152-155)."""
        return bool(self.dx < settings.xrange_zoom_level)

    # ---- reference orbit ------------------------------------------------------
    @property
    def FP_params(self):
        """The reference point's parameters, reloaded from the directory's
        stored orbit when none was computed in this session
        (perturbation.py:292-311)."""
        if self._FP_params is None:
            self.reload_ref_point()
        return self._FP_params

    @property
    def Zn_path(self):
        if self._Zn_path is None:
            self.reload_ref_point()
        return self._Zn_path

    def ref_point_file(self):
        """Base path of the stored reference orbit: <directory>/data/ref_pt
        (.json parameters and Xrange register, .npy orbit).  The reference
        pickles the same content to ref_pt.dat (perturbation.py:256-290); here
        nothing executable is stored or loaded."""
        if self.directory is None:
            return None
        return os.path.join(self.directory, "data", "ref_pt")

    def save_ref_point(self, FP_params, Zn_path):
        """Store the orbit and its parameters for later calculations of the
        same location (save_ref_point, perturbation.py:256-274)."""
        self._FP_params = FP_params
        self._Zn_path = Zn_path
        base = self.ref_point_file()
        if base is None:
            return
        os.makedirs(os.path.dirname(base), exist_ok=True)
        ref = FP_params["ref_point"]
        with mpmath.workprec(FP_params["prec"]):
            ref = mpmath.mpc(ref)
            # exact: signed mantissa and binary exponent of each part
            pt = []
            for v in (ref.real, ref.imag):
                sign, man, exp, _ = v._mpf_
                pt.append((-int(man) if sign else int(man), int(exp)))
        if self.holomorphic:
            xr = [[int(i), float(v[0]), float(v[1]), int(v[2])] for i, v in sorted(FP_params["xr"].items())]
        else:
            xr = [[int(i), float(v[0][0]), int(v[0][1]), float(v[1][0]), int(v[1][1])]
                  for i, v in sorted(FP_params["xr"].items())]
        meta = {k: v for k, v in FP_params.items() if k not in ("ref_point", "xr")}
        meta.update(ref_point=[[int(m), int(e)] for m, e in pt], xr=xr,
                    holomorphic=bool(self.holomorphic))
        np.save(base + ".npy.tmp.npy", np.ascontiguousarray(Zn_path, dtype=np.complex128),
                allow_pickle=False)
        os.replace(base + ".npy.tmp.npy", base + ".npy")
        with open(base + ".json.tmp", "w") as fh:
            json.dump(meta, fh)
        os.replace(base + ".json.tmp", base + ".json")

    def reload_ref_point(self):
        """Reload the stored orbit, if any (reload_ref_point,
        perturbation.py:277-290).  Returns True when one was found."""
        base = self.ref_point_file()
        if base is None or not (os.path.exists(base + ".json") and os.path.exists(base + ".npy")):
            return False
        with open(base + ".json") as fh:
            meta = json.load(fh)
        if meta.pop("holomorphic") != bool(self.holomorphic):
            return False
        (mr, er), (mi, ei) = meta.pop("ref_point")
        with mpmath.workprec(meta["prec"]):
            meta["ref_point"] = mpmath.mpc(mpmath.mpf((mr, er)), mpmath.mpf((mi, ei)))
        if self.holomorphic:
            meta["xr"] = {r[0]: (r[1], r[2], r[3]) for r in meta["xr"]}
        else:
            meta["xr"] = {r[0]: ((r[1], r[2]), (r[3], r[4])) for r in meta["xr"]}
        meta["init_kwargs"] = dict(meta["init_kwargs"])
        self._FP_params = meta
        self._Zn_path = np.load(base + ".npy", allow_pickle=False)
        return True

    def ref_point_matching(self):
        """Reuse the stored orbit when the precision, location, max_iter and
        model match (perturbation.py:211-253)."""
        fp = self.FP_params
        if fp is None:
            return False
        with mpmath.workprec(self.prec):
            drift = (self.x + 1j * self.y) - fp["ref_point"]
            loc_ok = abs(drift / self.dx) ** 2 < 1.e6
        return (self.prec <= fp["prec"] + 10 and loc_ok
                and fp["max_iter"] >= self.max_iter
                and _without_dir(fp["init_kwargs"]) == _without_dir(self.init_kwargs)
                and fp.get("xr") is not None
                and fp["xr_detect"] == self.xr_detect_activated)

    def get_FP_orbit(self, c0=None, newton="cv", order=None, max_newton=None):
        """Find the reference point (ball method + Newton) and compute its orbit
        (perturbation.py:651-768)."""
        if self.ref_point_matching():
            return
        if self.dx > settings.newton_zoom_level:
            self.compute_critical_orbit(self.critical_pt)
            return
        with mpmath.workprec(self.prec):
            self._get_FP_orbit(c0, newton, order, max_newton)

    def _get_FP_orbit(self, c0, newton, order, max_newton):
        if c0 is None:
            c0 = self.x + 1j * self.y
        if settings.no_newton or getattr(self, "no_newton", False) or newton is None or newton == "None":
            self.compute_FP_orbit(c0, None)
            return
        if order is None:
            order = self.ball_method(c0, self.dx * 1.0)
            if order is None:
                logger.warning("Ball method failed - Default to image center")
                self.compute_FP_orbit(c0, None)
                return
        prec = self.prec
        eps_pixel = self.dx * (1. / self.nx)
        try:
            newton_cv, nucleus = self.find_nucleus(c0, order, eps_pixel, prec, max_newton)
            no_div = True
        except NotImplementedError:
            newton_cv, nucleus = self.find_any_nucleus(c0, order, eps_pixel, prec, max_newton)
            no_div = False
        attempt = 1
        max_attempt = 2
        while not newton_cv and attempt <= max_attempt:
            attempt += 1
            prec = int(1.25 * prec)
            # the reference raises the global mpmath precision here, so the
            # orbit below is computed at the boosted precision too
            self.prec = prec
            try:
                newton_cv, nucleus = self.find_nucleus(c0, order, eps_pixel, prec, max_newton)
                no_div = True
            except NotImplementedError:
                newton_cv, nucleus = self.find_any_nucleus(c0, order, eps_pixel, prec, max_newton)
                no_div = False
            if no_div and not newton_cv and attempt == max_attempt:
                newton_cv, nucleus = self.find_any_nucleus(c0, order, eps_pixel, prec, max_newton)
        if not newton_cv:
            logger.warning("Newton descent failed - Default to image center")
            order = None
            nucleus = c0
        self.compute_FP_orbit(nucleus, order)

    def compute_critical_orbit(self, crit):
        """Shallow frames: a constant orbit at the critical point, order 100
        (perturbation.py:770-805)."""
        div_iter = 100
        self.save_ref_point({
            "ref_point": mpmath.mpc(crit), "prec": self.prec, "order": 100,
            "max_iter": self.max_iter, "FP_code": self.FP_code, "xr": {},
            "div_iter": div_iter, "init_kwargs": self.init_kwargs,
            "xr_detect": self.xr_detect_activated,
        }, np.full([div_iter + 1], complex(crit), dtype=np.complex128))

    def compute_FP_orbit(self, ref_point, order=None):
        """Arbitrary-precision orbit, wrapped at the cycle order
        (perturbation.py:808-854)."""
        max_iter = self.max_iter
        ref_orbit_len = max_iter + 1
        if order is not None:
            ref_orbit_len = min(order, ref_orbit_len)
        div_iter, Zn_path, xr = self.FP_loop(ref_orbit_len, ref_point)
        self.save_ref_point({
            "ref_point": mpmath.mpc(ref_point), "prec": self.prec, "order": order,
            "max_iter": max_iter, "FP_code": self.FP_code, "ref_orbit_len": ref_orbit_len,
            "xr": xr, "div_iter": div_iter, "init_kwargs": self.init_kwargs,
            "xr_detect": self.xr_detect_activated,
        }, Zn_path)

    def set_reference(self, ref_point, order, Zn_path, div_iter, xr):
        """Install an orbit computed elsewhere (a cache or another process)."""
        with mpmath.workprec(self.prec):
            ref_point = mpmath.mpc(ref_point)
        self._FP_params = {
            "ref_point": ref_point, "prec": self.prec, "order": order,
            "max_iter": self.max_iter, "FP_code": self.FP_code, "xr": xr,
            "div_iter": div_iter, "init_kwargs": self.init_kwargs,
            "xr_detect": self.xr_detect_activated,
        }
        self._Zn_path = Zn_path

    # ---- path data / kc -------------------------------------------------------
    def ref_point_kc(self):
        """Bound of |c - ref| over the frame corners, x 1.1 (perturbation.py:166-194)."""
        with mpmath.workprec(self.prec):
            w, h = self.projection.bounding_box(self.xy_ratio)
            mat = self.lin_mat
            ref = self.FP_params["ref_point"]
            c0 = self.x + 1j * self.y
            shift = ref - c0
            best = mpmath.mpf(0)
            for sx, sy in ((0.5, 0.5), (-0.5, 0.5), (-0.5, -0.5), (0.5, -0.5)):
                cx = mat[0, 0] * (sx * w) + mat[0, 1] * (sy * h)
                cy = mat[1, 0] * (sx * w) + mat[1, 1] * (sy * h)
                corner = mpmath.mpc(cx, cy) * self.dx
                d = abs(shift - corner)
                if d > best:
                    best = d
            return mpf_to_xr(best * 1.1)

    def get_path_data(self):
        """Orbit, Xrange register as sorted arrays, div_iter, order and the
        centre drift (perturbation.py:317-375)."""
        fp = self.FP_params
        ref_order = fp["order"]
        ref_div_iter = fp["div_iter"]
        if ref_order is not None:
            ref_div_iter = self.max_iter + 1
        idx = np.array(sorted(fp["xr"].keys()), dtype=np.int32)
        with mpmath.workprec(self.prec):
            if self.holomorphic:
                drift = mpc_to_xr((self.x + 1j * self.y) - fp["ref_point"])
                mant = np.zeros(2 * len(idx), dtype=np.float64)
                exp = np.zeros(len(idx), dtype=np.int32)
                for j, i in enumerate(idx):
                    m_re, m_im, e = fp["xr"][int(i)]
                    mant[2 * j], mant[2 * j + 1], exp[j] = m_re, m_im, e
                drift_m = (drift[0], drift[1])
                drift_e = (drift[2], 0)
            else:
                dxr = mpf_to_xr(self.x - fp["ref_point"].real)
                dyr = mpf_to_xr(self.y - fp["ref_point"].imag)
                mant = np.zeros(2 * len(idx), dtype=np.float64)
                exp = np.zeros(2 * len(idx), dtype=np.int32)
                for j, i in enumerate(idx):
                    (xm, xe), (ym, ye) = fp["xr"][int(i)]
                    mant[2 * j], exp[2 * j] = xm, xe
                    mant[2 * j + 1], exp[2 * j + 1] = ym, ye
                drift_m = (dxr[0], dyr[0])
                drift_e = (dxr[1], dyr[1])
        return {
            "Zn_path": np.ascontiguousarray(self.Zn_path, dtype=np.complex128),
            "xr_index": idx, "xr_mant": mant, "xr_exp": exp,
            "ref_div_iter": int(ref_div_iter),
            "ref_order": REF_ORDER_NONE if ref_order is None else int(ref_order),
            "drift_mant": drift_m, "drift_exp": drift_e,
        }

    # ---- engine descriptor ----------------------------------------------------
    def get_ref_desc(self, spec):
        """Descriptor fields + arrays of one calculation (the engine input and
        the oracle input are the same)."""
        pd = self.get_path_data()
        kc_m, kc_e = self.ref_point_kc()
        if xr_to_float(kc_m, kc_e) == 0. and not self.xr_detect_activated:
            raise RuntimeError("Resolution is too low for this zoom depth.")
        fields = {
            "kind": self.fsg_kind, "degree": getattr(self, "exponent", 2),
            "flavor": getattr(self, "flavor_int", 1),
            "xr_mode": int(self.xr_detect_activated),
            "ref_len": int(pd["Zn_path"].shape[0]),
            "ref_div_iter": pd["ref_div_iter"], "ref_order": pd["ref_order"],
            "dx_mant": self.dx_xr[0], "dx_exp": self.dx_xr[1],
            "drift_mant": pd["drift_mant"], "drift_exp": pd["drift_exp"],
            "lin_mat": tuple(float(v) for v in self.lin_mat.ravel()),
            "kc_mant": kc_m, "kc_exp": kc_e,
            "bla_on": int(spec["BLA_activated"]),
            "bla_eps": float(spec["BLA_eps"] or 0.),
            "max_iter": int(spec["max_iter"]),
            "M_divergence": float(spec["M_divergence"]),
            "epsilon_stationnary": float(spec.get("epsilon_stationnary", 0.) or 0.),
            "calc_dzndc": int(spec["calc_dzndc"]),
            "interior_detect": int(spec.get("calc_dzndz", False)),
            "precision": int(spec.get("precision", PRECISIONS[settings.gpu_precision])),
            "calc_orbit": int(bool(spec.get("calc_orbit", False))),
            "backshift": int(spec.get("backshift", 0) or 0) if spec.get("calc_orbit") else 0,
            "deriv_mant": 0., "deriv_exp": 0,
        }
        scale = self.projection.scale
        if scale != 1.:
            # derivative paths at dx * projection.scale (perturbation.py:459-467)
            with mpmath.workprec(self.prec):
                fields["deriv_mant"], fields["deriv_exp"] = mpf_to_xr(self.dx * scale)
        band = getattr(self.projection, "band_scale", 1)
        if band != 1:
            # a deep Expmap band: offsets relative to exp(h_band), which the
            # pixel scale takes over
            with mpmath.workprec(self.prec):
                fields["dx_mant"], fields["dx_exp"] = mpf_to_xr(self.dx * band)
        arrays = {
            "Zn_path": pd["Zn_path"].view(np.float64),
            "xr_index": pd["xr_index"], "xr_mant": pd["xr_mant"], "xr_exp": pd["xr_exp"],
        }
        return fields, arrays

    def get_cycle_indep_args(self, initialize, iterate):
        """Reference orbit, derivative paths and BLA tree resident on the GPU
        (perturbation.py:431-562)."""
        spec = iterate
        self.get_FP_orbit()
        fields, arrays = self.get_ref_desc(spec)
        self.kc = (fields["kc_mant"], fields["kc_exp"])
        proj = None if fsproj.is_cartesian(self.projection) else self.projection
        return CycleIndepArgs(fields, arrays, self._interrupted, projection=proj)

    @staticmethod
    def numba_cycle_call(cycle_dep_args, cycle_indep_args, grid_w=0):
        """One tile on the GPU; returns 0, or 1 when interrupted
        (perturbation.py:414-428).  grid_w > 0: c_pix holds rows of grid_w
        pixels, which the kernel takes in 8x8 blocks (same results)."""
        c_pix, Z, U, stop_reason, stop_iter = cycle_dep_args
        args = cycle_indep_args
        proj = args.projection
        c_in = c_pix if proj is None else np.ascontiguousarray(proj.f(c_pix))
        rc, stats = args.ctx.tile(c_in, Z, U, stop_reason, stop_iter, args.interrupted, grid_w=grid_w)
        args.last_stats = stats
        if proj is not None and proj.dzndc_modifier is not None:
            # perturbation.py:1387-1388, 1773-1777
            mod = proj.dzndc_modifier(c_pix)
            for row in args.dzndc_rows():
                Z[row] *= mod
        return rc

    def expmap_bands(self, calc_name, exp_zoom_step=None):
        """The bands of h of an exponential map (save_expdb_by_steps,
        core.py:891-963): yields (fields, arrays, chunk slices) per band, the
        descriptor with the BLA validity radii and derivative scale of the
        band's own |c| bound (reset_bla_tree, perturbation.py:565-580).  The
        projection is set to the band while the caller holds it, so that its
        f and dzndc_modifier are the band's."""
        proj = self.projection
        if not isinstance(proj, fsproj.Expmap):
            raise ValueError("steps apply to the Expmap projection")
        cd = self._calc_data[calc_name]
        state = cd["state"]
        if getattr(state, "subset", None) is not None:
            raise NotImplementedError("steps with a subset")
        nh = proj.nh(self)
        stp = int(exp_zoom_step or proj.nt(self))
        cs = settings.chunk_size
        try:
            for r in range(0, nh + 1, stp):
                i_max = min(r + stp, nh)
                i_min = max(r - cs, 0)
                proj.set_exp_zoom_step((proj.hmax * i_max + proj.hmin * (nh - i_max)) / nh,
                                       (proj.hmax * i_min + proj.hmin * (nh - i_min)) / nh)
                idx = 3 if proj.premul_1j else 1         # iyy (vertical) or ixx (horizontal)
                chunks = [c for c in self.chunk_slices() if r < c[idx] <= r + stp]
                if not chunks:
                    continue
                fields, arrays = self.get_ref_desc(cd["iterate"])
                yield fields, arrays, chunks
        finally:
            proj.del_exp_zoom_step()

    def calc_raw_by_steps(self, calc_name, exp_zoom_step=None, keep_bands=False):
        """An exponential map computed by bands of h (expmap_bands).  Every
        band shares the device orbit and BLA coefficients of the first (only
        the radii are rebuilt) and runs its tiles in one launch.  Returns
        calc_raw's dict (whole-frame arrays, row-major (y, x)); with
        keep_bands, also "bands": [(CycleIndepArgs, chunk slices)]."""
        proj = self.projection
        cd = self._calc_data[calc_name]
        state = cd["state"]
        base = cd["cycle_indep_args"]
        nx, ny = self.nx, self.ny
        n_Z, n_U = len(state.codes[0]), len(state.codes[1])
        Z = np.zeros((n_Z, ny, nx), dtype=state.complex_type)
        U = np.zeros((n_U, ny, nx), dtype=self.int_type)
        sr = -np.ones((1, ny, nx), dtype=self.termination_type)
        si = np.zeros((1, ny, nx), dtype=self.int_type)
        bands = []
        for fields, arrays, chunks in self.expmap_bands(calc_name, exp_zoom_step):
            args = CycleIndepArgs(fields, arrays, self._interrupted, projection=proj).shared_with(base.ctx)
            args.step_h = (proj.exp_step_hmax, proj.exp_step_hmin)
            c_pix = np.concatenate([np.ravel(self.chunk_pixel_pos(c)) for c in chunks])
            res = self._result_arrays(state, c_pix)
            if self.numba_cycle_call(res, args) != 0:
                raise RuntimeError("calculation interrupted")
            o = 0
            for ix, ixx, iy, iyy in chunks:
                k = (ixx - ix) * (iyy - iy)
                for dst, src in zip((Z, U, sr, si), res[1:]):
                    dst[:, iy:iyy, ix:ixx] = src[:, o:o + k].reshape(-1, iyy - iy, ixx - ix)
                o += k
            if keep_bands:
                bands.append((args, chunks))
            else:
                args.close()
        out = {"c_pix": np.ravel(self.chunk_pixel_pos((0, nx, 0, ny))),
               "Z": Z.reshape(n_Z, -1), "U": U.reshape(n_U, -1), "stop_reason": sr.reshape(1, -1),
               "stop_iter": si.reshape(1, -1), "codes": state.codes}
        if keep_bands:
            out["bands"] = bands
        return out

    def calc_raw(self, calc_name, jitter=False, supersampling=None):
        if getattr(self.projection, "deep", False):
            # a deep exponential map exists only by bands of h
            if jitter or supersampling:
                raise NotImplementedError("deep Expmap with jitter or supersampling")
            return self.calc_raw_by_steps(calc_name)
        args = self._calc_data[calc_name]["cycle_indep_args"]
        res = super().calc_raw(calc_name, jitter, supersampling)
        if res is not None:
            res["stats"] = args.last_stats
        return res

    # ---- period / nucleus helpers (model wrappers fill these in) ---------------
    def ball_method(self, c, px, kind=1, M_divergence=1.e5):
        return self._ball_method(c, px, self.max_iter, M_divergence)
