"""Post-processing formulas on the raw tile results: continuous iteration
count (potential) and distance estimate (DEM).

* continuous iteration: Continuous_iter_pp with kind "infinity"
  (src/fractalshades/postproc.py:352-406, formula 1001-1009):
  nu_frac = -log(log|z_n k| / log(M k)) / log d,  k = |a_d|^(1/(d-1)),
  then the fractional part is folded into (-1, 0] and n adjusted;
* DEM: DEM_pp (postproc.py:684-731): |z_n| ln|z_n| / |dz_n/dc|; for the
  burning ship |dz_n/dc| is the larger singular value of the Jacobian,
  (hypot(a + d, b - c) + hypot(a - d, b + c)) / 2;
* normal map: Attr_normal_pp kind "potential" with potential "infinity"
  (postproc.py:572-628): z_n / (dz_n/dc), for the burning ship J^T z_n,
  unskewed with the transposed skew (core.py:3147-3158), divided by its
  modulus.

Computed in float64 (the reference's default post-processing dtype is
float32, settings.py:82; the parity checks need float64).  Also: the error
report of the binary32 form against the binary64 drop-in, and the
descriptor of the same two layers computed on the GPU (fsg_post_desc).
"""
import numpy as np

from .storage import Fractal_array   # postproc.py:829-960  # noqa: F401


def continuous_iter(zn, stop_iter, d=2, a_d=1., M=1000., floor_iter=0):
    """Continuous iteration number nu for every pixel."""
    zn = np.asarray(zn, dtype=np.complex128)
    n = np.asarray(stop_iter).reshape(-1).astype(np.int64)
    k = np.abs(a_d) ** (1. / (d - 1.))
    with np.errstate(divide="ignore", invalid="ignore"):
        nu_frac = -(np.log(np.log(np.abs(zn * k)) / np.log(M * k)) / np.log(d))
    nu_div, nu_mod = np.divmod(-nu_frac, 1.)
    nu_frac = -nu_mod
    with np.errstate(invalid="ignore"):
        n = n - nu_div.astype(np.int64)
    return (n - floor_iter) + nu_frac


def dem(zn, dzndc=None, jac=None):
    """Distance estimate |z| ln|z| / |dz/dc| (holomorphic: dzndc; burning
    ship: jac = (dXdA, dXdB, dYdA, dYdB))."""
    abs_zn = np.abs(np.asarray(zn))
    if jac is not None:
        a, b, c, d = (np.asarray(v, dtype=np.float64) for v in jac)
        Q = np.hypot(a + d, b - c)
        R = np.hypot(a - d, b + c)
        abs_d = 0.5 * (Q + R)
    else:
        abs_d = np.abs(np.asarray(dzndc))
    with np.errstate(divide="ignore", invalid="ignore"):
        return abs_zn * np.log(abs_zn) / abs_d


def normal_map(zn, dzndc=None, jac=None, skew=None):
    """Unit normal of the potential map, as a complex number."""
    zn = np.asarray(zn, dtype=np.complex128)
    if jac is not None:
        dXdA, dXdB, dYdA, dYdB = (np.asarray(v, dtype=np.float64) for v in jac)
        normal = (dXdA * zn.real + dYdA * zn.imag) + 1j * (dXdB * zn.real + dYdB * zn.imag)
    else:
        normal = zn / np.asarray(dzndc, dtype=np.complex128)
    if skew is not None:
        nx, ny = normal.real, normal.imag
        normal = (skew[0, 0] * nx + skew[1, 0] * ny) + 1j * (skew[0, 1] * nx + skew[1, 1] * ny)
    return normal / np.abs(normal)


def screen_derivatives(projection, c_pix, dzndc=None, jac=None):
    """The derivatives as the layers use them (get_dzndc, postproc.py:184-206):
    dz/dc times the projection's df at the pixel (apply_df, postproc.py:
    972-978), or the burning-ship Jacobian times its dfBS matrix
    (apply_dfBS, postproc.py:980-998: dX/dpix_a = dX/dA m00 + dX/dB m10,
    dX/dpix_b = dX/dA m01 + dX/dB m11, the same for Y).  The identity for a
    Cartesian frame."""
    if projection is None or getattr(projection, "df_kind", 0) == 0:
        return dzndc if jac is None else jac
    if jac is None:
        return projection.df(c_pix) * np.asarray(dzndc, dtype=np.complex128)
    m00, m01, m10, m11 = projection.dfBS(c_pix)
    dXdA, dXdB, dYdA, dYdB = (np.asarray(v, dtype=np.float64) for v in jac)
    return (dXdA * m00 + dXdB * m10, dXdA * m01 + dXdB * m11,
            dYdA * m00 + dYdB * m10, dYdA * m01 + dYdB * m11)


def fields_of(Z, codes, holomorphic=True):
    """Split a raw Z array into named fields, following the codes."""
    names = codes[0]
    if holomorphic:
        return {k: Z[i] for i, k in enumerate(names)}
    out = {k: Z[i] for i, k in enumerate(names)}
    out["zn"] = Z[0] + 1j * Z[1]
    return out


def error_report(test, ref, holomorphic=True, d=2, M=1000.):
    """Per-output error of a result against a reference result of the same
    frame.  Each argument is (Z, U, stop_reason, stop_iter) in the raw
    layout.  Returns counts, mismatch fractions and relative errors of the
    continuous iteration (nu) and of the DEM on the escaped pixels."""
    Zt, _, srt, sit = test
    Zr, _, srr, sir = ref
    sit, sir = sit.reshape(-1), sir.reshape(-1)
    srt, srr = srt.reshape(-1), srr.reshape(-1)
    n = sit.size
    iter_mis = sit != sir
    reason_mis = srt != srr
    esc = (srr == 1) & (srt == 1) & ~iter_mis
    if holomorphic:
        zt, zr = Zt[0], Zr[0]
    else:
        zt, zr = Zt[0] + 1j * Zt[1], Zr[0] + 1j * Zr[1]
    out = {
        "pixels": int(n),
        "escape_count_mismatch": int(np.count_nonzero(iter_mis)),
        "escape_count_mismatch_frac": float(np.count_nonzero(iter_mis) / max(n, 1)),
        "escape_count_mismatch_pm1_frac": float(np.count_nonzero(np.abs(sit.astype(np.int64) - sir) > 1) / max(n, 1)),
        "stop_reason_mismatch": int(np.count_nonzero(reason_mis)),
        "non_finite_zn": int(np.count_nonzero(~np.isfinite(zt))),
        "unfinished": int(np.count_nonzero(srt < 0)),
    }
    if np.any(esc):
        nut = continuous_iter(zt[esc], sit[esc], d=d, M=M)
        nur = continuous_iter(zr[esc], sir[esc], d=d, M=M)
        rel = np.abs(nut - nur) / np.abs(nur)
        out["nu_rel_err_median"] = float(np.median(rel))
        out["nu_rel_err_p99"] = float(np.quantile(rel, 0.99))
        out["nu_rel_err_max"] = float(rel.max())
        if holomorphic and Zr.shape[0] > 1:
            dt = dem(zt[esc], Zt[-1][esc])
            dr = dem(zr[esc], Zr[-1][esc])
            rel = np.abs(dt - dr) / np.abs(dr)
            out["dem_rel_err_median"] = float(np.median(rel))
            out["dem_rel_err_p99"] = float(np.quantile(rel, 0.99))
            out["dem_rel_err_max"] = float(rel.max())
    return out


def post_desc(f, calc_name, floor_iter=0., px_snap=None):
    """fsg_post_desc of a calculation: potential parameters of the model
    (potential_d, potential_a_d, potential_M, as Continuous_iter_pp reads
    them, postproc.py:320-347), the row of dz/dc (dX/da for the burning
    ship) in its raw Z array, the frame's skew for the normal map and the
    projection's map of the derivatives (screen_derivatives), as it stands
    (whole map, or the current band of a map computed by bands)."""
    from ._native import PostDesc
    state = f._calc_data[calc_name]["state"]
    names = list(state.codes[0])
    holo = bool(getattr(f, "holomorphic", True))
    row = names.index("dzndc") if holo and "dzndc" in names else (
        names.index("dxnda") if (not holo and "dxnda" in names) else -1)
    d = float(getattr(f, "potential_d", 2))
    a_d = float(getattr(f, "potential_a_d", 1.))
    M = float(getattr(state, "potential_M", getattr(f, "potential_M_cutoff", 1000.)))
    pp = PostDesc(int(holo), int(row), d, a_d, M, float(floor_iter), float(px_snap or 0.))
    skew = getattr(f, "_skew", None)
    sk = np.ravel(skew) if skew is not None else (1., 0., 0., 1.)
    for i in range(4):
        pp.skew[i] = float(sk[i])
    proj = getattr(f, "projection", None)
    kind = int(getattr(proj, "df_kind", 0)) if proj is not None and not isinstance(proj, str) else 0
    pp.proj = kind
    if kind in (1, 2):
        # the whole map's offsets are exp(h_moy + t): exp(t) = w exp(-h_moy)
        pp.w_scale = float(np.exp(-proj.hmoy))
    mod, mod_scale = derivative_modifier(proj)
    pp.mod, pp.mod_scale = mod, mod_scale
    return pp


def derivative_modifier(proj):
    """The loop's dzndc_modifier (projection.py:205-217, 454-471) as a
    function of the pixel's projected offset w, for the device layers:
    (0, 0) none; (1, s): |w| s, the exponential map (w = exp(h_moy + t) over
    the whole map or by bands of a map within binary64, where the modifier is
    exp(Re t + h_moy - h_band); w = exp(h_moy - h_band + t) by bands of a
    deep map); (2, seam): |w + 1e-6| seam, the Cartesian expmap seam."""
    from .projection import Cartesian, Expmap
    if isinstance(proj, Expmap):
        if proj.use_step and not proj.deep:
            return 1, float(np.exp(-proj.exp_step_hmoy))
        return 1, 1.
    if isinstance(proj, Cartesian) and proj.expmap_seam is not None:
        return 2, float(proj.expmap_seam)
    return 0, 0.
