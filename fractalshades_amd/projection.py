"""Projections of the screen pixels before iteration
(src/fractalshades/projection.py): the Cartesian identity and the
exponential map.

A projection maps a screen pixel z_pix (x in [-0.5, 0.5], y scaled by
1 / xy_ratio) to f(z_pix); the perturbation loop then takes
c = dx * lin_mat . f(z_pix) + drift (ref_path_c_from_pix,
perturbation.py:2214-2230).  The derivative paths are scaled by
dx * projection.scale (perturbation.py:459-467) and, after the loop, dz/dc
is multiplied by the projection's ``dzndc_modifier`` of the pixel
(perturbation.py:1387-1388, 1773-1777).

Here f is applied to the pixel offsets on the host with numpy, before they
go to the GPU, and the modifier to the results after; the kernels only ever
see the projected offsets, so the GPU and the oracle consume the same
inputs.  (The reference evaluates f inside its numba loop: numba's complex
``exp`` and numpy's may differ in the last bit of some pixels.)
"""
import inspect

import mpmath
import numpy as np

from . import settings


class Projection:
    """A mapping of the screen pixels (projection.py:13-210)."""

    scale = 1.0

    @property
    def init_kwargs(self):
        return {name: getattr(self, name) for name in inspect.signature(self.__init__).parameters}

    def __eq__(self, other):
        return other.__class__ == self.__class__ and other.init_kwargs == self.init_kwargs

    def __hash__(self):
        return hash(self.__class__.__name__)

    def adjust_to_zoom(self, fractal):
        """Constraints the projection puts on the zoom (none by default)."""

    def f(self, pix):
        return pix

    dzndc_modifier = None

    def bounding_box(self, xy_ratio):
        raise NotImplementedError

    @property
    def min_local_scale(self):
        raise NotImplementedError

    # The map the post-processing layers apply to dz/dc (the Jacobian for the
    # burning ship) before the DEM and the normal (get_dzndc, postproc.py:
    # 184-206): none for the identity.  DF_* codes as in fsg_post_desc.proj.
    df_kind = 0

    def df(self, pix):
        """Complex factor of dz/dc at the pixels (None: no factor)."""
        return None

    def dfBS(self, pix):
        """(m00, m01, m10, m11) applied to the burning-ship Jacobian (None: none)."""
        return None


# df kinds (fsg_post_desc.proj): the factor of the holomorphic dz/dc and the
# matrix of the burning-ship Jacobian, t = pix_to_ht * pix
DF_NONE = 0          # Cartesian
DF_EXP = 1           # exp(t);            (-Im e^t, -Re e^t, Re e^t, -Im e^t)
DF_EXP_REAL = 2      # exp(Re t);         (e^Re t, 0, 0, e^Re t)
DF_ROT = 3           # exp(i Im t);       (cos Im t, -sin Im t, sin Im t, cos Im t)
DF_ONE = 4           # 1;                 identity


class Cartesian(Projection):
    """The identity (projection.py:212-266)."""

    def __init__(self, expmap_seam=None):
        self.scale = 1.0
        self.expmap_seam = expmap_seam
        if expmap_seam is not None:
            seam = float(expmap_seam)
            # projection.py:250-262
            self.dzndc_modifier = lambda cpix: np.abs(cpix + 1.e-6) * seam

    def bounding_box(self, xy_ratio):
        return 1., 1. / xy_ratio

    @property
    def min_local_scale(self):
        return 1.


class Expmap(Projection):
    """Exponential map z_pix -> exp(h_moy + dh z_pix), h from hmin to hmax
    along the screen's x axis ("horizontal") or y axis ("vertical")
    (projection.py:226-500).

    A deep map (exp(hmax) above 1e300: the reference holds h as Xrange
    there, projection.py:287-294) spans more than binary64 can hold as a
    pixel offset in units of dx.  It is computed by bands of h
    (PerturbationFractal.calc_raw_by_steps; calc_raw does so by itself): in
    a band, the offsets are exp(h - h_band + i t), within binary64, and the
    band's descriptor takes dx exp(h_band) as its (extended-range) pixel
    scale, which is also the derivative scale of the band
    (perturbation.py:459-467): the same c = dx exp(h + i t) for every pixel."""

    def __init__(self, hmin, hmax, rotates_df=True, orientation="horizontal"):
        if not 0 <= hmin < hmax:
            raise ValueError("Provide hmin, hmax with:  0 <= hmin < hmax for Expmap")
        # h itself stays within binary64 (h ~ 4605 at 1e2000); only exp(h) does not
        self.deep = bool(mpmath.exp(hmax) > 1. / settings.xrange_zoom_level)
        self.hmin = hmin
        self.hmax = hmax
        self.rotates_df = rotates_df
        self.orientation = orientation
        self.premul_1j = {"horizontal": False, "vertical": True}[orientation]
        self.hmoy = 0.5 * (float(hmin) + float(hmax))
        self.dh = float(hmax) - float(hmin)
        self.xy_ratio = None
        self.use_step = False

    def set_exp_zoom_step(self, exp_step_hmax, exp_step_hmin):
        """Compute the next band of h only (projection.py:311-331): the
        derivative scale, the modifier's shift and the bounding box follow
        the band; f changes only for a deep map (offsets relative to the
        band)."""
        self.exp_step_hmax = float(exp_step_hmax)
        self.exp_step_hmin = float(exp_step_hmin)
        self.exp_step_hmoy = 0.5 * (self.exp_step_hmax + self.exp_step_hmin)
        self.exp_step_dh = self.exp_step_hmax - self.exp_step_hmin
        self.use_step = True

    def del_exp_zoom_step(self):
        self.use_step = False

    def nh(self, fractal):
        return fractal.ny if self.premul_1j else fractal.nx

    def nt(self, fractal):
        return fractal.nx if self.premul_1j else fractal.ny

    @property
    def pix_to_ht(self):
        return (1j * self.dh * self.xy_ratio) if self.premul_1j else self.dh

    def adjust_to_zoom(self, fractal):
        """The screen's t axis spans 2 pi: xy_ratio (and nx in the vertical
        orientation) follow from dh (projection.py:339-360)."""
        if self.premul_1j:
            xy_ratio = (np.pi * 2.) / self.dh
            nx = int(fractal.nx * xy_ratio + 0.5)
        else:
            xy_ratio = self.dh / (np.pi * 2.)
            nx = fractal.nx
        fractal.xy_ratio = self.xy_ratio = xy_ratio
        fractal.nx = nx
        if hasattr(fractal, "zoom_kwargs"):
            fractal.zoom_kwargs["xy_ratio"] = xy_ratio
            fractal.zoom_kwargs["nx"] = nx

    def f(self, pix):
        """exp(h_moy + pix_to_ht * pix) (projection.py:363-372); for a deep
        map, relative to the current band: exp(h_moy - h_band + pix_to_ht * pix)
        (the band's pixel scale carries exp(h_band))."""
        pix = np.asarray(pix, dtype=np.complex128)
        if self.deep:
            if not self.use_step:
                raise ValueError("a deep Expmap (exp(hmax) > 1e300) is computed by bands of h "
                                 "(calc_raw_by_steps)")
            return np.exp((self.hmoy - self.exp_step_hmoy) + self.pix_to_ht * pix)
        return np.exp(self.hmoy + self.pix_to_ht * pix)

    @property
    def band_scale(self):
        """exp(h_band) of a deep map's current band, the factor between the
        offsets f returns and the frame's dx (1 otherwise)."""
        if self.deep and self.use_step:
            return mpmath.exp(self.exp_step_hmoy)
        return 1

    @property
    def df_kind(self):
        """The post-processing map of the derivatives (projection.py:375-453):
        by bands (use_step) the rotation exp(i Im t) only, the scale being
        applied by dzndc_modifier in the loop; over the whole map exp(t), or
        exp(Re t) without rotates_df.  (Over the whole map the reference's
        make_dfBS_impl stores its matrix in `df` as well, projection.py:435,
        which its holomorphic layers then cannot apply; here each model keeps
        its own map as make_df_impl / make_dfBS_impl write it.)"""
        if self.use_step:
            return DF_ROT if self.rotates_df else DF_ONE
        return DF_EXP if self.rotates_df else DF_EXP_REAL

    def df(self, pix):
        t = self.pix_to_ht * np.asarray(pix, dtype=np.complex128)
        k = self.df_kind
        if k == DF_EXP:
            return np.exp(t)
        if k == DF_EXP_REAL:
            return np.exp(np.real(t)).astype(np.complex128)
        if k == DF_ROT:
            return np.exp(1j * np.imag(t))
        return np.ones_like(t)

    def dfBS(self, pix):
        t = self.pix_to_ht * np.asarray(pix, dtype=np.complex128)
        k = self.df_kind
        if k == DF_EXP:
            r = np.exp(np.real(t))
            cr, sr = np.cos(np.imag(t)) * r, np.sin(np.imag(t)) * r
            return -sr, -cr, cr, -sr
        if k == DF_EXP_REAL:
            r = np.exp(np.real(t))
            z = np.zeros_like(r)
            return r, z, z, r
        if k == DF_ROT:
            c, s = np.cos(np.imag(t)), np.sin(np.imag(t))
            return c, -s, s, c
        o, z = np.ones(t.shape), np.zeros(t.shape)
        return o, z, z, o

    def dzndc_modifier(self, cpix):
        """exp(Re(pix_to_ht * cpix) + h_shift), h_shift = h_moy, or
        h_moy - the band's h_moy in steps (projection.py:454-471)."""
        hshift = (self.hmoy - self.exp_step_hmoy) if self.use_step else self.hmoy
        return np.exp(np.real(self.pix_to_ht * np.asarray(cpix, dtype=np.complex128)) + hshift)

    @property
    def scale(self):
        return mpmath.exp(self.exp_step_hmoy if self.use_step else self.hmoy)

    def bounding_box(self, xy_ratio):
        w = mpmath.exp(self.exp_step_hmax if self.use_step else self.hmax)
        return w, w

    @property
    def min_local_scale(self):
        return mpmath.exp(self.hmin)


def is_cartesian(projection):
    return projection is None or isinstance(projection, str) or (
        isinstance(projection, Cartesian) and projection.dzndc_modifier is None)
