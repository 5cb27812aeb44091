"""Fractal base class: zoom, tiling, pixel grid, calculation registry.

The parts of the reference's ``Fractal`` (src/fractalshades/core.py:1264-2808)
that sit on the perturbation path:

* ``zoom`` options and the linear pixel -> plane map (core.py:1400-1484);
* tiles of ``settings.chunk_size`` pixels (chunk_slices / chunks_count /
  chunk_rank / chunk_from_rank, core.py:1644-1700);
* ``chunk_pixel_pos``: pixel offsets in units of dx (core.py:1767-1830);
* ``calc_hook``: stores the calculation set up by ``calc_std_div``
  (core.py:1918-2009);
* ``get_cycling_dep_args``: per-tile input and result arrays (core.py:2044-2075);
* ``calc_raw``: computes every tile and returns the raw arrays.  The
  reference writes them to npy memmaps (core.py:2174-2512); here the frame is
  handed to the GPU in large batches and returned in memory.
"""
import types

import numpy as np

from . import settings
from .utils import zoom_options


class Fractal:
    """Base class of the escape-time models."""

    USER_INTERRUPTED = 1

    def __init__(self, directory):
        self.directory = directory
        self.termination_type = np.int8
        self.int_type = np.int32
        self._interrupted = np.array([0], dtype=np.uint8)
        self._calc_data = {}
        self._skew = None

    # ---- zoom / geometry -------------------------------------------------
    @zoom_options
    def zoom(self, *, x: float = 0., y: float = 0., dx: float = 8., nx: int = 800,
             xy_ratio: float = 1., theta_deg: float = 0., projection="cartesian",
             has_skew: bool = False, skew_00: float = 1., skew_01: float = 0.,
             skew_10: float = 0., skew_11: float = 1.):
        if has_skew:
            self._skew = np.array(((skew_00, skew_01), (skew_10, skew_11)), dtype=np.float64)
        self.lin_mat = self.get_lin_mat()

    @property
    def init_kwargs(self):
        return {"directory": self.directory}

    @property
    def ny(self):
        return int(self.nx / self.xy_ratio + 0.5)

    @property
    def dy(self):
        return self.dx / self.xy_ratio

    @property
    def float_type(self):
        return np.float64

    def get_lin_mat(self):
        """Rotation (theta_deg) then optional skew (core.py:1470-1484)."""
        theta = self.theta_deg / 180. * np.pi
        c = np.cos(theta)
        s = np.sin(theta)
        lin_mat = np.array(((c, -s), (s, c)), dtype=np.float64)
        if self._skew is not None:
            lin_mat = np.matmul(self._skew, lin_mat)
        return lin_mat

    # ---- tiles -------------------------------------------------------------
    def chunk_slices(self):
        cs = settings.chunk_size
        for ix in range(0, self.nx, cs):
            ixx = min(ix + cs, self.nx)
            for iy in range(0, self.ny, cs):
                iyy = min(iy + cs, self.ny)
                yield (ix, ixx, iy, iyy)

    @property
    def chunks_count(self):
        cs = settings.chunk_size
        cx, r = divmod(self.nx, cs)
        cx += (r != 0)
        cy, r = divmod(self.ny, cs)
        cy += (r != 0)
        return cx * cy

    def chunk_rank(self, chunk_slice):
        cs = settings.chunk_size
        ix, _, iy, _ = chunk_slice
        cy, r = divmod(self.ny, cs)
        cy += (r != 0)
        return (ix // cs) * cy + (iy // cs)

    def chunk_from_rank(self, rank):
        cs = settings.chunk_size
        cy, r = divmod(self.ny, cs)
        cy += (r != 0)
        kx, ky = divmod(rank, cy)
        ix, iy = kx * cs, ky * cs
        return (ix, min(ix + cs, self.nx), iy, min(iy + cs, self.ny))

    def chunk_pixel_pos(self, chunk_slice, jitter=False, supersampling=None):
        """Pixel offsets from the centre in units of dx, as complex128
        (core.py:1767-1830): x in [-0.5, 0.5], y flipped and divided by
        xy_ratio, row-major over (y, x)."""
        data_type = self.float_type
        nx, ny = self.nx, self.ny
        ix, ixx, iy, iyy = chunk_slice
        kx = 0.5 / (nx - 1)
        ky = 0.5 / (ny - 1)
        if supersampling is None:
            x_1d = np.linspace(kx * (2 * ix - nx + 1), kx * (2 * ixx - nx - 1),
                               num=(ixx - ix), dtype=data_type)
            y_1d = np.linspace(ky * (2 * iy - ny + 1), ky * (2 * iyy - ny - 1),
                               num=(iyy - iy), dtype=data_type)
        else:
            ssg = supersampling
            gap = (ssg - 1.) / ssg
            x_1d = np.linspace(kx * (2 * ix - nx + 1 - gap), kx * (2 * ixx - nx - 1 + gap),
                               num=(ixx - ix) * ssg, dtype=data_type)
            y_1d = np.linspace(ky * (2 * iy - ny + 1 - gap), ky * (2 * iyy - ny - 1 + gap),
                               num=(iyy - iy) * ssg, dtype=data_type)
        dx_screen, dy_screen = np.meshgrid(x_1d, -y_1d, indexing='xy')
        if jitter:
            rg = np.random.default_rng(0)
            rand_x = rg.random(dx_screen.shape, dtype=data_type)
            rand_y = rg.random(dy_screen.shape, dtype=data_type)
            k = 0.7071067811865476
            jitter_x = (0.5 - rand_x) * k / (nx - 1) * jitter
            jitter_y = (0.5 - rand_y) * k / (ny - 1) * jitter
            if supersampling is not None:
                jitter_x /= supersampling
                jitter_y /= supersampling
            dx_screen += jitter_x
            dy_screen += jitter_y
        dy_screen /= self.xy_ratio
        return dx_screen + 1j * dy_screen

    # ---- calculation registry ---------------------------------------------
    def calc_hook(self, calc_callable, calc_kwargs, return_dic):
        """Store a calculation set up by a @calc_options method."""
        calc_name = calc_kwargs["calc_name"]
        state = types.SimpleNamespace()
        for k, v in calc_kwargs.items():
            setattr(state, k, v)
            setattr(self, k, v)
        set_state = return_dic["set_state"]()
        set_state(state)
        set_state(self)
        initialize = return_dic["initialize"]()
        iterate = return_dic["iterate"]()
        cycle_indep_args = self.get_cycle_indep_args(initialize, iterate)
        self._calc_data[calc_name] = {
            "calc_class": type(self).__name__,
            "calc_callable": calc_callable,
            "calc_kwargs": calc_kwargs,
            "zoom_kwargs": dict(getattr(self, "zoom_kwargs", {})),
            "state": state,
            "cycle_indep_args": cycle_indep_args,
            "iterate": iterate,
        }

    def get_cycling_dep_args(self, calc_name, chunk_slice, jitter=False, supersampling=None):
        """(c_pix, Z, U, stop_reason, stop_iter) for one tile (core.py:2044-2075)."""
        c_pix = np.ravel(self.chunk_pixel_pos(chunk_slice, jitter, supersampling))
        state = self._calc_data[calc_name]["state"]
        subset = getattr(state, "subset", None)
        chunk_subset = None
        if subset is not None:
            chunk_subset = np.ravel(np.asarray(subset[chunk_slice], dtype=bool))
            c_pix = c_pix[chunk_subset]
        return self._result_arrays(state, c_pix), chunk_subset

    def _image_mask(self, subset):
        """A subset as a mask over the whole frame, row-major (y, x): an
        array is taken as it is; a tile-indexed subset (a Fractal_array)
        is assembled from its tiles."""
        if isinstance(subset, np.ndarray):
            return np.ravel(subset.astype(bool))
        mask = np.zeros((self.ny, self.nx), dtype=bool)
        for cs in self.chunk_slices():
            ix, ixx, iy, iyy = cs
            mask[iy:iyy, ix:ixx] = np.asarray(subset[cs], dtype=bool).reshape(iyy - iy, ixx - ix)
        return np.ravel(mask)

    def _result_arrays(self, state, c_pix):
        n_pts = c_pix.shape[0]
        n_Z, n_U, _ = (len(code) for code in state.codes)
        Z = np.zeros([n_Z, n_pts], dtype=state.complex_type)
        U = np.zeros([n_U, n_pts], dtype=self.int_type)
        stop_reason = -np.ones([1, n_pts], dtype=self.termination_type)
        stop_iter = np.zeros([1, n_pts], dtype=self.int_type)
        return (np.ascontiguousarray(c_pix), Z, U, stop_reason, stop_iter)

    def compute_tile(self, calc_name, chunk_slice, jitter=False, supersampling=None):
        """One tile, as the reference's compute_rawdata_dev (core.py:2515-2554)."""
        (c_pix, Z, U, stop_reason, stop_iter), subset = self.get_cycling_dep_args(
            calc_name, chunk_slice, jitter, supersampling)
        cycle_indep_args = self._calc_data[calc_name]["cycle_indep_args"]
        ix, ixx, _, _ = chunk_slice
        grid_w = 0 if subset is not None else (ixx - ix) * (supersampling or 1)
        ret = self.numba_cycle_call((c_pix, Z, U, stop_reason, stop_iter), cycle_indep_args,
                                    grid_w=grid_w)
        return ret, (c_pix, Z, U, stop_reason, stop_iter)

    def calc_raw(self, calc_name, jitter=False, supersampling=None):
        """All pixels of the frame in one batch.

        Returns a dict with c_pix [npts] and Z [n_Z, npts], U [1, npts],
        stop_reason [1, npts], stop_iter [1, npts], pixels in row-major (y, x)
        order of the whole image, and the engine's step counters.
        """
        if settings.skip_calc:
            return None
        state = self._calc_data[calc_name]["state"]
        c_pix = np.ravel(self.chunk_pixel_pos((0, self.nx, 0, self.ny), jitter, supersampling))
        subset = getattr(state, "subset", None)
        grid_w = self.nx * (supersampling or 1)
        if subset is not None:
            c_pix = c_pix[self._image_mask(subset)]
            grid_w = 0
        args = self._result_arrays(state, c_pix)
        ret = self.numba_cycle_call(args, self._calc_data[calc_name]["cycle_indep_args"],
                                    grid_w=grid_w)
        c_pix, Z, U, stop_reason, stop_iter = args
        return {"ret": ret, "c_pix": c_pix, "Z": Z, "U": U, "stop_reason": stop_reason,
                "stop_iter": stop_iter, "codes": state.codes,
                "stats": getattr(self, "last_stats", None)}
