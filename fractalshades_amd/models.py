"""The perturbation models: drop-in counterparts of the reference classes.

* ``Perturbation_mandelbrot`` -- z^2 + c (models/mandelbrot_M2.py:359-760)
* ``Perturbation_mandelbrot_N`` -- z^d + c (models/mandelbrot_Mn.py:387-856)
* ``Perturbation_burning_ship`` -- burning-ship family
  (models/burning_ship.py:863-1355)

``calc_std_div`` keeps the reference's keyword arguments, field codes
(``complex_codes`` / ``int_codes`` / ``stop_codes``) and returned factory
dict; the factories describe the GPU launch instead of building numba
closures.
"""
import math

import mpmath
import numpy as np

from . import settings
from . import fixed_orbit
from . import mpmath_loop as fsFP
from ._native import FSG_KIND_M2, FSG_KIND_MN, FSG_KIND_BS
from .perturbation import PerturbationFractal
from .utils import calc_options

BS_FLAVORS = ("Burning ship", "Perpendicular burning ship", "Shark fin", "Celtic", "Buffalo")


def _orbits():
    """The module that runs reference orbits (settings.orbit_backend)."""
    if settings.orbit_backend == "native":
        return fixed_orbit
    if settings.orbit_backend == "mpmath":
        return fsFP
    raise ValueError(f"unknown orbit_backend {settings.orbit_backend!r}")


def get_flavor_int(flavor):
    """burning_ship.py:78-81: flavours numbered from 1."""
    return BS_FLAVORS.index(flavor) + 1


class _HolomorphicBase(PerturbationFractal):

    def _holo_calc(self, *, calc_name, subset, max_iter, M_divergence, epsilon_stationnary,
                   BLA_eps, interior_detect, calc_dzndc, calc_orbit, backshift):
        if calc_orbit and backshift < 1:
            raise ValueError("calc_orbit needs backshift > 0")
        complex_codes = ["zn"]
        if interior_detect:
            complex_codes.append("dzndz")
        if calc_dzndc:
            complex_codes.append("dzndc")
        if calc_orbit:
            complex_codes.append("zn_orbit")      # mandelbrot_M2.py:501-506
        int_codes = ["ref_cycle_iter"]
        stop_codes = ["max_iter", "divergence", "stationnary"]
        BLA_activated = (BLA_eps is not None) and bool(self.dx < settings.newton_zoom_level)

        def set_state():
            def impl(instance):
                instance.complex_type = np.complex128
                instance.potential_M = M_divergence
                instance.codes = (complex_codes, int_codes, stop_codes)
                instance.calc_dZndz = interior_detect
                instance.calc_dZndc = calc_dzndc
                instance.backshift = backshift if calc_orbit else None
            return impl

        def initialize():
            return {"zn": 0, "dzndz": 1 if interior_detect else -1,
                    "dzndc": (1 + int(interior_detect)) if calc_dzndc else -1}

        def iterate():
            return {
                "max_iter": max_iter, "M_divergence": M_divergence,
                "epsilon_stationnary": epsilon_stationnary,
                "BLA_activated": BLA_activated, "BLA_eps": BLA_eps,
                "calc_dzndc": calc_dzndc, "calc_dzndz": interior_detect,
                "calc_orbit": calc_orbit, "backshift": backshift,
                "xr_detect_activated": self.xr_detect_activated,
            }

        return {"set_state": set_state, "initialize": initialize, "iterate": iterate}


class Perturbation_mandelbrot(_HolomorphicBase):
    """Arbitrary-precision power-2 Mandelbrot, z_{n+1} = z_n^2 + c."""

    fsg_kind = FSG_KIND_M2

    def __init__(self, directory):
        super().__init__(directory)
        self.potential_kind = "infinity"
        self.potential_d = 2
        self.potential_a_d = 1.
        self.potential_M_cutoff = 1000.
        self.critical_pt = 0.
        self.FP_code = "zn"
        self.holomorphic = True
        self.exponent = 2

    def FP_loop(self, ref_orbit_len, c0):
        return _orbits().holo_orbit(c0, 2, ref_orbit_len, self.M_divergence * 2,
                               self.xr_detect_activated, self.prec)

    @calc_options
    def calc_std_div(self, *, calc_name: str, subset=None, max_iter: int,
                     M_divergence: float, epsilon_stationnary: float,
                     BLA_eps: float = 1e-6, interior_detect: bool = False,
                     calc_dzndc: bool = True, calc_orbit: bool = False, backshift: int = 0):
        """Perturbation iterations for z^2 + c (mandelbrot_M2.py:432-588)."""
        return self._holo_calc(calc_name=calc_name, subset=subset, max_iter=max_iter,
                               M_divergence=M_divergence,
                               epsilon_stationnary=epsilon_stationnary, BLA_eps=BLA_eps,
                               interior_detect=interior_detect, calc_dzndc=calc_dzndc,
                               calc_orbit=calc_orbit, backshift=backshift)

    def _ball_method(self, c, px, maxiter, M_divergence):
        return _orbits().holo_ball_method(c, px, 2, maxiter, M_divergence, self.prec)

    def find_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        return _orbits().find_nucleus_m2(c, order, eps_pixel, prec or self.prec,
                                    max_newton or 80, eps_cv)

    def find_any_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        return _orbits().find_any_nucleus(c, order, 2, eps_pixel, prec or self.prec,
                                     max_newton or 80, eps_cv)

    def _nucleus_size_estimate(self, c0, order):
        s = abs(fsFP.nucleus_size_estimate(c0, order, 2, self.prec))
        return s, s ** 0.75


class Perturbation_mandelbrot_N(_HolomorphicBase):
    """Arbitrary-precision power-N Mandelbrot, z_{n+1} = z_n^N + c."""

    fsg_kind = FSG_KIND_MN

    def __init__(self, directory, exponent):
        super().__init__(directory)
        self.exponent = int(exponent)
        self.potential_kind = "infinity"
        self.potential_d = self.exponent
        self.potential_a_d = 1.
        self.potential_M_cutoff = 1000.
        self.critical_pt = 0.
        self.FP_code = "zn"
        self.holomorphic = True

    @property
    def init_kwargs(self):
        return {"directory": self.directory, "exponent": self.exponent}

    def FP_loop(self, ref_orbit_len, c0):
        return _orbits().holo_orbit(c0, self.exponent, ref_orbit_len, self.M_divergence * 2,
                               self.xr_detect_activated, self.prec)

    @calc_options
    def calc_std_div(self, *, calc_name: str, subset=None, max_iter: int,
                     M_divergence: float, epsilon_stationnary: float,
                     BLA_eps: float = 1e-6, interior_detect: bool = False,
                     calc_dzndc: bool = True, calc_orbit: bool = False, backshift: int = 0):
        """Perturbation iterations for z^N + c (mandelbrot_Mn.py:462-625)."""
        return self._holo_calc(calc_name=calc_name, subset=subset, max_iter=max_iter,
                               M_divergence=M_divergence,
                               epsilon_stationnary=epsilon_stationnary, BLA_eps=BLA_eps,
                               interior_detect=interior_detect, calc_dzndc=calc_dzndc,
                               calc_orbit=calc_orbit, backshift=backshift)

    def _ball_method(self, c, px, maxiter, M_divergence):
        return _orbits().holo_ball_method(c, px, self.exponent, maxiter, M_divergence, self.prec)

    def find_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        raise NotImplementedError(
            "Divide by undesired roots technique not implemented, Use 'find_any_nucleus'")

    def find_any_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        return _orbits().find_any_nucleus(c, order, self.exponent, eps_pixel, prec or self.prec,
                                     max_newton or 80, eps_cv)

    def _nucleus_size_estimate(self, c0, order):
        s = abs(fsFP.nucleus_size_estimate(c0, order, self.exponent, self.prec))
        d = self.exponent
        return s, s ** ((d + 1.) * (d - 1.) / d / d)


class Perturbation_burning_ship(PerturbationFractal):
    """Arbitrary-precision burning ship and its "abs" variants:
    x_{n+1} = x_n^2 - y_n^2 + a, y_{n+1} = 2 |x_n y_n| - b (flavour 1)."""

    fsg_kind = FSG_KIND_BS

    def __init__(self, directory, flavor="Burning ship"):
        super().__init__(directory)
        self.flavor = flavor
        self.flavor_int = get_flavor_int(flavor)
        self.potential_kind = "infinity"
        self.potential_d = 2
        self.potential_a_d = 1.
        self.potential_M_cutoff = 1000.
        self.critical_pt = 0.
        self.FP_code = ["xn", "yn"]
        self.holomorphic = False

    @property
    def init_kwargs(self):
        return {"directory": self.directory, "flavor": self.flavor}

    def FP_loop(self, ref_orbit_len, c0):
        return _orbits().bs_orbit(c0, self.flavor_int, ref_orbit_len, self.M_divergence * 2,
                             self.xr_detect_activated, self.prec)

    @calc_options
    def calc_std_div(self, *, calc_name: str, subset=None, max_iter: int,
                     M_divergence: float, BLA_eps: float = 1e-6,
                     calc_hessian: bool = True, calc_orbit: bool = False, backshift: int = 0):
        """Perturbation iterations for the burning-ship family
        (burning_ship.py:971-1130)."""
        if calc_orbit and backshift < 1:
            raise ValueError("calc_orbit needs backshift > 0")
        complex_codes = ["xn", "yn"]
        if calc_hessian:
            complex_codes += ["dxnda", "dxndb", "dynda", "dyndb"]
        if calc_orbit:
            complex_codes += ["xn_orbit", "yn_orbit"]   # burning_ship.py:1028-1034
        int_codes = ["ref_cycle_iter"]
        stop_codes = ["max_iter", "divergence"]
        BLA_activated = (BLA_eps is not None) and bool(self.dx < settings.newton_zoom_level)

        def set_state():
            def impl(instance):
                instance.complex_type = np.float64
                instance.potential_M = M_divergence
                instance.codes = (complex_codes, int_codes, stop_codes)
                instance.backshift = backshift if calc_orbit else None
            return impl

        def initialize():
            return {"xn": 0, "yn": 1}

        def iterate():
            return {
                "max_iter": max_iter, "M_divergence": M_divergence,
                "BLA_activated": BLA_activated, "BLA_eps": BLA_eps,
                "calc_dzndc": calc_hessian, "calc_dzndz": False,
                "calc_orbit": calc_orbit, "backshift": backshift,
                "xr_detect_activated": self.xr_detect_activated,
            }

        return {"set_state": set_state, "initialize": initialize, "iterate": iterate}

    def _ball_method(self, c, px, maxiter, M_divergence):
        return _orbits().bs_ball_method(c, px, self.flavor_int, maxiter, M_divergence, self.prec)

    def find_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        raise NotImplementedError("use find_any_nucleus")

    def find_any_nucleus(self, c, order, eps_pixel, prec=None, max_newton=None, eps_cv=None):
        return _orbits().bs_find_any_nucleus(c, order, self.flavor_int, eps_pixel,
                                        prec or self.prec, max_newton or 80, eps_cv)
