"""fractalshades_amd -- MI355X (gfx950) engine for the perturbation loop of
escape-time deep zooms, behind the Fractalshades model API.

The per-pixel loop runs in hand-written HIP kernels (libfsgpu.so, C ABI in
include/fsgpu.h); the arbitrary-precision reference orbit, period and nucleus
search run on the host, in the same library's multi-limb C++ loops.
"""
__version__ = "0.1.0"

from . import settings  # noqa: F401
from . import projection  # noqa: F401
from .models import (  # noqa: F401
    Perturbation_mandelbrot,
    Perturbation_mandelbrot_N,
    Perturbation_burning_ship,
)
