"""Module-level settings read by the perturbation path.

Same names and defaults as the reference's settings (src/fractalshades/
settings.py:7-40, 82).  The GPU engine replaces the thread pool, so
``enable_multithreading`` only selects whether tiles are batched into one
launch (True) or sent one by one (False).
"""
enable_multithreading: bool = True
skip_calc: bool = False
newton_zoom_level: float = 1.e-5
std_zoom_level: float = 1.e-8
xrange_zoom_level: float = 1.e-300
no_newton: bool = False
inspect_calc: bool = False
chunk_size: int = 200
BLA_compression: int = 3
postproc_dtype: str = "float32"

# Arithmetic of the GPU delta loop: "fp64" is the drop-in (the reference's
# binary64 / extended-range semantics, bit for bit); "fp32x" runs binary32
# mantissas with an int32 exponent (faster, approximate).
gpu_precision: str = "fp64"

# Reference orbit loop: "native" iterates in the library's multi-limb fixed
# point (fsg_ref_orbit); "mpmath" in mpmath at the same precision.
orbit_backend: str = "native"
