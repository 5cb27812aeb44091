"""Serving loop on the GPU: tile requests of mixed size, zoom depth and
max_iter, submitted as they arrive and continuously batched on the device
(the C side is fsg_server_* in include/fsgpu.h).

* ``TileServer`` wraps one native server.  ``submit`` returns at once; the
  server admits requests in arrival order as its device-memory budget
  allows and runs chunks of several requests at the same time.
* ``PinnedTile`` holds a request's buffers in pinned host memory, so that
  the input and result copies are asynchronous DMA transfers.
* ``Centre`` / ``serve_centres`` / ``serve_requests`` / ``build_request``
  make the seeded request stream: eight seeded minibrot nuclei, each with
  one reference orbit per arithmetic mode, computed once at the deepest
  precision of the stream, and per-request frames of random size, width and
  max_iter around them.

No reference counterpart: the reference renders the tiles of one frame with
a thread pool (mthreading.py:48-68).
"""
import ctypes
import json
import math
import os
import threading
import time
from dataclasses import dataclass

import mpmath
import numpy as np

from ._native import (FSG_PENDING, FsgError, RequestInfo, ServerOpts, check, load, n_z_fields)

SIZES = ((256, 256), (512, 512), (1024, 1024), (4096, 256))


class PinnedTile:
    """c_pix, Z, U, stop_reason, stop_iter of one request in pinned memory."""

    def __init__(self, n, nz, holomorphic):
        lib = load()
        zdt = np.complex128 if holomorphic else np.float64
        specs = [("c_pix", (n,), np.complex128), ("Z", (nz, n), zdt), ("U", (1, n), np.int32),
                 ("stop_reason", (1, n), np.int8), ("stop_iter", (1, n), np.int32)]
        self._ptrs = []
        for name, shape, dt in specs:
            nbytes = int(np.prod(shape)) * np.dtype(dt).itemsize
            p = lib.fsg_host_alloc(max(nbytes, 1))
            if not p:
                self.close()
                raise FsgError("fsg_host_alloc failed")
            self._ptrs.append(p)
            buf = (ctypes.c_char * max(nbytes, 1)).from_address(p)
            setattr(self, name, np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape))
        self.n = n

    def close(self):
        lib = load()
        for p in self._ptrs:
            lib.fsg_host_free(p)
        self._ptrs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LayerTile:
    """c_pix in, the post-processing layers (float32 nu, dem, and with
    normal=True the unit normal as float32 [n, 2]) with stop_reason and
    stop_iter out, in pinned memory."""

    def __init__(self, n, normal=False):
        lib = load()
        specs = [("c_pix", (n,), np.complex128), ("nu", (n,), np.float32), ("dem", (n,), np.float32),
                 ("stop_reason", (1, n), np.int8), ("stop_iter", (1, n), np.int32)]
        if normal:
            specs.append(("normal", (n, 2), np.float32))
        self.normal = None
        self._ptrs = []
        for name, shape, dt in specs:
            nbytes = int(np.prod(shape)) * np.dtype(dt).itemsize
            p = lib.fsg_host_alloc(max(nbytes, 1))
            if not p:
                self.close()
                raise FsgError("fsg_host_alloc failed")
            self._ptrs.append(p)
            buf = (ctypes.c_char * max(nbytes, 1)).from_address(p)
            setattr(self, name, np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape))
        self.n = n

    close = PinnedTile.close
    __del__ = PinnedTile.__del__


class TileServer:
    """One native serving loop (fsg_server_create)."""

    def __init__(self, lanes=3, chunk_px=65536, hbm_budget=4 << 30):
        lib = load()
        o = ServerOpts(lanes, chunk_px, hbm_budget)
        h = ctypes.c_void_p()
        check(lib.fsg_server_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._keep = {}
        self._lock = threading.Lock()

    def clock(self):
        ms = ctypes.c_double()
        check(load().fsg_server_clock(self._h, ctypes.byref(ms)))
        return ms.value

    def footprint(self, ctx, n):
        b = ctypes.c_int64()
        check(load().fsg_server_footprint(ctx.handle, n, ctypes.byref(b)))
        return b.value

    def submit(self, ctx, tile, grid_w=0):
        """Queue tile (a PinnedTile or any object with the five C-contiguous
        arrays) on context ctx (a GpuContext); returns the ticket.  grid_w >
        0: the pixels are rows of grid_w, iterated in 8x8 blocks (the same
        results)."""
        t = ctypes.c_uint64()
        arrs = (tile.c_pix, tile.Z, tile.U, tile.stop_reason, tile.stop_iter)
        for a in arrs:
            if not a.flags.c_contiguous:
                raise ValueError("request buffers must be C-contiguous")
        check(load().fsg_server_submit(self._h, ctx.handle, tile.c_pix.ctypes.data, tile.c_pix.shape[0],
                                       int(grid_w), tile.Z.ctypes.data, tile.U.ctypes.data, tile.stop_reason.ctypes.data,
                                       tile.stop_iter.ctypes.data, ctypes.byref(t)))
        with self._lock:
            self._keep[t.value] = (ctx, tile)   # alive until waited for
        return t.value

    def submit_layers(self, ctx, tile, post, grid_w=0):
        """Queue a request returning the nu / dem (/ normal) layers
        (LayerTile) instead of the raw fields; post is a PostDesc
        (postproc.post_desc); grid_w as for submit."""
        t = ctypes.c_uint64()
        nrm = tile.normal.ctypes.data if tile.normal is not None else None
        check(load().fsg_server_submit_layers(
            self._h, ctx.handle, tile.c_pix.ctypes.data, tile.c_pix.shape[0], int(grid_w), ctypes.byref(post),
            tile.nu.ctypes.data, tile.dem.ctypes.data, nrm, tile.stop_reason.ctypes.data,
            tile.stop_iter.ctypes.data, ctypes.byref(t)))
        with self._lock:
            self._keep[t.value] = (ctx, tile, post)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """Request info dict when done, None on timeout; raises on failure."""
        info = RequestInfo()
        rc = load().fsg_server_wait(self._h, ticket, int(timeout_ms), ctypes.byref(info))
        if rc == FSG_PENDING:
            return None
        with self._lock:
            self._keep.pop(ticket, None)
        check(rc)
        return info.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            load().fsg_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------
# The seeded request stream: eight centres, mixed sizes / widths / max_iter

@dataclass
class Centre:
    nucleus: object        # mpmath.mpc at `digits`
    order: int
    digits: int
    orbits: dict           # xr_mode -> (Zn_path, div_iter, xr register)


SERVE_CENTRES_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                  "serve_centres.json")


def _nuclei(n_centres, digits, seed0=100):
    """Seeded minibrot nuclei (period 50..500) refined by Newton to `digits`."""
    from . import configs, models
    out = []
    seed = seed0
    while len(out) < n_centres:
        w = configs.Workload(f"serve{seed}", "M2", "1e-20", 256, 256, 2000, 40,
                             period_range=(50, 500), seed=seed)
        seed += 1
        try:
            _, _, nucleus, order = configs.seeded_center(w)
        except RuntimeError:
            continue
        prec = mpmath.libmp.dps_to_prec(digits)
        with mpmath.workprec(prec):
            f = models.Perturbation_mandelbrot(None)
            f.prec = prec
            # convergence is |Newton step| <= 64 ulp; |z_order| at the
            # nucleus is then ~|dz/dc| ulp, so the validity bound is loose
            eps = mpmath.mpf(10) ** (-(digits // 2))
            ok, nuc = f.find_nucleus(mpmath.mpc(nucleus), order, eps, prec, 80)
        if ok:
            out.append((nuc, order))
    return out


def serve_centres(n_centres=8, digits=2100, cache=True):
    """The stream's centres: nuclei from the committed cache when present,
    with their reference orbits (binary64 and Xrange registers)."""
    from . import models
    nuclei = None
    if cache and os.path.exists(SERVE_CENTRES_FILE):
        with open(SERVE_CENTRES_FILE) as fh:
            d = json.load(fh)
        if d.get("digits") == digits and len(d.get("centres", [])) >= n_centres:
            prec = mpmath.libmp.dps_to_prec(digits)
            with mpmath.workprec(prec):
                nuclei = [(mpmath.mpc(c["x"], c["y"]), c["order"]) for c in d["centres"][:n_centres]]
    if nuclei is None:
        nuclei = _nuclei(n_centres, digits)
        if cache:
            os.makedirs(os.path.dirname(SERVE_CENTRES_FILE), exist_ok=True)
            with open(SERVE_CENTRES_FILE, "w") as fh:
                json.dump({"digits": digits, "centres": [
                    {"x": mpmath.nstr(n.real, digits + 5), "y": mpmath.nstr(n.imag, digits + 5), "order": o}
                    for n, o in nuclei]}, fh, indent=1)
    centres = []
    prec = mpmath.libmp.dps_to_prec(digits)
    for nuc, order in nuclei:
        orbits = {}
        for xr_mode in (False, True):
            f = models.Perturbation_mandelbrot(None)
            f.prec = prec
            f.M_divergence = 1.e3
            f.dx = mpmath.mpf("1e-400") if xr_mode else mpmath.mpf("1e-100")
            with mpmath.workprec(prec):
                div_iter, Zn_path, xr = f.FP_loop(order, nuc)
            orbits[xr_mode] = (Zn_path, div_iter, xr)
        centres.append(Centre(nuc, order, digits, orbits))
    return centres


def boundary_centre(ref_len=200001):
    """A centre on the boundary of the set rather than at a nucleus: the
    seeded 1e-2000 benchmark location (fractalshades_amd.configs, C2_2000:
    a point brought within 1e-2000 of the boundary by the potential-gradient
    descent, given to 2045 digits), with its orbit computed unwrapped from
    that point (no cycle; it escapes after ~7,200 steps).  Frames around it
    at any width down to 1e-2000 are exterior-dominated: their pixels escape
    after a number of iterations that grows with the depth, instead of
    stopping on the interior test of a nucleus's component."""
    from . import configs, models
    x, y, _, _ = configs.center_for("C2_2000", compute=False)
    digits = configs.WORKLOADS["C2_2000"].digits
    prec = mpmath.libmp.dps_to_prec(digits)
    with mpmath.workprec(prec):
        c = mpmath.mpc(x, y)
    f = models.Perturbation_mandelbrot(None)
    f.prec = prec
    f.M_divergence = 1.e3
    f.dx = mpmath.mpf("1e-400")
    with mpmath.workprec(prec):
        div_iter, Zn_path, xr = f.FP_loop(ref_len, c)
    return Centre(c, None, digits, {True: (Zn_path, div_iter, xr), False: (Zn_path, div_iter, {})})


def movie_centre(max_iter, centre=0, digits=2100):
    """The zoom sequence's centre: a cached nucleus with its orbit computed
    unwrapped up to max_iter (one resident orbit of max_iter + 1 points for
    every frame), binary64 and Xrange registers."""
    from . import models
    nuc, _ = _cached_nuclei(digits)[centre]
    prec = mpmath.libmp.dps_to_prec(digits)
    f = models.Perturbation_mandelbrot(None)
    f.prec = prec
    f.M_divergence = 1.e3
    f.dx = mpmath.mpf("1e-400")             # keeps the Xrange register
    with mpmath.workprec(prec):
        div_iter, Zn_path, xr = f.FP_loop(max_iter + 1, nuc)
    # the binary64 frames read the same rounded orbit without the register
    orbits = {True: (Zn_path, div_iter, xr), False: (Zn_path, div_iter, {})}
    return Centre(nuc, None, digits, orbits)


def movie_requests(n_frames, nx=480, ny=270, max_iter=500000, log10_width=(math.log10(4.), -2000.)):
    """A zoom sequence onto the centre: n_frames frames of nx x ny pixels,
    width geometric from 10^log10_width[0] down to 10^log10_width[1]."""
    out = []
    for k in range(n_frames):
        lw = log10_width[0] + (log10_width[1] - log10_width[0]) * k / max(n_frames - 1, 1)
        mant = 10 ** (lw - math.floor(lw))
        out.append(RequestSpec(0, nx, ny, f"{mant:.6f}e{int(math.floor(lw))}", max_iter, 0., 0.))
    return out


def _cached_nuclei(digits):
    with open(SERVE_CENTRES_FILE) as fh:
        d = json.load(fh)
    if d.get("digits") != digits:
        raise ValueError(f"no cached nuclei at {digits} digits")
    prec = mpmath.libmp.dps_to_prec(digits)
    with mpmath.workprec(prec):
        return [(mpmath.mpc(c["x"], c["y"]), c["order"]) for c in d["centres"]]


@dataclass
class RequestSpec:
    centre: int
    nx: int
    ny: int
    width: str
    max_iter: int
    ox: float
    oy: float


def serve_requests(n_requests=512, n_centres=8, seed=0, log10_width=(-20, -2000),
                   max_iter=(1e3, 1e7), sizes=SIZES):
    """Seeded stream: size uniform over `sizes`, width log-uniform,
    max_iter log-uniform, centre uniform, frame offset U(-0.3, 0.3) widths."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_requests):
        nx, ny = sizes[rng.integers(len(sizes))]
        lw = rng.uniform(log10_width[1], log10_width[0])
        mi = int(round(10 ** rng.uniform(math.log10(max_iter[0]), math.log10(max_iter[1]))))
        ox, oy = rng.uniform(-0.3, 0.3, 2)
        mant = 10 ** (lw - math.floor(lw))
        width = f"{mant:.6f}e{int(math.floor(lw))}"
        out.append(RequestSpec(int(rng.integers(n_centres)), nx, ny, width, mi, float(ox), float(oy)))
    return out


def build_request(spec, centres, precision="fp64"):
    """Model + calculation of one request on its centre's resident orbit.
    Returns (model, CycleIndepArgs, c_pix); the CycleIndepArgs carries
    ``resident_key``, the identity of the reference orbit it runs on, so
    that requests with the same key can share one device copy of it.

    Frames wider than settings.newton_zoom_level are computed around the
    critical orbit with BLA off, as the reference computes every shallow
    frame (perturbation.py:669-672, BLA_activated in mandelbrot_M2.py:521-524),
    not around the centre's nucleus."""
    from . import models, settings
    from .perturbation import PRECISIONS
    c = centres[spec.centre]
    width = mpmath.mpf(spec.width)
    with mpmath.workdps(30):
        digits = int(-mpmath.log10(width)) + 30
    prec = mpmath.libmp.dps_to_prec(digits)
    with mpmath.workprec(prec):
        x = c.nucleus.real + mpmath.mpf(spec.ox) * width
        y = c.nucleus.imag + mpmath.mpf(spec.oy) * width
        xs, ys = mpmath.nstr(x, digits + 5), mpmath.nstr(y, digits + 5)
    f = models.Perturbation_mandelbrot(None)
    f.zoom(precision=digits, x=xs, y=ys, dx=spec.width, nx=spec.nx, xy_ratio=spec.nx / spec.ny, theta_deg=0.)
    f.max_iter = spec.max_iter
    shallow = width > settings.newton_zoom_level
    if shallow:
        f.get_FP_orbit()                      # the critical orbit
    else:
        Zn_path, div_iter, xr = c.orbits[bool(f.xr_detect_activated)]
        f.set_reference(c.nucleus, c.order, Zn_path, div_iter, xr)
    f.calc_std_div(calc_name="serve", subset=None, max_iter=spec.max_iter, M_divergence=1.e3,
                   epsilon_stationnary=1.e-3, interior_detect=True, calc_dzndc=True, BLA_eps=1.e-6)
    cia = f._calc_data["serve"]["cycle_indep_args"]
    if PRECISIONS[precision] != cia.fields["precision"]:
        cia = cia.with_fields(precision=PRECISIONS[precision])
    cia.resident_key = ("critical" if shallow else spec.centre, cia.fields["xr_mode"])
    c_pix = np.ascontiguousarray(np.ravel(f.chunk_pixel_pos((0, spec.nx, 0, spec.ny))))
    return f, cia, c_pix


def n_fields(fields):
    return n_z_fields(fields)


def pinned_tiles(reqs):
    """One PinnedTile per prepared request [(cia, c_pix)], input filled in."""
    tiles = []
    for cia, c_pix in reqs:
        t = PinnedTile(c_pix.size, n_fields(cia.fields), cia.fields["kind"] != 2)
        t.c_pix[:] = c_pix
        tiles.append(t)
    return tiles


def layer_tiles(reqs):
    """One LayerTile per prepared request, input filled in."""
    tiles = []
    for cia, c_pix in reqs:
        t = LayerTile(c_pix.size)
        t.c_pix[:] = c_pix
        tiles.append(t)
    return tiles


def run_stream(server, reqs, rate_hz=None, seed=1, tiles=None, posts=None, widths=None):
    """Submit prepared requests [(cia, c_pix)] with Poisson arrivals at
    `rate_hz` (all at once when None); wait for all.  With `posts` (one
    PostDesc per request) the requests return the nu / dem layers instead of
    the raw fields; with `widths` (one row width per request) each request's
    pixels are iterated in 8x8 blocks of its rows.  Returns per-request
    dicts: latency_ms (arrival -> results in host memory), queue_ms, pixels,
    pixel_iters, tile."""
    rng = np.random.default_rng(seed)
    if tiles is None:
        tiles = layer_tiles(reqs) if posts is not None else pinned_tiles(reqs)
    gaps = rng.exponential(1. / rate_hz, len(reqs)) if rate_hz else np.zeros(len(reqs))
    arrivals = np.cumsum(gaps)
    tickets = []
    t0 = time.perf_counter()
    c0 = server.clock()
    for k, ((cia, _), tile, ta) in enumerate(zip(reqs, tiles, arrivals)):
        delay = ta - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        gw = widths[k] if widths is not None else 0
        if posts is not None:
            tickets.append(server.submit_layers(cia.ctx, tile, posts[k], grid_w=gw))
        else:
            tickets.append(server.submit(cia.ctx, tile, grid_w=gw))
    out = []
    for (cia, _), tile, tk, ta in zip(reqs, tiles, tickets, arrivals):
        info = server.wait(tk)
        arrival_ms = c0 + ta * 1e3
        out.append({
            "latency_ms": info["t_done_ms"] - min(arrival_ms, info["t_submit_ms"]),
            "queue_ms": info["t_first_ms"] - info["t_submit_ms"],
            "pixels": tile.n,
            "pixel_iters": int(tile.stop_iter.astype(np.int64).sum()),
            "t_done_ms": info["t_done_ms"], "t_submit_ms": info["t_submit_ms"],
            "tile": tile,
        })
    return out
