/*
 * fsgpu.h -- C ABI of libfsgpu.so, the MI355X (gfx950) perturbation engine.
 *
 * The library replaces the per-tile numba call of the reference,
 *   PerturbationFractal.numba_cycle_call(cycle_dep_args, cycle_indep_args)
 *   (src/fractalshades/perturbation.py:414-428),
 * and the per-calculation preparation behind cycle_indep_args
 *   (get_cycle_indep_args, perturbation.py:431-562).
 * Paths are relative to the Fractalshades source tree.
 *
 * Conventions: every function returns 0 on success and a negative FSG_E_*
 * code on failure (message from fsg_last_error()); USER_INTERRUPTED (1, as
 * in perturbation.py:984) is returned when the interrupt flag was set.  No
 * torch or HIP types appear in the signatures; `stream` arguments are
 * hipStream_t passed as void*.  Host arrays are C-ordered numpy buffers.
 * All entry points are thread-safe; calls on one context are serialised.
 */
#ifndef FSGPU_H
#define FSGPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSG_ABI_VERSION 6

/* model kinds */
#define FSG_KIND_M2 0   /* Perturbation_mandelbrot, models/mandelbrot_M2.py:359-627 */
#define FSG_KIND_MN 1   /* Perturbation_mandelbrot_N, models/mandelbrot_Mn.py:387-735 */
#define FSG_KIND_BS 2   /* Perturbation_burning_ship, models/burning_ship.py:863-1130 */

/* precision of the delta iteration */
#define FSG_PREC_FP64 0   /* reference semantics: binary64, Xrange when xr_mode */
#define FSG_PREC_FP32X 1  /* binary32 mantissa + int32 exponent (not the drop-in) */

/* error codes */
#define FSG_OK 0
#define FSG_INTERRUPTED 1
#define FSG_PENDING 2    /* fsg_server_wait: not done within the timeout */
#define FSG_E_INVALID (-1)
#define FSG_E_HIP (-2)
#define FSG_E_NOMEM (-3)
#define FSG_E_UNSUPPORTED (-4)
#define FSG_E_RANGE (-5)   /* an output buffer is too small (fsg_ref_orbit) */

/*
 * Everything the reference gathers in cycle_indep_args
 * (get_cycle_indep_args, perturbation.py:431-562, path data from
 * get_path_data, perturbation.py:317-375) and the options that
 * calc_std_div bakes into `iterate` (mandelbrot_M2.py:432-588,
 * mandelbrot_Mn.py:462-625, burning_ship.py:971-1130).
 *
 * Xrange values are (mantissa, int32 exponent) pairs: value = m * 2^e.
 */
typedef struct fsg_ref_desc {
    int32_t kind;            /* FSG_KIND_* */
    int32_t degree;          /* exponent d for FSG_KIND_MN (2..8), 2 otherwise */
    int32_t flavor;          /* burning-ship flavour 1..5 (burning_ship.py:63-76) */
    int32_t xr_mode;         /* xr_detect_activated: dx < 1e-300 (perturbation.py:152-155) */
    int64_t ref_len;         /* L = len(Zn_path) */
    int64_t ref_div_iter;    /* first invalid orbit index (max_iter + 1 for cycles) */
    int64_t ref_order;       /* cycle order; (1 << 62) when the orbit is not a cycle */
    const double *Zn_path;   /* [2 L] interleaved complex128 */
    int32_t n_xr;            /* size of the Xrange register (FP_params["xr"]) */
    int32_t pad0;
    const int32_t *xr_index; /* [n_xr] sorted orbit indices */
    const double *xr_mant;   /* [2 n_xr]: complex mantissa (holomorphic) or (x, y) */
    const int32_t *xr_exp;   /* [n_xr] (holomorphic) or [2 n_xr] (x, y) */
    double dx_mant;          /* dx_xr = lin_scale_xr = derivative scale */
    int32_t dx_exp;
    int32_t pad1;
    double drift_mant[2];    /* centre - reference point: complex (holomorphic, */
    int32_t drift_exp[2];    /*   drift_exp[0]) or x / y reals (burning ship)   */
    double lin_mat[4];       /* 2x2 rotation/skew matrix, row major (core.py:1470-1484) */
    double kc_mant;          /* kc = bound of |c| over the frame (perturbation.py:166-194) */
    int32_t kc_exp;
    int32_t bla_on;          /* BLA_activated (mandelbrot_M2.py:521-524) */
    double bla_eps;          /* BLA_eps */
    int64_t max_iter;        /* stop_iter is int32: max_iter + ref_len must stay below 2^31
                                (FSG_E_INVALID otherwise) */
    double M_divergence;
    double epsilon_stationnary;
    int32_t calc_dzndc;      /* holomorphic calc_dzndc; burning ship calc_hessian */
    int32_t interior_detect; /* holomorphic calc_dzndz (interior detection) */
    int32_t precision;       /* FSG_PREC_* */
    int32_t calc_orbit;      /* keep the orbit point at exit - backshift (perturbation.py:1096-1101,
                                1266-1273, 1390-1396): one more complex field zn_orbit (holomorphic)
                                or two float64 fields xn_orbit, yn_orbit (burning ship); FP64 only */
    int32_t backshift;       /* > 0 when calc_orbit */
    int32_t pad2;
    double deriv_mant;       /* scale of the dz/dc paths, dx * projection.scale (scale_deriv_xr, */
    int32_t deriv_exp;       /*   perturbation.py:459-467); deriv_mant == 0: dx */
    int32_t pad3;
} fsg_ref_desc;

/* plain / BLA step counters of the last call, for the FLOP accounting */
typedef struct fsg_stats {
    uint64_t plain_steps;    /* full perturbation iterations */
    uint64_t bla_steps;      /* accepted BLA jumps */
    uint64_t bla_iters;      /* iterations skipped by those jumps */
    uint64_t rebases;        /* rebases (reference divergence + dynamic glitch) */
    uint64_t bla_lookups;    /* BLA table lookups (attempts, accepted or not) */
    double kernel_ms;        /* device time of the tile kernel (HIP events) */
    uint64_t redo_pixels;    /* FSG_PREC_FP32X: pixels recomputed in binary64 because a
                                decision of the binary32 loop fell within its error bound */
} fsg_stats;

typedef struct fsg_ctx fsg_ctx;

/* Version of this ABI (FSG_ABI_VERSION). */
int fsg_abi_version(void);

/*
 * Prepare a reference: uploads the orbit and Xrange register, computes the
 * orbit derivative paths (numba_dZndc_path / numba_dZndz_path /
 * numba_dZndc_path_BS, perturbation.py:2282-2516) and builds the BLA tree on
 * the GPU (numba_make_BLA / numba_make_BLA_BS, perturbation.py:1819-2105).
 * Replaces get_cycle_indep_args (perturbation.py:431-562).  Input arrays are
 * copied; the caller may free them on return.
 */
int fsg_ctx_create(const fsg_ref_desc *desc, fsg_ctx **out);

/*
 * A context on the same reference orbit as `base` (same model, orbit and
 * Xrange register, checked by a hash of the host arrays in `desc`): the
 * device orbit, register and BLA coefficients of `base` are shared, not
 * copied; the derivative paths (which scale with dx) and the BLA radii
 * (which depend on kc and BLA_eps) are the new context's own.  This is how
 * frames of different width and max_iter around one centre keep one
 * resident orbit and BLA table.  The shared data lives until the last
 * context using it is destroyed.  Results are identical to fsg_ctx_create's
 * with the same descriptor.  The reference rebuilds everything per
 * calculation (get_cycle_indep_args, perturbation.py:431-562).
 */
int fsg_ctx_create_shared(const fsg_ref_desc *desc, fsg_ctx *base, fsg_ctx **out);

/* Release a context and its device memory. */
int fsg_ctx_destroy(fsg_ctx *ctx);

/*
 * One tile, host buffers in and out: the drop-in for numba_cycles_perturb /
 * numba_cycles_perturb_BS (perturbation.py:988-1045, 1406-1451) behind
 * numba_cycle_call (perturbation.py:414-428), cycle_dep_args =
 * (c_pix, Z, U, stop_reason, stop_iter) of core.py:2044-2075:
 *   c_pix       [2 n] complex128 pixel offsets in units of dx
 *   Z           holomorphic: [n_Z, n] complex128 (zn, [dzndz], [dzndc], [zn_orbit]);
 *               burning ship: [n_Z, n] float64 (xn, yn, [4 Jacobian], [xn_orbit, yn_orbit])
 *   U           [n] int32 (ref_cycle_iter)
 *   stop_reason [n] int8
 *   stop_iter   [n] int32
 *   interrupted optional pointer to the reference's _interrupted flag
 *               (core.py:1352), polled while the tile runs: once it is set,
 *               the kernel takes no new pixel (the pixels in flight finish)
 *               and the pixels never taken keep stop_reason -1, their other
 *               fields unspecified
 *   stats       optional
 * Synchronous.  Returns FSG_INTERRUPTED when *interrupted was set.
 */
int fsg_tile(fsg_ctx *ctx, const double *c_pix, int64_t n, double *Z, int32_t *U,
             int8_t *stop_reason, int32_t *stop_iter, const volatile uint8_t *interrupted,
             fsg_stats *stats);

/*
 * Same computation on device buffers, enqueued on `stream` (hipStream_t) and
 * returning immediately.  `work` is caller-owned device memory of at least
 * fsg_work_bytes() bytes; it is cleared by the call and holds the pixel
 * cursor and the step counters (read them with fsg_work_stats after the
 * stream has completed).  No device allocation happens in this call.
 * Replaces the same reference entry point as fsg_tile.
 */
int fsg_tile_device(fsg_ctx *ctx, const double *d_c_pix, int64_t n, double *d_Z,
                    int32_t *d_U, int8_t *d_stop_reason, int32_t *d_stop_iter,
                    void *d_work, void *stream);

/*
 * fsg_tile and fsg_tile_device over a pixel grid: c_pix holds rows of grid_w
 * pixels (the layout of core.py:1767-1830 for a tile of width grid_w; the
 * last row may be short) and the kernel takes them in 8x8 blocks, so that
 * the 64 lanes of a wave iterate neighbouring pixels.  Results are the same
 * as fsg_tile's, at the same indices; only the order in which pixels are
 * computed changes.  grid_w = 0 is fsg_tile / fsg_tile_device.  Replaces the
 * same reference entry point as fsg_tile (its caller knows the tile's
 * width from chunk_slice, core.py:2044-2075).
 */
int fsg_tile_grid(fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w, double *Z,
                  int32_t *U, int8_t *stop_reason, int32_t *stop_iter,
                  const volatile uint8_t *interrupted, fsg_stats *stats);
int fsg_tile_grid_device(fsg_ctx *ctx, const double *d_c_pix, int64_t n, int64_t grid_w,
                         double *d_Z, int32_t *d_U, int8_t *d_stop_reason, int32_t *d_stop_iter,
                         void *d_work, void *stream);

/*
 * fsg_tile_grid without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture elsewhere in the process, serving loops that forbid
 * allocation after start-up).  The device staging of the inputs and outputs
 * and the pixel cursor and counters live in a static per-device array of
 * FSG_NOALLOC_SCRATCH_BYTES (32 MiB), held by every device the library runs
 * on until the process ends; its address is taken once per device.  A tile
 * needing more staging than that (about 340,000 pixels with three complex
 * fields) is rejected with FSG_E_RANGE.  Calls on one device share the array
 * and take turns (the call is synchronous); a call leaves its data there and
 * the next call overwrites it.  Results equal fsg_tile_grid's.
 */
#define FSG_NOALLOC_SCRATCH_BYTES (32u << 20)
int fsg_tile_grid_noalloc(fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w, double *Z,
                          int32_t *U, int8_t *stop_reason, int32_t *stop_iter,
                          const volatile uint8_t *interrupted, fsg_stats *stats);

/* Size of the `work` buffer of fsg_tile_device. No reference counterpart. */
size_t fsg_work_bytes(void);

/* Copy the counters of a completed fsg_tile_device call to the host. No reference counterpart. */
int fsg_work_stats(const void *d_work, fsg_stats *stats);

/*
 * Copy a derived table back to the host (for parity tests): which =
 *   0: BLA M (M_bla, bla_len x 4 or 8 float64)   1: BLA radii (r_bla, bla_len)
 *   2: dZndc path mantissas (2 L)   3: dZndc exponents (L, int32)
 *   4: dZndz mantissas (2 (L + 1))  5: dZndz exponents (L + 1, int32)
 *   6..9: burning-ship paths dXnda, dXndb, dYnda, dYndb (L float64)
 *   10..13: their exponents (L int32)
 * `bytes` is the size of `out`; returns FSG_E_INVALID if too small.  The
 * tables are the reference's M_bla / r_bla / dZndc_path / dZndz_path
 * (perturbation.py:1881, 2336, 2516); an exponent whose mantissa is zero
 * reads -2^26 (the kernels' zero code).  fsg_ctx_bla_shape gives the BLA size.
 */
int fsg_ctx_read_table(fsg_ctx *ctx, int which, void *out, size_t bytes);
int fsg_ctx_bla_shape(fsg_ctx *ctx, int64_t *bla_len, int32_t *stages_bla);

/* Last error message of the calling thread. */
const char *fsg_last_error(void);

/*
 * ---- Post-processing layers --------------------------------------------------
 * The continuous iteration count and the distance estimate of each pixel,
 * from the raw results while they are on the device: Continuous_iter_pp with
 * kind "infinity" (postproc.py:352-406, 1001-1009) and DEM_pp
 * (postproc.py:684-731), computed in binary64 and stored as float32 (the
 * reference's default post-processing dtype, settings.py:82).  Pixels whose
 * stop reason is not divergence get NaN.  The reference computes these on
 * the host from the copied Z array; here 8 bytes per pixel leave the device
 * instead of the raw fields.
 */
typedef struct fsg_post_desc {
    int32_t holomorphic;     /* 1: Z rows are complex128; 0: float64 (xn, yn, Jacobian) */
    int32_t dzndc_row;       /* row of dz/dc (holomorphic) or of dX/da (burning ship) */
    double d;                /* potential_d (the model's degree) */
    double a_d;              /* potential_a_d */
    double M;                /* potential_M (the escape radius) */
    double floor_iter;       /* Continuous_iter_pp floor_iter */
    double px_snap;          /* DEM_pp px_snap; <= 0 for none */
    double skew[4];          /* the frame's skew (s00, s01, s10, s11); all zero = none */
    /* The projection's factors of the derivatives, applied before the DEM and
     * the normal.  First the loop's dzndc_modifier (perturbation.py:1387-1388,
     * 1773-1777; projection.py:205-217, 454-471), which the tile kernels leave
     * out: mod 0 none; 1 |w| * mod_scale (exponential map); 2 |w + 1e-6| *
     * mod_scale (the Cartesian expmap seam), w the pixel's projected offset.
     * Then the layers' map (get_dzndc / apply_df / apply_dfBS, postproc.py:
     * 184-206, 972-998; projection.py:375-453), with t = pix_to_ht * pix at the
     * screen pixel: proj 0 none; 1 exp(t); 2 exp(Re t); 3 exp(i Im t); 4 one;
     * the burning ship's Jacobian takes the matching 2x2 matrix.  The device
     * forms them from w = exp(h + t): exp(t) = w * w_scale, exp(i Im t) = w / |w|. */
    int32_t proj;
    int32_t mod;
    double w_scale;          /* exp(-h_moy) of the whole map (maps 1 and 2) */
    double mod_scale;
} fsg_post_desc;

/* Enqueue the layers of n pixels (Z rows of stride ld >= n) on stream;
 * nu, dem or normal may be null to skip that layer.  normal: float32 [n][2],
 * the unit normal of the potential map (Attr_normal_pp kind "potential",
 * postproc.py:572-628: z_n / (dz_n/dc), for the burning ship J^T z_n; then
 * unskewed with the transposed skew, core.py:3147-3158, and normalised).
 * d_c_pix: the tile's pixels as the tile kernel took them (complex128 [n]:
 * the projected offsets), read when pp->proj or pp->mod is not 0 (else may be
 * null). */
int fsg_postproc_device(const fsg_post_desc *pp, const double *d_Z, int64_t n, int64_t ld,
                        const double *d_c_pix, const int32_t *d_stop_iter, const int8_t *d_stop_reason,
                        float *d_nu, float *d_dem, float *d_normal, void *stream);

/*
 * ---- Serving loop ----------------------------------------------------------
 * No counterpart in the reference, which renders the tiles of one frame with
 * a thread pool (mthreading.py:48-68).  A server takes tile requests of mixed
 * size, context (centre, zoom depth, max_iter, model, precision) and arrival
 * time; it admits them in arrival order as its device-memory budget allows,
 * splits them into chunks and runs chunks of different requests concurrently
 * on `lanes` HIP streams.  Results are written to the caller's host buffers
 * (pinned memory from fsg_host_alloc makes the copies asynchronous).  A
 * request's results are identical to fsg_tile's for the same inputs.  The
 * contexts must stay alive until their requests are waited for.
 */
typedef struct fsg_server fsg_server;

typedef struct fsg_server_opts {
    int32_t lanes;        /* chunks in flight, one HIP stream each (1..16; 0 = 3) */
    int32_t chunk_px;     /* pixels per chunk launch (0 = 65536; rounded up to 256) */
    int64_t hbm_budget;   /* device bytes for admitted requests' buffers (0 = 4 GiB) */
} fsg_server_opts;

typedef struct fsg_request_info {
    int32_t status;       /* 0 or the FSG_E_* code of the failure */
    int32_t chunks;       /* launches the request was split into */
    int64_t footprint;    /* device bytes it held while admitted */
    double t_submit_ms;   /* times on the server clock (fsg_server_clock) */
    double t_admit_ms;    /* device buffers granted, input copy enqueued */
    double t_first_ms;    /* first chunk launched */
    double t_done_ms;     /* results in the caller's host buffers */
} fsg_request_info;

int fsg_server_create(const fsg_server_opts *opts, fsg_server **out);
/* Finishes every submitted request, then releases the server. */
int fsg_server_destroy(fsg_server *srv);
/* Device bytes a request of n pixels on ctx needs (admission footprint). */
int fsg_server_footprint(const fsg_ctx *ctx, int64_t n, int64_t *bytes);
/* Milliseconds since the server started. */
int fsg_server_clock(fsg_server *srv, double *ms);
/* Queue a request (same buffers as fsg_tile); returns at once with a ticket.
 * grid_w > 0: the pixels are rows of grid_w (as for fsg_tile_grid) and each
 * chunk is iterated in 8x8 blocks of its rows -- the same results at the
 * same indices, with less divergence; 0: index order.
 * FSG_E_NOMEM if the request alone exceeds the budget. */
int fsg_server_submit(fsg_server *srv, fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w,
                      double *Z, int32_t *U, int8_t *stop_reason, int32_t *stop_iter, uint64_t *ticket);
/* The same request returning the post-processing layers (nu, dem: float32
 * [n]; normal: float32 [n][2]; any may be null) with stop_reason and
 * stop_iter instead of the raw fields: the raw results stay on the device. */
int fsg_server_submit_layers(fsg_server *srv, fsg_ctx *ctx, const double *c_pix, int64_t n, int64_t grid_w,
                             const fsg_post_desc *pp, float *nu, float *dem, float *normal,
                             int8_t *stop_reason, int32_t *stop_iter, uint64_t *ticket);
/* Wait for a request (timeout_ms < 0: no limit).  Returns its status, or
 * FSG_PENDING on timeout; after a return other than FSG_PENDING the ticket
 * is released.  Several threads may wait on one ticket at the same time:
 * all of them return the request's status (and info); a wait that starts
 * after the ticket was released fails with FSG_E_INVALID. */
int fsg_server_wait(fsg_server *srv, uint64_t ticket, int32_t timeout_ms, fsg_request_info *info);

/* Pinned host memory for request buffers. */
void *fsg_host_alloc(size_t bytes);
int fsg_host_free(void *p);

/*
 * ---- Reference orbit ------------------------------------------------------
 * The arbitrary-precision orbit of the reference point, from the critical
 * point 0, rounded to complex128, with an Xrange copy of the iterates below
 * 1e-300.  Replaces
 *   perturbation_mandelbrotN_select_FP_loop (mpmath_utils/FP_loop.pyx:274-421)
 *   perturbation_nonholomorphic_FP_loop     (FP_loop.pyx:1828-1979)
 * which iterate in MPFR/MPC from decimal seed strings at seed_prec bits.
 * Here the caller passes c already rounded to its precision, as two
 * fixed-point values: a sign and a magnitude of n_limbs little-endian 64-bit
 * limbs, of which the low frac_limbs are the fraction (value = +-magnitude *
 * 2^(-64 frac_limbs)).  64 * frac_limbs >= prec + 128 keeps each operation's
 * error far below the reference's half ulp at prec bits; the integer limbs
 * must hold M^degree (2 M^2 for the burning ship).  Runs on the host.
 */
typedef struct fsg_orbit_desc {
    int32_t kind;          /* FSG_KIND_M2 / FSG_KIND_MN / FSG_KIND_BS */
    int32_t degree;        /* z^d + c (MN): 2..64 */
    int32_t flavor;        /* burning ship: 1..5 (burning_ship.py:63-76 order) */
    int32_t need_xr;       /* keep Xrange copies of iterates below 1e-300 */
    int32_t n_limbs;       /* limbs per value */
    int32_t frac_limbs;    /* fraction limbs among them */
    int32_t c_re_neg, c_im_neg;
    const uint64_t *c_re;  /* [n_limbs] */
    const uint64_t *c_im;  /* [n_limbs] */
    int64_t max_iter;      /* iterates 1..max_iter (FP_loop's max_iter) */
    double M;              /* escape radius: stop once |Z_i| > M */
    int32_t n_threads;     /* host threads sharing each product (0: automatic) */
    int32_t pad0;
} fsg_orbit_desc;

typedef struct fsg_orbit_out {
    double *Zn;            /* complex128 [max_iter + 1] (caller's); Zn[0] = 0, 0 past the escape */
    int64_t xr_cap;        /* capacity of the register arrays */
    int64_t *xr_index;     /* [xr_cap] iteration of each register entry */
    double *xr_mant;       /* [xr_cap][2]: holomorphic (re, im) on one exponent; BS (x, y) */
    int32_t *xr_exp;       /* [xr_cap][2]: holomorphic (e, e); BS (x_e, y_e) */
    int64_t n_xr;          /* out: entries of the register (FSG_E_RANGE when > xr_cap) */
    int64_t div_iter;      /* out: first invalid index (escape, or max_iter + 1) */
} fsg_orbit_out;

int fsg_ref_orbit(const fsg_orbit_desc *d, fsg_orbit_out *o);

/*
 * Period of the reference point by the ball method: the first i <= max_iter
 * with |z_i / (dz_i/dc)| < px (burning ship: the 2x2 solve of (x_i, y_i)
 * against the Jacobian in (a, b), both components and their hypot below px),
 * or -1 when |z_i| > M first or no i qualifies.  Replaces
 *   perturbation_mandelbrotN_select_ball_method (FP_loop.pyx:631-758)
 *   perturbation_nonholomorphic_ball_method     (FP_loop.pyx:2357-2462).
 * px = px_mant * 2^px_exp.  The orbit is iterated in fixed point as in
 * fsg_ref_orbit; the derivatives in block floating point (fixed-point
 * mantissas sharing a binary exponent); the ratio test in binary64 with a
 * wide exponent (the reference tests it at 54 bits).  d->M is the escape
 * radius M_divergence; d->max_iter the iteration limit.
 */
int fsg_ball_period(const fsg_orbit_desc *d, double px_mant, int64_t px_exp, int64_t *order);

/* Multi-precision values handed back to the caller:
 * value[i] = (neg[i] ? -1 : 1) * limbs[i] * 2^(exp[i] - 64 * frac_limbs),
 * limbs[i] the n_limbs little-endian words of the i-th magnitude. */
typedef struct fsg_mp_values {
    int64_t cap;           /* capacity, in values */
    uint64_t *limbs;       /* [cap][n_limbs] */
    int32_t *neg;          /* [cap] */
    int64_t *exp;          /* [cap] */
    int64_t n;             /* out: values written */
} fsg_mp_values;

/*
 * The orbit and its derivative at c after `order` steps, for one Newton step
 * of the nucleus search; the caller forms the step at its own precision.
 * Holomorphic: (Re z, Im z, Re dz/dc, Im dz/dc) after step `order`, and,
 * with `divisors`, first after every step i that divides order (the
 * divisor-removal Newton of perturbation_mandelbrot_find_nucleus,
 * FP_loop.pyx:900-1116); burning ship: (x, y, dx/da, dx/db, dy/da, dy/db).
 * Replaces the inner loops of those Newton searches
 * (FP_loop.pyx:900-1116, 1159-1329, 2564-2752).  FSG_E_RANGE when the orbit
 * leaves the fixed-point range.
 */
int fsg_newton_eval(const fsg_orbit_desc *d, int64_t order, int32_t divisors, fsg_mp_values *out);

#ifdef __cplusplus
}
#endif
#endif /* FSGPU_H */
