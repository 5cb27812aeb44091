// fsg_internal.hpp -- structures shared by the kernels and the C-ABI layer.
#pragma once
#include "xr.hpp"

namespace fsg {

enum Kind { KIND_M2 = 0, KIND_MN = 1, KIND_BS = 2 };

// Tables::zn_k codes below this one name an Xrange register slot
constexpr int32_t ZK_REGISTER = -(1 << 28);

// Workspace of a tile launch (KArgs::work), 64-bit words: [1..5] step
// counters, [6] unsure pixels, then the pixel cursors, one per XCD, each in
// a 128-byte line of its own: word WORK_CURSORS + 16 x for the first pass,
// WORK_CURSORS + 128 + 16 x for the binary64 pass over the unsure pixels.
constexpr int WORK_CURSORS = 16;
constexpr int WORK_WORDS = WORK_CURSORS + 2 * 8 * 16;
// LDS the tile loop declares itself (its counters' block sums), kept out of
// the budget for staged tables
constexpr int TILE_STATIC_LDS = 1024;

// Read-only per-reference data, resident in HBM for the life of a context.
struct Tables {
    const cplx *Zn;            // [L] reference orbit Z_n rounded to binary64
    int32_t L, div_iter, order, first_invalid;
    // holomorphic orbit derivatives: dZn/dc [L], dZn/dz [L+1]; in Xrange
    // mode the exponents sit in the parallel *_e arrays
    const cplx *dc; const int32_t *dc_e;
    const cplx *dz; const int32_t *dz_e;
    // Burning-ship Jacobian paths dX/da, dX/db, dY/da, dY/db [L]
    const double *bs[4]; const int32_t *bs_e[4];
    // sparse Xrange register of orbit points below 1e-300 (sorted indices)
    int32_t n_xr; int32_t zn_k_w;
    const int32_t *xr_idx; const double *xr_m; const int32_t *xr_e;
    // orbit points in the extended-range kernels' normalised form, zn_k_w
    // codes per point [zn_k_w L]: holomorphic, one per point (the complex
    // value); burning ship, two (x, then y).  A code k >= ZK_REGISTER: the
    // binary64 value scaled by 2^-k is the normalised mantissa (larger
    // component in [1, 2); k is the exponent of zero for a zero value);
    // k < ZK_REGISTER (in the first code of a point): the point is the
    // Xrange register's slot k - INT32_MIN
    const int32_t *zn_k;
    // BLA tree: 4 doubles (A, B complex) or 8 doubles (2x2 A, 2x2 B) per entry
    const double *M; const double *r;
    int32_t stages_bla; int32_t bla_on;
};

// The same tables in the binary32-mantissa + int32-exponent format of the
// FSG_PREC_FP32X kernels (12-byte {float re, float im, int32 exp} records).
struct TablesX {
    const void *Zx;            // [L] orbit
    const void *DCx;           // [L] dZn/dc
    const void *DZx;           // [L+1] dZn/dz
    const void *Ax, *Bx;       // [bla_len] BLA coefficients
    // burning ship: orbit as (x, y) pairs and the 4 Jacobian paths, each a
    // binary32 mantissa with an int32 exponent (8 bytes)
    const void *ZRx;           // [L] pairs
    const void *BSx[4];        // [L]
};

// Layout of the reference tables staged in LDS for one launch (byte
// offsets; -1 = not staged).  K orbit points [0, K) cover every index the
// loop reads (K = min(L, div_iter, order)); dZndz's wrapped value [L] sits
// at [K]; KB BLA entries cover every jump starting below K.
struct LdsPlan {
    int32_t K, KB, bytes, bla_dim;
    int32_t zn, dc, dc_e, dz, dz_e, bs[4], bs_e[4], zn_k, r, M;
};

// The binary32 form's tables staged in LDS (byte offsets, -1 = read from
// global memory): orbit (binary64, for the escape tests), BLA radii, and the
// xs orbit, derivative paths and BLA coefficients.  full: all of them; else
// the xs orbit and the BLA tables only.
struct LdsPlanX { int32_t K, KB, bytes, full, zn, r, zx, dcx, dzx, ax, bx; };

// One launch of the tile kernel.
struct KArgs {
    Tables t;
    TablesX tx;
    double lin[4];             // rotation/skew matrix (lin_mat)
    double dx_m; int32_t dx_e; int32_t pad1;
    double drift_m[2]; int32_t drift_e[2];
    int32_t max_iter; int32_t calc_dc; int32_t calc_dz; int32_t pad2;
    double M2, eps2;
    // pixels: complex128 screen offsets in units of dx ("c_pix")
    const double *pix;
    int64_t npts;
    // outputs, reference layout: Z [n_Z, ld] (row stride ld >= npts, so that
    // a launch can fill a column range of a larger tile), U/stop_iter int32,
    // stop_reason int8
    double *Z; int32_t *U; int8_t *stop_reason; int32_t *stop_iter;
    int64_t ld;
    // > 0: the pixels are rows of grid_w (the last row may be short), taken
    // in 8x8 blocks (grid_slot); 0: taken in index order
    int64_t grid_w;
    int32_t nz; int32_t max_blocks;   // grid cap: 0 fill the GPU, > 0 blocks, -k 1/k of the GPU
    int32_t calc_orbit, backshift;    // orbit point at exit - backshift in the last field(s)
    LdsPlan lds;                      // lds.bytes == 0: tables read from global memory
    LdsPlanX ldsx;                    // the same for the binary32 form
    // workspace (WORK_WORDS words): [1..5] step counters (plain, bla, bla
    // iters, rebase, bla lookups), [6] pixels the binary32 form left unsure,
    // the pixel cursors from WORK_CURSORS on
    unsigned long long *work;
    // 1: the binary64 pass over the pixels whose stop_reason the binary32
    // pass set to REASON_UNSURE (every other pixel is skipped)
    int32_t redo; int32_t pad3;
    // host-visible stop flag (pinned host memory) or null: once it is
    // nonzero, lanes take no new pixel; the pixels in flight finish
    const uint32_t *stop;
};

struct LaunchCfg { int kind, degree, flavor, xr, precision, orbit; };

hipError_t launch_tile(const LaunchCfg &cfg, const KArgs &a, hipStream_t stream);
// the calc_orbit instantiations (fsg_orbit.hip) and the LDS-staged ones
// (fsg_lds.hip, 1024-lane blocks)
typedef void (*tile_fn)(const KArgs);
tile_fn select_tile_orbit(const LaunchCfg &cfg);
// der: the derivatives the launch computes (bit 0 dz/dc, bit 1 dz/dz)
tile_fn select_tile_lds(const LaunchCfg &cfg, bool full, int der);
int tile_lds_block(const LaunchCfg &cfg);        // lanes per block of select_tile_lds's kernel
// true when the plan stages every table the launch reads
bool lds_plan_full(const Tables &t, const LdsPlan &p);
// radii_only: M already holds the coefficients (shared with another
// context of the same orbit); write the radii r only
hipError_t launch_bla_build(int kind, int degree, int flavor, const cplx *Zn, int32_t L,
                            double kc, double eps, double *M, double *r, int32_t stages,
                            hipStream_t stream, bool radii_only = false);
int tile_kernel_blocks(const LaunchCfg &cfg);
int grid_fill_blocks(const void *kernel);          // blocks filling the GPU at full occupancy
int fill_blocks(const void *kernel, int block, int lds_bytes);   // the same for any block / LDS size
// plan the LDS staging of a launch's tables (bytes = 0 when they do not fit)
LdsPlan plan_lds(const Tables &t, int bla_dim, int max_bytes);
int launch_blocks(int fill, const KArgs &a, int block = 256);   // capped by max_blocks and the pixel count
int max_lds_bytes();                // LDS bytes a block may use (FSG_LDS_BYTES caps it)

// FSG_PREC_FP32X kernels (fsg_fp32x.hip)
constexpr size_t XS_BYTES = 12;
hipError_t build_xs_tables(const Tables &t, int64_t bla_len, TablesX &x, hipStream_t stream);
hipError_t build_xsr_tables(const Tables &t, TablesX &x, hipStream_t stream);   // burning ship
constexpr size_t XSR_BYTES = 8;
hipError_t launch_tile_f32(const LaunchCfg &cfg, const KArgs &a, hipStream_t stream);
bool f32_supported(const LaunchCfg &cfg);

}  // namespace fsg
