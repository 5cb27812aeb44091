// fsg_kernels.hip -- gfx950 kernels for the per-pixel perturbation loop.
//
// k_tile: persistent, work-stealing tile kernel.  Each lane owns one pixel at
// a time and runs the reference's per-pixel loop (BLA jump, perturbation step,
// max_iter / interior / escape tests, rebasing); when its pixel stops it
// writes the results and takes the next pixel index from the pixel cursors.
// Refill is wave-aggregated: one atomic per wave hands out consecutive indices
// to all lanes that need work (ballot + mbcnt).  The arithmetic is scalar
// complex recurrences (no MFMA); the kernel is bound by the FP64 VALU rate.
//
// Reference loops restated here:
//   holomorphic: numba_iterate, src/fractalshades/perturbation.py:1065-1400
//   burning ship: numba_iterate_BS, perturbation.py:1468-1791
//   model kernels: models/mandelbrot_M2.py:599-622, mandelbrot_Mn.py:646-729,
//                  burning_ship.py:12-60, 440-859
//   BLA: numba_make_BLA* / combine_BLA* / ref_BLA_get, perturbation.py:1819-2170
#include "tile_models.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <atomic>
#include <unordered_map>

namespace fsg {

// --------------------------------------------------------------------------
// BLA build: level 0 + three compressed levels per block of 8 iterations,
// then one launch per stored level; every merge of a level is independent.

struct HEntry { cplx A, B; double r; };

// combine_BLA, perturbation.py:1983-2024: e1 covers [i, ii), e2 [ii, i + 2 step)
FSG_HD HEntry merge(const HEntry &e1, const HEntry &e2, double kc, double eps) {
    HEntry o;
    o.A = e2.A * e1.A;
    o.B = e2.A * e1.B + e2.B;
    double mA1 = hypot2(e1.A.re, e1.A.im), mB1 = hypot2(e1.B.re, e1.B.im);
    double num = e2.r - mB1 * kc;
    double den = (eps > mA1) ? eps : mA1;
    double q = num / den;
    double r2b = 0.95 * (q > 0. ? q : 0.);
    o.r = (r2b < e1.r) ? r2b : e1.r;
    return o;
}

struct BEntry { double M[8]; double r; };

// combine_BLA_BS, perturbation.py:2027-2105
FSG_HD BEntry merge(const BEntry &e1, const BEntry &e2, double kc, double eps) {
    const double *a = e1.M, *b = e2.M;
    BEntry o;
    o.M[0] = b[0] * a[0] + b[1] * a[2];
    o.M[1] = b[0] * a[1] + b[1] * a[3];
    o.M[2] = b[2] * a[0] + b[3] * a[2];
    o.M[3] = b[2] * a[1] + b[3] * a[3];
    o.M[4] = b[0] * a[4] + b[1] * a[6] + b[4];
    o.M[5] = b[0] * a[5] + b[1] * a[7] + b[5];
    o.M[6] = b[2] * a[4] + b[3] * a[6] + b[6];
    o.M[7] = b[2] * a[5] + b[3] * a[7] + b[7];
    double mA1 = fabs(a[0]), mB1 = fabs(a[4]);
    for (int q = 1; q < 4; q++) {
        mA1 = fabs(a[q]) > mA1 ? fabs(a[q]) : mA1;
        mB1 = fabs(a[4 + q]) > mB1 ? fabs(a[4 + q]) : mB1;
    }
    double num = e2.r - mB1 * kc;
    double den = (eps > mA1) ? eps : mA1;
    double q = num / den;
    double r2b = 0.95 * (q > 0. ? q : 0.);
    o.r = (r2b < e1.r) ? r2b : e1.r;
    return o;
}

__device__ __forceinline__ HEntry load_h(const double *M, const double *r, int64_t i) {
    HEntry e; e.A = mk(M[4 * i], M[4 * i + 1]); e.B = mk(M[4 * i + 2], M[4 * i + 3]); e.r = r[i];
    return e;
}
__device__ __forceinline__ void store_h(double *M, double *r, int64_t i, const HEntry &e) {
    M[4 * i] = e.A.re; M[4 * i + 1] = e.A.im; M[4 * i + 2] = e.B.re; M[4 * i + 3] = e.B.im; r[i] = e.r;
}
__device__ __forceinline__ BEntry load_b(const double *M, const double *r, int64_t i) {
    BEntry e;
    for (int q = 0; q < 8; q++) e.M[q] = M[8 * i + q];
    e.r = r[i];
    return e;
}
__device__ __forceinline__ void store_b(double *M, double *r, int64_t i, const BEntry &e) {
    for (int q = 0; q < 8; q++) M[8 * i + q] = e.M[q];
    r[i] = e.r;
}

// RADII_ONLY: the coefficients M are already in place (a context sharing
// another's orbit and BLA coefficients); only the radii, which depend on kc
// and eps, are written
template <int KIND, int DEG, bool RADII_ONLY>
__global__ void k_bla0_holo(const cplx *Zn, int32_t comp, double kc, double eps, double *M, double *r) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= comp) return;
    HEntry e[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        cplx A = dfdz<KIND, DEG>(Zn[8 * i + j]);
        e[j].A = A; e[j].B = mk(1., 0.); e[j].r = eps * hypot2(A.re, A.im);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
#pragma unroll
    for (int j = 0; j < 2; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    const HEntry o = merge(e[0], e[1], kc, eps);
    if constexpr (RADII_ONLY) r[2 * (int64_t)i] = o.r;
    else store_h(M, r, 2 * (int64_t)i, o);
}

template <int FL, bool RADII_ONLY>
__global__ void k_bla0_bs(const cplx *Zn, int32_t comp, double kc, double eps, double *M, double *r) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= comp) return;
    BEntry e[8];
    for (int j = 0; j < 8; j++) {
        cplx Z = Zn[8 * i + j];
        bs_jac<FL>(Z.re, Z.im, e[j].M);
        e[j].M[4] = 1.; e[j].M[5] = 0.; e[j].M[6] = 0.; e[j].M[7] = -1.;
        double ax = fabs(Z.re), ay = fabs(Z.im);
        e[j].r = eps * (ax < ay ? ax : ay);
    }
    for (int j = 0; j < 4; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    for (int j = 0; j < 2; j++) e[j] = merge(e[2 * j], e[2 * j + 1], kc, eps);
    const BEntry o = merge(e[0], e[1], kc, eps);
    if constexpr (RADII_ONLY) r[2 * (int64_t)i] = o.r;
    else store_b(M, r, 2 * (int64_t)i, o);
}

// one stored level: merge i in [0, floor(comp / 2^stg)) of stage stg
template <bool BS, bool RADII_ONLY>
__global__ void k_bla_merge(int32_t comp, int32_t stg, double kc, double eps, double *M, double *r) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t step = 1LL << stg;
    if (t >= comp / step) return;
    const int64_t i = t * step, ii = i + step / 2;
    const int64_t i1 = 2 * i + (step / 2 - 1), i2 = 2 * ii + (step / 2 - 1), ir = 2 * i + (step - 1);
    if constexpr (BS) {
        const BEntry o = merge(load_b(M, r, i1), load_b(M, r, i2), kc, eps);
        if constexpr (RADII_ONLY) r[ir] = o.r; else store_b(M, r, ir, o);
    } else {
        const HEntry o = merge(load_h(M, r, i1), load_h(M, r, i2), kc, eps);
        if constexpr (RADII_ONLY) r[ir] = o.r; else store_h(M, r, ir, o);
    }
}

// --------------------------------------------------------------------------
// dispatch
template <class Model> static tile_fn fn_of() { return k_tile<Model>; }

static tile_fn select_tile(const LaunchCfg &c) {
    if (c.orbit) return select_tile_orbit(c);
    if (c.kind == KIND_M2) return c.xr ? fn_of<HoloXr<KIND_M2, 2>>() : fn_of<HoloStd<KIND_M2, 2>>();
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return c.xr ? fn_of<HoloXr<KIND_MN, D>>() : fn_of<HoloStd<KIND_MN, D>>();
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    if (c.kind == KIND_BS) {
        switch (c.flavor) {
#define FSG_BS(F) case F: return c.xr ? fn_of<BsXr<F>>() : fn_of<BsStd<F>>();
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return nullptr;
        }
    }
    return nullptr;
}

// blocks that fill the GPU at the kernel's occupancy, cached per kernel
// Blocks of `block` lanes with `lds` bytes of dynamic LDS that fill the
// current device at full occupancy, or -1.  Cached per (kernel, device,
// block, LDS bytes) under a lock: the occupancy query costs far more than a
// launch of a small serving chunk, and callers come from several threads.
int fill_blocks(const void *fn, int block, int lds) {
    struct Key {
        const void *fn; int dev, block, lds;
        bool operator==(const Key &o) const { return fn == o.fn && dev == o.dev && block == o.block && lds == o.lds; }
    };
    struct Hash {
        size_t operator()(const Key &k) const {
            return std::hash<const void *>()(k.fn) ^ ((size_t)k.dev << 48) ^ ((size_t)k.block << 32) ^ (size_t)k.lds;
        }
    };
    static std::mutex mu;
    static std::unordered_map<Key, int, Hash> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    const Key key{fn, dev, block, lds};
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, (size_t)lds) != hipSuccess) return -1;
    // without dynamic LDS one block per CU always fits; with it, none may
    const int blocks = per_cu > 0 ? ncu * per_cu : (lds == 0 ? ncu : -1);
    cache[key] = blocks;
    return blocks;
}

int grid_fill_blocks(const void *fn) { return fill_blocks(fn, 256, 0); }

int tile_kernel_blocks(const LaunchCfg &cfg) {
    tile_fn f = select_tile(cfg);
    if (!f) return -1;
    return grid_fill_blocks(reinterpret_cast<const void *>(f));
}

int launch_blocks(int fill, const KArgs &a, int block) {
    int blocks = fill;
    if (a.max_blocks > 0 && a.max_blocks < blocks) blocks = a.max_blocks;
    if (a.max_blocks < 0) blocks = (fill + (-a.max_blocks) - 1) / (-a.max_blocks);   // a share of the GPU
    const int64_t need = (a.npts + block - 1) / block;
    if (need < blocks) blocks = (int)(need > 0 ? need : 1);
    return blocks;
}

int max_lds_bytes() {
    // per device, queried once (the first caller on a device stores it; a
    // concurrent first caller stores the same value)
    static std::atomic<int> per_dev[64] = {};
    int dev = 0, cached = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    cached = per_dev[dev].load(std::memory_order_relaxed) - 1;
    if (cached < 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 0;
        v = std::max(0, v - TILE_STATIC_LDS);     // the tile loop's own LDS
        per_dev[dev].store(v + 1, std::memory_order_relaxed);
        cached = v;
    }
    // FSG_LDS_BYTES caps the staging budget (to compare partial plans)
    const char *cap = std::getenv("FSG_LDS_BYTES");
    if (cap) {
        const long c = std::strtol(cap, nullptr, 10);
        if (c >= 0 && c < cached) return (int)c;
    }
    return cached;
}

// Byte offsets of the tables staged in LDS (-1: read from global memory).
// Entries 0 .. K-1 of each staged table (K = the iterations a pixel can
// reach, min(L, div_iter + 1, order)), with the wrapped dZndz value at [K].
// The tables are taken in the order the loop reads them most -- orbit,
// normalisation codes / Xrange slots, dZndz, dZndc (burning ship: its four paths), BLA radii and
// coefficients -- as long as they fit; bytes = 0 when not even the orbit does.
LdsPlan plan_lds(const Tables &t, int bla_dim, int max_bytes) {
    LdsPlan p;
    std::memset(&p, 0xff, sizeof(p));               // every offset -1
    int64_t K = std::min<int64_t>(t.L, std::min<int64_t>((int64_t)t.div_iter + 1, t.order));
    K = std::max<int64_t>(K, std::min<int64_t>(t.L, 2));   // Zn[1] (orbit finish)
    const int64_t bla_len = 2 * (int64_t)(t.L >> 3);
    const int64_t KB = t.bla_on ? std::min<int64_t>(bla_len, K / 4 + 2) : 0;
    int64_t off = 0;
    const auto al = [](int64_t b) { return (b + 15) & ~(int64_t)15; };
    // reserve a group of tables if it fits; returns false (nothing taken) otherwise
    const auto group = [&](std::initializer_list<int64_t> sizes) {
        int64_t total = 0;
        for (int64_t b : sizes) total += al(b);
        if (off + total > max_bytes || off + total >= INT32_MAX) return false;
        return true;
    };
    const auto take = [&](int64_t bytes) { int64_t o = off; off += al(bytes); return (int32_t)o; };
    if (!group({16 * K})) { p.K = (int32_t)K; p.KB = 0; p.bla_dim = bla_dim; p.bytes = 0; return p; }
    p.zn = take(16 * K);
    if (t.zn_k && group({4 * K * t.zn_k_w})) p.zn_k = take(4 * K * t.zn_k_w);
    if (t.dz && group({16 * (K + 1), t.dz_e ? 4 * (K + 1) : 0})) {
        p.dz = take(16 * (K + 1));
        if (t.dz_e) p.dz_e = take(4 * (K + 1));
    }
    if (t.dc && group({16 * K, t.dc_e ? 4 * K : 0})) {
        p.dc = take(16 * K);
        if (t.dc_e) p.dc_e = take(4 * K);
    }
    if (t.bs[0]) {
        const int64_t eb = t.bs_e[0] ? 4 * K : 0;
        if (group({8 * K, eb, 8 * K, eb, 8 * K, eb, 8 * K, eb})) {
            for (int q = 0; q < 4; q++) {
                if (t.bs[q]) p.bs[q] = take(8 * K);
                if (t.bs_e[q]) p.bs_e[q] = take(4 * K);
            }
        }
    }
    if (KB > 0 && group({8 * KB, 8 * bla_dim * KB})) { p.r = take(8 * KB); p.M = take(8 * bla_dim * KB); }
    // else the radii alone: the stage bisection reads them in a dependent
    // chain, the coefficients once per jump
    else if (KB > 0 && group({8 * KB})) p.r = take(8 * KB);
    p.K = (int32_t)K; p.KB = (int32_t)KB; p.bla_dim = bla_dim;
    p.bytes = (int32_t)off;
    return p;
}

bool lds_plan_full(const Tables &t, const LdsPlan &p) {
    bool ok = p.zn >= 0 && (!t.zn_k || p.zn_k >= 0) && (!t.dz || p.dz >= 0) &&
              (!t.dz_e || p.dz_e >= 0) && (!t.dc || p.dc >= 0) && (!t.dc_e || p.dc_e >= 0) &&
              (!t.bla_on || (p.r >= 0 && p.M >= 0));
    for (int q = 0; q < 4; q++) ok = ok && (!t.bs[q] || p.bs[q] >= 0) && (!t.bs_e[q] || p.bs_e[q] >= 0);
    return ok;
}

hipError_t launch_tile(const LaunchCfg &cfg, const KArgs &a, hipStream_t stream) {
    // FSG_TABLES=global keeps the tables in global memory (for comparisons)
    const char *tables = std::getenv("FSG_TABLES");
    const bool no_lds = tables && !std::strcmp(tables, "global");
    if (!cfg.orbit && !no_lds) {
        KArgs al = a;
        al.lds = plan_lds(a.t, cfg.kind == KIND_BS ? 8 : 4, max_lds_bytes());
        const int der = (a.calc_dc ? 1 : 0) | (a.calc_dz ? 2 : 0);
        tile_fn f = al.lds.bytes > 0 ? select_tile_lds(cfg, lds_plan_full(a.t, al.lds), der) : nullptr;
        const int block = tile_lds_block(cfg);
        const int fill = f ? fill_blocks(reinterpret_cast<const void *>(f), block, al.lds.bytes) : -1;
        if (fill > 0) {
            hipLaunchKernelGGL(f, dim3(launch_blocks(fill, al, block)), dim3(block), (size_t)al.lds.bytes,
                               stream, al);
            return hipGetLastError();
        }
    }
    tile_fn f = select_tile(cfg);
    if (!f) return hipErrorInvalidValue;
    const int fill = grid_fill_blocks(reinterpret_cast<const void *>(f));
    if (fill <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(f, dim3(launch_blocks(fill, a)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

template <bool RO>
static hipError_t bla_build_t(int kind, int degree, int flavor, const cplx *Zn, int32_t L,
                              double kc, double eps, double *M, double *r, int32_t stages,
                              hipStream_t stream) {
    const int32_t comp = L >> 3;
    if (comp <= 0) return hipErrorInvalidValue;
    const int tb = 128;
    dim3 g0((comp + tb - 1) / tb);
    if (kind == KIND_M2) hipLaunchKernelGGL((k_bla0_holo<KIND_M2, 2, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r);
    else if (kind == KIND_MN) {
        switch (degree) {
#define FSG_MN(D) case D: hipLaunchKernelGGL((k_bla0_holo<KIND_MN, D, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r); break;
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (flavor) {
#define FSG_BS(F) case F: hipLaunchKernelGGL((k_bla0_bs<F, RO>), g0, dim3(tb), 0, stream, Zn, comp, kc, eps, M, r); break;
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return hipErrorInvalidValue;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int32_t stg = 1; stg < stages - 3; stg++) {
        int64_t nm = comp >> stg;
        if (nm <= 0) break;
        dim3 g((unsigned)((nm + tb - 1) / tb));
        if (kind == KIND_BS) hipLaunchKernelGGL((k_bla_merge<true, RO>), g, dim3(tb), 0, stream, comp, stg, kc, eps, M, r);
        else hipLaunchKernelGGL((k_bla_merge<false, RO>), g, dim3(tb), 0, stream, comp, stg, kc, eps, M, r);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_bla_build(int kind, int degree, int flavor, const cplx *Zn, int32_t L,
                            double kc, double eps, double *M, double *r, int32_t stages,
                            hipStream_t stream, bool radii_only) {
    return radii_only ? bla_build_t<true>(kind, degree, flavor, Zn, L, kc, eps, M, r, stages, stream)
                      : bla_build_t<false>(kind, degree, flavor, Zn, L, kc, eps, M, r, stages, stream);
}

}  // namespace fsg
