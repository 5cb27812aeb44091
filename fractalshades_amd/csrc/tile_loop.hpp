// tile_loop.hpp -- the persistent, work-stealing tile kernel, generic over a
// pixel model (State / init / step / finish).  Each lane owns one pixel at a
// time; when it stops, the lane writes its outputs and takes the next index
// from a pixel cursor.  Lanes that need work in the same wave share one
// atomic: ballot + mbcnt hand out consecutive indices.
#pragma once
#include "fsg_internal.hpp"

namespace fsg {

struct Counters { unsigned long long plain, bla, bla_it, reb, tries; };

// Stop reason of a pixel whose binary32 result may differ from the binary64
// loop's (a decision taken within the pixel's error bound): set by the
// FSG_PREC_FP32X kernels, replaced by a second, binary64 pass over exactly
// those pixels (KArgs::redo).  Never left in a finished result.
constexpr int REASON_UNSURE = 3;


// --------------------------------------------------------------------------
// Pixel order over a grid of rows of w pixels (n = h w + a short last row).
// Slot s of the cursor maps to one pixel, a bijection of [0, n): first the
// 8x8 blocks of the top (h & ~7) rows and left (w & ~7) columns, block after
// block along the rows, each block row-major; then the right strip of those
// rows, row-major; then the remaining bottom rows.  A wave that refills 64
// consecutive slots then holds a square of neighbouring pixels, whose
// iteration counts and BLA paths are closer to each other than those of 64
// pixels of one row (2048x2048 frames: 3-7% less kernel time).
template <class T>
__host__ __device__ __forceinline__ T grid_slot_t(T s, T n, T w) {
    const T h8 = (n / w) & ~(T)7, w8 = w & ~(T)7;
    const T a = h8 * w8;
    if (s < a) {
        const T blk = s >> 6, r = s & 63, per_row = w8 >> 3;
        const T by = blk / per_row, bx = blk - by * per_row;
        return ((by << 3) + (r >> 3)) * w + (bx << 3) + (r & 7);
    }
    s -= a;
    const T rem = w - w8, strip = h8 * rem;
    if (s < strip) {
        const T row = s / rem;
        return row * w + w8 + (s - row * rem);
    }
    return h8 * w + (s - strip);
}

__device__ __forceinline__ int64_t grid_slot(int64_t s, int64_t n, int64_t w) {
    if (n <= 0xffffffffll)
        return (int64_t)grid_slot_t<uint32_t>((uint32_t)s, (uint32_t)n, (uint32_t)w);
    return (int64_t)grid_slot_t<uint64_t>((uint64_t)s, (uint64_t)n, (uint64_t)w);
}

// --------------------------------------------------------------------------
// BLA stage selection (ref_BLA_get, perturbation.py:2116-2170).  The tree
// keeps the stage s >= 3 jump starting at iteration w (a multiple of 2^s) at
// index 2 (w >> 3) + 2^(s-3) - 1.  The reference scans the stages from the
// highest one down and takes the first with 2^s < first_invalid - w and
// |z| < r.  Both conditions hold on a prefix of the stages: the first because
// 2^s grows with s, the second because a merged entry's radius is the
// minimum of its first half's radius and the new bound (combine_BLA,
// perturbation.py:2020-2023), so r never grows with the stage for a fixed
// start.  The highest valid stage is therefore found by bisection, with
// about log2(stages) dependent loads instead of one per stage.
// Returns the jump length (0 if no stage is valid) and the entry index.
// SCALED: |z| = az 2^ez (the modulus of an extended-range mantissa and its
// exponent); each radius is scaled by 2^-ez instead, which is exact.
template <bool SCALED = false>
__device__ __forceinline__ int32_t bla_find(const Tables &t, double az, int32_t w, int32_t &idx,
                                            int32_t ez = 0) {
    const int32_t it = w >> 3;
    const int32_t invalid_step = t.first_invalid - w;
    if (invalid_step <= 8) return 0;
    const int32_t tz = (it == 0) ? 31 : __builtin_ctz((uint32_t)it);
    const int32_t s_inv = 31 - __builtin_clz((uint32_t)(invalid_step - 1));   // 2^s < invalid_step
    int32_t hi = min(min(3 + tz, t.stages_bla - 1), s_inv);
    int32_t lo = 3, best = 0;
    if (hi >= 3) {
        // the lowest stage first: invalid there, invalid at every stage (the
        // prefix property), which settles 40 % of the lookups of the 1e-300
        // frame in one read instead of about four (C2 -2.8 %)
        const double r3 = t.r[2 * it];
        if (!(az < (SCALED ? ldexp(r3, -ez) : r3))) return 0;
        best = 3; lo = 4;
    }
    while (lo <= hi) {
        const int32_t mid = (lo + hi) >> 1;
        const double r = t.r[2 * it + (1 << (mid - 3)) - 1];
        if (az < (SCALED ? ldexp(r, -ez) : r)) { best = mid; lo = mid + 1; }
        else hi = mid - 1;
    }
    if (best == 0) return 0;
    idx = 2 * it + (1 << (best - 3)) - 1;
    return 1 << best;
}

// --------------------------------------------------------------------------
// The persistent tile kernel
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// A lane whose pixel stops goes idle; the results are written and new
// pixels taken only once REFILL_BATCH lanes of the wave are idle (or all of
// them).  Writing a pixel's results and starting the next one costs about as
// much as a few iterations plus two dependent memory round trips (the
// cursor atomic, then the pixel offsets), and under SIMT a wave pays it for
// every trip on which any lane switches -- about one switch per trip when
// each lane switches on its own.  Batching the switches trades idle lanes for
// that cost; measured on the 2048x2048 1e-300 frame: batch 8 -> 2.51 ms,
// 16 -> 2.19, 24 -> 2.09, 32 -> 2.02, 40 -> 1.96, 48 -> 1.96.  Re-measured
// once the refill's atomics were spread over the XCDs and the step's tests
// taken together, against 40 (1e-300 frame / interior-dominated frame at
// 5e-601 / z^3 + c at 1e-800): 24: +1 / +15 / +5 %, 32: +1 / +7 / +2 %,
// 48: -2 / -5 / -1 %, 56: -2 / -9 / -2.5 %, 60: -3 / -10 / -2 %; 62 and 64
// (a refill only once every lane is idle) within 1 % of 60.
constexpr int REFILL_BATCH = 60;

template <class Model>
__device__ __forceinline__ void tile_loop(const KArgs &a) {
    typename Model::State s;
    Counters k = {0, 0, 0, 0, 0};
    unsigned long long unsure = 0;
    int64_t idx = -1;
    int reason = 0;
    // Pixel cursors, one per XCD.  A launch's blocks go to the 8 XCDs round
    // robin, and an atomic on one address from every XCD is serialised across
    // the dies: with one cursor all 4096 waves of a launch queue on it at the
    // start (about 0.3 ms) and at every refill after.  The slots form groups
    // of 2^gsh (at most 16384 -- 256 8x8 blocks -- and at least 64, with 8
    // groups or more when the launch has 512 slots or more); cursor x hands
    // out the groups g with g % 8 == x, in order.  A wave draws from its own
    // XCD's cursor, then, once that one is drained, from the next ones in turn
    // (work stealing), so that every slot is taken exactly once whatever the
    // block-to-XCD placement.  cur is uniform over the wave.  Group sizes
    // measured against one cursor: 64 slots -5 %, 4096 -7 %, 16384 -7 % on
    // the 2048x2048 frame at 1e-300, -15 % on z^3 + c at 1e-800; the
    // interior-dominated frame lost 9 %, 3.5 % and 0.8 %.
    unsigned long long *const cursors = a.work + WORK_CURSORS + (a.redo ? 128 : 0);
    const unsigned xcd = blockIdx.x & 7u;
    const int gsh = min(14, max(6, 63 - __clzll((unsigned long long)max<int64_t>(a.npts >> 3, 1))));
    unsigned cur = 0;
    bool active = false;      // iterating a pixel
    bool pending = false;     // holding a stopped pixel whose results are not written yet
    bool exhausted = false;   // every cursor is drained for this lane
    for (;;) {
        const unsigned long long act = __ballot(active);
        // idle lanes with something to do at a refill: a stopped pixel to
        // write, or cursors not all found drained.  Once there are none, the
        // wave only steps its remaining pixels (a launch's last, longest
        // ones) without the refill's ballots, and leaves the loop when none
        // is left: every result is then written.
        const unsigned long long want = __ballot(!active && (pending || !exhausted));
        if (act == 0ull && want == 0ull) break;
        if (want != 0ull && (act == 0ull || 64 - __popcll(act) >= REFILL_BATCH)) {
            bool drained = false;     // this lane found the current cursor drained
            if (!active) {
                if (pending) {
                    Model::finish(s, a, idx, reason);
                    unsure += (reason == REASON_UNSURE);
                    pending = false;
                }
                if (!exhausted) {
                    // idle lanes take consecutive indices from one atomic
                    const unsigned long long mask = __ballot(1);
                    const int leader = __ffsll((long long)mask) - 1;
                    const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                        (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
                    const unsigned cx = (xcd + cur) & 7u;
                    unsigned long long base = 0;
                    unsigned stop = 0;
                    if ((int)__lane_id() == leader) {
                        base = atomicAdd(cursors + 16 * cx, (unsigned long long)__popcll(mask));
                        // the interrupt flag, read beside the cursor (its latency overlaps)
                        if (a.stop) stop = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    }
                    const unsigned lo = __builtin_amdgcn_readlane((unsigned)base, leader);
                    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(base >> 32), leader);
                    const int64_t loc = (int64_t)(((unsigned long long)hi << 32) | lo) + rank;
                    idx = ((((loc >> gsh) << 3) + (int64_t)cx) << gsh) + (loc & ((1ll << gsh) - 1));
                    if (__builtin_amdgcn_readlane(stop, leader) != 0u) {
                        exhausted = true;          // interrupted: no new pixel
                    } else if (idx < a.npts) {
                        if (a.grid_w > 0) idx = grid_slot(idx, a.npts, a.grid_w);
                        // the binary64 pass takes only the pixels the binary32 one left unsure
                        if (!a.redo || a.stop_reason[idx] == REASON_UNSURE) {
                            Model::init(s, a, a.pix[2 * idx], a.pix[2 * idx + 1]);
                            active = true;
                        }
                    } else {
                        drained = true;
                        // cur >= 7: the wave has found every cursor drained
                        // (cur counts the cursors found drained, in turn)
                        exhausted = (cur >= 7u);
                    }
                }
            }
            if (__ballot(drained) != 0ull) cur++;
            // lanes still idle after the refill (the binary64 pass can draw a
            // whole batch of slots without an unsure pixel) refill again on
            // the next trip
        }
        if (active) {
            reason = Model::step(s, a, k);
            if (reason >= 0) {
                active = false;
                pending = true;
            }
        }
    }
    // All lanes have left the loop.  The block's waves add up their counters
    // in LDS and six lanes add the block's sums to the workspace: one atomic
    // per counter per block.  One per counter per wave queued 4096 waves on
    // six addresses across the dies as they finished (about 0.15 ms of a
    // 1.4 ms launch).  The waves of a persistent launch have nothing left to
    // do once they leave the loop, so waiting at the barrier costs nothing.
    __shared__ unsigned long long part[16][6];
    static_assert(sizeof(part) <= TILE_STATIC_LDS, "static LDS of the tile loop");
    const unsigned long long c[6] = {wave_sum(k.plain), wave_sum(k.bla), wave_sum(k.bla_it),
                                     wave_sum(k.reb), wave_sum(k.tries), wave_sum(unsure)};
    const int wv = (int)(threadIdx.x >> 6);
    if (__lane_id() == 0) {
#pragma unroll
        for (int q = 0; q < 6; q++) part[wv][q] = c[q];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        unsigned long long sum = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) sum += part[w][threadIdx.x];
        if (sum) atomicAdd(a.work + 1 + threadIdx.x, sum);
    }
}


template <class Model>
__global__ void __launch_bounds__(256) k_tile(const KArgs a) {
    tile_loop<Model>(a);
}

// The same loop with the reference tables staged in LDS: the loop is bound
// by the latency of its table reads (orbit point, derivative paths, BLA
// radius and coefficients at the lane's own iteration index), and an LDS
// read answers in a fraction of an L2 hit's time.  One 1024-lane block per
// CU (or more when they fit) copies the tables once, then runs the loop.
// FULL: the plan stages every table the launch has, so each table pointer
// is either null or an LDS address, and the compiler reads the tables with
// LDS instructions instead of generic (flat) ones
template <bool FULL, class T>
__device__ __forceinline__ const T *stage(unsigned char *lds, int32_t off, const T *src, int64_t n) {
    if constexpr (FULL) {
        if (src == nullptr || off < 0) return nullptr;   // off < 0: a table the launch never reads
    } else {
        if (off < 0 || src == nullptr) return src;
    }
    T *dst = reinterpret_cast<T *>(lds + off);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
    return dst;
}

template <class Model, bool FULL, int BLOCK = 1024>
__global__ void __launch_bounds__(BLOCK) k_tile_lds(const KArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const LdsPlan &p = a.lds;
    KArgs al = a;
    Tables &t = al.t;
    const int64_t K = p.K, KB = p.KB;
    t.Zn = stage<FULL>(lds, p.zn, a.t.Zn, K);
    t.dc = stage<FULL>(lds, p.dc, a.t.dc, K);
    t.dc_e = stage<FULL>(lds, p.dc_e, a.t.dc_e, K);
    if ((FULL || p.dz >= 0) && a.t.dz) {
        cplx *dz = reinterpret_cast<cplx *>(lds + p.dz);
        for (int64_t i = threadIdx.x; i < K; i += blockDim.x) dz[i] = a.t.dz[i];
        if (threadIdx.x == 0) dz[K] = a.t.dz[a.t.L];
        t.dz = dz;
    } else if constexpr (FULL) {
        t.dz = nullptr;
    }
    if ((FULL || p.dz_e >= 0) && a.t.dz_e) {
        int32_t *dze = reinterpret_cast<int32_t *>(lds + p.dz_e);
        for (int64_t i = threadIdx.x; i < K; i += blockDim.x) dze[i] = a.t.dz_e[i];
        if (threadIdx.x == 0) dze[K] = a.t.dz_e[a.t.L];
        t.dz_e = dze;
    } else if constexpr (FULL) {
        t.dz_e = nullptr;
    }
    for (int q = 0; q < 4; q++) {
        t.bs[q] = stage<FULL>(lds, p.bs[q], a.t.bs[q], K);
        t.bs_e[q] = stage<FULL>(lds, p.bs_e[q], a.t.bs_e[q], K);
    }
    t.zn_k = stage<FULL>(lds, p.zn_k, a.t.zn_k, K * a.t.zn_k_w);
    t.r = stage<FULL>(lds, p.r, a.t.r, KB);
    t.M = stage<FULL>(lds, p.M, a.t.M, KB * p.bla_dim);
    if ((FULL || p.dz >= 0) && a.t.dz) t.L = (int32_t)K;   // the wrapped dZndz value is staged at [K]
    __syncthreads();
    tile_loop<Model>(al);
}

}  // namespace fsg
