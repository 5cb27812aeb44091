// tile_loop.hpp -- the persistent, work-stealing tile kernel, generic over a
// pixel model (State / init / step / finish).  Each lane owns one pixel at a
// time; when it stops, the lane writes its outputs and takes the next index
// from the global pixel cursor.  Lanes that need work in the same wave share
// one atomic: ballot + mbcnt hand out consecutive indices.
#pragma once
#include "fsg_internal.hpp"

namespace fsg {

struct Counters { unsigned long long plain, bla, bla_it, reb, tries; };

// --------------------------------------------------------------------------
// The persistent tile kernel
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <class Model>
__global__ void __launch_bounds__(256) k_tile(const KArgs a) {
    typename Model::State s;
    Counters k = {0, 0, 0, 0, 0};
    int64_t idx = 0;
    bool need = true;
    for (;;) {
        if (need) {
            // lanes that need a pixel take consecutive indices from one atomic
            const unsigned long long mask = __ballot(1);
            const int leader = __ffsll((long long)mask) - 1;
            const unsigned rank = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
            unsigned long long base = 0;
            if ((int)__lane_id() == leader)
                base = atomicAdd(a.work, (unsigned long long)__popcll(mask));
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)base, leader);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(base >> 32), leader);
            idx = (int64_t)(((unsigned long long)hi << 32) | lo) + rank;
            if (idx >= a.npts) break;
            Model::init(s, a, idx);
            need = false;
        }
        const int reason = Model::step(s, a, k);
        if (reason >= 0) {
            Model::finish(s, a, idx, reason);
            need = true;
        }
    }
    // all lanes have left the loop: one atomic per wave per counter
    unsigned long long c0 = wave_sum(k.plain), c1 = wave_sum(k.bla);
    unsigned long long c2 = wave_sum(k.bla_it), c3 = wave_sum(k.reb), c4 = wave_sum(k.tries);
    if (__lane_id() == 0) {
        atomicAdd(a.work + 1, c0); atomicAdd(a.work + 2, c1);
        atomicAdd(a.work + 3, c2); atomicAdd(a.work + 4, c3);
        atomicAdd(a.work + 5, c4);
    }
}

}  // namespace fsg
