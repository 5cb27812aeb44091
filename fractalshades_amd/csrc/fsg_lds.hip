// fsg_lds.hip -- the tile kernels with the reference tables staged in LDS
// (k_tile_lds), in a translation unit of their own so that they compile in
// parallel with the others.
#include "tile_models.hpp"

namespace fsg {

template <class Model> static tile_fn fn_lds(bool full) {
    return full ? k_tile_lds<Model, true> : k_tile_lds<Model, false>;
}
// z^d + c of degree 6..8 with extended-range deltas needs more than the 128
// VGPRs a 1024-lane block allows (it spilled): 512-lane blocks
template <class Model> static tile_fn fn_lds512(bool full) {
    return full ? k_tile_lds<Model, true, 512> : k_tile_lds<Model, false, 512>;
}

int tile_lds_block(const LaunchCfg &c) { return (c.kind == KIND_MN && c.xr && c.degree >= 6) ? 512 : 1024; }

// z^2 + c in extended range, the headline case: one instantiation per
// combination of derivatives (der: bit 0 dz/dc, bit 1 dz/dz)
static tile_fn fn_lds_m2xr(bool full, int der) {
    switch (der) {
        case 0: return fn_lds<HoloXr<KIND_M2, 2, false, 0>>(full);
        case 1: return fn_lds<HoloXr<KIND_M2, 2, false, 1>>(full);
        case 2: return fn_lds<HoloXr<KIND_M2, 2, false, 2>>(full);
        default: return fn_lds<HoloXr<KIND_M2, 2, false, 3>>(full);
    }
}

// the same with binary64 deltas
static tile_fn fn_lds_m2std(bool full, int der) {
    switch (der) {
        case 0: return fn_lds<HoloStd<KIND_M2, 2, false, 0>>(full);
        case 1: return fn_lds<HoloStd<KIND_M2, 2, false, 1>>(full);
        case 2: return fn_lds<HoloStd<KIND_M2, 2, false, 2>>(full);
        default: return fn_lds<HoloStd<KIND_M2, 2, false, 3>>(full);
    }
}

tile_fn select_tile_lds(const LaunchCfg &c, bool full, int der) {
    if (tile_lds_block(c) == 512) {
        switch (c.degree) {
            case 6: return fn_lds512<HoloXr<KIND_MN, 6>>(full);
            case 7: return fn_lds512<HoloXr<KIND_MN, 7>>(full);
            case 8: return fn_lds512<HoloXr<KIND_MN, 8>>(full);
            default: return nullptr;
        }
    }
    if (c.kind == KIND_M2) return c.xr ? fn_lds_m2xr(full, der) : fn_lds_m2std(full, der);
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return c.xr ? fn_lds<HoloXr<KIND_MN, D>>(full) : fn_lds<HoloStd<KIND_MN, D>>(full);
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    if (c.kind == KIND_BS) {
        switch (c.flavor) {
#define FSG_BS(F) case F: return c.xr ? fn_lds<BsXr<F>>(full) : fn_lds<BsStd<F>>(full);
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return nullptr;
        }
    }
    return nullptr;
}

}  // namespace fsg
