// fsg_orbit.hip -- the calc_orbit instantiations of the tile kernel
// (perturbation.py:1096-1101, 1266-1273, 1390-1396), in a translation unit
// of their own so that they compile in parallel with the others.
#include "tile_models.hpp"

namespace fsg {

template <class Model> static tile_fn fn_of() { return k_tile<Model>; }

tile_fn select_tile_orbit(const LaunchCfg &c) {
    if (c.kind == KIND_M2) return c.xr ? fn_of<HoloXr<KIND_M2, 2, true>>() : fn_of<HoloStd<KIND_M2, 2, true>>();
    if (c.kind == KIND_MN) {
        switch (c.degree) {
#define FSG_MN(D) case D: return c.xr ? fn_of<HoloXr<KIND_MN, D, true>>() : fn_of<HoloStd<KIND_MN, D, true>>();
            FSG_MN(2) FSG_MN(3) FSG_MN(4) FSG_MN(5) FSG_MN(6) FSG_MN(7) FSG_MN(8)
#undef FSG_MN
            default: return nullptr;
        }
    }
    if (c.kind == KIND_BS) {
        switch (c.flavor) {
#define FSG_BS(F) case F: return c.xr ? fn_of<BsXr<F, true>>() : fn_of<BsStd<F, true>>();
            FSG_BS(1) FSG_BS(2) FSG_BS(3) FSG_BS(4) FSG_BS(5)
#undef FSG_BS
            default: return nullptr;
        }
    }
    return nullptr;
}

}  // namespace fsg
