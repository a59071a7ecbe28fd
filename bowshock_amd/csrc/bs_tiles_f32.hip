// The FP32 state's stencil and wall-flux kernels (bs_tiles.hpp).
#include "bs_tiles.hpp"

namespace bs {
TileFnT<float> tile_fn_f32(const bs_config& c, int mode) { return pick_tile<float>(c, mode); }
WallFnT<float> wall_fn_f32(const bs_config& c) { return pick_wall<float>(c); }
}  // namespace bs
