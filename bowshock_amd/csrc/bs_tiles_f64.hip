// The FP64 state's stencil and wall-flux kernels (bs_tiles.hpp).
#include "bs_tiles.hpp"

namespace bs {
TileFnT<double> tile_fn_f64(const bs_config& c, int mode) { return pick_tile<double>(c, mode); }
WallFnT<double> wall_fn_f64(const bs_config& c) { return pick_wall<double>(c); }
}  // namespace bs
