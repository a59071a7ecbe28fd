// The batch's large-tile (64 x 8) stencil kernels (bs_batch_tiles.hpp).
#include "bs_batch_tiles.hpp"

namespace bs {
BatchTileFn batch_tile_fn_large(const bs_config& c, int mode) { return pick_batch_tile<BTX_L, BTY_L>(c, mode); }
}  // namespace bs
