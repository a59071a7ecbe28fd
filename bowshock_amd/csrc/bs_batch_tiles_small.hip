// The batch's small-tile (32 x 4) stencil kernels (bs_batch_tiles.hpp).
#include "bs_batch_tiles.hpp"

namespace bs {
BatchTileFn batch_tile_fn_small(const bs_config& c, int mode) { return pick_batch_tile<BTX_S, BTY_S>(c, mode); }
}  // namespace bs
