// The batch's small-tile stencil kernels of an FP32 state (bs_batch_tiles.hpp).
#include "bs_batch_tiles.hpp"

namespace bs {
BatchTileFnT<float> batch_tile_fn_small(const bs_config& c, int mode) {
  return pick_batch_tile<float, BatchTile<float>::sx, BatchTile<float>::sy>(c, mode);
}
}  // namespace bs
