// The batch's large-tile stencil kernels of an FP32 state (bs_batch_tiles.hpp).
#include "bs_batch_tiles.hpp"

namespace bs {
BatchTileFnT<float> batch_tile_fn_large(const bs_config& c, int mode) {
  return pick_batch_tile<float, BatchTile<float>::lx, BatchTile<float>::ly>(c, mode);
}
}  // namespace bs
