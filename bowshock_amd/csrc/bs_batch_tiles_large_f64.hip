// The batch's large-tile stencil kernels of an FP64 state (bs_batch_tiles.hpp).
#include "bs_batch_tiles.hpp"

namespace bs {
BatchTileFnT<double> batch_tile_fn_large_f64(const bs_config& c, int mode) {
  return pick_batch_tile<double, BatchTile<double>::lx, BatchTile<double>::ly>(c, mode);
}
}  // namespace bs
