// Tile classes of the batched stencil launches and their kernel tables:
// one launch per stage covers every active case, each block a (case, tile)
// work item. Cases of many cells take the large tile, small grids the
// small one, where the large tile would leave most threads idle. Each
// class's kernels are instantiated in their own unit
// (bs_batch_tiles_large.hip, bs_batch_tiles_small.hip).
#pragma once

#include "bs_internal.hpp"

namespace bs {

constexpr int BTX_L = 64, BTY_L = 8;  // large tile
constexpr int BTX_S = 32, BTY_S = 4;  // small tile

using BatchTileFn = void (*)(const TileArgs*, const WorkItem*);

BatchTileFn batch_tile_fn_large(const bs_config& c, int mode);
BatchTileFn batch_tile_fn_small(const bs_config& c, int mode);

namespace {

template <int TX_, int TY_, int REC, int LIM, int FLUX>
BatchTileFn pick_batch_mode(int mode) {
  if (mode == TILE_STAGE) return batch_tile_kernel<TX_, TY_, REC, LIM, FLUX, TILE_STAGE>;
  if (mode == TILE_IMPLICIT) return batch_tile_kernel<TX_, TY_, REC, LIM, FLUX, TILE_IMPLICIT>;
  return batch_tile_kernel<TX_, TY_, REC, LIM, FLUX, TILE_RESIDUAL>;
}

template <int TX_, int TY_, int FLUX>
BatchTileFn pick_batch_rec(int rec, int lim, int mode) {
  if (rec == SHARED) return pick_batch_mode<TX_, TY_, SHARED, VAN_ALBADA, FLUX>(mode);
  if (rec == SHARED_SCALED) return pick_batch_mode<TX_, TY_, SHARED_SCALED, VAN_ALBADA, FLUX>(mode);
  switch (lim) {
    case FIRST_ORDER: return pick_batch_mode<TX_, TY_, PER_VARIABLE, FIRST_ORDER, FLUX>(mode);
    case MINMOD: return pick_batch_mode<TX_, TY_, PER_VARIABLE, MINMOD, FLUX>(mode);
    case VAN_LEER: return pick_batch_mode<TX_, TY_, PER_VARIABLE, VAN_LEER, FLUX>(mode);
    case VAN_ALBADA: return pick_batch_mode<TX_, TY_, PER_VARIABLE, VAN_ALBADA, FLUX>(mode);
    case MONOTONIZED_CENTRAL: return pick_batch_mode<TX_, TY_, PER_VARIABLE, MONOTONIZED_CENTRAL, FLUX>(mode);
    default: return pick_batch_mode<TX_, TY_, PER_VARIABLE, UNLIMITED_CENTRAL, FLUX>(mode);
  }
}

template <int TX_, int TY_>
BatchTileFn pick_batch_tile(const bs_config& c, int mode) {
  return c.flux == BS_FLUX_RUSANOV ? pick_batch_rec<TX_, TY_, RUSANOV>(c.reconstruction, c.limiter, mode)
                                   : pick_batch_rec<TX_, TY_, HLLC>(c.reconstruction, c.limiter, mode);
}

}  // namespace
}  // namespace bs
