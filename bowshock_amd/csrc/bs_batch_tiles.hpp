// Tile classes of the batched stencil launches and their kernel tables:
// one launch per stage covers every active case, each block a (case, tile)
// work item. Cases of many cells take the large tile, small grids the
// small one, where the large tile would leave most threads idle. Each
// class's kernels are instantiated, per state precision, in a unit of
// their own (bs_batch_tiles_{large,small}{,_f64}.hip).
#pragma once

#include "bs_internal.hpp"

namespace bs {

// Tile classes per state precision; the large FP64 tile is the single
// FP64 solver's.
template <typename T>
struct BatchTile {
  static constexpr int lx = 32, ly = 8;  // large tile (the single FP32 solver's)
  static constexpr int sx = 32, sy = 4;  // small tile
};
template <>
struct BatchTile<double> {
  static constexpr int lx = 32, ly = 8;
  static constexpr int sx = 32, sy = 4;
};

template <typename T>
using BatchTileFnT = void (*)(const TileArgsT<T>*, const WorkItem*);

BatchTileFnT<float> batch_tile_fn_large(const bs_config& c, int mode);
BatchTileFnT<float> batch_tile_fn_small(const bs_config& c, int mode);
BatchTileFnT<double> batch_tile_fn_large_f64(const bs_config& c, int mode);
BatchTileFnT<double> batch_tile_fn_small_f64(const bs_config& c, int mode);

template <typename T>
BatchTileFnT<T> batch_tile_fn(const bs_config& c, int cls, int mode) {
  if constexpr (std::is_same<T, float>::value)
    return cls == 0 ? batch_tile_fn_large(c, mode) : batch_tile_fn_small(c, mode);
  else
    return cls == 0 ? batch_tile_fn_large_f64(c, mode) : batch_tile_fn_small_f64(c, mode);
}

namespace {

template <typename T, int TX_, int TY_, int REC, int LIM, int FLUX>
BatchTileFnT<T> pick_batch_mode(int mode) {
  if constexpr (std::is_same<T, float>::value) {
    if (mode == TILE_STAGE) return batch_tile_kernel_occ8<TX_, TY_, REC, LIM, FLUX, TILE_STAGE>;
  } else {
    if (mode == TILE_STAGE) return batch_tile_kernel<T, TX_, TY_, REC, LIM, FLUX, TILE_STAGE>;
  }
  if (mode == TILE_IMPLICIT) return batch_tile_kernel<T, TX_, TY_, REC, LIM, FLUX, TILE_IMPLICIT>;
  return batch_tile_kernel<T, TX_, TY_, REC, LIM, FLUX, TILE_RESIDUAL>;
}

template <typename T, int TX_, int TY_, int FLUX>
BatchTileFnT<T> pick_batch_rec(int rec, int lim, int mode) {
  if (rec == SHARED) return pick_batch_mode<T, TX_, TY_, SHARED, VAN_ALBADA, FLUX>(mode);
  if (rec == SHARED_SCALED) return pick_batch_mode<T, TX_, TY_, SHARED_SCALED, VAN_ALBADA, FLUX>(mode);
  switch (lim) {
    case FIRST_ORDER: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, FIRST_ORDER, FLUX>(mode);
    case MINMOD: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, MINMOD, FLUX>(mode);
    case VAN_LEER: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, VAN_LEER, FLUX>(mode);
    case VAN_ALBADA: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, VAN_ALBADA, FLUX>(mode);
    case MONOTONIZED_CENTRAL: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, MONOTONIZED_CENTRAL, FLUX>(mode);
    default: return pick_batch_mode<T, TX_, TY_, PER_VARIABLE, UNLIMITED_CENTRAL, FLUX>(mode);
  }
}

template <typename T, int TX_, int TY_>
BatchTileFnT<T> pick_batch_tile(const bs_config& c, int mode) {
  return c.flux == BS_FLUX_RUSANOV ? pick_batch_rec<T, TX_, TY_, RUSANOV>(c.reconstruction, c.limiter, mode)
                                   : pick_batch_rec<T, TX_, TY_, HLLC>(c.reconstruction, c.limiter, mode);
}

}  // namespace
}  // namespace bs
