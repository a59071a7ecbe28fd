// Kernel tables of the single-grid stencil and wall-flux kernels: one
// instance per scheme (reconstruction, limiter, flux) and tile mode.
// Included by bs_tiles_f32.hip and bs_tiles_f64.hip, each of which
// instantiates the kernels of one state precision.
#pragma once

#include "bs_internal.hpp"

namespace bs {
namespace {

template <typename T, int REC, int LIM, int FLUX>
TileFnT<T> pick_mode(int mode) {
  constexpr int X = Tile<T>::x, Y = Tile<T>::y;
  if constexpr (std::is_same<T, float>::value) {
    if (mode == TILE_STAGE) return tile_kernel_occ8<X, Y, REC, LIM, FLUX, TILE_STAGE>;
  } else {
    if (mode == TILE_STAGE) return tile_kernel<T, X, Y, REC, LIM, FLUX, TILE_STAGE>;
  }
  if (mode == TILE_IMPLICIT) return tile_kernel<T, X, Y, REC, LIM, FLUX, TILE_IMPLICIT>;
  return tile_kernel<T, X, Y, REC, LIM, FLUX, TILE_RESIDUAL>;
}

template <typename T, int FLUX>
TileFnT<T> pick_rec(int rec, int lim, int mode) {
  if (rec == SHARED) return pick_mode<T, SHARED, VAN_ALBADA, FLUX>(mode);
  if (rec == SHARED_SCALED) return pick_mode<T, SHARED_SCALED, VAN_ALBADA, FLUX>(mode);
  switch (lim) {
    case FIRST_ORDER: return pick_mode<T, PER_VARIABLE, FIRST_ORDER, FLUX>(mode);
    case MINMOD: return pick_mode<T, PER_VARIABLE, MINMOD, FLUX>(mode);
    case VAN_LEER: return pick_mode<T, PER_VARIABLE, VAN_LEER, FLUX>(mode);
    case VAN_ALBADA: return pick_mode<T, PER_VARIABLE, VAN_ALBADA, FLUX>(mode);
    case MONOTONIZED_CENTRAL: return pick_mode<T, PER_VARIABLE, MONOTONIZED_CENTRAL, FLUX>(mode);
    default: return pick_mode<T, PER_VARIABLE, UNLIMITED_CENTRAL, FLUX>(mode);
  }
}

template <typename T>
TileFnT<T> pick_tile(const bs_config& c, int mode) {
  return c.flux == BS_FLUX_RUSANOV ? pick_rec<T, RUSANOV>(c.reconstruction, c.limiter, mode)
                                   : pick_rec<T, HLLC>(c.reconstruction, c.limiter, mode);
}

template <typename S>
WallFnT<S> pick_wall(const bs_config& c) {
  if (c.flux == BS_FLUX_RUSANOV) {
    if (c.reconstruction == SHARED) return wall_flux_kernel<S, SHARED, RUSANOV>;
    if (c.reconstruction == SHARED_SCALED) return wall_flux_kernel<S, SHARED_SCALED, RUSANOV>;
    return wall_flux_kernel<S, PER_VARIABLE, RUSANOV>;
  }
  if (c.reconstruction == SHARED) return wall_flux_kernel<S, SHARED, HLLC>;
  if (c.reconstruction == SHARED_SCALED) return wall_flux_kernel<S, SHARED_SCALED, HLLC>;
  return wall_flux_kernel<S, PER_VARIABLE, HLLC>;
}

}  // namespace
}  // namespace bs
