// Kernel table of the persistent batch kernel (bs_persistent.hpp): one
// instance per scheme (reconstruction, limiter, flux) and tile shape.
#include "bs_persistent.hpp"

namespace bs {
namespace {

template <int TX_, int TY_, int FLUX>
PersistFn pick_persist_rec(int rec, int lim) {
  if (rec == SHARED) return batch_persistent_kernel<TX_, TY_, SHARED, VAN_ALBADA, FLUX>;
  if (rec == SHARED_SCALED) return batch_persistent_kernel<TX_, TY_, SHARED_SCALED, VAN_ALBADA, FLUX>;
  switch (lim) {
    case FIRST_ORDER: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, FIRST_ORDER, FLUX>;
    case MINMOD: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, MINMOD, FLUX>;
    case VAN_LEER: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, VAN_LEER, FLUX>;
    case VAN_ALBADA: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, VAN_ALBADA, FLUX>;
    case MONOTONIZED_CENTRAL: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, MONOTONIZED_CENTRAL, FLUX>;
    default: return batch_persistent_kernel<TX_, TY_, PER_VARIABLE, UNLIMITED_CENTRAL, FLUX>;
  }
}

template <int TX_, int TY_>
PersistFn pick_persist(const bs_config& c) {
  return c.flux == BS_FLUX_RUSANOV ? pick_persist_rec<TX_, TY_, RUSANOV>(c.reconstruction, c.limiter)
                                   : pick_persist_rec<TX_, TY_, HLLC>(c.reconstruction, c.limiter);
}

}  // namespace

PersistFn batch_persistent_fn(const bs_config& c, int tile) {
  return tile == 1 ? pick_persist<32, 4>(c) : pick_persist<32, 8>(c);
}

void persistent_tile_shape(int tile, int* tx, int* ty) {
  *tx = 32;
  *ty = tile == 1 ? 4 : 8;
}

}  // namespace bs
