// Kernel table of the dataflow stage kernel (bs_dataflow.hpp): one instance
// per scheme (reconstruction, limiter, flux) on the FP32 stencil tile.
#include "bs_dataflow.hpp"

namespace bs {
namespace {

template <int FLUX>
DataflowFn pick_dataflow_rec(int rec, int lim) {
  if (rec == SHARED) return dataflow_stage_kernel<TX, TY, SHARED, VAN_ALBADA, FLUX>;
  if (rec == SHARED_SCALED) return dataflow_stage_kernel<TX, TY, SHARED_SCALED, VAN_ALBADA, FLUX>;
  switch (lim) {
    case FIRST_ORDER: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, FIRST_ORDER, FLUX>;
    case MINMOD: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, MINMOD, FLUX>;
    case VAN_LEER: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, VAN_LEER, FLUX>;
    case VAN_ALBADA: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, VAN_ALBADA, FLUX>;
    case MONOTONIZED_CENTRAL: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, MONOTONIZED_CENTRAL, FLUX>;
    default: return dataflow_stage_kernel<TX, TY, PER_VARIABLE, UNLIMITED_CENTRAL, FLUX>;
  }
}

}  // namespace

DataflowFn dataflow_fn(const bs_config& c) {
  return c.flux == BS_FLUX_RUSANOV ? pick_dataflow_rec<RUSANOV>(c.reconstruction, c.limiter)
                                   : pick_dataflow_rec<HLLC>(c.reconstruction, c.limiter);
}

}  // namespace bs
