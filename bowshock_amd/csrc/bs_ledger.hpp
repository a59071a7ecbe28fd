// The step ledger (StepLedger.account, src/engine-gpu/time-loop.ts:125-150):
// what a batch of steps left in the control block (its status word and step
// count) checked against the steps accounted for before it and the steps
// submitted. A set non-physical bit is NonPhysicalStateError; a count other
// than accounted + submitted means the device rejected the work, ran out of
// memory or was lost (GpuExecutionError), which would otherwise read back as
// a flow that has stopped changing. Plain host C++ without HIP, so the rules
// are tested on the CPU with fabricated readings. (The reference's drift
// check of its compensated clock has no counterpart: the clock here is FP64.)
#pragma once

#include <cstddef>
#include <cstdio>

#include "../../include/bowshock.h"

namespace bs {

// `submitted` < 0: the step count is not checked (a batch of unknown size).
// On BS_OK the ledger moves to `steps`; otherwise `msg` says why.
inline bs_status ledger_account(long long& accounted, unsigned status, long long steps, long long submitted,
                                char* msg, size_t len) {
  if (status & BS_STATUS_NON_PHYSICAL) {
    snprintf(msg, len, "non-physical state on the GPU between steps %lld and %lld", accounted, steps);
    return BS_ERR_NONPHYSICAL;
  }
  if (submitted >= 0 && steps != accounted + submitted) {
    snprintf(msg, len,
             "the GPU ran %lld of the %lld steps submitted: it rejected the work, ran out of memory or lost the "
             "device",
             steps - accounted, submitted);
    return BS_ERR_EXECUTION;
  }
  accounted = steps;
  return BS_OK;
}

}  // namespace bs
