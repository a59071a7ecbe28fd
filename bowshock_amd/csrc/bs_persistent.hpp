// The persistent batch kernel: many explicit steps of every active case of
// an FP32 batch in one launch.
//
// The host deals the cases to the 8 XCDs (every tile of a case on one XCD,
// so the blocks of a case share that XCD's L2) as one list of tiles per
// XCD. A block reads its XCD from the hardware register and claims the
// next tile of that XCD's list with an atomic; blocks that find their
// list taken return. The dispatcher deals workgroups to the XCDs
// round-robin, so a launch of 8 x (longest list) + 64 blocks gives every
// XCD at least as many blocks as its list has tiles. Between two stages the blocks of a
// case meet at a barrier of their own: each thread waits for its stores to
// complete (they are then in the L2), the block arrives with a relaxed
// device-scope atomic add on the case's counter and polls it with relaxed
// device-scope loads. The next stage reads the states other blocks wrote
// with device-scope loads (tile_body<..., COH = true>), which read the L2
// and never a stale line of the CU's L1. No L2 writeback or invalidate is
// issued inside the launch: that is what makes a device-wide barrier cost
// 20-90 us on this chip, against ~1 us for this one.
//
// Every wait is bounded: a block whose barrier does not complete within
// `spin_limit` polls sets PERSIST_TIMEOUT in the fault word and returns,
// and a block whose XCD number is out of range sets PERSIST_MISPLACED and
// returns before any store; the host reports either as an execution
// error.
#pragma once

#include "bs_internal.hpp"

namespace bs {

constexpr unsigned PERSIST_TIMEOUT = 1u;
constexpr unsigned PERSIST_MISPLACED = 2u;
constexpr int PERSIST_XCDS = 8;
// Ints between two cases' barrier counters: one 256-byte line each, so
// that polls of one case's counter do not contend with another's.
constexpr int PERSIST_BAR_STRIDE = 64;

// One tile of an XCD's list: case c, its tile (bx, by), and the number of
// tiles of case c.
struct PersistItem {
  int c;
  int bx, by;
  int members;
};

// The XCD this wave runs on (HW_REG_XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15; }

// Arrive at the case's barrier and wait until `target` arrivals have been
// counted; false when the launch is to be abandoned (this wait timed out,
// or another block's fault was seen while waiting).
__device__ __forceinline__ bool group_barrier(int* ctr, int target, unsigned* fault, int spin_limit) {
  __shared__ int go;
  // Every store of this thread completes (reaches the L2) before the block
  // arrives; the signal fences keep the compiler from moving memory
  // accesses across the wait.
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int n = 0; __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; n++) {
      if (n >= spin_limit) {
        __hip_atomic_fetch_or(fault, PERSIST_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      if ((n & 31) == 31 && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    go = ok;
  }
  __syncthreads();
  return go != 0;
}

// `steps` explicit steps of `ns` stages. The stage records are the batch's
// TileArgs table: record (3 * cur + s) * count + c is stage s of case c
// with `cur` the current state buffer (fixed for two- and three-stage
// schemes, advancing by one per step for forward Euler).
// XCD x's list is items[x * stride, x * stride + lens[x]); claims[x]
// counts the blocks of XCD x that have taken a tile.
template <int TX, int TY, int REC, int LIM, int FLUX>
__global__ __launch_bounds__(TX * TY) void batch_persistent_kernel(const TileArgsT<float>* __restrict__ recs, int count,
                                                                   const PersistItem* __restrict__ items,
                                                                   const int* __restrict__ lens,
                                                                   int* __restrict__ claims, int stride, int steps,
                                                                   int ns, int cur0, int* __restrict__ bars,
                                                                   unsigned* __restrict__ fault, int spin_limit) {
  __shared__ int slot;
  if (threadIdx.x == 0) {
    const int x = xcc_id();
    if (x >= PERSIST_XCDS) {
      __hip_atomic_fetch_or(fault, PERSIST_MISPLACED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      slot = -1;
    } else {
      const int k = __hip_atomic_fetch_add(claims + x, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      slot = k < lens[x] ? x * stride + k : -1;
    }
  }
  __syncthreads();
  if (slot < 0) return;
  const PersistItem it = items[slot];
  int* ctr = bars + (size_t)it.c * PERSIST_BAR_STRIDE;
  int epoch = 0;
  for (int step = 0; step < steps; step++) {
    const int cur = ns == 1 ? (cur0 + step) % 3 : cur0;
    for (int s = 0; s < ns; s++) {
      if (epoch > 0 && !group_barrier(ctr, epoch * it.members, fault, spin_limit)) return;
      tile_body<float, TX, TY, REC, LIM, FLUX, TILE_STAGE, true>(recs[(size_t)(3 * cur + s) * count + it.c], it.bx,
                                                                 it.by, 0);
      epoch++;
    }
  }
}

// Zeroes the barrier counters, the claim counters and the fault word
// before a launch, on the launch's stream (a kernel, not a memset, so that
// the pair can sit in a captured graph).
static __global__ __launch_bounds__(256) void clear_counters(int* __restrict__ bars, int n, int* __restrict__ claims,
                                                             unsigned* __restrict__ fault) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < n) bars[(size_t)k * PERSIST_BAR_STRIDE] = 0;
  if (k < PERSIST_XCDS) claims[k] = 0;
  if (k == 0) *fault = 0u;
}

using PersistFn = void (*)(const TileArgsT<float>*, int, const PersistItem*, const int*, int*, int, int, int, int,
                           int*, unsigned*, int);

// The persistent kernel of a scheme for tile shape `tile` (1: 32 x 4,
// 2: 32 x 8), and that shape.
PersistFn batch_persistent_fn(const bs_config& c, int tile);
void persistent_tile_shape(int tile, int* tx, int* ty);

}  // namespace bs
