// The dataflow stage kernel: many explicit steps of one FP32 grid in one
// launch, with no barrier between stages.
//
// The launch's work items are (epoch, tile) pairs in epoch-major order,
// epoch e being stage e % ns of step e / ns; blocks take them from one
// atomic counter. A tile's stage reads its own and its eight neighbours'
// cells of the previous epoch (the stencil reaches two cells, a tile is at
// least that wide), so item (e, t) waits only until t and its neighbours
// have finished epoch e - 1: every tile keeps the number of epochs it has
// finished in a flag, written after the tile's stores have completed.
// Items are taken in order and wait only on earlier items, which running
// blocks hold, so the launch cannot deadlock however many blocks are
// resident; the fill and drain between dependent stages, and the
// kernel-boundary writeback of the L2s, overlap with the work of other
// tiles.
//
// Neighbouring tiles may run on different XCDs, whose L2s are not coherent
// with each other: the states and time steps written in the launch go
// through device-scope stores and loads (COH_DEVICE), which write through
// and read past the XCD's L2; flags and the counter are device-scope
// atomics. The metrics, read-only, are cached as usual.
//
// The three state buffers rotate as in the per-stage launches. A tile runs
// epoch e only after its neighbours finished e - 1, and a neighbour runs
// e + 1 only after the tile finished e, so neighbours are at most one epoch
// apart; the buffer a stage writes is never the one the same or the
// previous epoch reads at a neighbour's cells.
//
// Every wait is bounded: a flag that does not arrive within `spin_limit`
// polls sets DATAFLOW_TIMEOUT in the fault word and the block returns; the
// host reports it as an execution error.
#pragma once

#include "bs_internal.hpp"

namespace bs {

constexpr unsigned DATAFLOW_TIMEOUT = 1u;
// Ints between two tiles' flags: one 256-byte line each.
constexpr int DATAFLOW_FLAG_STRIDE = 64;

struct DataflowArgs {
  const TileArgsT<float>* recs;  // record 3 cur + s: stage s with current state buffer cur
  int tiles_x, tiles_y;
  int ns;                        // stages per step
  int steps;
  int cur0;                      // current buffer of the first step
  int wrap_i, wrap_j;            // periodic sides: the neighbours wrap around
  int* flags;                    // per tile: epochs finished
  int* next;                     // the work-item counter
  unsigned* fault;
  int spin_limit;
};

// Waits until tile (bx, by) and its neighbours have finished `epochs`
// epochs; called by the first wavefront, lane l < 9 polling neighbour
// (l % 3 - 1, l / 3 - 1), so the nine flags are read in one round trip.
// False (on every lane) when the wait is abandoned.
__device__ __forceinline__ bool wait_neighbours(const DataflowArgs& D, int bx, int by, int epochs) {
  const int lane = threadIdx.x & 63;
  const int* f = nullptr;
  if (lane < 9) {
    int nx = bx + lane % 3 - 1, ny = by + lane / 3 - 1;
    bool ok = true;
    if (nx < 0 || nx >= D.tiles_x) {
      ok = D.wrap_i != 0;
      nx = (nx + D.tiles_x) % D.tiles_x;
    }
    if (ny < 0 || ny >= D.tiles_y) {
      ok = ok && D.wrap_j != 0;
      ny = (ny + D.tiles_y) % D.tiles_y;
    }
    if (ok) f = D.flags + (size_t)(nx + ny * D.tiles_x) * DATAFLOW_FLAG_STRIDE;
  }
  for (int n = 0;; n++) {
    const bool ready = f == nullptr || __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epochs;
    if (__ballot(!ready) == 0) return true;
    if (n >= D.spin_limit) {
      if (lane == 0) __hip_atomic_fetch_or(D.fault, DATAFLOW_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((n & 31) == 31 && __hip_atomic_load(D.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Not held to 64 VGPRs: the loop's state and the device-scope accesses'
// 64-bit addresses would spill (108 VGPRs, four tiles per CU).
template <int TX, int TY, int REC, int LIM, int FLUX>
__global__ __launch_bounds__(TX * TY) void dataflow_stage_kernel(DataflowArgs D) {
  __shared__ int item_s, go_s;
  const int tiles = D.tiles_x * D.tiles_y;
  const int total = D.steps * D.ns * tiles;
  if (threadIdx.x == 0) item_s = __hip_atomic_fetch_add(D.next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  int k = item_s;
  while (k < total) {
    const int e = k / tiles, t = k - e * tiles;
    const int bx = t % D.tiles_x, by = t / D.tiles_x;
    int next = 0;
    if (threadIdx.x < 64) {
      // The next item is taken before this one's wait: a block holding a
      // later item only ever waits on earlier ones.
      if (threadIdx.x == 0) next = __hip_atomic_fetch_add(D.next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool go = e == 0 || wait_neighbours(D, bx, by, e);
      if (threadIdx.x == 0) go_s = go;
    }
    __syncthreads();
    if (!go_s) return;
    const int step = e / D.ns, s = e - step * D.ns;
    const int cur = D.ns == 1 ? (D.cur0 + step) % 3 : D.cur0;
    tile_body<float, TX, TY, REC, LIM, FLUX, TILE_STAGE, COH_DEVICE>(D.recs[3 * cur + s], bx, by, 0);
    // Every store of the tile completes before its flag is raised; the
    // signal fences keep the compiler from moving memory accesses across
    // the wait.
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
      __hip_atomic_store(D.flags + (size_t)t * DATAFLOW_FLAG_STRIDE, e + 1, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      item_s = next;
    }
    __syncthreads();
    k = item_s;
  }
}

// Zeroes the flags, the counter and the fault word before a launch.
static __global__ __launch_bounds__(256) void dataflow_clear(int* __restrict__ flags, int tiles, int* __restrict__ next,
                                                             unsigned* __restrict__ fault) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < tiles) flags[(size_t)k * DATAFLOW_FLAG_STRIDE] = 0;
  if (k == 0) {
    *next = 0;
    *fault = 0u;
  }
}

using DataflowFn = void (*)(DataflowArgs);
DataflowFn dataflow_fn(const bs_config& c);

}  // namespace bs
