// Host-side internals shared by the library's translation units: error
// reporting, device allocation, captured step graphs, the Runge-Kutta
// tables, the FP64 metrics, state packing, and the stencil kernel tables
// (instantiated once per state precision in their own units, so the
// kernels compile in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>

#include "../../include/bowshock.h"
#include "bs_kernels.hpp"

struct bs_ctx {
  int device;
};

namespace bs {

// Records the calling thread's last error message (bs_last_error) and
// returns `code`.
bs_status fail(bs_status code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
bs_status hip_fail(hipError_t e, const char* what);

#define HIP_TRY(expr, what)                            \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return ::bs::hip_fail(_e, what); \
  } while (0)

constexpr int BATCH = 256;       // steps between status checks (time-loop.ts:13)
constexpr int GRAPH_STEPS = 12;  // steps per captured graph (a multiple of 3)

// Tile of the FP32 stencil kernels: TX cells along i by TY along j (a
// wavefront covers 64 / TX rows of TX cells); set at build time. 32 x 8
// measured fastest for the stage kernel at 1280 x 640 (15.7 us; 64 x 8
// 16.1, 32 x 16 16.0, 32 x 12 16.4, 48 x 8 17.4, 32 x 4 18.8): its 3,200
// tiles, eight to a CU, leave a finer last round than 1,600 of 64 x 8.
#ifndef BS_TILE_X
#define BS_TILE_X 32
#endif
#ifndef BS_TILE_Y
#define BS_TILE_Y 8
#endif
constexpr int TX = BS_TILE_X, TY = BS_TILE_Y;
static_assert(TX * TY <= 1024 && TX * TY >= TX + TY && (TX * TY) % 64 == 0, "tile shape");

// Tile per state precision. The FP64 tile has the FP32 tile's 256
// threads; at 32 x 8 it stages fewer halo cells per cell than at 64 x 4
// (1280 x 640, FP64 implicit prepare 49.7 against 54.5 us). Both hold
// four FP64 tiles to a CU.
template <typename T>
struct Tile {
  static constexpr int x = TX, y = TY;
};
template <>
struct Tile<double> {
  static constexpr int x = 32, y = 8;
};

template <typename T>
bs_status dev_alloc(T** p, size_t count, const char* what) {
  hipError_t e = hipMalloc((void**)p, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ERR_MEMORY, "allocating %s (%zu bytes): %s", what, count * sizeof(T), hipGetErrorString(e));
  }
  return BS_OK;
}

// A captured run of GRAPH_STEPS steps starting from state buffer `cur`;
// the run returns to the same buffer (RK stages write the last stage in
// place; single-stage and implicit steps cycle the three buffers, and
// GRAPH_STEPS is a multiple of 3).
struct StepGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int cur = -1;
  int sweeps = 0;
  double omega = 0, growth = 0, cfl_max = 0;
  // The remainder of a call after its whole graphs: `tail_len` steps from
  // buffer `tail_cur`, ending on buffer `tail_end` (kept while calls of one
  // length repeat, so a call of any length replays graphs only).
  hipGraph_t tail_graph = nullptr;
  hipGraphExec_t tail_exec = nullptr;
  int tail_len = 0, tail_cur = -1, tail_end = -1;
};

inline void drop_tail(StepGraph& g) {
  if (g.tail_exec) (void)hipGraphExecDestroy(g.tail_exec);
  if (g.tail_graph) (void)hipGraphDestroy(g.tail_graph);
  g.tail_exec = nullptr;
  g.tail_graph = nullptr;
  g.tail_len = 0;
  g.tail_cur = g.tail_end = -1;
}

// A blocking host-device copy on the object's own stream: ordered after the
// work already queued there, and allowed while another thread captures a
// graph on its stream (a copy on the legacy stream is refused then, even
// with thread-local capture and non-blocking streams).
inline hipError_t copy_on(hipStream_t stream, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(stream);
}

inline void drop_graph(StepGraph& g) {
  drop_tail(g);
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g = StepGraph();
}

// Captures the `n` steps that `encode` enqueues on `stream` into one
// instantiated graph. Nothing runs during the capture, so the current
// buffer index `cur` is restored afterwards; `cur_end` is the buffer the
// steps end on.
template <typename Encode>
bs_status capture_graph(hipStream_t stream, int& cur, int n, Encode encode, hipGraph_t* graph_out,
                        hipGraphExec_t* exec_out, int* cur_end) {
  const int cur0 = cur;
  HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "graph capture");
  bs_status st = BS_OK;
  for (int k = 0; k < n && st == BS_OK; k++) st = encode();
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(stream, &graph);
  *cur_end = cur;
  cur = cur0;
  if (st != BS_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (st != BS_OK) return st;
    return hip_fail(e, "graph capture");
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "graph instantiate");
  }
  *graph_out = graph;
  *exec_out = exec;
  return BS_OK;
}

// `count` steps from buffer `cur` as whole captured graphs of GRAPH_STEPS
// steps (which return to the buffer they start from) and one tail graph
// for the rest, capturing what is not cached yet.
template <typename Encode>
bs_status replay_steps(hipStream_t stream, int& cur, StepGraph& g, int count, Encode encode) {
  if (!(g.exec && g.cur == cur)) {
    drop_graph(g);
    int end = -1;
    bs_status st = capture_graph(stream, cur, GRAPH_STEPS, encode, &g.graph, &g.exec, &end);
    if (st) return st;
    if (end != cur) {
      drop_graph(g);
      return fail(BS_ERR_EXECUTION, "captured steps do not return to their state buffer");
    }
    g.cur = cur;
  }
  int k = 0;
  for (; k + GRAPH_STEPS <= count; k += GRAPH_STEPS) HIP_TRY(hipGraphLaunch(g.exec, stream), "graph launch");
  const int rest = count - k;
  if (rest > 0) {
    if (!(g.tail_exec && g.tail_cur == cur && g.tail_len == rest)) {
      drop_tail(g);
      bs_status st = capture_graph(stream, cur, rest, encode, &g.tail_graph, &g.tail_exec, &g.tail_end);
      if (st) return st;
      g.tail_cur = cur;
      g.tail_len = rest;
    }
    HIP_TRY(hipGraphLaunch(g.tail_exec, stream), "graph launch");
    cur = g.tail_end;
  }
  return BS_OK;
}

// Runge-Kutta coefficients (runge-kutta.ts:17-28) in the state's precision.
template <typename T>
int rk_table(int integrator, const T** a, const T** b) {
  static const T ea[] = {T(0)}, eb[] = {T(1)};
  static const T a2[] = {T(0), T(0.5)}, b2[] = {T(1), T(0.5)};
  static const T a3[] = {T(0), T(0.75), T(1.0 / 3.0)}, b3[] = {T(1), T(0.25), T(2.0 / 3.0)};
  if (integrator == BS_RK_SSPRK2) { *a = a2; *b = b2; return 2; }
  if (integrator == BS_RK_SSPRK3) { *a = a3; *b = b3; return 3; }
  *a = ea; *b = eb;
  return 1;
}

// structured.ts:94-168 computeMetrics, FP64 on the host; any output may be
// null.
bs_status metrics(int ni, int nj, const double* x, const double* y, int axisym, double* inx, double* iny,
                  double* ia, double* jnx, double* jny, double* ja, double* vol, double* planar, double* cxo,
                  double* cyo);

// Conserved state of a primitive one with the gamma `g` the state is
// formed with (the FP32 gamma for an FP32 state, as the reference GPU
// engine packs states, euler2d-gpu.ts:412-432), rounded to T at the end.
template <typename T>
void pack_cons(const double* w, double g, T out[4]) {
  out[0] = (T)w[0];
  out[1] = (T)(w[0] * w[1]);
  out[2] = (T)(w[0] * w[2]);
  out[3] = (T)(w[3] / (g - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

inline bool physical(const double* w) { return w[0] > 0 && w[3] > 0 && std::isfinite(w[1]) && std::isfinite(w[2]); }

// ---- stencil kernel tables (bs_tiles_f32.hip, bs_tiles_f64.hip) ----------

template <typename T>
using TileFnT = void (*)(TileArgsT<T>);
template <typename S>
using WallFnT = void (*)(WallArgsT<S>, WallBoundaryT<S>);

// The tile kernel of a scheme (reconstruction, limiter, flux) for a tile
// mode (TileMode), for an FP32 or an FP64 state.
TileFnT<float> tile_fn_f32(const bs_config& c, int mode);
TileFnT<double> tile_fn_f64(const bs_config& c, int mode);
// The FP64 wall-flux kernel of a scheme reading an FP32 or FP64 state.
WallFnT<float> wall_fn_f32(const bs_config& c);
WallFnT<double> wall_fn_f64(const bs_config& c);

template <typename T>
TileFnT<T> tile_fn(const bs_config& c, int mode) {
  if constexpr (std::is_same<T, float>::value) return tile_fn_f32(c, mode);
  else return tile_fn_f64(c, mode);
}

template <typename S>
WallFnT<S> wall_fn(const bs_config& c) {
  if constexpr (std::is_same<S, float>::value) return wall_fn_f32(c);
  else return wall_fn_f64(c);
}

}  // namespace bs
