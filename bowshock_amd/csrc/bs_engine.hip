// C ABI of the gfx950 Euler engine (include/bowshock.h): solver objects,
// device memory, kernel launches, batches of steps and their status checks.
//
// Host semantics follow the reference's GPU engine
// (src/engine-gpu/euler2d-gpu.ts) and its time loop
// (src/engine-gpu/time-loop.ts): steps run in batches of at most 256
// between readbacks of a small control block (status bits, steps run,
// FP64 clock); a set status bit or a short batch is an error.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "bs_internal.hpp"
#include "bs_ledger.hpp"
#include "bs_anderson.hpp"
#include "bs_linegrid.hpp"

using namespace bs;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace bs {

bs_status fail(bs_status code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

bs_status hip_fail(hipError_t e, const char* what) {
  bs_status code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? BS_ERR_MEMORY : BS_ERR_EXECUTION;
  return fail(code, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace bs


// Device arrays of one solver. The untyped arrays hold values of the
// solver's state precision (float, or double for an FP64 solver).
struct DeviceBuffers {
  void* state[3] = {nullptr, nullptr, nullptr};  // 4 * n each, SoA
  void *inx = nullptr, *iny = nullptr, *ia = nullptr, *jnx = nullptr, *jny = nullptr, *ja = nullptr;
  void *vol = nullptr, *planar = nullptr, *dt = nullptr, *dt_global = nullptr, *dt_partials = nullptr;
  double* partials = nullptr;
  double* scalar = nullptr;
  Control* control = nullptr;
  void* bnd = nullptr;  // the BoundaryDevT table the stencil kernels read
  void* face_cons[4] = {nullptr, nullptr, nullptr, nullptr};
  void* face_flux[4] = {nullptr, nullptr, nullptr, nullptr};
  double* face_geo[4] = {nullptr, nullptr, nullptr, nullptr};
  double* wall_out = nullptr;
  double* shock_out = nullptr;  // fitted shock: speed, Mach number, signal per jMax face
  // the grid on lines (bs_set_line_grid): per line x, y, dx, dy; fractions; heights; last moves
  double *line = nullptr, *eta = nullptr, *height = nullptr, *move = nullptr;
  // implicit work
  float *res = nullptr, *diag = nullptr, *lam_i = nullptr, *lam_j = nullptr, *dq = nullptr;
};

struct Timing {
  bool enabled = false;
  bool region = false;     // one pair per batch of explicit local steps
  bool suppress = false;   // inside such a batch: no per-launch pairs
  std::vector<hipEvent_t> pool;
  std::vector<int> kinds;  // kind per recorded pair
  std::vector<long long> weights;  // launches each pair brackets
  int used = 0;
  double total_ms[4] = {0, 0, 0, 0};
  long long launches[4] = {0, 0, 0, 0};
};

struct bs_solver {
  bs_config cfg;
  int device;
  int ni, nj, n;
  float gamma32;
  bool fp64 = false;  // state, metrics and stencil arithmetic in FP64
  bool relax_linear = false;  // BS_RELAX_LINEARIZED (always on an FP64 solver)
  bool fitted = false;        // jMax is a fitted shock ahead of which the free stream is shock_free
  bool lines = false;         // the grid lies on lines with moving heights
  bool symmetry = false;      // the lines' first node is normal to the axis
  double shock_free[4] = {0, 0, 0, 0};
  hipStream_t stream = nullptr;
  DeviceBuffers d;
  int cur = 0;
  bool imposed[4] = {false, false, false, false};
  std::vector<double> x, y;
  int tiles_x, tiles_y;
  std::atomic<bool> busy{false};
  Timing timing;
  long long steps_accounted = 0;
  bool graphs = true;
  StepGraph explicit_graph, implicit_graph, anderson_graph, fitted_graph;
  long long anderson_for = 0;  // id of the history the anderson graph was captured with
  int fitted_for = 0;          // steps per move the fitted graph was captured with
};

namespace {

struct Guard {
  bs_solver* s;
  bool ok;
  explicit Guard(bs_solver* solver) : s(solver) {
    bool expected = false;
    ok = s->busy.compare_exchange_strong(expected, true);
  }
  ~Guard() {
    if (ok) s->busy.store(false);
  }
};

#define SOLVER_GUARD(s)                                                                        \
  if ((s) == nullptr) return fail(BS_ERR_RANGE, "no solver");                                  \
  Guard _guard(s);                                                                             \
  if (!_guard.ok) return fail(BS_ERR_BUSY, "bs_solver: await the previous operation first"); \
  (void)hipSetDevice((s)->device)

int along(const bs_solver* s, int side) { return side < 2 ? s->nj : s->ni; }

// Bytes of one state value of the solver.
size_t real_size(const bs_solver* s) { return s->fp64 ? sizeof(double) : sizeof(float); }

// The gamma the state's conserved variables are formed with: the FP32
// gamma for an FP32 state (euler2d-gpu.ts:412-432), the configured one for
// an FP64 state.
double state_gamma(const bs_solver* s) { return s->fp64 ? s->cfg.gamma : (double)s->gamma32; }

template <typename T>
StateDevT<T> state_of(bs_solver* s, int b) {
  StateDevT<T> st;
  T* base = static_cast<T*>(s->d.state[b]);
  for (int q = 0; q < 4; q++) st.q[q] = base + (size_t)q * s->n;
  return st;
}

template <typename T>
GridDevT<T> grid_of(bs_solver* s) {
  GridDevT<T> G;
  G.ni = s->ni;
  G.nj = s->nj;
  G.inx = static_cast<const T*>(s->d.inx); G.iny = static_cast<const T*>(s->d.iny);
  G.ia = static_cast<const T*>(s->d.ia);
  G.jnx = static_cast<const T*>(s->d.jnx); G.jny = static_cast<const T*>(s->d.jny);
  G.ja = static_cast<const T*>(s->d.ja);
  G.vol = static_cast<const T*>(s->d.vol); G.planar = static_cast<const T*>(s->d.planar);
  return G;
}

template <typename T>
BoundaryDevT<T> boundary_of(bs_solver* s) {
  BoundaryDevT<T> B;
  for (int k = 0; k < 4; k++) {
    const bs_boundary& b = s->cfg.boundaries[k];
    B.kind[k] = b.kind;
    B.state[k][0] = (T)b.rho;
    B.state[k][1] = (T)b.u;
    B.state[k][2] = (T)b.v;
    B.state[k][3] = (T)b.p;
    B.face_cons[k] = static_cast<const T*>(s->d.face_cons[k]);
    B.face_flux[k] = s->imposed[k] ? static_cast<const T*>(s->d.face_flux[k]) : nullptr;
  }
  return B;
}

// The boundary table lives in device memory so kernels index it by a
// runtime side; re-uploaded whenever an imposed flux is switched. The copy
// is ordered on the solver's stream before any later launch.
template <typename T>
bs_status upload_boundary_t(bs_solver* s) {
  BoundaryDevT<T> B = boundary_of<T>(s);
  HIP_TRY(hipMemcpyAsync(s->d.bnd, &B, sizeof B, hipMemcpyHostToDevice, s->stream), "boundary table");
  HIP_TRY(hipStreamSynchronize(s->stream), "boundary table");
  return BS_OK;
}

bs_status upload_boundary(bs_solver* s) {
  return s->fp64 ? upload_boundary_t<double>(s) : upload_boundary_t<float>(s);
}

// ---- timing -------------------------------------------------------------------

hipEvent_t* timing_pair(bs_solver* s, int kind, long long weight) {
  Timing& t = s->timing;
  if (!t.enabled) return nullptr;
  if ((size_t)(2 * t.used + 2) > t.pool.size()) {
    for (int k = 0; k < 64; k++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      t.pool.push_back(e);
    }
  }
  hipEvent_t* pair = &t.pool[2 * t.used];
  t.kinds.resize(t.used + 1);
  t.weights.resize(t.used + 1);
  t.kinds[t.used] = kind;
  t.weights[t.used] = weight;
  t.used++;
  (void)hipEventRecord(pair[0], s->stream);
  return pair;
}

hipEvent_t* timing_begin(bs_solver* s, int kind) {
  if (s->timing.suppress) return nullptr;
  return timing_pair(s, kind, 1);
}

void timing_end(bs_solver* s, hipEvent_t* pair) {
  if (pair) (void)hipEventRecord(pair[1], s->stream);
}

// After a synchronisation: fold the recorded pairs into the totals.
void timing_collect(bs_solver* s) {
  Timing& t = s->timing;
  for (int k = 0; k < t.used; k++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.pool[2 * k], t.pool[2 * k + 1]) == hipSuccess) {
      t.total_ms[t.kinds[k]] += ms;
      t.launches[t.kinds[k]] += t.weights[k];
    }
  }
  t.used = 0;
}

// ---- launches -----------------------------------------------------------------

template <typename T>
TileArgsT<T> tile_args(bs_solver* s) {
  TileArgsT<T> A;
  memset(&A, 0, sizeof A);
  A.G = grid_of<T>(s);
  A.B = static_cast<const BoundaryDevT<T>*>(s->d.bnd);
  A.gamma = (T)state_gamma(s);
  A.axisymmetric = s->cfg.geometry == BS_GEOMETRY_AXISYMMETRIC;
  A.limiter = s->cfg.limiter;
  A.cfl = (T)s->cfg.cfl;
  A.dt = static_cast<T*>(s->d.dt);
  A.dt_global = static_cast<const T*>(s->d.dt_global);
  A.control = s->d.control;
  return A;
}

template <typename T>
bs_status launch_tile(bs_solver* s, int mode, const TileArgsT<T>& A, int kind) {
  dim3 grid(s->tiles_x * s->tiles_y);
  hipEvent_t* ev = timing_begin(s, kind);
  hipLaunchKernelGGL(tile_fn<T>(s->cfg, mode), grid, dim3(Tile<T>::x * Tile<T>::y), 0, s->stream, A);
  timing_end(s, ev);
  HIP_TRY(hipGetLastError(), "tile kernel launch");
  return BS_OK;
}

// One explicit step: the time step (global stepping: per-cell steps, their
// minimum clamped to t_end, then the clock), then the stages. Stage k reads
// `src` with the stencil and u0 at its own cell only, so the last stage of a
// multi-stage scheme writes u0 in place (euler2d.ts:1146-1157).
template <typename T>
bs_status encode_step_t(bs_solver* s) {
  const T *ca, *cb;
  const int ns = rk_table(s->cfg.integrator, &ca, &cb);
  const bool global = s->cfg.time_stepping == BS_STEPPING_GLOBAL;
  if (s->fitted) {
    // The fitted shock's faces from the solution at the start of the step,
    // held through its stages (euler2d-gpu.ts:906-912).
    ShockArgsT<T> S;
    S.ni = s->ni;
    S.nj = s->nj;
    S.s = state_of<T>(s, s->cur);
    S.face_geo = s->d.face_geo[JMAX];
    S.cons = static_cast<T*>(s->d.face_cons[JMAX]);
    S.flux = static_cast<T*>(s->d.face_flux[JMAX]);
    S.out = s->d.shock_out;
    for (int k = 0; k < 4; k++) S.free[k] = s->shock_free[k];
    S.gamma = state_gamma(s);
    hipLaunchKernelGGL(shock_boundary_kernel<T>, dim3((s->ni + 255) / 256), dim3(256), 0, s->stream, S);
  }
  if (global) {
    const int nb = (s->n + 255) / 256;
    hipLaunchKernelGGL(time_step_kernel<T>, dim3(nb), dim3(256), 0, s->stream, grid_of<T>(s),
                       state_of<T>(s, s->cur), (T)state_gamma(s), (T)s->cfg.cfl, static_cast<T*>(s->d.dt_partials));
    hipLaunchKernelGGL(finalize_time_step<T>, dim3(1), dim3(256), 0, s->stream,
                       static_cast<const T*>(s->d.dt_partials), nb, s->d.control, static_cast<T*>(s->d.dt_global));
  }
  const int cur = s->cur;
  const int other[2] = {(cur + 1) % 3, (cur + 2) % 3};
  int prev = cur;
  int last = cur;
  for (int k = 0; k < ns; k++) {
    const int dst = (k == ns - 1 && k > 0) ? cur : other[k % 2];
    TileArgsT<T> A = tile_args<T>(s);
    A.u0 = state_of<T>(s, cur);
    A.src = state_of<T>(s, prev);
    A.dst = state_of<T>(s, dst);
    A.a = ca[k];
    A.b = cb[k];
    A.dt_mode = global ? DT_GLOBAL : (k == 0 ? DT_COMPUTE_LOCAL : DT_READ);
    A.count_step = (!global && k == ns - 1) ? 1 : 0;
    bs_status st = launch_tile(s, TILE_STAGE, A, 0);
    if (st) return st;
    prev = dst;
    last = dst;
  }
  if (global) hipLaunchKernelGGL(advance_clock, dim3(1), dim3(1), 0, s->stream, s->d.control);
  s->cur = last;
  return BS_OK;
}

bs_status encode_step(bs_solver* s) { return s->fp64 ? encode_step_t<double>(s) : encode_step_t<float>(s); }

// `count` steps: captured graphs (replay_steps) for a call of at least
// GRAPH_STEPS steps, a shorter call launched kernel by kernel. Per-launch
// kernel timing needs the events between launches, so it runs without
// graphs.
template <typename Encode>
bs_status run_steps(bs_solver* s, StepGraph& g, int count, Encode encode) {
  if (s->graphs && (!s->timing.enabled || s->timing.suppress) && count >= GRAPH_STEPS)
    return replay_steps(s->stream, s->cur, g, count, encode);
  for (int k = 0; k < count; k++) {
    bs_status st = encode();
    if (st) return st;
  }
  return BS_OK;
}

bs_status read_control(bs_solver* s, Control* c) {
  HIP_TRY(hipMemcpyAsync(c, s->d.control, sizeof(Control), hipMemcpyDeviceToHost, s->stream), "control readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "stream synchronize");
  timing_collect(s);
  return BS_OK;
}

void fill_report(const Control& c, bs_report* r) {
  if (!r) return;
  r->steps = c.steps;
  r->status = c.status;
  r->time = c.time;
  r->dt = c.dt;
}

// StepLedger.account (time-loop.ts:125-150), bs_ledger.hpp.
bs_status account(bs_solver* s, const Control& c, long long submitted) {
  char msg[256];
  const bs_status st = ledger_account(s->steps_accounted, c.status, c.steps, submitted, msg, sizeof msg);
  return st ? fail(st, "%s", msg) : BS_OK;
}

bs_status reset_control(bs_solver* s) {
  Control c;
  HIP_TRY(copy_on(s->stream, &c, s->d.control, sizeof c, hipMemcpyDeviceToHost), "control readback");
  double cfl = c.cfl;
  memset(&c, 0, sizeof c);
  c.cfl = cfl > 0 ? cfl : 1.0;
  c.t_end = 3e38;
  HIP_TRY(copy_on(s->stream, s->d.control, &c, sizeof c, hipMemcpyHostToDevice), "control upload");
  s->steps_accounted = 0;
  return BS_OK;
}

void free_all(bs_solver* s) {
  drop_graph(s->explicit_graph);
  drop_graph(s->implicit_graph);
  drop_graph(s->anderson_graph);
  drop_graph(s->fitted_graph);
  DeviceBuffers& d = s->d;
  void* all[] = {d.state[0], d.state[1], d.state[2], d.inx, d.iny, d.ia, d.jnx, d.jny, d.ja, d.vol, d.planar,
                 d.dt, d.dt_global, d.dt_partials, d.partials, d.scalar, d.control, d.bnd, d.wall_out, d.shock_out,
                 d.line, d.eta, d.height, d.move, d.res,
                 d.diag, d.lam_i, d.lam_j, d.dq};
  for (void* p : all)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 4; k++) {
    if (d.face_cons[k]) (void)hipFree(d.face_cons[k]);
    if (d.face_flux[k]) (void)hipFree(d.face_flux[k]);
    if (d.face_geo[k]) (void)hipFree(d.face_geo[k]);
  }
  d = DeviceBuffers();
  for (hipEvent_t e : s->timing.pool) (void)hipEventDestroy(e);
  s->timing.pool.clear();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

// `count` values of the solver's state precision.
bs_status real_alloc(const bs_solver* s, void** p, size_t count, const char* what) {
  const size_t bytes = count * real_size(s);
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 16);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ERR_MEMORY, "allocating %s (%zu bytes): %s", what, bytes, hipGetErrorString(e));
  }
  return BS_OK;
}

// Copies FP64 host values to `dst` in the solver's state precision.
bs_status copy_real(const bs_solver* s, void* dst, const double* src, size_t count, const char* what) {
  if (s->fp64) {
    HIP_TRY(copy_on(s->stream, dst, src, count * sizeof(double), hipMemcpyHostToDevice), what);
    return BS_OK;
  }
  std::vector<float> f(src, src + count);
  HIP_TRY(copy_on(s->stream, dst, f.data(), count * sizeof(float), hipMemcpyHostToDevice), what);
  return BS_OK;
}

bs_status upload_real(const bs_solver* s, void** dst, const std::vector<double>& src, const char* what) {
  bs_status st = real_alloc(s, dst, src.size(), what);
  if (st) return st;
  return copy_real(s, *dst, src.data(), src.size(), what);
}

}  // namespace

namespace bs {

bs_status metrics(int ni, int nj, const double* x, const double* y, int axisym, double* inx, double* iny,
                  double* ia, double* jnx, double* jny, double* ja, double* vol, double* planar, double* cxo,
                  double* cyo) {
  const size_t nn = (size_t)(ni + 1) * (nj + 1);
  if (axisym)
    for (size_t k = 0; k < nn; k++)
      if (y[k] < 0) return fail(BS_ERR_RANGE, "axisymmetric grids must have y >= 0 (y is the radius)");
  auto node = [&](int i, int j) { return i + j * (ni + 1); };
  auto face = [&](double* NX, double* NY, double* AR, size_t k, int a, int b) {
    double dx = x[b] - x[a], dy = y[b] - y[a];
    double len = std::hypot(dx, dy);
    if (NX) NX[k] = len > 0 ? dy / len : 0.0;
    if (NY) NY[k] = len > 0 ? -dx / len : 0.0;
    if (AR) AR[k] = axisym ? len * 0.5 * (y[a] + y[b]) : len;
  };
  for (int j = 0; j < nj; j++)
    for (int i = 0; i <= ni; i++) face(inx, iny, ia, i + (size_t)j * (ni + 1), node(i, j), node(i, j + 1));
  for (int j = 0; j <= nj; j++)
    for (int i = 0; i < ni; i++) face(jnx, jny, ja, i + (size_t)j * ni, node(i + 1, j), node(i, j));
  for (int j = 0; j < nj; j++)
    for (int i = 0; i < ni; i++) {
      const int corners[4] = {node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)};
      double area = 0, cx = 0, cy = 0;
      for (int c = 0; c < 4; c++) {
        int p = corners[c], q = corners[(c + 1) % 4];
        double cross = x[p] * y[q] - x[q] * y[p];
        area += cross;
        cx += (x[p] + x[q]) * cross;
        cy += (y[p] + y[q]) * cross;
      }
      area *= 0.5;
      if (!(area > 0)) return fail(BS_ERR_RANGE, "cell (%d, %d) has non-positive area %g", i, j, area);
      size_t k = i + (size_t)j * ni;
      double gx = cx / (6 * area), gy = cy / (6 * area);
      if (planar) planar[k] = area;
      if (cxo) cxo[k] = gx;
      if (cyo) cyo[k] = gy;
      if (vol) vol[k] = axisym ? area * gy : area;
    }
  return BS_OK;
}

}  // namespace bs

// =============================================================================
// C ABI
// =============================================================================

extern "C" {

const char* bs_last_error(void) { return g_last_error.c_str(); }

const char* bs_version(void) { return "bowshock-mi355x 0.1 (gfx950)"; }

bs_status bs_ctx_create(int device, bs_ctx** out) {
  if (!out) return fail(BS_ERR_RANGE, "no output pointer");
  *out = nullptr;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) return fail(BS_ERR_RANGE, "device %d out of range (%d devices)", device, count);
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  *out = new bs_ctx{device};
  return BS_OK;
}

bs_status bs_ctx_destroy(bs_ctx* ctx) {
  delete ctx;
  return BS_OK;
}

bs_status bs_compute_metrics(int32_t ni, int32_t nj, const double* x, const double* y, int32_t geometry,
                             double* i_nx, double* i_ny, double* i_area, double* j_nx, double* j_ny,
                             double* j_area, double* volume, double* planar_area, double* centroid_x,
                             double* centroid_y) {
  if (ni < 1 || nj < 1 || !x || !y) return fail(BS_ERR_RANGE, "a grid needs at least one cell in each direction");
  return metrics(ni, nj, x, y, geometry == BS_GEOMETRY_AXISYMMETRIC, i_nx, i_ny, i_area, j_nx, j_ny, j_area, volume,
                 planar_area, centroid_x, centroid_y);
}

static bs_status solver_create(bs_ctx* ctx, const bs_config* c, bool fp64, bs_solver** out) {
  if (!ctx || !c || !out) return fail(BS_ERR_RANGE, "null argument");
  *out = nullptr;
  // config.ts:35-55 validateGrid2DConfig and the GPU engine's own checks.
  if (!(c->cfl > 0 && c->cfl <= 1)) return fail(BS_ERR_RANGE, "cfl must be in (0, 1]");
  if (c->ni < 2 || c->nj < 2) return fail(BS_ERR_RANGE, "the grid needs at least 2 x 2 cells");
  // The stencil kernels address an array as base + 32-bit byte offset.
  if ((size_t)(c->ni + 1) * (size_t)(c->nj + 1) > ((size_t)1 << 28))
    return fail(BS_ERR_RANGE, "the grid exceeds 2^28 nodes");
  if ((c->reconstruction == BS_RECON_SHARED || c->reconstruction == BS_RECON_SHARED_SCALED) &&
      c->limiter != BS_LIMITER_VAN_ALBADA)
    return fail(BS_ERR_RANGE, "shared reconstruction requires the 'vanAlbada' limiter");
  if (c->reconstruction < 0 || c->reconstruction > 2 || c->limiter < 0 || c->limiter > 5)
    return fail(BS_ERR_RANGE, "unknown reconstruction or limiter");
  if (c->flux != BS_FLUX_HLLC && c->flux != BS_FLUX_RUSANOV)
    return fail(BS_ERR_RANGE, "flux %d is not available on the GPU", c->flux);
  if (c->integrator < 0 || c->integrator > 2) return fail(BS_ERR_RANGE, "unknown time integrator");
  for (int k = 0; k < 4; k++)
    if (c->boundaries[k].kind < 0 || c->boundaries[k].kind > 5) return fail(BS_ERR_RANGE, "unknown boundary kind");
  if ((c->boundaries[0].kind == BS_BC_PERIODIC) != (c->boundaries[1].kind == BS_BC_PERIODIC))
    return fail(BS_ERR_RANGE, "periodic boundaries must be used on both iMin and iMax");
  if ((c->boundaries[2].kind == BS_BC_PERIODIC) != (c->boundaries[3].kind == BS_BC_PERIODIC))
    return fail(BS_ERR_RANGE, "periodic boundaries must be used on both jMin and jMax");
  if (!c->x || !c->y) return fail(BS_ERR_RANGE, "the grid has no nodes");

  bs_solver* s = new bs_solver();
  s->cfg = *c;
  s->device = ctx->device;
  s->ni = c->ni;
  s->nj = c->nj;
  s->n = c->ni * c->nj;
  s->gamma32 = (float)c->gamma;
  s->fp64 = fp64;
  s->relax_linear = fp64;
  const int ni = s->ni, nj = s->nj, n = s->n;
  const size_t nn = (size_t)(ni + 1) * (nj + 1);
  s->x.assign(c->x, c->x + nn);
  s->y.assign(c->y, c->y + nn);
  s->cfg.x = s->x.data();
  s->cfg.y = s->y.data();
  const int tx = fp64 ? Tile<double>::x : Tile<float>::x, ty = fp64 ? Tile<double>::y : Tile<float>::y;
  s->tiles_x = (ni + tx - 1) / tx;
  s->tiles_y = (nj + ty - 1) / ty;

  const size_t fi = (size_t)(ni + 1) * nj, fj = (size_t)ni * (nj + 1);
  std::vector<double> inx(fi), iny(fi), ia(fi), jnx(fj), jny(fj), ja(fj), vol(n), planar(n);
  bs_status st = metrics(ni, nj, s->x.data(), s->y.data(), c->geometry == BS_GEOMETRY_AXISYMMETRIC, inx.data(),
                         iny.data(), ia.data(), jnx.data(), jny.data(), ja.data(), vol.data(), planar.data(), nullptr,
                         nullptr);
  if (st) {
    delete s;
    return st;
  }
  (void)hipSetDevice(s->device);
  auto bail = [&](bs_status code) {
    free_all(s);
    delete s;
    return code;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "hipStreamCreate failed"));
  for (int b = 0; b < 3; b++)
    if ((st = real_alloc(s, &s->d.state[b], 4 * (size_t)n, "state"))) return bail(st);
  if ((st = upload_real(s, &s->d.inx, inx, "i-face normals")) ||
      (st = upload_real(s, &s->d.iny, iny, "i-face normals")) || (st = upload_real(s, &s->d.ia, ia, "i-face areas")) ||
      (st = upload_real(s, &s->d.jnx, jnx, "j-face normals")) ||
      (st = upload_real(s, &s->d.jny, jny, "j-face normals")) || (st = upload_real(s, &s->d.ja, ja, "j-face areas")) ||
      (st = upload_real(s, &s->d.vol, vol, "volumes")) || (st = upload_real(s, &s->d.planar, planar, "planar areas")))
    return bail(st);
  const int blocks = s->tiles_x * s->tiles_y;
  const int nb256 = (n + 255) / 256;
  if ((st = real_alloc(s, &s->d.dt, (size_t)n, "time steps")) || (st = real_alloc(s, &s->d.dt_global, 1, "dt")) ||
      (st = real_alloc(s, &s->d.dt_partials, (size_t)nb256, "dt partials")) ||
      (st = dev_alloc(&s->d.partials, (size_t)(blocks > nb256 ? blocks : nb256) + 64, "partials")) ||
      (st = dev_alloc(&s->d.scalar, 4, "scalars")) || (st = dev_alloc(&s->d.control, 1, "control")) ||
      (st = dev_alloc((char**)&s->d.bnd, fp64 ? sizeof(BoundaryDevT<double>) : sizeof(BoundaryDevT<float>),
                      "boundary table")) ||
      (st = dev_alloc(&s->d.wall_out, 9 * (size_t)(ni > nj ? ni : nj), "wall fluxes")))
    return bail(st);
  // On the solver's own (non-blocking) stream: a memset on the null stream
  // is not ordered with its work.
  HIP_TRY(hipMemsetAsync(s->d.control, 0, sizeof(Control), s->stream), "control reset");
  HIP_TRY(hipStreamSynchronize(s->stream), "control reset");
  // Per-side FP64 face geometry: midpoint, normal towards increasing
  // index, area (euler2d.ts:1430-1469 boundaryFace).
  for (int side = 0; side < 4; side++) {
    const int m = along(s, side);
    std::vector<double> geo(5 * (size_t)m);
    for (int t = 0; t < m; t++) {
      int n0, n1;
      double fnx, fny, far;
      if (side < 2) {
        int i = side == 0 ? 0 : ni;
        size_t f = i + (size_t)t * (ni + 1);
        n0 = (int)f;
        n1 = (int)f + ni + 1;
        fnx = inx[f]; fny = iny[f]; far = ia[f];
      } else {
        int j = side == 2 ? 0 : nj;
        size_t f = t + (size_t)j * ni;
        n0 = t + j * (ni + 1);
        n1 = n0 + 1;
        fnx = jnx[f]; fny = jny[f]; far = ja[f];
      }
      geo[5 * t + 0] = 0.5 * (s->x[n0] + s->x[n1]);
      geo[5 * t + 1] = 0.5 * (s->y[n0] + s->y[n1]);
      geo[5 * t + 2] = fnx;
      geo[5 * t + 3] = fny;
      geo[5 * t + 4] = far;
    }
    if ((st = dev_alloc(&s->d.face_geo[side], geo.size(), "boundary geometry"))) return bail(st);
    HIP_TRY(copy_on(s->stream, s->d.face_geo[side], geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice),
            "boundary geometry");
    if (c->boundaries[side].kind == BS_BC_FACE_STATES) {
      if ((st = real_alloc(s, &s->d.face_cons[side], 4 * (size_t)m, "face states")) ||
          (st = real_alloc(s, &s->d.face_flux[side], 4 * (size_t)m, "face fluxes")))
        return bail(st);
      const double w[4] = {c->boundaries[side].rho, c->boundaries[side].u, c->boundaries[side].v,
                           c->boundaries[side].p};
      std::vector<double> packed(4 * (size_t)m);
      for (int t = 0; t < m; t++) pack_cons(w, state_gamma(s), &packed[4 * t]);
      if ((st = copy_real(s, s->d.face_cons[side], packed.data(), packed.size(), "face states"))) return bail(st);
    }
  }
  Control init;
  memset(&init, 0, sizeof init);
  init.cfl = 1.0;
  init.t_end = 3e38;
  HIP_TRY(copy_on(s->stream, s->d.control, &init, sizeof init, hipMemcpyHostToDevice), "control init");
  if ((st = upload_boundary(s))) return bail(st);
  *out = s;
  return BS_OK;
}

bs_status bs_solver_create(bs_ctx* ctx, const bs_config* c, bs_solver** out) {
  return solver_create(ctx, c, false, out);
}

bs_status bs_solver_create_fp64(bs_ctx* ctx, const bs_config* c, bs_solver** out) {
  return solver_create(ctx, c, true, out);
}

bs_status bs_solver_precision(bs_solver* s, int32_t* precision) {
  if (!s || !precision) return fail(BS_ERR_RANGE, "null argument");
  *precision = s->fp64 ? BS_PRECISION_FP64 : BS_PRECISION_FP32;
  return BS_OK;
}

bs_status bs_solver_destroy(bs_solver* s) {
  if (!s) return BS_OK;
  // Taken and never released: the solver is freed below.
  bool idle = false;
  if (!s->busy.compare_exchange_strong(idle, true))
    return fail(BS_ERR_BUSY, "bs_solver: await the previous operation first (destroy while a call runs)");
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
  return BS_OK;
}

bs_status bs_set_primitives(bs_solver* s, const double* prim) {
  SOLVER_GUARD(s);
  if (!prim) return fail(BS_ERR_RANGE, "no states");
  const int n = s->n;
  std::vector<double> packed(4 * (size_t)n);
  for (int c = 0; c < n; c++) {
    const double* w = prim + 4 * (size_t)c;
    if (!physical(w)) return fail(BS_ERR_NONPHYSICAL, "non-physical initial state in cell %d", c);
    double q[4];
    pack_cons(w, state_gamma(s), q);
    for (int k = 0; k < 4; k++) packed[(size_t)k * n + c] = q[k];
  }
  s->cur = 0;
  // An FP32 state rounds the FP64 conserved values once.
  bs_status st = copy_real(s, s->d.state[0], packed.data(), packed.size(), "state upload");
  if (st) return st;
  return reset_control(s);
}

// C++ helpers of the entry points below.
extern "C++" {
// Uploads four arrays of n values of type V as the current state.
template <typename V>
static bs_status set_state_from(bs_solver* s, const V* const src[4]) {
  const size_t n = s->n;
  s->cur = 0;
  const bool same = (sizeof(V) == sizeof(double)) == s->fp64;
  char* base = static_cast<char*>(s->d.state[0]);
  for (int q = 0; q < 4; q++) {
    void* dst = base + q * n * real_size(s);
    if (same) {
      HIP_TRY(copy_on(s->stream, dst, src[q], n * sizeof(V), hipMemcpyHostToDevice), "state upload");
    } else if (s->fp64) {
      std::vector<double> v(src[q], src[q] + n);
      HIP_TRY(copy_on(s->stream, dst, v.data(), n * sizeof(double), hipMemcpyHostToDevice), "state upload");
    } else {
      std::vector<float> v(src[q], src[q] + n);
      HIP_TRY(copy_on(s->stream, dst, v.data(), n * sizeof(float), hipMemcpyHostToDevice), "state upload");
    }
  }
  return reset_control(s);
}

// Reads the current state into four arrays of type V (null ones skipped).
template <typename V>
static bs_status read_state_into(bs_solver* s, V* const dst[4], bs_report* report) {
  const size_t n = s->n;
  const bool same = (sizeof(V) == sizeof(double)) == s->fp64;
  std::vector<char> stage(same ? 0 : 4 * n * real_size(s));
  const char* base = static_cast<const char*>(s->d.state[s->cur]);
  for (int q = 0; q < 4; q++)
    if (dst[q]) {
      void* to = same ? (void*)dst[q] : (void*)(stage.data() + q * n * real_size(s));
      HIP_TRY(hipMemcpyAsync(to, base + q * n * real_size(s), n * real_size(s), hipMemcpyDeviceToHost, s->stream),
              "state readback");
    }
  Control c;
  bs_status st = read_control(s, &c);
  if (st) return st;
  if (!same)
    for (int q = 0; q < 4; q++)
      if (dst[q]) {
        const char* from = stage.data() + q * n * real_size(s);
        for (size_t k = 0; k < n; k++)
          dst[q][k] = s->fp64 ? (V)reinterpret_cast<const double*>(from)[k] : (V)reinterpret_cast<const float*>(from)[k];
      }
  fill_report(c, report);
  if (c.status & STATUS_NON_PHYSICAL) return fail(BS_ERR_NONPHYSICAL, "the GPU solution is non-physical");
  return BS_OK;
}
}  // extern "C++"

bs_status bs_set_state(bs_solver* s, const float* rho, const float* mu, const float* mv, const float* energy,
                       int32_t count) {
  SOLVER_GUARD(s);
  if (!rho || !mu || !mv || !energy) return fail(BS_ERR_RANGE, "four arrays are needed");
  if (count != s->n) return fail(BS_ERR_RANGE, "the state needs %d values per array, not %d", s->n, count);
  const float* const src[4] = {rho, mu, mv, energy};
  return set_state_from(s, src);
}

bs_status bs_set_state_f64(bs_solver* s, const double* rho, const double* mu, const double* mv,
                           const double* energy, int32_t count) {
  SOLVER_GUARD(s);
  if (!rho || !mu || !mv || !energy) return fail(BS_ERR_RANGE, "four arrays are needed");
  if (count != s->n) return fail(BS_ERR_RANGE, "the state needs %d values per array, not %d", s->n, count);
  const double* const src[4] = {rho, mu, mv, energy};
  return set_state_from(s, src);
}

bs_status bs_read_state(bs_solver* s, float* rho, float* mu, float* mv, float* energy, bs_report* report) {
  SOLVER_GUARD(s);
  float* const dst[4] = {rho, mu, mv, energy};
  return read_state_into(s, dst, report);
}

bs_status bs_read_state_f64(bs_solver* s, double* rho, double* mu, double* mv, double* energy, bs_report* report) {
  SOLVER_GUARD(s);
  double* const dst[4] = {rho, mu, mv, energy};
  return read_state_into(s, dst, report);
}

bs_status bs_set_face_states(bs_solver* s, int32_t side, const double* prim, int32_t count) {
  SOLVER_GUARD(s);
  if (side < 0 || side > 3 || s->cfg.boundaries[side].kind != BS_BC_FACE_STATES)
    return fail(BS_ERR_RANGE, "the boundary does not take face states");
  if (count != along(s, side)) return fail(BS_ERR_RANGE, "the boundary needs %d face states", along(s, side));
  std::vector<double> packed(4 * (size_t)count);
  for (int t = 0; t < count; t++) pack_cons(prim + 4 * (size_t)t, state_gamma(s), &packed[4 * t]);
  return copy_real(s, s->d.face_cons[side], packed.data(), packed.size(), "face states");
}

bs_status bs_set_face_fluxes(bs_solver* s, int32_t side, const double* flux, int32_t count) {
  SOLVER_GUARD(s);
  if (side < 0 || side > 3 || s->cfg.boundaries[side].kind != BS_BC_FACE_STATES)
    return fail(BS_ERR_RANGE, "the boundary does not take face fluxes");
  if (!flux) {
    s->imposed[side] = false;
    return upload_boundary(s);
  }
  if (count != 4 * along(s, side)) return fail(BS_ERR_RANGE, "the boundary needs %d flux values", 4 * along(s, side));
  bs_status st = copy_real(s, s->d.face_flux[side], flux, (size_t)count, "face fluxes");
  if (st) return st;
  s->imposed[side] = true;
  return upload_boundary(s);
}

bs_status bs_iterate(bs_solver* s, int32_t count, bs_report* report) {
  SOLVER_GUARD(s);
  if (count < 0) return fail(BS_ERR_RANGE, "the step count cannot be negative");
  Control c;
  if (s->cfg.time_stepping == BS_STEPPING_GLOBAL) {
    // iterate() with global steps runs without an end time.
    HIP_TRY(hipMemcpyAsync(&c, s->d.control, sizeof c, hipMemcpyDeviceToHost, s->stream), "control");
    HIP_TRY(hipStreamSynchronize(s->stream), "sync");
    c.t_end = 3e38;
    HIP_TRY(hipMemcpyAsync(s->d.control, &c, sizeof c, hipMemcpyHostToDevice, s->stream), "control");
  }
  // Region timing: a batch of explicit local steps launches nothing but
  // the stage kernel, so one event pair around the batch, split over its
  // stage launches, times that kernel with graphs left on.
  const double *ca, *cb;
  const int ns = rk_table(s->cfg.integrator, &ca, &cb);
  const bool region = s->timing.enabled && s->timing.region && s->cfg.time_stepping == BS_STEPPING_LOCAL;
  int remaining = count;
  while (remaining > 0) {
    const int batch = remaining < BATCH ? remaining : BATCH;
    hipEvent_t* pair = region ? timing_pair(s, 0, (long long)batch * ns) : nullptr;
    s->timing.suppress = region;
    bs_status st = run_steps(s, s->explicit_graph, batch, [s]() { return encode_step(s); });
    s->timing.suppress = false;
    timing_end(s, pair);
    if (st) return st;
    st = read_control(s, &c);
    if (st) return st;
    fill_report(c, report);
    if ((st = account(s, c, batch))) return st;
    remaining -= batch;
  }
  if (count == 0) {
    bs_status st = read_control(s, &c);
    if (st) return st;
    fill_report(c, report);
  }
  return BS_OK;
}

bs_status bs_advance_to(bs_solver* s, double t_end, int32_t max_steps, bs_report* report) {
  SOLVER_GUARD(s);
  if (s->cfg.time_stepping != BS_STEPPING_GLOBAL)
    return fail(BS_ERR_RANGE, "advanceTo needs global time stepping; use iterate() for local stepping");
  Control c;
  HIP_TRY(hipMemcpyAsync(&c, s->d.control, sizeof c, hipMemcpyDeviceToHost, s->stream), "control");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  c.t_end = t_end;
  HIP_TRY(hipMemcpyAsync(s->d.control, &c, sizeof c, hipMemcpyHostToDevice, s->stream), "control");
  const long long first = c.steps;
  double dt_est = 0;
  // advanceInBatches (time-loop.ts:158-177): one step first, then batches
  // sized from the latest time step, until the clock arrives.
  while (true) {
    if (c.time >= t_end) break;
    if (c.steps - first >= max_steps) return fail(BS_ERR_RANGE, "advanceTo(%g) needed more than %d steps", t_end, max_steps);
    long long batch = 1;
    if (dt_est > 0) {
      double need = std::ceil((t_end - c.time) / dt_est) + 1;
      batch = need < 1 ? 1 : (need > BATCH ? BATCH : (long long)need);
    }
    long long left = max_steps - (c.steps - first);
    if (batch > left) batch = left;
    for (long long k = 0; k < batch; k++) {
      bs_status st = encode_step(s);
      if (st) return st;
    }
    bs_status st = read_control(s, &c);
    if (st) return st;
    fill_report(c, report);
    if (c.status & STATUS_NON_PHYSICAL)
      return fail(BS_ERR_NONPHYSICAL, "non-physical state on the GPU between steps %lld and %lld",
                  s->steps_accounted, c.steps);
    if (c.steps == s->steps_accounted && c.time < t_end)
      return fail(BS_ERR_EXECUTION, "the GPU clock stopped short of t_end");
    s->steps_accounted = c.steps;
    dt_est = c.dt;
  }
  fill_report(c, report);
  return BS_OK;
}

// C++ helpers of the entry points below.
extern "C++" {
template <typename T>
static bs_status residual_norm_t(bs_solver* s, double* norm) {
  TileArgsT<T> A = tile_args<T>(s);
  A.src = state_of<T>(s, s->cur);
  A.partials = s->d.partials;
  bs_status st = launch_tile(s, TILE_RESIDUAL, A, 0);
  if (st) return st;
  const int blocks = s->tiles_x * s->tiles_y;
  hipLaunchKernelGGL(finalize_sum, dim3(1), dim3(256), 0, s->stream, s->d.partials, blocks, (double)s->n, 1,
                     s->d.scalar);
  HIP_TRY(hipMemcpyAsync(norm, s->d.scalar, sizeof(double), hipMemcpyDeviceToHost, s->stream), "norm readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  timing_collect(s);
  return BS_OK;
}
}  // extern "C++"

bs_status bs_residual_norm(bs_solver* s, double* norm) {
  SOLVER_GUARD(s);
  if (!norm) return fail(BS_ERR_RANGE, "no output");
  return s->fp64 ? residual_norm_t<double>(s, norm) : residual_norm_t<float>(s, norm);
}

// C++ helpers of the entry points below.
extern "C++" {
template <typename T>
static bs_status residual_vector_t(bs_solver* s, double* residual) {
  TileArgsT<T> A = tile_args<T>(s);
  A.src = state_of<T>(s, s->cur);
  // The spare buffer (neither the current state nor needed after) holds it.
  const int spare = (s->cur + 1) % 3;
  StateDevT<T> out = state_of<T>(s, spare);
  for (int q = 0; q < 4; q++) A.rvec[q] = out.q[q];
  bs_status st = launch_tile(s, TILE_RESIDUAL, A, 0);
  if (st) return st;
  std::vector<T> r(4 * (size_t)s->n);
  HIP_TRY(hipMemcpyAsync(r.data(), s->d.state[spare], r.size() * sizeof(T), hipMemcpyDeviceToHost, s->stream),
          "residual readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  timing_collect(s);
  for (int c = 0; c < s->n; c++)
    for (int q = 0; q < 4; q++) residual[4 * (size_t)c + q] = r[(size_t)q * s->n + c];
  return BS_OK;
}
}  // extern "C++"

bs_status bs_residual_vector(bs_solver* s, double* residual) {
  SOLVER_GUARD(s);
  if (!residual) return fail(BS_ERR_RANGE, "no output");
  return s->fp64 ? residual_vector_t<double>(s, residual) : residual_vector_t<float>(s, residual);
}

// The bar of the precision check (gpu-transonic.ts:108-117 RESIDUAL_PRECISION):
// the largest difference between the device's residual and a double one
// on the same state, relative to the largest residual.
static constexpr double RESIDUAL_PRECISION = 1e-8;

// C++ helpers of the precision check.
extern "C++" {
// Copies `count` values of an FP32 device array into an FP64 one through
// the host (a check made once per solve).
static bs_status widen_copy(hipStream_t stream, const void* src, void* dst, size_t count, const char* what) {
  std::vector<float> f(count);
  std::vector<double> d(count);
  HIP_TRY(copy_on(stream, f.data(), src, count * sizeof(float), hipMemcpyDeviceToHost), what);
  for (size_t k = 0; k < count; k++) d[k] = f[k];
  HIP_TRY(copy_on(stream, dst, d.data(), count * sizeof(double), hipMemcpyHostToDevice), what);
  return BS_OK;
}

// The FP32 solver's state, face states and imposed fluxes, widened, on an
// FP64 solver of the same configuration; its residual vector (the FP64
// stencil kernel, FP64 metrics) is the double one the check compares with.
static bs_status double_residual(bs_solver* s, std::vector<double>& out) {
  bs_ctx ctx{s->device};
  bs_solver* twin = nullptr;
  bs_status st = solver_create(&ctx, &s->cfg, true, &twin);
  if (st) return st;
  auto done = [&](bs_status code) {
    bs_solver_destroy(twin);
    (void)hipSetDevice(s->device);
    return code;
  };
  HIP_TRY(hipStreamSynchronize(s->stream), "precision check");
  const size_t n = s->n;
  if ((st = widen_copy(s->stream, s->d.state[s->cur], twin->d.state[0], 4 * n, "precision check state"))) return done(st);
  twin->cur = 0;
  for (int side = 0; side < 4; side++) {
    if (s->cfg.boundaries[side].kind != BS_BC_FACE_STATES) continue;
    const size_t m = 4 * (size_t)along(s, side);
    if ((st = widen_copy(s->stream, s->d.face_cons[side], twin->d.face_cons[side], m, "precision check faces")))
      return done(st);
    if (s->imposed[side]) {
      if ((st = widen_copy(s->stream, s->d.face_flux[side], twin->d.face_flux[side], m, "precision check fluxes")))
        return done(st);
      twin->imposed[side] = true;
    }
  }
  if ((st = upload_boundary(twin))) return done(st);
  out.assign(4 * n, 0.0);
  st = residual_vector_t<double>(twin, out.data());
  return done(st);
}
}  // extern "C++"

bs_status bs_check_precision(bs_solver* s, double* difference) {
  SOLVER_GUARD(s);
  if (difference) *difference = 0.0;
  Control c;
  bs_status st = read_control(s, &c);
  if (st) return st;
  if (c.status & STATUS_NON_PHYSICAL) return fail(BS_ERR_NONPHYSICAL, "the GPU solution is non-physical");
  // An FP64 solver's residual is the double one (transonic-drag.ts:60-62:
  // the engine in double precision passes at once).
  if (s->fp64) return BS_OK;
  std::vector<double> gpu(4 * (size_t)s->n), ref;
  if ((st = residual_vector_t<float>(s, gpu.data()))) return st;
  if ((st = double_residual(s, ref))) return st;
  // residualPrecision (gpu-transonic.ts:118-131).
  double worst = 0.0, largest = 0.0;
  for (size_t k = 0; k < ref.size(); k++) {
    const double d = std::fabs(gpu[k] - ref[k]);
    worst = std::isnan(d) ? INFINITY : (d > worst ? d : worst);
    largest = std::fabs(ref[k]) > largest ? std::fabs(ref[k]) : largest;
  }
  const double rel = largest > 0 ? worst / largest : (worst > 0 ? INFINITY : 0.0);
  if (difference) *difference = rel;
  if (worst <= RESIDUAL_PRECISION * largest) return BS_OK;
  return fail(BS_ERR_PRECISION,
              "this GPU's residual differs from the double-precision one by %.2e of the largest (at most %g "
              "allowed): an FP32 solver loses the low parts; use the FP64 solver",
              rel, RESIDUAL_PRECISION);
}

// C++ helpers of the entry points below.
extern "C++" {
template <typename S>
static bs_status wall_launch_t(bs_solver* s, int side, double* out_dev, double p_inf, double* partials, int* blocks) {
  WallArgsT<S> A;
  memset(&A, 0, sizeof A);
  A.G = grid_of<S>(s);
  A.s = state_of<S>(s, s->cur);
  A.side = side;
  A.count = along(s, side);
  A.kind = s->cfg.boundaries[side].kind;
  A.face_geo = s->d.face_geo[side];
  A.gamma = s->cfg.gamma;
  A.gamma32 = state_gamma(s);
  A.rec = s->cfg.reconstruction;
  A.limiter = s->cfg.limiter;
  A.flux = s->cfg.flux;
  A.out = out_dev;
  A.p_inf = p_inf;
  A.partials = partials;
  WallBoundaryT<S> WB;
  for (int k = 0; k < 4; k++) {
    const bs_boundary& b = s->cfg.boundaries[k];
    WB.kind[k] = b.kind;
    WB.state[k][0] = b.rho; WB.state[k][1] = b.u; WB.state[k][2] = b.v; WB.state[k][3] = b.p;
    WB.cons[k] = static_cast<const S*>(s->d.face_cons[k]);
  }
  *blocks = (A.count + 255) / 256;
  hipLaunchKernelGGL(wall_fn<S>(s->cfg), dim3(*blocks), dim3(256), 0, s->stream, A, WB);
  HIP_TRY(hipGetLastError(), "wall flux launch");
  return BS_OK;
}
}  // extern "C++"

static bs_status wall_launch(bs_solver* s, int side, double* out_dev, double p_inf, double* partials, int* blocks) {
  return s->fp64 ? wall_launch_t<double>(s, side, out_dev, p_inf, partials, blocks)
                 : wall_launch_t<float>(s, side, out_dev, p_inf, partials, blocks);
}

bs_status bs_wall_fluxes(bs_solver* s, int32_t side, double* faces, int32_t count) {
  SOLVER_GUARD(s);
  if (side < 0 || side > 3) return fail(BS_ERR_RANGE, "no such side");
  if (!faces || count != along(s, side)) return fail(BS_ERR_RANGE, "the side has %d faces", along(s, side));
  if (s->imposed[side]) return fail(BS_ERR_UNSUPPORTED, "the side's fluxes are imposed");
  int blocks;
  bs_status st = wall_launch(s, side, s->d.wall_out, 0.0, nullptr, &blocks);
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(faces, s->d.wall_out, 9 * (size_t)count * sizeof(double), hipMemcpyDeviceToHost, s->stream),
          "wall flux readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  return BS_OK;
}

bs_status bs_forebody_drag(bs_solver* s, int32_t side, double p_inf, double q_inf, double base_radius, double* cd) {
  SOLVER_GUARD(s);
  if (side < 0 || side > 3 || !cd) return fail(BS_ERR_RANGE, "bad side or output");
  int blocks;
  bs_status st = wall_launch(s, side, nullptr, p_inf, s->d.partials, &blocks);
  if (st) return st;
  hipLaunchKernelGGL(finalize_sum, dim3(1), dim3(256), 0, s->stream, s->d.partials, blocks, 1.0, 0, s->d.scalar);
  double force = 0;
  HIP_TRY(hipMemcpyAsync(&force, s->d.scalar, sizeof(double), hipMemcpyDeviceToHost, s->stream), "drag readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  *cd = (2.0 * force) / (q_inf * base_radius * base_radius);
  return BS_OK;
}

bs_status bs_set_implicit_cfl(bs_solver* s, double cfl) {
  SOLVER_GUARD(s);
  if (!(cfl > 0)) return fail(BS_ERR_RANGE, "the CFL number must be positive");
  Control c;
  HIP_TRY(hipMemcpyAsync(&c, s->d.control, sizeof c, hipMemcpyDeviceToHost, s->stream), "control");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  c.cfl = cfl;
  HIP_TRY(hipMemcpyAsync(s->d.control, &c, sizeof c, hipMemcpyHostToDevice, s->stream), "control");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  return BS_OK;
}

static bs_status ensure_implicit(bs_solver* s) {
  if (s->d.res) return BS_OK;
  const size_t n = s->n;
  bs_status st;
  if ((st = dev_alloc(&s->d.res, 4 * n, "implicit residual")) || (st = dev_alloc(&s->d.diag, n, "implicit diagonal")) ||
      (st = dev_alloc(&s->d.lam_i, (size_t)(s->ni + 1) * s->nj, "lambda")) ||
      (st = dev_alloc(&s->d.lam_j, (size_t)s->ni * (s->nj + 1), "lambda")) ||
      (st = dev_alloc(&s->d.dq, 4 * n, "implicit dq")))
    return st;
  return BS_OK;
}

// C++ helpers of the implicit entry point.
extern "C++" {
// One implicit step: prepare (residual, lambda A, diagonal), the
// red-black half-sweeps (the first red one formed from R and D by the
// first launch), the update into the next buffer, which the last
// relaxation launch writes itself: q + alpha dQ from the window's FP32
// state, or, on an FP64 state (relaxed in FP32 on the rounded state and
// metrics), from the FP64 state in HBM.
template <typename T>
static bs_status encode_implicit(bs_solver* s, const StateDevT<float>& res, const StateDevT<float>& dq,
                                 double cfl_max, int sweeps, double omega, double growth, bool dq_only = false) {
  const int src = s->cur, dst = (s->cur + 1) % 3;
  TileArgsT<T> A = tile_args<T>(s);
  A.src = state_of<T>(s, src);
  A.omega = (T)omega;
  for (int q = 0; q < 4; q++) {
    A.res[q] = res.q[q];
    A.dq[q] = dq.q[q];
  }
  A.diag = s->d.diag;
  A.lam_i = s->d.lam_i;
  A.lam_j = s->d.lam_j;
  bs_status r = launch_tile(s, TILE_IMPLICIT, A, 1);
  if (r) return r;
  RelaxArgsT<T> R;
  R.G = grid_of<T>(s);
  R.q = state_of<T>(s, src);
  for (int q = 0; q < 4; q++) {
    R.res[q] = res.q[q];
    R.dq[q] = dq.q[q];
  }
  R.diag = s->d.diag;
  R.lam_i = s->d.lam_i;
  R.lam_j = s->d.lam_j;
  R.gamma = s->gamma32;
  R.omega = (float)omega;
  UpdateArgsT<T> U;
  U.src = state_of<T>(s, src);
  U.dst = state_of<T>(s, dst);
  U.dq = dq;
  U.n = s->n;
  U.gamma = (T)state_gamma(s);
  U.control = s->d.control;
  U.cfl_max = cfl_max;
  U.growth = growth;
  // dq_only: the correction dQ alone (the accelerated steps mix it first);
  // otherwise the last relaxation launch writes the updated state.
  const bool fused = !dq_only;
  // The 2 sweeps - 1 half-sweeps after the first red one, up to three per
  // launch (relax_tile_kernel), starting with black.
  // [form][half-sweeps - 1][first launch]
  using K = void (*)(RelaxArgsT<T>, int, UpdateArgsT<T>, int);
  K kernels[2][3][2] = {};
  if constexpr (std::is_same<T, float>::value) {  // an FP64 state relaxes linearized only
    kernels[0][0][0] = relax_tile_kernel<T, 1, false, false>;
    kernels[0][1][0] = relax_tile_kernel<T, 2, false, false>;
    kernels[0][2][0] = relax_tile_kernel<T, 3, false, false>;
    kernels[0][0][1] = relax_tile_kernel<T, 1, false, true>;
    kernels[0][1][1] = relax_tile_kernel<T, 2, false, true>;
    kernels[0][2][1] = relax_tile_kernel<T, 3, false, true>;
  }
  kernels[1][0][0] = relax_tile_kernel<T, 1, true, false>;
  kernels[1][1][0] = relax_tile_kernel<T, 2, true, false>;
  kernels[1][2][0] = relax_tile_kernel<T, 3, true, false>;
  kernels[1][0][1] = relax_tile_kernel<T, 1, true, true>;
  kernels[1][1][1] = relax_tile_kernel<T, 2, true, true>;
  kernels[1][2][1] = relax_tile_kernel<T, 3, true, true>;
  const int form = (s->relax_linear || !std::is_same<T, float>::value) ? 1 : 0;
  dim3 rgrid(((s->ni + RTX - 1) / RTX) * ((s->nj + RTY - 1) / RTY));
  for (int done = 0, colour = 1, halves = 2 * sweeps - 1; done < halves;) {
    const int nh = halves - done < 3 ? halves - done : 3;
    const int init = done == 0, last = fused && done + nh == halves;
    hipEvent_t* ev = timing_begin(s, 2);
    hipLaunchKernelGGL(kernels[form][nh - 1][init], rgrid, dim3(RNT), 0, s->stream, R, colour, U, last);
    timing_end(s, ev);
    colour = (colour + nh) & 1;
    done += nh;
  }
  if (dq_only) {
    HIP_TRY(hipGetLastError(), "implicit launch");
    return BS_OK;
  }
  HIP_TRY(hipGetLastError(), "implicit launch");
  s->cur = dst;
  return BS_OK;
}
}  // extern "C++"

bs_status bs_iterate_implicit(bs_solver* s, int32_t count, double cfl_max, int32_t sweeps, double omega, double growth,
                              bs_report* report) {
  SOLVER_GUARD(s);
  if (!(cfl_max > 0 && omega >= 1 && growth >= 1))
    return fail(BS_ERR_RANGE, "implicit steps need cfl > 0, omega >= 1 and growth >= 1");
  if (sweeps < 1) return fail(BS_ERR_RANGE, "the relaxation needs a whole number of sweeps, at least one");
  if (s->fitted) return fail(BS_ERR_UNSUPPORTED, "implicit steps do not move a fitted shock");
  bs_status st = ensure_implicit(s);
  if (st) return st;
  const size_t n = s->n;
  StateDevT<float> res, dq;
  for (int q = 0; q < 4; q++) {
    res.q[q] = s->d.res + q * n;
    dq.q[q] = s->d.dq + q * n;
  }
  auto encode = [&]() -> bs_status {
    return s->fp64 ? encode_implicit<double>(s, res, dq, cfl_max, sweeps, omega, growth)
                   : encode_implicit<float>(s, res, dq, cfl_max, sweeps, omega, growth);
  };
  StepGraph& g = s->implicit_graph;
  if (g.exec && (g.sweeps != sweeps || g.omega != omega || g.growth != growth || g.cfl_max != cfl_max))
    drop_graph(g);
  Control c;
  int remaining = count;
  while (remaining > 0) {
    const int batch = remaining < 64 ? remaining : 64;
    if ((st = run_steps(s, g, batch, encode))) return st;
    g.sweeps = sweeps;
    g.omega = omega;
    g.growth = growth;
    g.cfl_max = cfl_max;
    if ((st = read_control(s, &c))) return st;
    fill_report(c, report);
    if ((st = account(s, c, batch))) return st;
    remaining -= batch;
  }
  if (count == 0) {
    if ((st = read_control(s, &c))) return st;
    fill_report(c, report);
  }
  return BS_OK;
}

// ---- Anderson acceleration ---------------------------------------------------

}  // extern "C"

struct bs_anderson {
  long long id;  // unique per history: a captured graph names its buffers
  int device;
  long long size;
  int depth;
  int blocks;  // of the push and update kernels (fixed: the partials' order)
  hipStream_t stream = nullptr;
  double *dg = nullptr, *df = nullptr, *last_x = nullptr, *last_f = nullptr, *partials = nullptr;
  AndersonCtl* ctl = nullptr;
  double *x_dev = nullptr, *f_dev = nullptr, *out_dev = nullptr;  // bs_anderson_next staging
  std::atomic<bool> busy{false};
};

namespace {

// The push kernel whose accumulators cover the depth.
template <typename TX, typename TF>
void (*push_kernel(int depth))(const TX*, const TF*, AndersonArgs) {
  if (depth <= 1) return anderson_push<TX, TF, 1>;
  if (depth <= 2) return anderson_push<TX, TF, 2>;
  if (depth <= 4) return anderson_push<TX, TF, 4>;
  if (depth <= 8) return anderson_push<TX, TF, 8>;
  return anderson_push<TX, TF, 16>;
}

AndersonArgs anderson_args(bs_anderson* a) {
  AndersonArgs A;
  A.dg = a->dg;
  A.df = a->df;
  A.last_x = a->last_x;
  A.last_f = a->last_f;
  A.partials = a->partials;
  A.ctl = a->ctl;
  A.size = a->size;
  A.depth = a->depth;
  return A;
}

void anderson_free(bs_anderson* a) {
  void* all[] = {a->dg, a->df, a->last_x, a->last_f, a->partials, a->ctl, a->x_dev, a->f_dev, a->out_dev};
  for (void* p : all)
    if (p) (void)hipFree(p);
  if (a->stream) (void)hipStreamDestroy(a->stream);
}

struct AndersonGuard {
  bs_anderson* a;
  bool ok;
  explicit AndersonGuard(bs_anderson* x) : a(x) {
    bool expected = false;
    ok = a->busy.compare_exchange_strong(expected, true);
  }
  ~AndersonGuard() {
    if (ok) a->busy.store(false);
  }
};

// One accelerated implicit step (transonic-flow.ts:184-208): the
// correction f at x, the reset rule, the mixed iterate x + f - (dX + dF)
// gamma, the plain step where that is not physical, the CFL ramp.
template <typename T>
bs_status encode_anderson_step(bs_solver* s, bs_anderson* a, const StateDevT<float>& res,
                               const StateDevT<float>& dq, double cfl_max, int sweeps, double omega, double growth) {
  bs_status st = encode_implicit<T>(s, res, dq, cfl_max, sweeps, omega, growth, true);
  if (st) return st;
  const int src = s->cur, dst = (s->cur + 1) % 3;
  const T* x = static_cast<const T*>(s->d.state[src]);
  T* out = static_cast<T*>(s->d.state[dst]);
  const float* f = s->d.dq;
  const AndersonArgs A = anderson_args(a);
  hipEvent_t* ev = timing_begin(s, 3);
  hipLaunchKernelGGL((push_kernel<T, float>(a->depth)), dim3(a->blocks), dim3(256), 0, s->stream, x, f, A);
  hipLaunchKernelGGL(anderson_solve, dim3(1), dim3(AA_SOLVE_THREADS), 0, s->stream, A, a->blocks, 1);
  const int cb = (s->n + 255) / 256;
  hipLaunchKernelGGL(anderson_update_cells<T>, dim3(cb), dim3(256), 0, s->stream, x, f, out, s->n,
                     (T)state_gamma(s), A);
  UpdateArgsT<T> U;
  memset(&U, 0, sizeof U);
  U.n = s->n;
  U.gamma = (T)state_gamma(s);
  U.control = s->d.control;
  U.cfl_max = cfl_max;
  U.growth = growth;
  hipLaunchKernelGGL(anderson_fallback<T>, dim3(cb), dim3(256), 0, s->stream, x, f, out, s->n, U, A);
  timing_end(s, ev);
  HIP_TRY(hipGetLastError(), "anderson launch");
  s->cur = dst;
  return BS_OK;
}

}  // namespace

extern "C" {

bs_status bs_anderson_create(bs_ctx* ctx, int64_t size, int32_t depth, bs_anderson** out) {
  if (!ctx || !out) return fail(BS_ERR_RANGE, "null argument");
  *out = nullptr;
  if (size < 1) return fail(BS_ERR_RANGE, "the vectors need at least one component");
  if (depth < 1 || depth > AA_MAX_DEPTH) return fail(BS_ERR_RANGE, "depth must be 1 to %d", AA_MAX_DEPTH);
  (void)hipSetDevice(ctx->device);
  static std::atomic<long long> next_id{1};
  bs_anderson* a = new bs_anderson();
  a->id = next_id++;
  a->device = ctx->device;
  a->size = size;
  a->depth = depth;
  const long long want = (size + 255) / 256;
  a->blocks = (int)(want < 2 * AA_SOLVE_THREADS ? want : 2 * AA_SOLVE_THREADS);  // anderson_solve reads two per thread
  auto bail = [&](bs_status code) {
    anderson_free(a);
    delete a;
    return code;
  };
  bs_status st;
  if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "hipStreamCreate failed"));
  const size_t n = (size_t)size;
  if ((st = dev_alloc(&a->dg, (size_t)depth * n, "anderson dX + dF")) ||
      (st = dev_alloc(&a->df, (size_t)depth * n, "anderson dF")) ||
      (st = dev_alloc(&a->last_x, n, "anderson x")) || (st = dev_alloc(&a->last_f, n, "anderson f")) ||
      (st = dev_alloc(&a->partials, (size_t)a->blocks * AA_VALUES, "anderson partials")) ||
      (st = dev_alloc(&a->ctl, 1, "anderson control")))
    return bail(st);
  AndersonCtl c;
  memset(&c, 0, sizeof c);
  c.best = 1.0e308;
  if (copy_on(a->stream, a->ctl, &c, sizeof c, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "anderson control upload"));
  *out = a;
  return BS_OK;
}

bs_status bs_anderson_destroy(bs_anderson* a) {
  if (!a) return BS_OK;
  bool idle = false;
  if (!a->busy.compare_exchange_strong(idle, true))
    return fail(BS_ERR_BUSY, "bs_anderson: await the previous operation first (destroy while a call runs)");
  (void)hipSetDevice(a->device);
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  anderson_free(a);
  delete a;
  return BS_OK;
}

bs_status bs_anderson_reset(bs_anderson* a) {
  if (!a) return fail(BS_ERR_RANGE, "no history");
  AndersonGuard g(a);
  if (!g.ok) return fail(BS_ERR_BUSY, "bs_anderson: await the previous operation first");
  (void)hipSetDevice(a->device);
  hipLaunchKernelGGL(anderson_reset_kernel, dim3(1), dim3(1), 0, a->stream, a->ctl, 1);
  HIP_TRY(hipStreamSynchronize(a->stream), "anderson reset");
  return BS_OK;
}

bs_status bs_anderson_next(bs_anderson* a, const double* x, const double* f, double* out) {
  if (!a || !x || !f || !out) return fail(BS_ERR_RANGE, "null argument");
  AndersonGuard g(a);
  if (!g.ok) return fail(BS_ERR_BUSY, "bs_anderson: await the previous operation first");
  (void)hipSetDevice(a->device);
  const size_t n = (size_t)a->size;
  bs_status st;
  if (!a->x_dev && ((st = dev_alloc(&a->x_dev, n, "anderson staging")) ||
                    (st = dev_alloc(&a->f_dev, n, "anderson staging")) ||
                    (st = dev_alloc(&a->out_dev, n, "anderson staging"))))
    return st;
  HIP_TRY(hipMemcpyAsync(a->x_dev, x, n * sizeof(double), hipMemcpyHostToDevice, a->stream), "anderson x");
  HIP_TRY(hipMemcpyAsync(a->f_dev, f, n * sizeof(double), hipMemcpyHostToDevice, a->stream), "anderson f");
  const AndersonArgs A = anderson_args(a);
  hipLaunchKernelGGL((push_kernel<double, double>(a->depth)), dim3(a->blocks), dim3(256), 0, a->stream, a->x_dev, a->f_dev,
                     A);
  hipLaunchKernelGGL(anderson_solve, dim3(1), dim3(AA_SOLVE_THREADS), 0, a->stream, A, a->blocks, 0);
  hipLaunchKernelGGL((anderson_update<double, double, double>), dim3(a->blocks), dim3(256), 0, a->stream, a->x_dev,
                     a->f_dev, a->out_dev, A);
  HIP_TRY(hipGetLastError(), "anderson launch");
  HIP_TRY(hipMemcpyAsync(out, a->out_dev, n * sizeof(double), hipMemcpyDeviceToHost, a->stream), "anderson out");
  HIP_TRY(hipStreamSynchronize(a->stream), "anderson next");
  return BS_OK;
}

bs_status bs_anderson_state(bs_anderson* a, int32_t* columns, int32_t* used, int32_t* restarts, double* gamma) {
  if (!a) return fail(BS_ERR_RANGE, "no history");
  AndersonGuard g(a);
  if (!g.ok) return fail(BS_ERR_BUSY, "bs_anderson: await the previous operation first");
  (void)hipSetDevice(a->device);
  AndersonCtl c;
  HIP_TRY(copy_on(a->stream, &c, a->ctl, sizeof c, hipMemcpyDeviceToHost), "anderson state");
  if (columns) *columns = c.count;
  if (used) *used = c.used;
  if (restarts) *restarts = c.restarts;
  if (gamma)  // oldest column first
    for (int k = 0; k < a->depth; k++) gamma[k] = k < c.count ? c.gamma[(c.oldest + k) % a->depth] : 0.0;
  return BS_OK;
}

bs_status bs_iterate_anderson(bs_solver* s, bs_anderson* a, int32_t count, double cfl_max, int32_t sweeps,
                              double omega, double growth, bs_report* report) {
  SOLVER_GUARD(s);
  if (!a) return fail(BS_ERR_RANGE, "no history");
  AndersonGuard ag(a);
  if (!ag.ok) return fail(BS_ERR_BUSY, "bs_anderson: await the previous operation first");
  if (a->size != 4LL * s->n) return fail(BS_ERR_RANGE, "the history holds %lld values, the state %lld", a->size,
                                         4LL * s->n);
  if (a->device != s->device) return fail(BS_ERR_RANGE, "the history is on another device");
  if (!(cfl_max > 0 && omega >= 1 && growth >= 1))
    return fail(BS_ERR_RANGE, "implicit steps need cfl > 0, omega >= 1 and growth >= 1");
  if (sweeps < 1) return fail(BS_ERR_RANGE, "the relaxation needs a whole number of sweeps, at least one");
  if (s->fitted) return fail(BS_ERR_UNSUPPORTED, "implicit steps do not move a fitted shock");
  bs_status st = ensure_implicit(s);
  if (st) return st;
  // The history's own stream may still run a reset or a next.
  HIP_TRY(hipStreamSynchronize(a->stream), "anderson stream");
  const size_t n = s->n;
  StateDevT<float> res, dq;
  for (int q = 0; q < 4; q++) {
    res.q[q] = s->d.res + q * n;
    dq.q[q] = s->d.dq + q * n;
  }
  auto encode = [&]() -> bs_status {
    return s->fp64 ? encode_anderson_step<double>(s, a, res, dq, cfl_max, sweeps, omega, growth)
                   : encode_anderson_step<float>(s, a, res, dq, cfl_max, sweeps, omega, growth);
  };
  StepGraph& g = s->anderson_graph;
  if (g.exec && (g.sweeps != sweeps || g.omega != omega || g.growth != growth || g.cfl_max != cfl_max ||
                 s->anderson_for != a->id))
    drop_graph(g);
  Control c;
  int remaining = count;
  while (remaining > 0) {
    const int batch = remaining < 64 ? remaining : 64;
    if ((st = run_steps(s, g, batch, encode))) return st;
    g.sweeps = sweeps;
    g.omega = omega;
    g.growth = growth;
    g.cfl_max = cfl_max;
    s->anderson_for = a->id;
    if ((st = read_control(s, &c))) return st;
    fill_report(c, report);
    if ((st = account(s, c, batch))) return st;
    remaining -= batch;
  }
  if (count == 0) {
    if ((st = read_control(s, &c))) return st;
    fill_report(c, report);
  }
  return BS_OK;
}

bs_status bs_set_fitted_shock(bs_solver* s, const double* free_stream) {
  SOLVER_GUARD(s);
  if (free_stream) {
    if (s->cfg.boundaries[JMAX].kind != BS_BC_FACE_STATES)
      return fail(BS_ERR_RANGE, "a fitted shock needs a 'faceStates' jMax boundary");
    if (s->nj < 2) return fail(BS_ERR_RANGE, "a fitted shock needs two rows of cells below it");
    if (!physical(free_stream)) return fail(BS_ERR_NONPHYSICAL, "the free stream is not physical");
    if (!s->d.shock_out) {
      bs_status st = dev_alloc(&s->d.shock_out, 3 * (size_t)s->ni, "shock");
      if (st) return st;
      HIP_TRY(hipMemsetAsync(s->d.shock_out, 0, 3 * (size_t)s->ni * sizeof(double), s->stream), "shock");
      HIP_TRY(hipStreamSynchronize(s->stream), "shock");
    }
    for (int k = 0; k < 4; k++) s->shock_free[k] = free_stream[k];
  }
  s->fitted = free_stream != nullptr;
  s->imposed[JMAX] = s->fitted;
  drop_graph(s->explicit_graph);
  drop_graph(s->fitted_graph);
  return upload_boundary(s);
}

bs_status bs_read_shock(bs_solver* s, double* speed, double* mach, double* signal, int32_t count) {
  SOLVER_GUARD(s);
  if (!s->fitted) return fail(BS_ERR_RANGE, "there is no fitted shock");
  if (count != s->ni) return fail(BS_ERR_RANGE, "the shock has %d faces", s->ni);
  std::vector<double> out(3 * (size_t)s->ni);
  HIP_TRY(hipMemcpyAsync(out.data(), s->d.shock_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream),
          "shock readback");
  HIP_TRY(hipStreamSynchronize(s->stream), "sync");
  for (int i = 0; i < s->ni; i++) {
    if (speed) speed[i] = out[3 * (size_t)i];
    if (mach) mach[i] = out[3 * (size_t)i + 1];
    if (signal) signal[i] = out[3 * (size_t)i + 2];
  }
  return BS_OK;
}

}  // extern "C"

namespace {

LineGridDev line_grid_of(bs_solver* s) {
  LineGridDev L;
  L.ni = s->ni;
  L.nj = s->nj;
  L.axisymmetric = s->cfg.geometry == BS_GEOMETRY_AXISYMMETRIC;
  L.symmetry = s->symmetry;
  L.lines = s->d.line;
  L.eta = s->d.eta;
  L.h = s->d.height;
  L.move = s->d.move;
  return L;
}

template <typename T>
void launch_line_metrics(bs_solver* s) {
  MetricOut<T> M;
  M.inx = static_cast<T*>(s->d.inx); M.iny = static_cast<T*>(s->d.iny); M.ia = static_cast<T*>(s->d.ia);
  M.jnx = static_cast<T*>(s->d.jnx); M.jny = static_cast<T*>(s->d.jny); M.ja = static_cast<T*>(s->d.ja);
  M.vol = static_cast<T*>(s->d.vol); M.planar = static_cast<T*>(s->d.planar);
  for (int k = 0; k < 4; k++) M.geo[k] = s->d.face_geo[k];
  const long long total = (long long)(s->ni + 1) * s->nj + (long long)s->ni * (s->nj + 1) + (long long)s->n;
  hipLaunchKernelGGL(line_metrics_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s->stream,
                     line_grid_of(s), M);
}

void line_metrics(bs_solver* s) { s->fp64 ? launch_line_metrics<double>(s) : launch_line_metrics<float>(s); }

// The flow in the current buffer moved with the lines into the spare one.
template <typename T>
void launch_line_remap(bs_solver* s) {
  const int spare = (s->cur + 1) % 3;
  hipLaunchKernelGGL(line_remap_kernel<T>, dim3((s->n + 255) / 256), dim3(256), 0, s->stream, line_grid_of(s),
                     state_of<T>(s, s->cur), state_of<T>(s, spare), state_gamma(s));
  s->cur = spare;
}

void line_remap(bs_solver* s) { s->fp64 ? launch_line_remap<double>(s) : launch_line_remap<float>(s); }

// One move of the fitted shock (GpuLineGrid.encodeShockMove): the nodes by
// the last step's shock speeds, the flow remapped, the moved metrics.
void encode_shock_move(bs_solver* s) {
  hipLaunchKernelGGL(shock_moves_kernel, dim3(1), dim3(256), 0, s->stream, line_grid_of(s), s->d.shock_out);
  line_remap(s);
  line_metrics(s);
}

}  // namespace

extern "C" {

bs_status bs_set_line_grid(bs_solver* s, const double* lines, const double* fractions, int32_t first_node_symmetry) {
  SOLVER_GUARD(s);
  if (!lines || !fractions) return fail(BS_ERR_RANGE, "the lines and their fractions are needed");
  const int ni = s->ni, nj = s->nj;
  if (first_node_symmetry && ni < 2) return fail(BS_ERR_RANGE, "a first node normal to the axis needs three lines");
  for (int i = 0; i <= ni; i++) {
    const double* l = lines + 4 * i;
    if (!(std::fabs(std::hypot(l[2], l[3]) - 1.0) < 1e-12)) return fail(BS_ERR_RANGE, "line %d's direction is not a unit vector", i);
    for (int j = 0; j < nj; j++)
      if (!(fractions[i + (size_t)(j + 1) * (ni + 1)] > fractions[i + (size_t)j * (ni + 1)]))
        return fail(BS_ERR_RANGE, "the fractions of line %d do not increase", i);
  }
  bs_status st;
  const size_t nl = (size_t)ni + 1;
  if (!s->d.line && ((st = dev_alloc(&s->d.line, 4 * nl, "lines")) ||
                     (st = dev_alloc(&s->d.eta, nl * (nj + 1), "fractions")) ||
                     (st = dev_alloc(&s->d.height, nl, "heights")) || (st = dev_alloc(&s->d.move, nl, "moves"))))
    return st;
  drop_graph(s->fitted_graph);  // the symmetry flag is captured with the moves
  HIP_TRY(copy_on(s->stream, s->d.line, lines, 4 * nl * sizeof(double), hipMemcpyHostToDevice), "lines");
  HIP_TRY(copy_on(s->stream, s->d.eta, fractions, nl * (nj + 1) * sizeof(double), hipMemcpyHostToDevice), "fractions");
  s->lines = true;
  s->symmetry = first_node_symmetry != 0;
  return BS_OK;
}

bs_status bs_set_heights(bs_solver* s, const double* heights, int32_t count) {
  SOLVER_GUARD(s);
  if (!s->lines) return fail(BS_ERR_RANGE, "set the lines first (bs_set_line_grid)");
  if (!heights) return fail(BS_ERR_RANGE, "no heights");
  if (count != s->ni + 1) return fail(BS_ERR_RANGE, "expected %d heights, got %d", s->ni + 1, count);
  const size_t nl = (size_t)s->ni + 1;
  for (size_t i = 0; i < nl; i++)
    if (!(heights[i] > 0)) return fail(BS_ERR_RANGE, "height %zu is not positive", i);
  std::vector<double> zero(nl, 0.0);
  HIP_TRY(copy_on(s->stream, s->d.height, heights, nl * sizeof(double), hipMemcpyHostToDevice), "heights");
  HIP_TRY(copy_on(s->stream, s->d.move, zero.data(), nl * sizeof(double), hipMemcpyHostToDevice), "moves");
  line_metrics(s);
  HIP_TRY(hipGetLastError(), "line metrics");
  HIP_TRY(hipStreamSynchronize(s->stream), "line metrics");
  return BS_OK;
}

bs_status bs_read_heights(bs_solver* s, double* heights, int32_t count) {
  SOLVER_GUARD(s);
  if (!s->lines || !heights) return fail(BS_ERR_RANGE, "no lines or no output");
  if (count != s->ni + 1) return fail(BS_ERR_RANGE, "expected room for %d heights, got %d", s->ni + 1, count);
  HIP_TRY(hipMemcpyAsync(heights, s->d.height, ((size_t)s->ni + 1) * sizeof(double), hipMemcpyDeviceToHost,
                         s->stream), "heights");
  HIP_TRY(hipStreamSynchronize(s->stream), "heights");
  return BS_OK;
}

bs_status bs_move_grid(bs_solver* s, const double* heights, int32_t count) {
  SOLVER_GUARD(s);
  if (!s->lines) return fail(BS_ERR_RANGE, "set the lines first (bs_set_line_grid)");
  if (!heights) return fail(BS_ERR_RANGE, "no heights");
  if (count != s->ni + 1) return fail(BS_ERR_RANGE, "expected %d heights, got %d", s->ni + 1, count);
  const size_t nl = (size_t)s->ni + 1;
  std::vector<double> now(nl), move(nl);
  HIP_TRY(hipMemcpyAsync(now.data(), s->d.height, nl * sizeof(double), hipMemcpyDeviceToHost, s->stream), "heights");
  HIP_TRY(hipStreamSynchronize(s->stream), "heights");
  for (size_t i = 0; i < nl; i++) {
    if (!(heights[i] > 0)) return fail(BS_ERR_RANGE, "height %zu is not positive", i);
    move[i] = heights[i] - now[i];
  }
  HIP_TRY(copy_on(s->stream, s->d.height, heights, nl * sizeof(double), hipMemcpyHostToDevice), "heights");
  HIP_TRY(copy_on(s->stream, s->d.move, move.data(), nl * sizeof(double), hipMemcpyHostToDevice), "moves");
  line_remap(s);
  line_metrics(s);
  HIP_TRY(hipGetLastError(), "grid move");
  HIP_TRY(hipStreamSynchronize(s->stream), "grid move");
  return BS_OK;
}

bs_status bs_iterate_fitted(bs_solver* s, int32_t moves, int32_t steps_per_move, bs_report* report) {
  SOLVER_GUARD(s);
  if (!s->fitted) return fail(BS_ERR_RANGE, "there is no fitted shock");
  if (!s->lines) return fail(BS_ERR_RANGE, "set the lines first (bs_set_line_grid)");
  if (moves < 0) return fail(BS_ERR_RANGE, "the move count cannot be negative");
  if (steps_per_move < 1 || steps_per_move > BATCH)
    return fail(BS_ERR_RANGE, "steps per move must be an integer from 1 to %d", BATCH);
  const int per_batch = BATCH / steps_per_move;
  // One move: its steps, then the shock move. Whole graphs of moves are
  // replayed where they fit (each move advances the state buffer by the
  // remap's one, so GRAPH_STEPS moves return to it).
  auto encode_move = [s, steps_per_move]() -> bs_status {
    for (int k = 0; k < steps_per_move; k++) {
      bs_status st = encode_step(s);
      if (st) return st;
    }
    encode_shock_move(s);
    return BS_OK;
  };
  if (s->fitted_for != steps_per_move) {
    drop_graph(s->fitted_graph);
    s->fitted_for = steps_per_move;
  }
  Control c;
  int remaining = moves;
  while (remaining > 0) {
    const int batch = remaining < per_batch ? remaining : per_batch;
    bs_status st0 = run_steps(s, s->fitted_graph, batch, encode_move);
    if (st0) return st0;
    HIP_TRY(hipGetLastError(), "fitted steps");
    bs_status st = read_control(s, &c);
    if (st) return st;
    fill_report(c, report);
    if ((st = account(s, c, (long long)batch * steps_per_move))) return st;
    remaining -= batch;
  }
  if (moves == 0) {
    bs_status st = read_control(s, &c);
    if (st) return st;
    fill_report(c, report);
  }
  return BS_OK;
}

bs_status bs_set_relaxation(bs_solver* s, int32_t form) {
  SOLVER_GUARD(s);
  if (form != BS_RELAX_FLUX_DIFFERENCE && form != BS_RELAX_LINEARIZED)
    return fail(BS_ERR_RANGE, "unknown relaxation form %d", form);
  if (s->fp64 && form != BS_RELAX_LINEARIZED)
    return fail(BS_ERR_UNSUPPORTED, "an FP64 solver relaxes with the linearized neighbour term only");
  const bool linear = form == BS_RELAX_LINEARIZED;
  if (linear != s->relax_linear) {
    drop_graph(s->implicit_graph);
    drop_graph(s->anderson_graph);
  }
  s->relax_linear = linear;
  return BS_OK;
}

bs_status bs_enable_graphs(bs_solver* s, int32_t enable) {
  SOLVER_GUARD(s);
  s->graphs = enable != 0;
  if (!s->graphs) {
    drop_graph(s->explicit_graph);
    drop_graph(s->implicit_graph);
    drop_graph(s->anderson_graph);
    drop_graph(s->fitted_graph);
  }
  return BS_OK;
}

bs_status bs_enable_timing(bs_solver* s, int32_t enable) {
  SOLVER_GUARD(s);
  s->timing.enabled = enable != 0;
  s->timing.region = enable == 2;
  for (int k = 0; k < 4; k++) {
    s->timing.total_ms[k] = 0;
    s->timing.launches[k] = 0;
  }
  return BS_OK;
}

bs_status bs_kernel_timing(bs_solver* s, int32_t kernel, double* ms_per_launch, int64_t* launches) {
  SOLVER_GUARD(s);
  if (kernel < 0 || kernel > 3) return fail(BS_ERR_RANGE, "no such kernel");
  long long L = s->timing.launches[kernel];
  if (ms_per_launch) *ms_per_launch = L > 0 ? s->timing.total_ms[kernel] / (double)L : 0.0;
  if (launches) *launches = L;
  s->timing.total_ms[kernel] = 0;
  s->timing.launches[kernel] = 0;
  return BS_OK;
}

}  // extern "C"

