// Batched solves: many grids of uneven size advanced together, one launch
// per stage for every active case (bs_batch_* in include/bowshock.h).
//
// Layout: every per-cell and per-face array of all cases is one arena in
// HBM, case c owning a contiguous range (its four state components
// contiguous inside its range). The kernels are the single-grid stencil,
// relaxation and update bodies; a block reads its work item (case, tile,
// partial slot) and the case's argument record from device memory. Records
// are role-major (a role being a stage of the scheme for a given current
// state buffer), so a launch picks its role by the base pointer it passes.
//
// A batch holds its state, metrics and time steps in FP32, or in FP64
// (bs_batch_create_fp64): the FP64 solver's scheme, whose implicit steps
// form the residual, diagonal and update in FP64 around an FP32 relaxation
// with the linearized neighbour term. The host functions below are
// templates on the state type; the arenas are untyped.
//
// Cases are split into two size classes by cell count, with a large tile
// for big grids and a small one for small grids where the large tile would
// leave most of its threads idle; the threshold is set by the caller's
// autotuner (bs_batch_set_tiling). Each step of the active set is captured
// as HIP graphs; the work lists and graphs are rebuilt when the active set
// or the tiling changes.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "bs_batch_tiles.hpp"
#include "bs_persistent.hpp"

using namespace bs;

namespace {

struct BatchCase {
  int ni, nj, n;
  size_t cell, iface, jface, wall;  // arena offsets
  int tiles_x[2], tiles_y[2];       // per tile class
  int slot0, slots;                 // partial slots of its stencil launches
  int wall_slot0, wall_blocks;
  int cls;                          // tile class: 0 large, 1 small
};

enum BatchRole {
  ROLE_STAGE0 = 0,   // + 3 * cur + k for stage k (9 roles)
  ROLE_RESIDUAL = 9, // + cur
  ROLE_PREPARE = 12, // + cur
  N_TILE_ROLES = 15
};

template <typename S>
using BatchWallFnT = void (*)(const WallArgsT<S>*, const WallBoundaryT<S>*, const WorkItem*);

template <typename S, int REC, int FLUX>
__global__ __launch_bounds__(256) void batch_wall_kernel(const WallArgsT<S>* __restrict__ args,
                                                         const WallBoundaryT<S>* __restrict__ wb,
                                                         const WorkItem* __restrict__ work) {
  const WorkItem w = work[blockIdx.x];
  wall_body<S, REC, FLUX>(args[w.args], wb[w.by], w.bx, w.slot);
}

template <typename S>
BatchWallFnT<S> batch_wall_fn(const bs_config& c) {
  if (c.flux == BS_FLUX_RUSANOV) {
    if (c.reconstruction == SHARED) return batch_wall_kernel<S, SHARED, RUSANOV>;
    if (c.reconstruction == SHARED_SCALED) return batch_wall_kernel<S, SHARED_SCALED, RUSANOV>;
    return batch_wall_kernel<S, PER_VARIABLE, RUSANOV>;
  }
  if (c.reconstruction == SHARED) return batch_wall_kernel<S, SHARED, HLLC>;
  if (c.reconstruction == SHARED_SCALED) return batch_wall_kernel<S, SHARED_SCALED, HLLC>;
  return batch_wall_kernel<S, PER_VARIABLE, HLLC>;
}

// Sum of each case's range of partials, in a fixed order (one block per
// case): out[c] = sum, or sqrt(sum / cells) when `rms`.
__global__ __launch_bounds__(256) void segment_sums(const double* __restrict__ partials, const int* __restrict__ first,
                                                    const int* __restrict__ count, const int* __restrict__ cells,
                                                    const int* __restrict__ cases, int rms, double* __restrict__ out) {
  __shared__ double red[4];
  const int c = cases[blockIdx.x];
  double v = 0.0;
  for (int k = threadIdx.x; k < count[c]; k += 256) v += partials[first[c] + k];
  double s = block_sum<256>(v, red);
  if (threadIdx.x == 0) out[c] = rms ? sqrt(s / (double)cells[c]) : s;
}

}  // namespace

struct bs_batch {
  int device;
  hipStream_t stream = nullptr;  // every kernel and copy of the batch
  std::map<const void*, size_t> capacity;  // elements allocated per work-list array
  bs_config scheme;
  float gamma32;
  bool fp64 = false;  // state, metrics and time steps in FP64
  int count;
  std::vector<BatchCase> cases;
  size_t cells = 0, ifaces = 0, jfaces = 0, walls = 0;
  int slots = 0, wall_slots = 0;
  // arenas of the state's precision
  void* state[3] = {nullptr, nullptr, nullptr};
  void *inx = nullptr, *iny = nullptr, *ia = nullptr, *jnx = nullptr, *jny = nullptr, *ja = nullptr;
  void *vol = nullptr, *planar = nullptr, *dt = nullptr;
  // implicit work, FP32 in both precisions (the relaxation's)
  float *res = nullptr, *diag = nullptr, *lam_i = nullptr, *lam_j = nullptr, *dq = nullptr;
  double* wall_geo = nullptr;
  double* partials = nullptr;
  double* wall_partials = nullptr;
  double* sums = nullptr;
  Control* control = nullptr;
  void* bnd = nullptr;       // BoundaryDevT per case
  void* wall_bnd = nullptr;  // WallBoundaryT per case
  // argument records (TileArgsT, RelaxArgsT, UpdateArgsT, WallArgsT of the
  // state's type) and work lists (device)
  void* tile_args = nullptr;    // N_TILE_ROLES x count
  void* relax_args = nullptr;   // 3 (cur) x count
  void* update_args = nullptr;  // 3 (cur) x count
  void* wall_args = nullptr;    // 3 (cur) x count
  WorkItem* work_tiles[2] = {nullptr, nullptr};
  WorkItem* work_relax = nullptr;
  WorkItem* work_wall = nullptr;
  int n_tiles[2] = {0, 0}, n_relax = 0, n_wall = 0;
  bool relax_linear = false;  // BS_RELAX_LINEARIZED (always on an FP64 batch)
  int relax_fusion = 3;       // half-sweeps per relaxation launch at most (bs_batch_set_relax_fusion)
  int *seg_first = nullptr, *seg_count = nullptr, *seg_cells = nullptr, *seg_cases = nullptr;
  int *wall_first = nullptr, *wall_count = nullptr;
  int n_seg = 0;
  std::vector<char> active;
  int small_below = 0;  // cases with fewer cells take the small tile
  bool lists_dirty = true;
  bool implicit_ready = false;
  double upd_cfl_max = -1, upd_growth = -1;
  int cur = 0;
  StepGraph explicit_graph, implicit_graph;
  // bs_batch_set_persistent: explicit steps in persistent launches of up to
  // PERSIST_STEPS steps (bs_persistent.hpp) when the active cases' tiles fit
  // on the chip at once; 0 off, 1 32 x 4 tiles, 2 32 x 8.
  int persist_tile = 0;
  bool persist_dirty = true;
  int persist_blocks = 0;      // blocks of a launch; 0: the active set does not fit
  int persist_stride = 0;      // the longest XCD list
  int persist_capacity = 0;    // co-resident blocks per XCD
  int persist_used = 0;        // the last bs_batch_iterate ran persistent launches
  PersistItem* persist_items = nullptr;  // 8 lists of persist_stride tiles
  int* persist_lens = nullptr;           // tiles per list
  int* persist_claims = nullptr;         // per XCD: blocks that have taken a tile
  int* persist_bars = nullptr;           // one barrier counter per case
  unsigned* persist_fault = nullptr;
  // bs_batch_enable_timing: HIP-event pairs around each stencil and
  // relaxation launch (kind 0 stage, 1 prepare, 2 relaxation); the steps
  // then run without graphs.
  struct {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<int> kinds;
    std::vector<int> weights;  // passes per pair (a persistent launch holds many)
    int used = 0;
    double total_ms[3] = {0, 0, 0};
    long long launches[3] = {0, 0, 0};
  } timing;
  std::atomic<bool> busy{false};
};

namespace {

struct BatchGuard {
  bs_batch* b;
  bool ok;
  explicit BatchGuard(bs_batch* batch) : b(batch) {
    bool expected = false;
    ok = b->busy.compare_exchange_strong(expected, true);
  }
  ~BatchGuard() {
    if (ok) b->busy.store(false);
  }
};

#define BATCH_GUARD(b)                                                                        \
  if ((b) == nullptr) return fail(BS_ERR_RANGE, "no batch");                                  \
  BatchGuard _bguard(b);                                                                      \
  if (!_bguard.ok) return fail(BS_ERR_BUSY, "bs_batch: await the previous operation first"); \
  (void)hipSetDevice((b)->device)

// The gamma the state's conserved variables are formed with: FP32 for an
// FP32 state (euler2d-gpu.ts:412-432), the configured one for FP64.
double state_gamma(const bs_batch* b) { return b->fp64 ? b->scheme.gamma : (double)b->gamma32; }
size_t real_size(const bs_batch* b) { return b->fp64 ? sizeof(double) : sizeof(float); }

template <typename T>
T* arena(void* base, size_t offset) {
  return static_cast<T*>(base) + offset;
}

template <typename T>
StateDevT<T> batch_state(bs_batch* b, int buf, const BatchCase& k) {
  StateDevT<T> s;
  for (int q = 0; q < 4; q++) s.q[q] = arena<T>(b->state[buf], 4 * k.cell + (size_t)q * k.n);
  return s;
}

template <typename T>
GridDevT<T> batch_grid(bs_batch* b, const BatchCase& k) {
  GridDevT<T> G;
  G.ni = k.ni;
  G.nj = k.nj;
  G.inx = arena<T>(b->inx, k.iface); G.iny = arena<T>(b->iny, k.iface); G.ia = arena<T>(b->ia, k.iface);
  G.jnx = arena<T>(b->jnx, k.jface); G.jny = arena<T>(b->jny, k.jface); G.ja = arena<T>(b->ja, k.jface);
  G.vol = arena<T>(b->vol, k.cell); G.planar = arena<T>(b->planar, k.cell);
  return G;
}

bs_status real_alloc(const bs_batch* b, void** p, size_t count, const char* what) {
  const size_t bytes = (count ? count : 1) * real_size(b);
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ERR_MEMORY, "allocating %s (%zu bytes): %s", what, bytes, hipGetErrorString(e));
  }
  return BS_OK;
}

bs_status byte_alloc(void** p, size_t bytes, const char* what) {
  const hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(BS_ERR_MEMORY, "allocating %s (%zu bytes): %s", what, bytes, hipGetErrorString(e));
  }
  return BS_OK;
}

void batch_free(bs_batch* b) {
  drop_graph(b->explicit_graph);
  drop_graph(b->implicit_graph);
  void* all[] = {b->state[0], b->state[1], b->state[2], b->inx, b->iny, b->ia, b->jnx, b->jny, b->ja, b->vol,
                 b->planar, b->dt, b->res, b->diag, b->lam_i, b->lam_j, b->dq, b->wall_geo, b->partials,
                 b->wall_partials, b->sums, b->control, b->bnd, b->wall_bnd, b->tile_args, b->relax_args,
                 b->update_args, b->wall_args, b->work_tiles[0], b->work_tiles[1], b->work_relax,
                 b->work_wall, b->seg_first, b->seg_count, b->seg_cells, b->seg_cases, b->wall_first,
                 b->wall_count, b->persist_items, b->persist_lens, b->persist_claims, b->persist_bars,
                 b->persist_fault};
  for (void* p : all)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : b->timing.pool) (void)hipEventDestroy(e);
  b->timing.pool.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

// Index of a new event pair of `kind` with its start recorded, or -1 when
// timing is off (or no event could be made).
int timing_begin(bs_batch* b, int kind, int weight = 1) {
  auto& t = b->timing;
  if (!t.enabled || kind < 0) return -1;
  if ((size_t)(2 * t.used + 2) > t.pool.size()) {
    for (int k = 0; k < 64; k++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      t.pool.push_back(e);
    }
  }
  t.kinds.resize(t.used + 1);
  t.weights.resize(t.used + 1);
  t.kinds[t.used] = kind;
  t.weights[t.used] = weight;
  (void)hipEventRecord(t.pool[2 * t.used], b->stream);
  return t.used++;
}

void timing_end(bs_batch* b, int pair) {
  if (pair >= 0) (void)hipEventRecord(b->timing.pool[2 * pair + 1], b->stream);
}

// After a synchronisation: the recorded pairs into the totals.
void timing_collect(bs_batch* b) {
  auto& t = b->timing;
  for (int k = 0; k < t.used; k++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.pool[2 * k], t.pool[2 * k + 1]) == hipSuccess) {
      t.total_ms[t.kinds[k]] += ms;
      t.launches[t.kinds[k]] += t.weights[k];
    }
  }
  t.used = 0;
}

// Argument records of every role for every case (uploaded once; the
// implicit records when first needed).
template <typename T>
bs_status batch_upload_tile_args(bs_batch* b) {
  const int n = b->count;
  std::vector<TileArgsT<T>> recs((size_t)N_TILE_ROLES * n);
  const T *ca, *cb;
  const int ns = rk_table(b->scheme.integrator, &ca, &cb);
  for (int c = 0; c < n; c++) {
    const BatchCase& k = b->cases[c];
    TileArgsT<T> base;
    memset(&base, 0, sizeof base);
    base.G = batch_grid<T>(b, k);
    base.B = static_cast<const BoundaryDevT<T>*>(b->bnd) + c;
    base.gamma = (T)state_gamma(b);
    base.axisymmetric = b->scheme.geometry == BS_GEOMETRY_AXISYMMETRIC;
    base.limiter = b->scheme.limiter;
    base.cfl = (T)b->scheme.cfl;
    base.dt = arena<T>(b->dt, k.cell);
    base.control = b->control + c;
    for (int cur = 0; cur < 3; cur++) {
      const int other[2] = {(cur + 1) % 3, (cur + 2) % 3};
      int prev = cur;
      for (int s = 0; s < 3; s++) {
        TileArgsT<T> A = base;
        if (s < ns) {
          const int dst = (s == ns - 1 && s > 0) ? cur : other[s % 2];
          A.u0 = batch_state<T>(b, cur, k);
          A.src = batch_state<T>(b, prev, k);
          A.dst = batch_state<T>(b, dst, k);
          A.a = ca[s];
          A.b = cb[s];
          A.dt_mode = s == 0 ? DT_COMPUTE_LOCAL : DT_READ;
          A.count_step = s == ns - 1;
          prev = dst;
        }
        recs[(size_t)(ROLE_STAGE0 + 3 * cur + s) * n + c] = A;
      }
      TileArgsT<T> R = base;
      R.src = batch_state<T>(b, cur, k);
      R.partials = b->partials;
      recs[(size_t)(ROLE_RESIDUAL + cur) * n + c] = R;
      TileArgsT<T> P = base;
      P.src = batch_state<T>(b, cur, k);
      if (b->implicit_ready) {
        for (int q = 0; q < 4; q++) {
          P.res[q] = b->res + 4 * k.cell + (size_t)q * k.n;
          P.dq[q] = b->dq + 4 * k.cell + (size_t)q * k.n;
        }
        P.diag = b->diag + k.cell;
        P.lam_i = b->lam_i + k.iface;
        P.lam_j = b->lam_j + k.jface;
      }
      recs[(size_t)(ROLE_PREPARE + cur) * n + c] = P;
    }
  }
  HIP_TRY(copy_on(b->stream, b->tile_args, recs.data(), recs.size() * sizeof(TileArgsT<T>), hipMemcpyHostToDevice),
          "batch records");
  return BS_OK;
}

template <typename T>
bs_status batch_upload_implicit_args(bs_batch* b, double omega, double cfl_max, double growth) {
  const int n = b->count;
  std::vector<RelaxArgsT<T>> rel((size_t)3 * n);
  std::vector<UpdateArgsT<T>> upd((size_t)3 * n);
  std::vector<TileArgsT<T>> prep((size_t)3 * n);
  TileArgsT<T>* tile_args = static_cast<TileArgsT<T>*>(b->tile_args);
  HIP_TRY(copy_on(b->stream, prep.data(), tile_args + (size_t)ROLE_PREPARE * n, prep.size() * sizeof(TileArgsT<T>),
                    hipMemcpyDeviceToHost),
          "batch records");
  for (int c = 0; c < n; c++) {
    const BatchCase& k = b->cases[c];
    for (int cur = 0; cur < 3; cur++) {
      RelaxArgsT<T> R;
      R.G = batch_grid<T>(b, k);
      R.q = batch_state<T>(b, cur, k);
      for (int q = 0; q < 4; q++) {
        R.res[q] = b->res + 4 * k.cell + (size_t)q * k.n;
        R.dq[q] = b->dq + 4 * k.cell + (size_t)q * k.n;
      }
      R.diag = b->diag + k.cell;
      R.lam_i = b->lam_i + k.iface;
      R.lam_j = b->lam_j + k.jface;
      R.gamma = b->gamma32;
      R.omega = (float)omega;
      rel[(size_t)cur * n + c] = R;
      UpdateArgsT<T> U;
      U.src = batch_state<T>(b, cur, k);
      U.dst = batch_state<T>(b, (cur + 1) % 3, k);
      for (int q = 0; q < 4; q++) U.dq.q[q] = b->dq + 4 * k.cell + (size_t)q * k.n;
      U.n = k.n;
      U.gamma = (T)state_gamma(b);
      U.control = b->control + c;
      U.cfl_max = cfl_max;
      U.growth = growth;
      upd[(size_t)cur * n + c] = U;
      prep[(size_t)cur * n + c].omega = (T)omega;
    }
  }
  HIP_TRY(copy_on(b->stream, b->relax_args, rel.data(), rel.size() * sizeof(RelaxArgsT<T>), hipMemcpyHostToDevice),
          "records");
  HIP_TRY(copy_on(b->stream, b->update_args, upd.data(), upd.size() * sizeof(UpdateArgsT<T>), hipMemcpyHostToDevice),
          "records");
  HIP_TRY(copy_on(b->stream, tile_args + (size_t)ROLE_PREPARE * n, prep.data(), prep.size() * sizeof(TileArgsT<T>),
                    hipMemcpyHostToDevice),
          "records");
  return BS_OK;
}

// Uploads a host list into one of the batch's device arrays on the batch's
// stream. The array is reallocated only when the list outgrows it: the work
// lists shrink as cases converge, so a batch driven from its own thread
// frees no memory while another batch's thread captures a graph.
template <typename T>
bs_status upload_vec(bs_batch* b, T** dst, const std::vector<T>& v, const char* what) {
  size_t& cap = b->capacity[(const void*)dst];
  if (!*dst || cap < v.size()) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    cap = 0;
    const size_t n = v.size() ? v.size() : 1;
    bs_status st = dev_alloc(dst, n, what);
    if (st) return st;
    cap = n;
  }
  if (!v.empty()) HIP_TRY(copy_on(b->stream, *dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
  return BS_OK;
}

// Work lists of the active cases: stencil tiles per class (ordered by case,
// so neighbouring blocks share a grid), relaxation tiles (whose last
// launch of a step also writes the updated state), wall
// blocks, and the segments of the per-case reductions.
bs_status batch_build_lists(bs_batch* b) {
  std::vector<WorkItem> tiles[2], relax, wall;
  std::vector<int> seg_cases;
  for (int c = 0; c < b->count; c++) {
    if (!b->active[c]) continue;
    BatchCase& k = b->cases[c];
    k.cls = k.n < b->small_below ? 1 : 0;
    seg_cases.push_back(c);
    const int tx = k.tiles_x[k.cls], ty = k.tiles_y[k.cls];
    for (int by = 0; by < ty; by++)
      for (int bx = 0; bx < tx; bx++) tiles[k.cls].push_back({c, bx, by, k.slot0 + bx + by * tx});
    const int rx = (k.ni + RTX - 1) / RTX, ry = (k.nj + RTY - 1) / RTY;
    for (int by = 0; by < ry; by++)
      for (int bx = 0; bx < rx; bx++) relax.push_back({c, bx, by, 0});
    for (int blk = 0; blk < k.wall_blocks; blk++) wall.push_back({c, blk, c, k.wall_slot0 + blk});
  }
  // Slots and segments depend on the class of each case.
  std::vector<int> first(b->count), cnt(b->count), cells(b->count), wfirst(b->count), wcnt(b->count);
  for (int c = 0; c < b->count; c++) {
    const BatchCase& k = b->cases[c];
    first[c] = k.slot0;
    cnt[c] = k.tiles_x[k.cls] * k.tiles_y[k.cls];
    cells[c] = k.n;
    wfirst[c] = k.wall_slot0;
    wcnt[c] = k.wall_blocks;
  }
  bs_status st;
  for (int cl = 0; cl < 2; cl++)
    if ((st = upload_vec(b, &b->work_tiles[cl], tiles[cl], "work list"))) return st;
  if ((st = upload_vec(b, &b->work_relax, relax, "work list")) ||
      (st = upload_vec(b, &b->work_wall, wall, "work list")) || (st = upload_vec(b, &b->seg_first, first, "segments")) ||
      (st = upload_vec(b, &b->seg_count, cnt, "segments")) || (st = upload_vec(b, &b->seg_cells, cells, "segments")) ||
      (st = upload_vec(b, &b->seg_cases, seg_cases, "segments")) || (st = upload_vec(b, &b->wall_first, wfirst, "segments")) ||
      (st = upload_vec(b, &b->wall_count, wcnt, "segments")))
    return st;
  b->n_tiles[0] = (int)tiles[0].size();
  b->n_tiles[1] = (int)tiles[1].size();
  b->n_relax = (int)relax.size();
  b->n_wall = (int)wall.size();
  b->n_seg = (int)seg_cases.size();
  b->lists_dirty = false;
  b->persist_dirty = true;
  drop_graph(b->explicit_graph);
  drop_graph(b->implicit_graph);
  return BS_OK;
}

// The persistent launch's tile lists for the active cases: every case's
// tiles on one XCD (cases dealt largest first to the XCD with the fewest
// tiles). A launch takes 8 x (longest list) + 64 blocks, which the
// dispatcher's round-robin spreads evenly over the XCDs. persist_blocks
// stays 0 when an XCD would get more blocks than it holds at once (the
// launch would then wait on blocks that cannot start).
bs_status batch_build_persistent(bs_batch* b) {
  b->persist_dirty = false;
  b->persist_blocks = 0;
  int tx, ty;
  persistent_tile_shape(b->persist_tile, &tx, &ty);
  if (b->persist_capacity == 0) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, b->device), "device properties");
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void*>(batch_persistent_fn(b->scheme, b->persist_tile)), tx * ty, 0),
            "occupancy");
    b->persist_capacity = prop.multiProcessorCount % PERSIST_XCDS == 0
                              ? prop.multiProcessorCount / PERSIST_XCDS * per_cu : 0;
  }
  std::vector<int> order;
  for (int c = 0; c < b->count; c++)
    if (b->active[c]) order.push_back(c);
  auto tiles = [&](int c) {
    const BatchCase& k = b->cases[c];
    return ((k.ni + tx - 1) / tx) * ((k.nj + ty - 1) / ty);
  };
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return tiles(p) > tiles(q); });
  std::vector<std::vector<PersistItem>> lists(PERSIST_XCDS);
  for (int c : order) {
    size_t x = 0;
    for (size_t y = 1; y < lists.size(); y++)
      if (lists[y].size() < lists[x].size()) x = y;
    const BatchCase& k = b->cases[c];
    const int nx = (k.ni + tx - 1) / tx, ny = (k.nj + ty - 1) / ty;
    for (int by = 0; by < ny; by++)
      for (int bx = 0; bx < nx; bx++) lists[x].push_back({c, bx, by, nx * ny});
  }
  size_t longest = 0;
  for (auto& l : lists) longest = l.size() > longest ? l.size() : longest;
  constexpr int SPARE = 64;  // blocks beyond 8 x longest
  if (longest == 0 || (int)longest + SPARE / PERSIST_XCDS > b->persist_capacity) return BS_OK;
  std::vector<PersistItem> items(PERSIST_XCDS * longest, PersistItem{-1, 0, 0, 0});
  std::vector<int> lens(PERSIST_XCDS);
  for (int x = 0; x < PERSIST_XCDS; x++) {
    lens[x] = (int)lists[x].size();
    for (size_t k = 0; k < lists[x].size(); k++) items[x * longest + k] = lists[x][k];
  }
  bs_status st;
  if ((st = upload_vec(b, &b->persist_items, items, "persistent tiles")) ||
      (st = upload_vec(b, &b->persist_lens, lens, "persistent tiles")))
    return st;
  b->persist_stride = (int)longest;
  b->persist_blocks = PERSIST_XCDS * (int)longest + SPARE;
  return BS_OK;
}

constexpr int PERSIST_STEPS = 60;        // steps per persistent launch
constexpr int PERSIST_SPIN = 1 << 16;    // polls before a barrier gives up

// `steps` explicit steps in persistent launches; the fault word is read
// after each launch.
bs_status batch_persistent_run(bs_batch* b, int steps) {
  const float *ca, *cb;
  const int ns = rk_table(b->scheme.integrator, &ca, &cb);
  const PersistFn fn = batch_persistent_fn(b->scheme, b->persist_tile);
  int tx, ty;
  persistent_tile_shape(b->persist_tile, &tx, &ty);
  const TileArgsT<float>* recs = static_cast<const TileArgsT<float>*>(b->tile_args) + (size_t)ROLE_STAGE0 * b->count;
  for (int done = 0; done < steps;) {
    const int n = steps - done < PERSIST_STEPS ? steps - done : PERSIST_STEPS;
    const int cur0 = ns == 1 ? (b->cur + done) % 3 : b->cur;
    hipLaunchKernelGGL(clear_counters, dim3((b->count + 255) / 256), dim3(256), 0, b->stream, b->persist_bars,
                       b->count, b->persist_claims, b->persist_fault);
    const int pair = timing_begin(b, 0, n * ns);
    hipLaunchKernelGGL(fn, dim3(b->persist_blocks), dim3(tx * ty), 0, b->stream, recs, b->count, b->persist_items,
                       b->persist_lens, b->persist_claims, b->persist_stride, n, ns, cur0, b->persist_bars,
                       b->persist_fault, PERSIST_SPIN);
    timing_end(b, pair);
    HIP_TRY(hipGetLastError(), "persistent launch");
    unsigned fault = 0;
    HIP_TRY(hipMemcpyAsync(&fault, b->persist_fault, sizeof fault, hipMemcpyDeviceToHost, b->stream),
            "persistent fault word");
    HIP_TRY(hipStreamSynchronize(b->stream), "persistent launch");
    if (fault) {
      b->persist_tile = 0;
      return fail(BS_ERR_EXECUTION,
                  "persistent launch abandoned (%s); persistent launches are off for this batch and the states "
                  "of its cases are indeterminate",
                  fault & PERSIST_MISPLACED ? "a block read an XCD number out of range" : "a barrier timed out");
    }
    done += n;
  }
  if (ns == 1) b->cur = (b->cur + steps) % 3;
  timing_collect(b);
  return BS_OK;
}

template <typename T>
bs_status batch_launch_tiles(bs_batch* b, int mode, int role) {
  const TileArgsT<T>* recs = static_cast<const TileArgsT<T>*>(b->tile_args) + (size_t)role * b->count;
  // Both tile classes' launches count as one launch of the pass.
  const int pair = timing_begin(b, mode == TILE_STAGE ? 0 : mode == TILE_IMPLICIT ? 1 : -1);
  for (int cl = 0; cl < 2; cl++) {
    if (b->n_tiles[cl] == 0) continue;
    const int threads = cl == 0 ? BatchTile<T>::lx * BatchTile<T>::ly : BatchTile<T>::sx * BatchTile<T>::sy;
    hipLaunchKernelGGL(batch_tile_fn<T>(b->scheme, cl, mode), dim3(b->n_tiles[cl]), dim3(threads), 0, b->stream, recs,
                       b->work_tiles[cl]);
  }
  timing_end(b, pair);
  HIP_TRY(hipGetLastError(), "batch stencil launch");
  return BS_OK;
}

bs_status launch_tiles(bs_batch* b, int mode, int role) {
  return b->fp64 ? batch_launch_tiles<double>(b, mode, role) : batch_launch_tiles<float>(b, mode, role);
}

template <typename T>
bs_status batch_encode_explicit(bs_batch* b) {
  const T *ca, *cb;
  const int ns = rk_table(b->scheme.integrator, &ca, &cb);
  for (int s = 0; s < ns; s++) {
    bs_status st = batch_launch_tiles<T>(b, TILE_STAGE, ROLE_STAGE0 + 3 * b->cur + s);
    if (st) return st;
  }
  if (ns == 1) b->cur = (b->cur + 1) % 3;
  return BS_OK;
}

template <typename T>
bs_status batch_encode_implicit(bs_batch* b, int sweeps) {
  const int cur = b->cur;
  bs_status st = batch_launch_tiles<T>(b, TILE_IMPLICIT, ROLE_PREPARE + cur);
  if (st) return st;
  const RelaxArgsT<T>* rel = static_cast<const RelaxArgsT<T>*>(b->relax_args) + (size_t)cur * b->count;
  // The first red half-sweep (formed by the first launch) and the
  // half-sweeps after it, up to three per launch; the last launch writes
  // the updated states (from the FP64 state in HBM on an FP64 batch).
  const UpdateArgsT<T>* upd = static_cast<const UpdateArgsT<T>*>(b->update_args) + (size_t)cur * b->count;
  // [form][half-sweeps - 1][first launch]
  using K = void (*)(const RelaxArgsT<T>*, const WorkItem*, int, const UpdateArgsT<T>*);
  K kernels[2][3][2] = {};
  if constexpr (std::is_same<T, float>::value) {  // an FP64 state relaxes linearized only
    kernels[0][0][0] = batch_relax_tile_kernel<T, 1, false, false>;
    kernels[0][1][0] = batch_relax_tile_kernel<T, 2, false, false>;
    kernels[0][2][0] = batch_relax_tile_kernel<T, 3, false, false>;
    kernels[0][0][1] = batch_relax_tile_kernel<T, 1, false, true>;
    kernels[0][1][1] = batch_relax_tile_kernel<T, 2, false, true>;
    kernels[0][2][1] = batch_relax_tile_kernel<T, 3, false, true>;
  }
  kernels[1][0][0] = batch_relax_tile_kernel<T, 1, true, false>;
  kernels[1][1][0] = batch_relax_tile_kernel<T, 2, true, false>;
  kernels[1][2][0] = batch_relax_tile_kernel<T, 3, true, false>;
  kernels[1][0][1] = batch_relax_tile_kernel<T, 1, true, true>;
  kernels[1][1][1] = batch_relax_tile_kernel<T, 2, true, true>;
  kernels[1][2][1] = batch_relax_tile_kernel<T, 3, true, true>;
  const int form = (b->relax_linear || !std::is_same<T, float>::value) ? 1 : 0;
  for (int done = 0, colour = 1, halves = 2 * sweeps - 1; done < halves && b->n_relax;) {
    const int nh = halves - done < b->relax_fusion ? halves - done : b->relax_fusion;
    const int init = done == 0;
    const UpdateArgsT<T>* u = done + nh == halves ? upd : nullptr;
    const int pair = timing_begin(b, 2);
    hipLaunchKernelGGL(kernels[form][nh - 1][init], dim3(b->n_relax), dim3(RNT), 0, b->stream, rel, b->work_relax,
                       colour, u);
    timing_end(b, pair);
    colour = (colour + nh) & 1;
    done += nh;
  }
  HIP_TRY(hipGetLastError(), "batch implicit launch");
  b->cur = (cur + 1) % 3;
  return BS_OK;
}

template <typename Encode>
bs_status batch_run(bs_batch* b, StepGraph& g, int count, Encode encode) {
  if (count >= GRAPH_STEPS && !b->timing.enabled) {
    bs_status st = replay_steps(b->stream, b->cur, g, count, encode);
    if (st) return st;
  } else {
    for (int k = 0; k < count; k++) {
      bs_status st = encode();
      if (st) return st;
    }
  }
  HIP_TRY(hipStreamSynchronize(b->stream), "batch steps");
  timing_collect(b);
  return BS_OK;
}

// Zero the status word and step count of cases [first, first + count),
// keeping their implicit CFL number (a new start, as bs_set_primitives).
bs_status batch_reset_control(bs_batch* b, int first, int count) {
  std::vector<Control> ctl(count);
  HIP_TRY(copy_on(b->stream, ctl.data(), b->control + first, count * sizeof(Control), hipMemcpyDeviceToHost), "control");
  for (auto& c : ctl) {
    const double cfl = c.cfl;
    memset(&c, 0, sizeof c);
    c.cfl = cfl > 0 ? cfl : 1.0;
    c.t_end = 3e38;
  }
  HIP_TRY(copy_on(b->stream, b->control + first, ctl.data(), count * sizeof(Control), hipMemcpyHostToDevice), "control");
  return BS_OK;
}

// Uploads `packed` (4 x n values of the case, conserved) as case c's
// current state.
template <typename T>
bs_status upload_case_state(bs_batch* b, const BatchCase& k, const std::vector<T>& packed) {
  HIP_TRY(copy_on(b->stream, arena<T>(b->state[b->cur], 4 * k.cell), packed.data(), packed.size() * sizeof(T),
                    hipMemcpyHostToDevice),
          "state upload");
  return BS_OK;
}

template <typename T>
bs_status batch_set_uniform(bs_batch* b, const double* states) {
  for (int c = 0; c < b->count; c++) {
    const BatchCase& k = b->cases[c];
    const double* w = states + 4 * (size_t)c;
    if (!physical(w)) return fail(BS_ERR_NONPHYSICAL, "non-physical initial state of case %d", c);
    T q[4];
    pack_cons(w, state_gamma(b), q);
    std::vector<T> v(4 * (size_t)k.n);
    for (int qq = 0; qq < 4; qq++) std::fill(v.begin() + (size_t)qq * k.n, v.begin() + (size_t)(qq + 1) * k.n, q[qq]);
    bs_status st = upload_case_state<T>(b, k, v);
    if (st) return st;
  }
  return batch_reset_control(b, 0, b->count);
}

template <typename T>
bs_status batch_set_primitives(bs_batch* b, int index, const double* prim) {
  const BatchCase& k = b->cases[index];
  std::vector<T> v(4 * (size_t)k.n);
  for (int c = 0; c < k.n; c++) {
    const double* w = prim + 4 * (size_t)c;
    if (!physical(w)) return fail(BS_ERR_NONPHYSICAL, "non-physical state in cell %d of case %d", c, index);
    T q[4];
    pack_cons(w, state_gamma(b), q);
    for (int qq = 0; qq < 4; qq++) v[(size_t)qq * k.n + c] = q[qq];
  }
  bs_status st = upload_case_state<T>(b, k, v);
  if (st) return st;
  return batch_reset_control(b, index, 1);
}

// Reads case `index`'s state into four arrays of type V (null ones skipped).
template <typename T, typename V>
bs_status batch_read_state(bs_batch* b, int index, V* const dst[4]) {
  const BatchCase& k = b->cases[index];
  std::vector<T> v(k.n);
  for (int q = 0; q < 4; q++) {
    if (!dst[q]) continue;
    const T* src = arena<T>(b->state[b->cur], 4 * k.cell + (size_t)q * k.n);
    if (std::is_same<T, V>::value) {
      HIP_TRY(copy_on(b->stream, dst[q], src, k.n * sizeof(T), hipMemcpyDeviceToHost), "state readback");
    } else {
      HIP_TRY(copy_on(b->stream, v.data(), src, k.n * sizeof(T), hipMemcpyDeviceToHost), "state readback");
      for (int c = 0; c < k.n; c++) dst[q][c] = (V)v[c];
    }
  }
  return BS_OK;
}

}  // namespace

// The warm start's coarse fields to the cases' cells: case c's coarse
// primitives (4 x cnj x cni, FP64, row j0 = 0 first) times the case's scale
// per variable, interpolated linearly between the cell centres of the two
// uniform grids in (i / ni, j / nj), clamped at the ends; density and
// pressure held at least `floor` (NaN kept, and flagged); stored conserved
// in the state's precision with the gamma the solver uses. The arithmetic
// is FP64 in the order of the host's resampling, without contraction, so
// the state equals the one set from the same field cell by cell on the
// host.
template <typename T>
struct CoarseCase {
  T* state;  // the case's 4 x n conserved values
  int n, ni;
  int nj;
  const double* coarse;
  double scale[4];
};

__device__ __forceinline__ void coarse_axis(int k, int nd, int ns, int& a0, int& a1, double& w) {
#pragma clang fp contract(off)
  double s = (k + 0.5) / nd * ns - 0.5;
  s = fmin(fmax(s, 0.0), ns - 1.0);
  a0 = min((int)floor(s), max(ns - 2, 0));
  w = s - a0;
  a1 = min(a0 + 1, ns - 1);
}

template <typename T>
__global__ __launch_bounds__(256) void batch_coarse_kernel(CoarseCase<T> K, int cni, int cnj, double floor_, double g,
                                                           int* bad) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K.n) return;
  const int i = c % K.ni, j = c / K.ni;
  int i0, i1, j0, j1;
  double wi, wj;
  coarse_axis(i, K.ni, cni, i0, i1, wi);
  coarse_axis(j, K.nj, cnj, j0, j1, wj);
  double w[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const double* f = K.coarse + (size_t)q * cnj * cni;
    const double sq = K.scale[q];
    const double a = (f[j0 * cni + i0] * sq) * (1.0 - wj) + (f[j1 * cni + i0] * sq) * wj;
    const double b = (f[j0 * cni + i1] * sq) * (1.0 - wj) + (f[j1 * cni + i1] * sq) * wj;
    w[q] = a * (1.0 - wi) + b * wi;
  }
  if (w[0] < floor_) w[0] = floor_;  // a NaN stays, and fails the check below
  if (w[3] < floor_) w[3] = floor_;
  if (!(w[0] > 0.0 && w[3] > 0.0 && isfinite(w[1]) && isfinite(w[2]))) atomicOr(bad, 1);
  K.state[c] = (T)w[0];
  K.state[(size_t)K.n + c] = (T)(w[0] * w[1]);
  K.state[2 * (size_t)K.n + c] = (T)(w[0] * w[2]);
  K.state[3 * (size_t)K.n + c] = (T)(w[3] / (g - 1) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

// The coarse-to-fine start (transonic-flow.ts:157-168): fine cell (i, j)
// takes the state of the coarse cell whose index range holds its centre in
// (i / ni, j / nj), i.e. (i / r, j / r) on grids that nest by r.
template <typename TF, typename TC>
struct InjectCase {
  TF* fine;
  const TC* coarse;
  int ni, nj, n, cni, cnj, cn;
};

template <typename TF, typename TC>
__global__ __launch_bounds__(256) void inject_kernel(InjectCase<TF, TC> K) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= K.n) return;
  const int i = c % K.ni, j = c / K.ni;
  const int ci = (int)(((long long)(2 * i + 1) * K.cni) / (2LL * K.ni));
  const int cj = (int)(((long long)(2 * j + 1) * K.cnj) / (2LL * K.nj));
  const int cc = ci + cj * K.cni;
#pragma unroll
  for (int q = 0; q < 4; q++) K.fine[(size_t)q * K.n + c] = (TF)K.coarse[(size_t)q * K.cn + cc];
}

// ---- the allocation-free warm start (bs_batch_set_coarse_noalloc) --------

// Its staging: a file-scope device array, one copy per device the library
// runs on, held until the process ends. Bytes [0, 256) hold the flag of a
// non-physical field, the rest the coarse fields of as many cases at a time
// as fit. Every call on a device shares it, so calls that overlap on one
// device race.
static constexpr size_t COARSE_SCRATCH_BYTES = BS_COARSE_SCRATCH_BYTES;
static constexpr size_t COARSE_FLAG_BYTES = 256;
static __device__ __align__(256) unsigned char scratch[COARSE_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on `device` (which is current), taken once
// per device: host code passes this address to kernels and copies, never
// the array's name, which on the host is only a stand-in.
bs_status scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(BS_ERR_RANGE, "device %d has no scratch slot", device);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "scratch address");
    if (p == nullptr) return fail(BS_ERR_EXECUTION, "scratch address: null");
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[device];
  return BS_OK;
}

__global__ void clear_flag_kernel(int* flag) { *flag = 0; }

// The coarse fields of every case through a staging area of `room` bytes
// (whole cases at a time, in order), the fields' flag in `bad` (cleared
// first by a kernel), then the flag read back; allocates nothing.
template <typename T>
bs_status coarse_through(bs_batch* b, const double* coarse, int cni, int cnj, const double* scale,
                         double floor_value, double* staging, size_t room, int* bad) {
  const size_t per = 4 * (size_t)cni * cnj;
  const size_t at_once = room / (per * sizeof(double));
  if (at_once == 0)
    return fail(BS_ERR_RANGE, "a coarse field of %zu bytes exceeds the %zu bytes of staging", per * sizeof(double),
                room);
  hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, b->stream, bad);
  for (int c0 = 0; c0 < b->count; c0 += (int)at_once) {
    const size_t left = (size_t)(b->count - c0);
    const int n = (int)(left < at_once ? left : at_once);
    // Stream order: these fields overwrite the staging only after the
    // previous group's kernels have read it.
    HIP_TRY(hipMemcpyAsync(staging, coarse + per * c0, per * n * sizeof(double), hipMemcpyHostToDevice, b->stream),
            "coarse fields upload");
    for (int c = c0; c < c0 + n; c++) {
      const BatchCase& k = b->cases[c];
      CoarseCase<T> K;
      K.state = arena<T>(b->state[b->cur], 4 * k.cell);
      K.n = k.n;
      K.ni = k.ni;
      K.nj = k.nj;
      K.coarse = staging + per * (c - c0);
      for (int q = 0; q < 4; q++) K.scale[q] = scale[4 * (size_t)c + q];
      hipLaunchKernelGGL(batch_coarse_kernel<T>, dim3((k.n + 255) / 256), dim3(256), 0, b->stream, K, cni, cnj,
                         floor_value, state_gamma(b), bad);
    }
    HIP_TRY(hipGetLastError(), "coarse fields launch");
  }
  int flag = 0;
  HIP_TRY(hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, b->stream), "coarse fields flag");
  HIP_TRY(hipStreamSynchronize(b->stream), "coarse fields");
  if (flag) return fail(BS_ERR_NONPHYSICAL, "a coarse field gives a non-physical state");
  return batch_reset_control(b, 0, b->count);
}

bs_status coarse_into(bs_batch* b, const double* coarse, int cni, int cnj, const double* scale, double floor_value,
                      double* staging, size_t room, int* bad) {
  return b->fp64 ? coarse_through<double>(b, coarse, cni, cnj, scale, floor_value, staging, room, bad)
                 : coarse_through<float>(b, coarse, cni, cnj, scale, floor_value, staging, room, bad);
}

// Metrics, boundary tables, controls and wall records of every case.
template <typename T>
bs_status batch_setup(bs_batch* b, const bs_config* configs) {
  const int count = b->count;
  std::vector<BoundaryDevT<T>> bnds(count);
  std::vector<WallBoundaryT<T>> wbs(count);
  std::vector<Control> ctl(count);
  std::vector<WallArgsT<T>> wargs((size_t)3 * count);
  for (int c = 0; c < count; c++) {
    const bs_config& cf = configs[c];
    const BatchCase& k = b->cases[c];
    const int ni = k.ni, nj = k.nj;
    const size_t fi = (size_t)(ni + 1) * nj, fj = (size_t)ni * (nj + 1);
    std::vector<double> inx(fi), iny(fi), ia(fi), jnx(fj), jny(fj), ja(fj), vol(k.n), planar(k.n);
    bs_status st = metrics(ni, nj, cf.x, cf.y, cf.geometry == BS_GEOMETRY_AXISYMMETRIC, inx.data(), iny.data(),
                           ia.data(), jnx.data(), jny.data(), ja.data(), vol.data(), planar.data(), nullptr, nullptr);
    if (st) return fail(st, "case %d: %s", c, bs_last_error());
    auto put = [&](void* base, size_t off, const std::vector<double>& v) -> bs_status {
      std::vector<T> f(v.begin(), v.end());
      HIP_TRY(hipMemcpyAsync(arena<T>(base, off), f.data(), f.size() * sizeof(T), hipMemcpyHostToDevice, b->stream),
              "metrics upload");
      HIP_TRY(hipStreamSynchronize(b->stream), "metrics upload");
      return BS_OK;
    };
    if ((st = put(b->inx, k.iface, inx)) || (st = put(b->iny, k.iface, iny)) || (st = put(b->ia, k.iface, ia)) ||
        (st = put(b->jnx, k.jface, jnx)) || (st = put(b->jny, k.jface, jny)) || (st = put(b->ja, k.jface, ja)) ||
        (st = put(b->vol, k.cell, vol)) || (st = put(b->planar, k.cell, planar)))
      return st;
    // The jMin side's FP64 face geometry (the body wall).
    std::vector<double> geo(5 * (size_t)ni);
    for (int t = 0; t < ni; t++) {
      const int n0 = t, n1 = t + 1;
      geo[5 * t + 0] = 0.5 * (cf.x[n0] + cf.x[n1]);
      geo[5 * t + 1] = 0.5 * (cf.y[n0] + cf.y[n1]);
      geo[5 * t + 2] = jnx[t];
      geo[5 * t + 3] = jny[t];
      geo[5 * t + 4] = ja[t];
    }
    HIP_TRY(copy_on(b->stream, b->wall_geo + 5 * k.wall, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice),
            "wall geometry");
    BoundaryDevT<T>& B = bnds[c];
    WallBoundaryT<T>& W = wbs[c];
    for (int sd = 0; sd < 4; sd++) {
      const bs_boundary& bb = cf.boundaries[sd];
      B.kind[sd] = bb.kind;
      B.state[sd][0] = (T)bb.rho; B.state[sd][1] = (T)bb.u;
      B.state[sd][2] = (T)bb.v; B.state[sd][3] = (T)bb.p;
      B.face_cons[sd] = nullptr;
      B.face_flux[sd] = nullptr;
      W.kind[sd] = bb.kind;
      W.state[sd][0] = bb.rho; W.state[sd][1] = bb.u; W.state[sd][2] = bb.v; W.state[sd][3] = bb.p;
      W.cons[sd] = nullptr;
    }
    memset(&ctl[c], 0, sizeof(Control));
    ctl[c].cfl = 1.0;
    ctl[c].t_end = 3e38;
    for (int cur = 0; cur < 3; cur++) {
      WallArgsT<T>& A = wargs[(size_t)cur * count + c];
      memset(&A, 0, sizeof A);
      A.G = batch_grid<T>(b, k);
      A.s = batch_state<T>(b, cur, k);
      A.side = JMIN;
      A.count = ni;
      A.kind = cf.boundaries[JMIN].kind;
      A.face_geo = b->wall_geo + 5 * k.wall;
      A.gamma = cf.gamma;
      A.gamma32 = state_gamma(b);
      A.rec = cf.reconstruction;
      A.limiter = cf.limiter;
      A.flux = cf.flux;
      A.partials = b->wall_partials;
    }
  }
  HIP_TRY(copy_on(b->stream, b->bnd, bnds.data(), bnds.size() * sizeof(BoundaryDevT<T>), hipMemcpyHostToDevice),
          "boundaries");
  HIP_TRY(copy_on(b->stream, b->wall_bnd, wbs.data(), wbs.size() * sizeof(WallBoundaryT<T>), hipMemcpyHostToDevice),
          "boundaries");
  HIP_TRY(copy_on(b->stream, b->control, ctl.data(), ctl.size() * sizeof(Control), hipMemcpyHostToDevice), "control");
  HIP_TRY(copy_on(b->stream, b->wall_args, wargs.data(), wargs.size() * sizeof(WallArgsT<T>), hipMemcpyHostToDevice),
          "records");
  return batch_upload_tile_args<T>(b);
}

bs_status batch_create(bs_ctx* ctx, const bs_config* configs, int32_t count, bool fp64, bs_batch** out) {
  if (!ctx || !configs || !out || count < 1) return fail(BS_ERR_RANGE, "a batch needs at least one case");
  *out = nullptr;
  const bs_config& s0 = configs[0];
  for (int c = 0; c < count; c++) {
    const bs_config& k = configs[c];
    if (k.geometry != s0.geometry || k.gamma != s0.gamma || k.flux != s0.flux || k.limiter != s0.limiter ||
        k.reconstruction != s0.reconstruction || k.integrator != s0.integrator || k.cfl != s0.cfl ||
        k.time_stepping != s0.time_stepping)
      return fail(BS_ERR_RANGE, "case %d: the cases of a batch share one scheme", c);
    if (k.time_stepping != BS_STEPPING_LOCAL)
      return fail(BS_ERR_RANGE, "batched solves use local time steps");
    // Boundary kinds may differ between cases (an H grid's inflow iMin, a
    // C grid's axis): every kernel reads them from the case's own record.
    for (int sd = 0; sd < 4; sd++) {
      if (k.boundaries[sd].kind < 0 || k.boundaries[sd].kind > 5)
        return fail(BS_ERR_RANGE, "case %d: unknown boundary kind", c);
      if (k.boundaries[sd].kind == BS_BC_FACE_STATES)
        return fail(BS_ERR_RANGE, "case %d: faceStates boundaries are not batched", c);
    }
    if ((k.boundaries[0].kind == BS_BC_PERIODIC) != (k.boundaries[1].kind == BS_BC_PERIODIC) ||
        (k.boundaries[2].kind == BS_BC_PERIODIC) != (k.boundaries[3].kind == BS_BC_PERIODIC))
      return fail(BS_ERR_RANGE, "case %d: periodic boundaries come in pairs", c);
    if (!(k.cfl > 0 && k.cfl <= 1)) return fail(BS_ERR_RANGE, "cfl must be in (0, 1]");
    if (k.ni < 2 || k.nj < 2 || !k.x || !k.y) return fail(BS_ERR_RANGE, "case %d: the grid needs 2 x 2 cells", c);
    if ((size_t)(k.ni + 1) * (size_t)(k.nj + 1) > ((size_t)1 << 28))
      return fail(BS_ERR_RANGE, "case %d: the grid exceeds 2^28 nodes", c);
    if ((k.reconstruction == BS_RECON_SHARED || k.reconstruction == BS_RECON_SHARED_SCALED) &&
        k.limiter != BS_LIMITER_VAN_ALBADA)
      return fail(BS_ERR_RANGE, "shared reconstruction requires the 'vanAlbada' limiter");
    if (k.flux != BS_FLUX_HLLC && k.flux != BS_FLUX_RUSANOV) return fail(BS_ERR_RANGE, "flux not available");
  }
  bs_batch* b = new bs_batch();
  b->device = ctx->device;
  b->scheme = s0;
  b->scheme.x = b->scheme.y = nullptr;
  b->gamma32 = (float)s0.gamma;
  b->fp64 = fp64;
  b->relax_linear = fp64;
  b->count = count;
  b->cases.resize(count);
  b->active.assign(count, 1);
  (void)hipSetDevice(b->device);
  auto bail = [&](bs_status code) {
    batch_free(b);
    delete b;
    return code;
  };
  const int lx = fp64 ? BatchTile<double>::lx : BatchTile<float>::lx;
  const int ly = fp64 ? BatchTile<double>::ly : BatchTile<float>::ly;
  const int sx = fp64 ? BatchTile<double>::sx : BatchTile<float>::sx;
  const int sy = fp64 ? BatchTile<double>::sy : BatchTile<float>::sy;
  // Offsets.
  for (int c = 0; c < count; c++) {
    BatchCase& k = b->cases[c];
    k.ni = configs[c].ni;
    k.nj = configs[c].nj;
    k.n = k.ni * k.nj;
    k.cell = b->cells;
    k.iface = b->ifaces;
    k.jface = b->jfaces;
    k.wall = b->walls;
    b->cells += k.n;
    b->ifaces += (size_t)(k.ni + 1) * k.nj;
    b->jfaces += (size_t)k.ni * (k.nj + 1);
    b->walls += k.ni;
    k.tiles_x[0] = (k.ni + lx - 1) / lx;
    k.tiles_y[0] = (k.nj + ly - 1) / ly;
    k.tiles_x[1] = (k.ni + sx - 1) / sx;
    k.tiles_y[1] = (k.nj + sy - 1) / sy;
    k.slot0 = b->slots;
    const int t0 = k.tiles_x[0] * k.tiles_y[0], t1 = k.tiles_x[1] * k.tiles_y[1];
    k.slots = t0 > t1 ? t0 : t1;
    b->slots += k.slots;
    k.wall_slot0 = b->wall_slots;
    k.wall_blocks = (k.ni + 255) / 256;
    b->wall_slots += k.wall_blocks;
    k.cls = 0;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(BS_ERR_EXECUTION, "hipStreamCreate failed"));
  bs_status st;
  for (int q = 0; q < 3; q++)
    if ((st = real_alloc(b, &b->state[q], 4 * b->cells, "batch state"))) return bail(st);
  const size_t tile_bytes = fp64 ? sizeof(TileArgsT<double>) : sizeof(TileArgsT<float>);
  const size_t relax_bytes = fp64 ? sizeof(RelaxArgsT<double>) : sizeof(RelaxArgsT<float>);
  const size_t upd_bytes = fp64 ? sizeof(UpdateArgsT<double>) : sizeof(UpdateArgsT<float>);
  const size_t wall_bytes = fp64 ? sizeof(WallArgsT<double>) : sizeof(WallArgsT<float>);
  const size_t bnd_bytes = fp64 ? sizeof(BoundaryDevT<double>) : sizeof(BoundaryDevT<float>);
  const size_t wb_bytes = fp64 ? sizeof(WallBoundaryT<double>) : sizeof(WallBoundaryT<float>);
  if ((st = real_alloc(b, &b->inx, b->ifaces, "metrics")) || (st = real_alloc(b, &b->iny, b->ifaces, "metrics")) ||
      (st = real_alloc(b, &b->ia, b->ifaces, "metrics")) || (st = real_alloc(b, &b->jnx, b->jfaces, "metrics")) ||
      (st = real_alloc(b, &b->jny, b->jfaces, "metrics")) || (st = real_alloc(b, &b->ja, b->jfaces, "metrics")) ||
      (st = real_alloc(b, &b->vol, b->cells, "metrics")) || (st = real_alloc(b, &b->planar, b->cells, "metrics")) ||
      (st = real_alloc(b, &b->dt, b->cells, "time steps")) || (st = dev_alloc(&b->wall_geo, 5 * b->walls, "wall")) ||
      (st = dev_alloc(&b->partials, (size_t)b->slots, "partials")) ||
      (st = dev_alloc(&b->wall_partials, (size_t)b->wall_slots, "partials")) ||
      (st = dev_alloc(&b->sums, (size_t)count, "sums")) || (st = dev_alloc(&b->control, (size_t)count, "control")) ||
      (st = byte_alloc(&b->bnd, bnd_bytes * count, "boundaries")) ||
      (st = byte_alloc(&b->wall_bnd, wb_bytes * count, "boundaries")) ||
      (st = byte_alloc(&b->tile_args, tile_bytes * N_TILE_ROLES * count, "records")) ||
      (st = byte_alloc(&b->relax_args, relax_bytes * 3 * count, "records")) ||
      (st = byte_alloc(&b->update_args, upd_bytes * 3 * count, "records")) ||
      (st = byte_alloc(&b->wall_args, wall_bytes * 3 * count, "records")))
    return bail(st);
  if ((st = fp64 ? batch_setup<double>(b, configs) : batch_setup<float>(b, configs))) return bail(st);
  *out = b;
  return BS_OK;
}

template <typename TF, typename TC>
bs_status batch_inject(bs_batch* b, bs_batch* from) {
  HIP_TRY(hipStreamSynchronize(from->stream), "coarse batch");
  for (int c = 0; c < b->count; c++) {
    const BatchCase& k = b->cases[c];
    const BatchCase& kc = from->cases[c];
    InjectCase<TF, TC> K;
    K.fine = arena<TF>(b->state[b->cur], 4 * k.cell);
    K.coarse = arena<TC>(from->state[from->cur], 4 * kc.cell);
    K.ni = k.ni; K.nj = k.nj; K.n = k.n;
    K.cni = kc.ni; K.cnj = kc.nj; K.cn = kc.n;
    hipLaunchKernelGGL((inject_kernel<TF, TC>), dim3((k.n + 255) / 256), dim3(256), 0, b->stream, K);
  }
  HIP_TRY(hipGetLastError(), "injection launch");
  HIP_TRY(hipStreamSynchronize(b->stream), "injection");
  return batch_reset_control(b, 0, b->count);
}

template <typename T>
bs_status batch_drags(bs_batch* b, const double* p_inf, const double* q_inf, const double* base_radius, double* cd) {
  // p_inf enters the wall kernel per case: set it in the records.
  WallArgsT<T>* recs = static_cast<WallArgsT<T>*>(b->wall_args) + (size_t)b->cur * b->count;
  std::vector<WallArgsT<T>> wargs((size_t)b->count);
  HIP_TRY(copy_on(b->stream, wargs.data(), recs, wargs.size() * sizeof(WallArgsT<T>), hipMemcpyDeviceToHost), "records");
  for (int c = 0; c < b->count; c++) wargs[c].p_inf = p_inf[c];
  HIP_TRY(copy_on(b->stream, recs, wargs.data(), wargs.size() * sizeof(WallArgsT<T>), hipMemcpyHostToDevice), "records");
  if (b->n_wall)
    hipLaunchKernelGGL(batch_wall_fn<T>(b->scheme), dim3(b->n_wall), dim3(256), 0, b->stream, recs,
                       static_cast<const WallBoundaryT<T>*>(b->wall_bnd), b->work_wall);
  if (b->n_seg)
    hipLaunchKernelGGL(segment_sums, dim3(b->n_seg), dim3(256), 0, b->stream, b->wall_partials, b->wall_first,
                       b->wall_count, b->seg_cells, b->seg_cases, 0, b->sums);
  std::vector<double> force(b->count);
  HIP_TRY(hipMemcpyAsync(force.data(), b->sums, force.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream),
          "drag readback");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  for (int c = 0; c < b->count; c++)
    cd[c] = b->active[c] ? (2.0 * force[c]) / (q_inf[c] * base_radius[c] * base_radius[c]) : NAN;
  return BS_OK;
}

}  // namespace

extern "C" {

bs_status bs_batch_create(bs_ctx* ctx, const bs_config* configs, int32_t count, bs_batch** out) {
  return batch_create(ctx, configs, count, false, out);
}

bs_status bs_batch_create_fp64(bs_ctx* ctx, const bs_config* configs, int32_t count, bs_batch** out) {
  return batch_create(ctx, configs, count, true, out);
}

bs_status bs_batch_precision(bs_batch* b, int32_t* precision) {
  if (!b || !precision) return fail(BS_ERR_RANGE, "null argument");
  *precision = b->fp64 ? BS_PRECISION_FP64 : BS_PRECISION_FP32;
  return BS_OK;
}

bs_status bs_batch_destroy(bs_batch* b) {
  if (!b) return BS_OK;
  bool idle = false;
  if (!b->busy.compare_exchange_strong(idle, true))
    return fail(BS_ERR_BUSY, "bs_batch: await the previous operation first (destroy while a call runs)");
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  delete b;
  return BS_OK;
}

bs_status bs_batch_set_uniform(bs_batch* b, const double* states) {
  BATCH_GUARD(b);
  if (!states) return fail(BS_ERR_RANGE, "no states");
  return b->fp64 ? batch_set_uniform<double>(b, states) : batch_set_uniform<float>(b, states);
}

bs_status bs_batch_set_primitives(bs_batch* b, int32_t index, const double* prim) {
  BATCH_GUARD(b);
  if (index < 0 || index >= b->count || !prim) return fail(BS_ERR_RANGE, "no such case");
  return b->fp64 ? batch_set_primitives<double>(b, index, prim) : batch_set_primitives<float>(b, index, prim);
}

bs_status bs_batch_set_coarse(bs_batch* b, const double* coarse, int32_t coarse_ni, int32_t coarse_nj,
                              const double* scale, double floor_value) {
  BATCH_GUARD(b);
  if (!coarse || !scale || coarse_ni < 1 || coarse_nj < 1) return fail(BS_ERR_RANGE, "no coarse fields");
  if (!(floor_value > 0)) return fail(BS_ERR_RANGE, "the floor must be positive");
  const size_t bytes = 4 * (size_t)coarse_ni * coarse_nj * sizeof(double) * b->count;
  double* d_coarse = nullptr;
  int* d_bad = nullptr;
  auto release = [&]() {
    if (d_coarse) (void)hipFree(d_coarse);
    if (d_bad) (void)hipFree(d_bad);
  };
  if (hipMalloc((void**)&d_coarse, bytes) != hipSuccess || hipMalloc((void**)&d_bad, sizeof(int)) != hipSuccess) {
    release();
    return fail(BS_ERR_MEMORY, "coarse fields: out of device memory");
  }
  // Every case's field in one group through the allocated staging.
  const bs_status st = coarse_into(b, coarse, coarse_ni, coarse_nj, scale, floor_value, d_coarse, bytes, d_bad);
  release();
  return st;
}

bs_status bs_batch_set_coarse_noalloc(bs_batch* b, const double* coarse, int32_t coarse_ni, int32_t coarse_nj,
                                      const double* scale, double floor_value) {
  BATCH_GUARD(b);
  if (!coarse || !scale || coarse_ni < 1 || coarse_nj < 1) return fail(BS_ERR_RANGE, "no coarse fields");
  if (!(floor_value > 0)) return fail(BS_ERR_RANGE, "the floor must be positive");
  const size_t per = 4 * (size_t)coarse_ni * coarse_nj * sizeof(double);
  if (per > COARSE_SCRATCH_BYTES - COARSE_FLAG_BYTES)
    return fail(BS_ERR_RANGE, "a %d x %d coarse field takes %zu bytes; the scratch holds at most %zu", coarse_ni,
                coarse_nj, per, COARSE_SCRATCH_BYTES - COARSE_FLAG_BYTES);
  unsigned char* base = nullptr;
  bs_status st = scratch_address(b->device, &base);
  if (st) return st;
  return coarse_into(b, coarse, coarse_ni, coarse_nj, scale, floor_value,
                     reinterpret_cast<double*>(base + COARSE_FLAG_BYTES), COARSE_SCRATCH_BYTES - COARSE_FLAG_BYTES,
                     reinterpret_cast<int*>(base));
}

bs_status bs_batch_set_from_coarser(bs_batch* b, bs_batch* coarse) {
  BATCH_GUARD(b);
  if (!coarse || coarse == b) return fail(BS_ERR_RANGE, "the start needs another batch");
  BatchGuard cguard(coarse);
  if (!cguard.ok) return fail(BS_ERR_BUSY, "bs_batch: await the coarse batch's previous operation first");
  if (coarse->count != b->count) return fail(BS_ERR_RANGE, "the batches hold %d and %d cases", b->count, coarse->count);
  if (coarse->device != b->device) return fail(BS_ERR_RANGE, "the batches are on different devices");
  for (int c = 0; c < b->count; c++)
    if (coarse->cases[c].ni > b->cases[c].ni || coarse->cases[c].nj > b->cases[c].nj)
      return fail(BS_ERR_RANGE, "case %d: a start must come from a coarser grid (%d x %d over %d x %d)", c,
                  coarse->cases[c].ni, coarse->cases[c].nj, b->cases[c].ni, b->cases[c].nj);
  if (b->fp64) return coarse->fp64 ? batch_inject<double, double>(b, coarse) : batch_inject<double, float>(b, coarse);
  return coarse->fp64 ? batch_inject<float, double>(b, coarse) : batch_inject<float, float>(b, coarse);
}

bs_status bs_batch_read_state(bs_batch* b, int32_t index, float* rho, float* mu, float* mv, float* energy) {
  BATCH_GUARD(b);
  if (index < 0 || index >= b->count) return fail(BS_ERR_RANGE, "no such case");
  float* const dst[4] = {rho, mu, mv, energy};
  return b->fp64 ? batch_read_state<double, float>(b, index, dst) : batch_read_state<float, float>(b, index, dst);
}

bs_status bs_batch_read_state_f64(bs_batch* b, int32_t index, double* rho, double* mu, double* mv, double* energy) {
  BATCH_GUARD(b);
  if (index < 0 || index >= b->count) return fail(BS_ERR_RANGE, "no such case");
  double* const dst[4] = {rho, mu, mv, energy};
  return b->fp64 ? batch_read_state<double, double>(b, index, dst) : batch_read_state<float, double>(b, index, dst);
}

bs_status bs_batch_set_active(bs_batch* b, const uint8_t* active, int32_t count) {
  BATCH_GUARD(b);
  if (!active) return fail(BS_ERR_RANGE, "no mask");
  if (count != b->count) return fail(BS_ERR_RANGE, "the mask needs %d values (one per case), not %d", b->count, count);
  bool changed = false;
  for (int c = 0; c < b->count; c++) {
    const char a = active[c] ? 1 : 0;
    changed |= a != b->active[c];
    b->active[c] = a;
  }
  if (changed) b->lists_dirty = true;
  return BS_OK;
}

bs_status bs_batch_enable_timing(bs_batch* b, int32_t enable) {
  BATCH_GUARD(b);
  b->timing.enabled = enable != 0;
  for (int k = 0; k < 3; k++) {
    b->timing.total_ms[k] = 0;
    b->timing.launches[k] = 0;
  }
  return BS_OK;
}

bs_status bs_batch_kernel_timing(bs_batch* b, int32_t kernel, double* ms_per_launch, int64_t* launches) {
  BATCH_GUARD(b);
  if (kernel < 0 || kernel > 2) return fail(BS_ERR_RANGE, "no such kernel");
  const long long L = b->timing.launches[kernel];
  if (ms_per_launch) *ms_per_launch = L > 0 ? b->timing.total_ms[kernel] / (double)L : 0.0;
  if (launches) *launches = L;
  b->timing.total_ms[kernel] = 0;
  b->timing.launches[kernel] = 0;
  return BS_OK;
}

bs_status bs_batch_set_tiling(bs_batch* b, int32_t small_tile_below_cells) {
  BATCH_GUARD(b);
  if (small_tile_below_cells != b->small_below) b->lists_dirty = true;
  b->small_below = small_tile_below_cells < 0 ? 0 : small_tile_below_cells;
  return BS_OK;
}

bs_status bs_batch_set_relax_fusion(bs_batch* b, int32_t half_sweeps) {
  BATCH_GUARD(b);
  if (half_sweeps < 1 || half_sweeps > 3) return fail(BS_ERR_RANGE, "half-sweeps per launch: 1, 2 or 3");
  if (half_sweeps != b->relax_fusion) drop_graph(b->implicit_graph);
  b->relax_fusion = half_sweeps;
  return BS_OK;
}

bs_status bs_batch_set_persistent(bs_batch* b, int32_t tile) {
  BATCH_GUARD(b);
  if (tile < 0 || tile > 2) return fail(BS_ERR_RANGE, "persistent tile: 0 (off), 1 (32 x 4) or 2 (32 x 8)");
  if (tile != 0 && b->fp64) return fail(BS_ERR_RANGE, "persistent launches run FP32 batches");
  if (tile != b->persist_tile) {
    b->persist_tile = tile;
    b->persist_dirty = true;
    b->persist_capacity = 0;
  }
  if (tile != 0 && b->persist_fault == nullptr) {
    bs_status st;
    if ((st = dev_alloc(&b->persist_bars, (size_t)b->count * PERSIST_BAR_STRIDE, "barrier counters")) ||
        (st = dev_alloc(&b->persist_claims, (size_t)PERSIST_XCDS, "claim counters")) ||
        (st = dev_alloc(&b->persist_fault, (size_t)1, "fault word")))
      return st;
  }
  return BS_OK;
}

bs_status bs_batch_persistent_info(bs_batch* b, int32_t* used, int32_t* blocks, int32_t* capacity) {
  BATCH_GUARD(b);
  bs_status st;
  if (b->persist_tile && !b->fp64) {
    if (b->lists_dirty && (st = batch_build_lists(b))) return st;
    if (b->persist_dirty && (st = batch_build_persistent(b))) return st;
  }
  if (used) *used = b->persist_used;
  if (blocks) *blocks = b->persist_tile ? b->persist_blocks : 0;
  if (capacity) *capacity = b->persist_tile ? b->persist_capacity * PERSIST_XCDS : 0;
  return BS_OK;
}

bs_status bs_batch_iterate(bs_batch* b, int32_t steps) {
  BATCH_GUARD(b);
  if (steps < 0) return fail(BS_ERR_RANGE, "the step count cannot be negative");
  bs_status st;
  if (b->lists_dirty && (st = batch_build_lists(b))) return st;
  b->persist_used = 0;
  if (b->persist_tile && !b->fp64) {
    if (b->persist_dirty && (st = batch_build_persistent(b))) return st;
    if (b->persist_blocks > 0) {
      b->persist_used = 1;
      return batch_persistent_run(b, steps);
    }
  }
  if (b->fp64) return batch_run(b, b->explicit_graph, steps, [b]() { return batch_encode_explicit<double>(b); });
  return batch_run(b, b->explicit_graph, steps, [b]() { return batch_encode_explicit<float>(b); });
}

bs_status bs_batch_set_implicit_cfl(bs_batch* b, double cfl) {
  BATCH_GUARD(b);
  if (!(cfl > 0)) return fail(BS_ERR_RANGE, "the CFL number must be positive");
  std::vector<Control> ctl(b->count);
  HIP_TRY(copy_on(b->stream, ctl.data(), b->control, ctl.size() * sizeof(Control), hipMemcpyDeviceToHost), "control");
  for (auto& c : ctl) c.cfl = cfl;
  HIP_TRY(copy_on(b->stream, b->control, ctl.data(), ctl.size() * sizeof(Control), hipMemcpyHostToDevice), "control");
  return BS_OK;
}

bs_status bs_batch_set_relaxation(bs_batch* b, int32_t form) {
  BATCH_GUARD(b);
  if (form != BS_RELAX_FLUX_DIFFERENCE && form != BS_RELAX_LINEARIZED)
    return fail(BS_ERR_RANGE, "unknown relaxation form %d", form);
  if (b->fp64 && form != BS_RELAX_LINEARIZED)
    return fail(BS_ERR_UNSUPPORTED, "an FP64 batch relaxes with the linearized neighbour term only");
  const bool linear = form == BS_RELAX_LINEARIZED;
  if (linear != b->relax_linear) drop_graph(b->implicit_graph);
  b->relax_linear = linear;
  return BS_OK;
}

bs_status bs_batch_set_implicit_cfls(bs_batch* b, const double* cfl) {
  BATCH_GUARD(b);
  if (!cfl) return fail(BS_ERR_RANGE, "no CFL numbers");
  for (int c = 0; c < b->count; c++)
    if (!(cfl[c] > 0)) return fail(BS_ERR_RANGE, "the CFL number of case %d must be positive", c);
  std::vector<Control> ctl(b->count);
  HIP_TRY(copy_on(b->stream, ctl.data(), b->control, ctl.size() * sizeof(Control), hipMemcpyDeviceToHost), "control");
  for (int c = 0; c < b->count; c++) ctl[c].cfl = cfl[c];
  HIP_TRY(copy_on(b->stream, b->control, ctl.data(), ctl.size() * sizeof(Control), hipMemcpyHostToDevice), "control");
  return BS_OK;
}

bs_status bs_batch_iterate_implicit(bs_batch* b, int32_t steps, double cfl_max, int32_t sweeps, double omega,
                                    double growth) {
  BATCH_GUARD(b);
  if (!(cfl_max > 0 && omega >= 1 && growth >= 1))
    return fail(BS_ERR_RANGE, "implicit steps need cfl > 0, omega >= 1 and growth >= 1");
  if (sweeps < 1) return fail(BS_ERR_RANGE, "the relaxation needs a whole number of sweeps, at least one");
  bs_status st;
  if (!b->implicit_ready) {
    if ((st = dev_alloc(&b->res, 4 * b->cells, "implicit residual")) ||
        (st = dev_alloc(&b->diag, b->cells, "implicit diagonal")) ||
        (st = dev_alloc(&b->lam_i, b->ifaces, "lambda")) || (st = dev_alloc(&b->lam_j, b->jfaces, "lambda")) ||
        (st = dev_alloc(&b->dq, 4 * b->cells, "implicit dq")))
      return st;
    b->implicit_ready = true;
    if ((st = b->fp64 ? batch_upload_tile_args<double>(b) : batch_upload_tile_args<float>(b))) return st;
    b->upd_cfl_max = -1;
  }
  StepGraph& g = b->implicit_graph;
  if (b->upd_cfl_max != cfl_max || b->upd_growth != growth || g.omega != omega || g.sweeps != sweeps) {
    if ((st = b->fp64 ? batch_upload_implicit_args<double>(b, omega, cfl_max, growth)
                      : batch_upload_implicit_args<float>(b, omega, cfl_max, growth)))
      return st;
    b->upd_cfl_max = cfl_max;
    b->upd_growth = growth;
    drop_graph(g);
  }
  if (b->lists_dirty && (st = batch_build_lists(b))) return st;
  if (b->fp64) st = batch_run(b, g, steps, [b, sweeps]() { return batch_encode_implicit<double>(b, sweeps); });
  else st = batch_run(b, g, steps, [b, sweeps]() { return batch_encode_implicit<float>(b, sweeps); });
  g.sweeps = sweeps;
  g.omega = omega;
  return st;
}

bs_status bs_batch_residual_norms(bs_batch* b, double* norms) {
  BATCH_GUARD(b);
  if (!norms) return fail(BS_ERR_RANGE, "no output");
  bs_status st;
  if (b->lists_dirty && (st = batch_build_lists(b))) return st;
  if ((st = launch_tiles(b, TILE_RESIDUAL, ROLE_RESIDUAL + b->cur))) return st;
  if (b->n_seg)
    hipLaunchKernelGGL(segment_sums, dim3(b->n_seg), dim3(256), 0, b->stream, b->partials, b->seg_first, b->seg_count,
                       b->seg_cells, b->seg_cases, 1, b->sums);
  std::vector<double> all(b->count);
  HIP_TRY(hipMemcpyAsync(all.data(), b->sums, all.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream),
          "norm readback");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  for (int c = 0; c < b->count; c++) norms[c] = b->active[c] ? all[c] : NAN;
  return BS_OK;
}

bs_status bs_batch_drags(bs_batch* b, const double* p_inf, const double* q_inf, const double* base_radius,
                         double* cd) {
  BATCH_GUARD(b);
  if (!p_inf || !q_inf || !base_radius || !cd) return fail(BS_ERR_RANGE, "null argument");
  bs_status st;
  if (b->lists_dirty && (st = batch_build_lists(b))) return st;
  return b->fp64 ? batch_drags<double>(b, p_inf, q_inf, base_radius, cd)
                 : batch_drags<float>(b, p_inf, q_inf, base_radius, cd);
}

bs_status bs_batch_status(bs_batch* b, uint32_t* status, int64_t* steps) {
  BATCH_GUARD(b);
  std::vector<Control> ctl(b->count);
  HIP_TRY(copy_on(b->stream, ctl.data(), b->control, ctl.size() * sizeof(Control), hipMemcpyDeviceToHost), "control");
  for (int c = 0; c < b->count; c++) {
    if (status) status[c] = ctl[c].status;
    if (steps) steps[c] = ctl[c].steps;
  }
  return BS_OK;
}

}  // extern "C"
