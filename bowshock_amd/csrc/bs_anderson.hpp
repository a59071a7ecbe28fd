// Anderson acceleration on the device: the reference's AndersonMixing
// (src/solver/anderson.ts:20-123; D. G. Anderson, J. ACM 12 (1965) 547-560,
// in the type II form of H. F. Walker and P. Ni, SIAM J. Numer. Anal. 49
// (2011) 1715-1735): with the differences of the last m iterates, dX, and
// of their corrections, dF,
//
//   gamma = argmin |f - dF gamma|,  x' = x + f - (dX + dF) gamma,
//
// the oldest differences dropped while the columns are close to dependent.
//
// The history lives in HBM in FP64 (m columns of dF and of dX + dF, the
// only combination the update reads, and the last x and f). Each step is
// four kernels with no host round trip, so a solver's accelerated steps
// are captured in HIP graphs like its plain ones:
//   anderson_push    forms the new columns and, in the same pass, their
//                    dot products with the kept columns, with f and f . f
//                    (per-block partials);
//   anderson_solve   one block: sums the partials in a fixed order, keeps
//                    the Gram matrix G = dF^T dF by slot, and solves the
//                    least-squares problem by the Cholesky factor of G
//                    (R^T R = G, the R of the reference's modified
//                    Gram-Schmidt), dropping the oldest columns while a
//                    pivot is not above the dependence threshold;
//   anderson_update  x + f - (dX + dF) gamma;
//   anderson_fallback (solver steps) where the mixed iterate left a cell
//                    non-physical, the plain step instead and a reset
//                    (transonic-flow.ts:191-205).
// The Gram matrix resolves a pivot only down to about 3e-8 of its column's
// norm (FP64 rounding of G_cc), where modified Gram-Schmidt resolves
// 1e-10; columns are therefore dropped below 1e-7 of their norm instead of
// the reference's 1e-10. The choice changes the path of the iteration,
// never its fixed point (f = 0).

namespace bs {

constexpr int AA_MAX_DEPTH = 16;
constexpr int AA_VALUES = 2 * AA_MAX_DEPTH + 1;  // partials per block: G column, dF^T f, f . f
constexpr double AA_DEPENDENCE = 1e-7;

// The history's bookkeeping, in device memory.
struct AndersonCtl {
  int count;      // columns held
  int oldest;     // slot of the oldest column
  int has_last;   // last_x / last_f hold the previous iterate
  int restarts;   // resets by the solver loop's rules
  unsigned failed;  // the mixed iterate of this step left a cell non-physical
  int used;       // columns the last solve used (count - dropped)
  double best;    // smallest |f|^2 since the last reset (solver loop)
  double fnorm2;  // |f|^2 of this step
  double gamma[AA_MAX_DEPTH];                 // by slot
  double gram[AA_MAX_DEPTH * AA_MAX_DEPTH];   // G by slot pair
  double rhs[AA_MAX_DEPTH];                   // dF^T f by slot
};

struct AndersonArgs {
  double* dg;       // depth x size: dX + dF
  double* df;       // depth x size
  double* last_x;   // size
  double* last_f;   // size
  double* partials; // blocks x AA_VALUES
  AndersonCtl* ctl;
  long long size;
  int depth;
};

// The slot the new column goes to (the oldest one when the history is
// full), or -1 without a previous iterate.
__device__ __forceinline__ int aa_new_slot(int count, int oldest, int has_last, int depth) {
  if (!has_last) return -1;
  return count < depth ? (oldest + count) % depth : oldest;
}

// The slots of `count` columns from `oldest`, as a bit mask.
__device__ __forceinline__ unsigned aa_mask(int count, int oldest, int depth) {
  unsigned m = 0;
  for (int k = 0; k < count; k++) m |= 1u << ((oldest + k) % depth);
  return m;
}

// Slots kept once the new column is in.
__device__ __forceinline__ unsigned aa_kept_after_push(int count, int oldest, int has_last, int depth) {
  if (has_last) {
    if (count < depth) count++;
    else oldest = (oldest + 1) % depth;
  }
  return aa_mask(count, oldest, depth);
}

// gamma by slot into registers (zero beyond the depth).
__device__ __forceinline__ void aa_load_gamma(const AndersonCtl* c, int depth, double gm[AA_MAX_DEPTH]) {
#pragma unroll
  for (int s = 0; s < AA_MAX_DEPTH; s++) gm[s] = s < depth ? c->gamma[s] : 0.0;
}

// MAXD: a power of two at least the depth, so the accumulators stay in
// registers without the full AA_MAX_DEPTH of them. Values of slots beyond
// MAXD are written as zeros.
template <typename TX, typename TF, int MAXD>
__global__ __launch_bounds__(256) void anderson_push(const TX* __restrict__ x, const TF* __restrict__ f,
                                                      AndersonArgs A) {
  __shared__ double red[AA_VALUES][4];
  const int count = A.ctl->count, oldest = A.ctl->oldest, has_last = A.ctl->has_last;
  const int depth = A.depth;
  const int ns = aa_new_slot(count, oldest, has_last, depth);
  const unsigned kept = aa_kept_after_push(count, oldest, has_last, depth);
  double g[MAXD], b[MAXD], ff = 0.0;
#pragma unroll
  for (int s = 0; s < MAXD; s++) g[s] = b[s] = 0.0;
  const long long n = A.size;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
    const double xv = (double)x[k], fv = (double)f[k];
    double d[MAXD];
#pragma unroll
    for (int s = 0; s < MAXD; s++) d[s] = ((kept >> s) & 1u) && s != ns ? A.df[(long long)s * n + k] : 0.0;
    double dfn = 0.0;
    if (ns >= 0) {
      dfn = fv - A.last_f[k];
      A.dg[(long long)ns * n + k] = (xv - A.last_x[k]) + dfn;
      A.df[(long long)ns * n + k] = dfn;
    }
    A.last_x[k] = xv;
    A.last_f[k] = fv;
    ff += fv * fv;
#pragma unroll
    for (int s = 0; s < MAXD; s++) {
      const double ds = s == ns ? dfn : d[s];
      b[s] += ds * fv;
      g[s] += ds * dfn;
    }
  }
  // Block sums, one wave reduction each.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto wave_sum = [&](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  };
#pragma unroll
  for (int s = 0; s < MAXD; s++) {
    const double gs = wave_sum(g[s]), bs_ = wave_sum(b[s]);
    if (lane == 0) {
      red[s][wave] = gs;
      red[AA_MAX_DEPTH + s][wave] = bs_;
    }
  }
  const double fs = wave_sum(ff);
  if (lane == 0) red[2 * AA_MAX_DEPTH][wave] = fs;
  __syncthreads();
  if (threadIdx.x < AA_VALUES) {
    const int v = threadIdx.x;
    const bool live = v == 2 * AA_MAX_DEPTH || (v % AA_MAX_DEPTH) < MAXD;
    A.partials[(long long)v * gridDim.x + blockIdx.x] =
        live ? ((red[v][0] + red[v][1]) + red[v][2]) + red[v][3] : 0.0;
  }
}

// One block of 1024 threads: the partials (value-major, `blocks` per
// value) summed across the block in a fixed order; the Gram matrix and
// the right-hand side staged in LDS; the bookkeeping and the solve on one
// thread; the new Gram row written back by the block. `restart`: apply the
// solver loop's reset rule (|f|^2 above 100 times the best since the last
// reset, transonic-flow.ts:187-193) before the least-squares solve.
constexpr int AA_SOLVE_THREADS = 1024;
__global__ __launch_bounds__(AA_SOLVE_THREADS) void anderson_solve(AndersonArgs A, int blocks, int restart) {
  constexpr int WAVES = AA_SOLVE_THREADS / 64;
  __shared__ double part[AA_VALUES][WAVES];
  __shared__ double sum[AA_VALUES];
  __shared__ double gram[AA_MAX_DEPTH * AA_MAX_DEPTH];
  __shared__ double r[AA_MAX_DEPTH][AA_MAX_DEPTH];
  __shared__ double y[AA_MAX_DEPTH];
  __shared__ double gam[AA_MAX_DEPTH];
  __shared__ int slot[AA_MAX_DEPTH];
  __shared__ int new_slot;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  AndersonCtl& c = *A.ctl;
  // The bookkeeping scalars, loaded first so their latency overlaps the sums.
  int count = 0, oldest = 0, has_last = 0, restarts = 0;
  double best = 0.0;
  if (t == 0) {
    count = c.count;
    oldest = c.oldest;
    has_last = c.has_last;
    restarts = c.restarts;
    best = c.best;
  }
  if (t < AA_MAX_DEPTH * AA_MAX_DEPTH) gram[t] = c.gram[t];
  // Every partial in flight at once: blocks <= 2 AA_SOLVE_THREADS, two per
  // value and thread.
  double acc[AA_VALUES];
  const int t2 = t + AA_SOLVE_THREADS;
#pragma unroll
  for (int v = 0; v < AA_VALUES; v++) {
    const double* p = A.partials + (long long)v * blocks;
    acc[v] = (t < blocks ? p[t] : 0.0) + (t2 < blocks ? p[t2] : 0.0);
  }
#pragma unroll
  for (int v = 0; v < AA_VALUES; v++) {
    double a = acc[v];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) part[v][wave] = a;
  }
  __syncthreads();
  if (t < AA_VALUES) {
    double v = 0.0;
    for (int w = 0; w < WAVES; w++) v += part[t][w];
    sum[t] = v;
  }
  __syncthreads();
  if (t == 0) {
    const int depth = A.depth;
    const int ns = aa_new_slot(count, oldest, has_last, depth);
    new_slot = ns;
    if (has_last) {
      if (count < depth) count++;
      else oldest = oldest + 1 == depth ? 0 : oldest + 1;
      const unsigned kept = aa_mask(count, oldest, depth);
      for (int s = 0; s < depth; s++)
        if ((kept >> s) & 1u) gram[ns * AA_MAX_DEPTH + s] = gram[s * AA_MAX_DEPTH + ns] = sum[s];
    }
    const double ff = sum[2 * AA_MAX_DEPTH];
    if (restart) {
      if (ff > 100.0 * best) {
        count = 0;
        restarts++;
        best = ff;
      }
      best = ff < best ? ff : best;
    }
    c.count = count;
    c.oldest = oldest;
    c.has_last = 1;
    c.restarts = restarts;
    c.best = best;
    c.fnorm2 = ff;
    c.failed = 0u;
    for (int s = 0; s < AA_MAX_DEPTH; s++) gam[s] = 0.0;
    int used = 0;
    // Columns oldest first; drop from the oldest until the Cholesky factor
    // of the rest has every pivot above the dependence threshold.
    const int m = count;
    for (int k = 0, sl = oldest; k < m; k++, sl = sl + 1 == depth ? 0 : sl + 1) slot[k] = sl;
    for (int first = 0; first < m; first++) {
      const int cols = m - first;
      bool ok = true;
      for (int j = 0; j < cols && ok; j++) {
        const int sj = slot[first + j];
        for (int p = 0; p < j; p++) {
          double v = gram[slot[first + p] * AA_MAX_DEPTH + sj];
          for (int k = 0; k < p; k++) v -= r[k][p] * r[k][j];
          r[p][j] = v / r[p][p];
        }
        const double g0 = gram[sj * AA_MAX_DEPTH + sj];
        double d = g0;
        for (int p = 0; p < j; p++) d -= r[p][j] * r[p][j];
        const double norm0 = sqrt(g0 > 0.0 ? g0 : 0.0);
        if (!(d > 0.0) || norm0 == 0.0 || !(sqrt(d) > AA_DEPENDENCE * norm0)) {
          ok = false;
          break;
        }
        r[j][j] = sqrt(d);
      }
      if (!ok) continue;
      // R^T y = dF^T f, then R gamma = y.
      for (int j = 0; j < cols; j++) {
        double v = sum[AA_MAX_DEPTH + slot[first + j]];
        for (int p = 0; p < j; p++) v -= r[p][j] * y[p];
        y[j] = v / r[j][j];
      }
      for (int j = cols - 1; j >= 0; j--) {
        double v = y[j];
        for (int p = j + 1; p < cols; p++) v -= r[j][p] * gam[slot[first + p]];
        gam[slot[first + j]] = v / r[j][j];
      }
      used = cols;
      break;
    }
    c.used = used;
  }
  __syncthreads();
  if (t < AA_MAX_DEPTH) {
    c.gamma[t] = gam[t];
    c.rhs[t] = sum[AA_MAX_DEPTH + t];
  }
  const int ns = new_slot;
  if (ns >= 0 && t < AA_MAX_DEPTH) {
    c.gram[ns * AA_MAX_DEPTH + t] = gram[ns * AA_MAX_DEPTH + t];
    c.gram[t * AA_MAX_DEPTH + ns] = gram[t * AA_MAX_DEPTH + ns];
  }
}

// out = x + f - sum over slots of gamma (dX + dF), element by element.
template <typename TX, typename TF, typename TO>
__global__ __launch_bounds__(256) void anderson_update(const TX* __restrict__ x, const TF* __restrict__ f,
                                                        TO* __restrict__ out, AndersonArgs A) {
  double gm[AA_MAX_DEPTH];
  aa_load_gamma(A.ctl, A.depth, gm);
  const long long n = A.size;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
    double v = (double)x[k] + (double)f[k];
#pragma unroll
    for (int s = 0; s < AA_MAX_DEPTH; s++)
      if (gm[s] != 0.0) v -= gm[s] * A.dg[(long long)s * n + k];
    out[k] = (TO)v;
  }
}

// A solver's mixed iterate, cell by cell (the state is four arrays of
// `cells`): written to `out`, with ctl->failed set if a cell is not
// physical (setConserved's check, transonic-flow.ts:197-205).
template <typename T>
__global__ __launch_bounds__(256) void anderson_update_cells(const T* __restrict__ x, const float* __restrict__ f,
                                                             T* __restrict__ out, int cells, T gamma,
                                                             AndersonArgs A) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  double gm[AA_MAX_DEPTH];
  aa_load_gamma(A.ctl, A.depth, gm);
  const long long n = A.size;
  bool bad = false;
  if (c < cells) {
    T o[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const long long k = (long long)q * cells + c;
      double v = (double)x[k] + (double)f[k];
#pragma unroll
      for (int s = 0; s < AA_MAX_DEPTH; s++)
        if (gm[s] != 0.0) v -= gm[s] * A.dg[(long long)s * n + k];
      o[q] = (T)v;
      out[k] = o[q];
    }
    bad = !is_physical(Vec4<T>{o[0], o[1], o[2], o[3]}, gamma);
  }
  const unsigned long long m = __ballot(bad);
  if (bad && (int)(__lane_id()) == __ffsll((long long)m) - 1) atomicOr(&A.ctl->failed, 1u);
}

// After anderson_update_cells: where the mixed iterate failed, the plain
// implicit update of x instead (q + alpha dQ with alpha halving, as the
// reference's fallback step recomputes the same correction at x) and a
// reset of the history; then the step's CFL ramp and count.
template <typename T>
__global__ __launch_bounds__(256) void anderson_fallback(const T* __restrict__ x, const float* __restrict__ f,
                                                         T* __restrict__ out, int cells, UpdateArgsT<T> U,
                                                         AndersonArgs A) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const bool failed = A.ctl->failed != 0u;
  bool bad = false;
  if (failed && c < cells) {
    const Vec4<T> o = update_cell<T>(x[c], x[cells + c], x[2 * (long long)cells + c], x[3 * (long long)cells + c],
                                     (T)f[c], (T)f[cells + c], (T)f[2 * (long long)cells + c],
                                     (T)f[3 * (long long)cells + c], U.gamma);
    out[c] = o.x;
    out[cells + c] = o.y;
    out[2 * (long long)cells + c] = o.z;
    out[3 * (long long)cells + c] = o.w;
    bad = !is_physical(o, U.gamma);
  }
  flag_non_physical(bad, U.control);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (failed) {
      A.ctl->count = 0;
      A.ctl->has_last = 0;
      A.ctl->restarts++;
    }
    advance_control(U);
  }
}

// Forget the history (AndersonMixing.reset); `full` also clears the
// solver loop's best norm and restart count.
__global__ void anderson_reset_kernel(AndersonCtl* c, int full) {
  c->count = 0;
  c->oldest = 0;
  c->has_last = 0;
  c->failed = 0u;
  c->used = 0;
  if (full) {
    c->restarts = 0;
    c->best = 1.0e308;
  }
}

}  // namespace bs
