// A grid on straight lines from the body whose heights move with a fitted
// shock: the reference's GpuLineGrid (src/engine-gpu/line-grid.ts,
// shaders/line-grid.wgsl) for the solver's own metric arrays. Node (i, j)
// lies at line i's origin plus h_i eta_ij times its direction; the
// fractions eta are fixed (blunt-grid.ts withFixedFractions), the heights
// h change. Everything here is FP64 (the reference computes in FP32 from
// differences taken on the host to keep the precision); the metrics are
// stored in the solver's precision.
//
//   line_metrics_kernel  the metrics of structured.ts:94-168 for the
//                        current heights, and the per-side FP64 face
//                        geometry the boundary kernels read;
//   line_remap_kernel    the flow moved with the lines: each cell's new
//                        centre lies within half a cell of its old one, so
//                        its primitive state is interpolated between the
//                        old cell and its neighbour above or below
//                        (remapAlongLines, shock-fitting.ts:305-338, for
//                        moves under half the end cells' height;
//                        line-grid.wgsl line_grid_remap);
//   shock_moves_kernel   one block: each node moved by the speed of the
//                        shock face on its axis side over half its top
//                        cell's height per signal speed, the first node
//                        normal to the axis or held (fitShock,
//                        shock-fitting.ts:252-281; line-grid.wgsl
//                        line_grid_shock_moves).

namespace bs {

struct LineGridDev {
  int ni, nj;
  int axisymmetric;
  int symmetry;         // first node normal to the axis (else held)
  const double* lines;  // per line: x, y, dx, dy
  const double* eta;    // node fractions, i + j (ni + 1)
  double* h;            // heights, ni + 1
  double* move;         // the last move of each height, ni + 1
};

__device__ __forceinline__ void line_node(const LineGridDev& L, int i, int j, double& x, double& y) {
  const double* l = L.lines + 4 * i;
  const double s = L.h[i] * L.eta[i + j * (L.ni + 1)];
  x = l[0] + s * l[2];
  y = l[1] + s * l[3];
}

// Metric arrays of the solver, written.
template <typename T>
struct MetricOut {
  T *inx, *iny, *ia, *jnx, *jny, *ja, *vol, *planar;
  double* geo[4];  // per side: x, y, nx, ny, area per face
};

// One thread per i-face, j-face or cell (in that order of the index).
template <typename T>
__global__ __launch_bounds__(256) void line_metrics_kernel(LineGridDev L, MetricOut<T> M) {
  const int ni = L.ni, nj = L.nj;
  const long long fi = (long long)(ni + 1) * nj, fj = (long long)ni * (nj + 1), cells = (long long)ni * nj;
  long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  auto face = [&](int ia_i, int ia_j, int ib_i, int ib_j, double& nx, double& ny, double& area, double& mx,
                  double& my) {
    double xa, ya, xb, yb;
    line_node(L, ia_i, ia_j, xa, ya);
    line_node(L, ib_i, ib_j, xb, yb);
    const double dx = xb - xa, dy = yb - ya;
    const double len = sqrt(dx * dx + dy * dy);
    nx = len > 0 ? dy / len : 0.0;
    ny = len > 0 ? -dx / len : 0.0;
    area = L.axisymmetric ? len * 0.5 * (ya + yb) : len;
    mx = 0.5 * (xa + xb);
    my = 0.5 * (ya + yb);
  };
  if (t < fi) {
    const int i = (int)(t % (ni + 1)), j = (int)(t / (ni + 1));
    double nx, ny, area, mx, my;
    face(i, j, i, j + 1, nx, ny, area, mx, my);
    M.inx[t] = (T)nx;
    M.iny[t] = (T)ny;
    M.ia[t] = (T)area;
    if (i == 0 || i == ni) {
      double* g = M.geo[i == 0 ? IMIN : IMAX] + 5 * j;
      g[0] = mx; g[1] = my; g[2] = nx; g[3] = ny; g[4] = area;
    }
    return;
  }
  t -= fi;
  if (t < fj) {
    const int i = (int)(t % ni), j = (int)(t / ni);
    double nx, ny, area, mx, my;
    face(i + 1, j, i, j, nx, ny, area, mx, my);
    M.jnx[t] = (T)nx;
    M.jny[t] = (T)ny;
    M.ja[t] = (T)area;
    if (j == 0 || j == nj) {
      double* g = M.geo[j == 0 ? JMIN : JMAX] + 5 * i;
      g[0] = mx; g[1] = my; g[2] = nx; g[3] = ny; g[4] = area;
    }
    return;
  }
  t -= fj;
  if (t < cells) {
    const int i = (int)(t % ni), j = (int)(t / ni);
    double x[4], y[4];
    line_node(L, i, j, x[0], y[0]);
    line_node(L, i + 1, j, x[1], y[1]);
    line_node(L, i + 1, j + 1, x[2], y[2]);
    line_node(L, i, j + 1, x[3], y[3]);
    double area = 0, cy = 0;  // the centroid's radius is all the volume needs
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int d = (c + 1) & 3;
      const double cross = x[c] * y[d] - x[d] * y[c];
      area += cross;
      cy += (y[c] + y[d]) * cross;
    }
    area *= 0.5;
    const double gy = cy / (6 * area);
    M.planar[t] = (T)area;
    M.vol[t] = (T)(L.axisymmetric ? area * gy : area);
  }
}

// Distance along the lines between the old centres of cells (i, j) and
// (i, j + 1): the mean over the column's two lines of the old height times
// half the fraction steps.
__device__ __forceinline__ double old_centre_spacing(const LineGridDev& L, int i, int j) {
  double sum = 0.0;
  for (int k = i; k <= i + 1; k++) {
    const double h_old = L.h[k] - L.move[k];
    const double* e = L.eta + k;
    const int w = L.ni + 1;
    const double lower = 0.5 * (e[(j + 1) * w] - e[j * w]), upper = 0.5 * (e[(j + 2) * w] - e[(j + 1) * w]);
    sum += h_old * (lower + upper);
  }
  return 0.5 * sum;
}

template <typename T>
__global__ __launch_bounds__(256) void line_remap_kernel(LineGridDev L, StateDevT<T> before, StateDevT<T> after,
                                                         double g) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int ni = L.ni, nj = L.nj;
  if (c >= ni * nj) return;
  const int i = c % ni, j = c / ni;
  const int w = ni + 1;
  // The centre's move along the column: the mean of its two lines' moves
  // at the cell's mid fraction.
  double shift = 0.0;
  for (int k = i; k <= i + 1; k++) shift += L.move[k] * 0.5 * (L.eta[k + j * w] + L.eta[k + (j + 1) * w]);
  shift *= 0.5;
  Prim<double> a = prim64<double>(before, c, g);
  Prim<double> out = a;
  if (shift > 0.0 && j + 1 < nj) {
    const double s = old_centre_spacing(L, i, j);
    const double t = fmin(shift / s, 1.0);
    const Prim<double> b = prim64<double>(before, c + ni, g);
    out = {a.rho + t * (b.rho - a.rho), a.u + t * (b.u - a.u), a.v + t * (b.v - a.v), a.p + t * (b.p - a.p)};
  } else if (shift < 0.0 && j > 0) {
    const double s = old_centre_spacing(L, i, j - 1);
    const double t = fmin(-shift / s, 1.0);
    const Prim<double> b = prim64<double>(before, c - ni, g);
    out = {a.rho + t * (b.rho - a.rho), a.u + t * (b.u - a.u), a.v + t * (b.v - a.v), a.p + t * (b.p - a.p)};
  }
  after.q[0][c] = (T)out.rho;
  after.q[1][c] = (T)(out.rho * out.u);
  after.q[2][c] = (T)(out.rho * out.v);
  after.q[3][c] = (T)(out.p / (g - 1.0) + 0.5 * out.rho * (out.u * out.u + out.v * out.v));
}

// The move of node k >= 1 from the speed of the shock face on its axis
// side (k - 1) and the signal of the cell below node k's top cell.
__device__ __forceinline__ double node_move(const LineGridDev& L, const double* shock, int k) {
  const int ni = L.ni, nj = L.nj;
  const int f = k - 1;
  double x0, y0, x1, y1;
  line_node(L, f + 1, nj, x1, y1);
  line_node(L, f, nj, x0, y0);
  // The j-face from node (f + 1, nj) to node (f, nj), normal (dy, -dx) / len.
  const double dx = x0 - x1, dy = y0 - y1;
  const double len = sqrt(dx * dx + dy * dy);
  const double nx = dy / len, ny = -dx / len;
  const double* l = L.lines + 4 * k;
  const double* e = L.eta + k;
  const double size = L.h[k] * (e[nj * (ni + 1)] - e[(nj - 1) * (ni + 1)]) * sqrt(l[2] * l[2] + l[3] * l[3]);
  const int below = k < ni - 1 ? k : ni - 1;
  const double dt = 0.5 * size / shock[3 * below + 2];
  return shock[3 * f] * dt / (l[2] * nx + l[3] * ny);
}

// One block: the moves of every node (the first from nodes 1 and 2 when it
// is normal to the axis: the parabola x = x0 + c y^2 through them), then
// the heights moved.
__global__ __launch_bounds__(256) void shock_moves_kernel(LineGridDev L, const double* shock) {
  const int ni = L.ni, nj = L.nj;
  for (int k = 1 + threadIdx.x; k <= ni; k += 256) L.move[k] = node_move(L, shock, k);
  __syncthreads();
  if (threadIdx.x == 0) {
    double m0 = 0.0;
    if (L.symmetry) {
      auto moved = [&](int k, double& x, double& y) {
        const double* l = L.lines + 4 * k;
        const double s = (L.h[k] + L.move[k]) * L.eta[k + nj * (ni + 1)];
        x = l[0] + s * l[2];
        y = l[1] + s * l[3];
      };
      double x1, y1, x2, y2;
      moved(1, x1, y1);
      moved(2, x2, y2);
      const double c = (x2 - x1) / (y2 * y2 - y1 * y1);
      const double* l0 = L.lines;
      const double next = (x1 - c * y1 * y1 - l0[0]) / (l0[2] * L.eta[nj * (ni + 1)]);
      m0 = next - L.h[0];
    }
    L.move[0] = m0;
  }
  __syncthreads();
  for (int k = threadIdx.x; k <= ni; k += 256) L.h[k] += L.move[k];
}

}  // namespace bs
