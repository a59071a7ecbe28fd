// HIP kernels of the Euler path for gfx950 (64-wide wavefronts).
//
// State: three FP32 structure-of-arrays buffers (rho, rho u, rho v, E), each
// component a contiguous array of ni * nj floats indexed i + j * ni, so a
// wavefront's loads along i are coalesced. Metrics: FP32 unit normals and
// areas per face set and volume / planar area per cell, computed in FP64 on
// the host and uploaded once.
//
// The explicit stage kernel works on a TX x TY tile of cells: it stages the
// primitives of the tile and a two-cell halo (ghost cells evaluated on the
// fly from the boundary conditions) in LDS, computes every face flux of the
// tile exactly once (each thread its cell's low i- and j-faces, one wave
// the tile's high edges), keeps flux times area in LDS, and then updates
// each cell: residual, axisymmetric source, local time step (stage 1),
// the SSP-RK combination, the physical check and, when asked, the FP64
// partial sum of the squared density residual.
#pragma once

#include <type_traits>

#include "bs_numerics.hpp"

namespace bs {

constexpr unsigned STATUS_NON_PHYSICAL = 1u;

// Device structures are templates on the state's floating-point type:
// FP32 (the default) or FP64 (solvers created with BS_PRECISION_FP64).
template <typename T>
struct GridDevT {
  int ni, nj;
  const T *inx, *iny, *ia;  // i-faces, (ni + 1) * nj
  const T *jnx, *jny, *ja;  // j-faces, ni * (nj + 1)
  const T *vol, *planar;    // cells
};
using GridDev = GridDevT<float>;

template <typename T>
struct BoundaryDevT {
  int kind[4];
  T state[4][4];         // primitive (rho, u, v, p) outside inflow / far-field sides
  const T *face_cons[4]; // FACE_STATES sides: conserved (4 values) per face
  const T *face_flux[4]; // imposed flux per unit area (4 values) per face, or null
};
using BoundaryDev = BoundaryDevT<float>;

struct Control {
  unsigned status;
  unsigned pad;
  long long steps;
  double time;
  double dt;
  double cfl;
  double t_end;
};

template <typename T>
struct StateDevT {
  T *q[4];
};
using StateDev = StateDevT<float>;

// Element idx of a state array: a plain load, or (COH) a device-scope
// relaxed atomic load, which reads the L2 and never a line the CU's L1
// still holds from before another block's store.
template <bool COH, typename T>
__device__ __forceinline__ T ld(const T* base, int idx) {
  if constexpr (COH) return __hip_atomic_load(const_cast<T*>(base) + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return base[idx];
}

// The conserved state of cell (i, j), or of a ghost cell up to two layers
// outside one side (never a corner): euler2d.wgsl load_cell, euler2d.ts
// fillGhosts 1361-1403. Ghost layer g mirrors interior layer g (outflow,
// wall with the momentum reflected about the boundary face normal), wraps
// (periodic), or holds the inflow / face / far-field state.
//
// COH: the state is read with device-scope loads, which bypass the CU's L1
// cache (the persistent batch kernel, whose blocks read states that other
// blocks wrote during the same launch).
template <typename T, bool COH = false>
__device__ __forceinline__ Vec4<T> load_cell(const StateDevT<T>& s, const GridDevT<T>& G, const BoundaryDevT<T>& B,
                                             int i, int j, T g) {
  const int ni = G.ni, nj = G.nj;
  if (i >= 0 && i < ni && j >= 0 && j < nj) {
    int c = i + j * ni;
    return {ld<COH>(s.q[0], c), ld<COH>(s.q[1], c), ld<COH>(s.q[2], c), ld<COH>(s.q[3], c)};
  }
  int side, t, mi, mj, wi, wj;
  T nx, ny;
  if (i < 0) {
    side = IMIN; t = j; mi = -1 - i; mj = j; wi = i + ni; wj = j;
    int f = j * (ni + 1);
    nx = G.inx[f]; ny = G.iny[f];
  } else if (i >= ni) {
    side = IMAX; t = j; mi = 2 * ni - 1 - i; mj = j; wi = i - ni; wj = j;
    int f = ni + j * (ni + 1);
    nx = G.inx[f]; ny = G.iny[f];
  } else if (j < 0) {
    side = JMIN; t = i; mi = i; mj = -1 - j; wi = i; wj = j + nj;
    nx = G.jnx[i]; ny = G.jny[i];
  } else {
    side = JMAX; t = i; mi = i; mj = 2 * nj - 1 - j; wi = i; wj = j - nj;
    int f = i + nj * ni;
    nx = G.jnx[f]; ny = G.jny[f];
  }
  const int kind = B.kind[side];
  if (kind == FACE_STATES) {
    const T* b = B.face_cons[side] + 4 * t;
    return {b[0], b[1], b[2], b[3]};
  }
  if (kind == INFLOW) {
    Prim<T> w = {B.state[side][0], B.state[side][1], B.state[side][2], B.state[side][3]};
    return to_cons(w, g);
  }
  if (kind == FAR_FIELD) {
    int ii = side == IMIN ? 0 : side == IMAX ? ni - 1 : i;
    int jj = side == JMIN ? 0 : side == JMAX ? nj - 1 : j;
    int c = ii + jj * ni;
    Prim<T> in = to_prim(ld<COH>(s.q[0], c), ld<COH>(s.q[1], c), ld<COH>(s.q[2], c), ld<COH>(s.q[3], c), g);
    Prim<T> fr = {B.state[side][0], B.state[side][1], B.state[side][2], B.state[side][3]};
    T o = (side == IMIN || side == JMIN) ? -T(1.) : T(1.);
    return to_cons(far_field_state(in, fr, o * nx, o * ny, g), g);
  }
  int c = kind == PERIODIC ? wi + wj * ni : mi + mj * ni;
  Vec4<T> v = {ld<COH>(s.q[0], c), ld<COH>(s.q[1], c), ld<COH>(s.q[2], c), ld<COH>(s.q[3], c)};
  if (kind == WALL) {
    T mn = v.y * nx + v.z * ny;
    v.y -= T(2.) * mn * nx;
    v.z -= T(2.) * mn * ny;
  }
  return v;
}

// Sets the status word's non-physical bit with one atomic per wavefront
// (the lowest flagging lane's), however many of its lanes flag a cell.
__device__ __forceinline__ void flag_non_physical(bool bad, Control* ctl) {
  const unsigned long long m = __ballot(bad);
  if (bad && (int)(__lane_id()) == __ffsll((long long)m) - 1) atomicOr(&ctl->status, STATUS_NON_PHYSICAL);
}

template <typename T>
__device__ __forceinline__ bool is_physical(const Vec4<T>& q, T g) {
  Prim<T> w = to_prim(q.x, q.y, q.z, q.w, g);
  // Comparisons with NaN are false; the bounds reject infinities
  // (euler2d.wgsl is_physical).
  return w.rho > T(0) && w.rho < T(3.0e38) && w.p > T(0) && w.p < T(3.0e38) && bs_abs(w.u) < T(3.0e38) &&
         bs_abs(w.v) < T(3.0e38);
}

enum DtMode : int { DT_READ = 0, DT_COMPUTE_LOCAL = 1, DT_GLOBAL = 2 };
enum TileMode : int { TILE_STAGE = 0, TILE_IMPLICIT = 1, TILE_RESIDUAL = 2 };

template <typename T>
struct TileArgsT {
  GridDevT<T> G;
  const BoundaryDevT<T>* B;       // device-resident: indexed by a runtime side without spilling the arguments
  StateDevT<T> u0, src, dst;
  T gamma;
  int axisymmetric;
  int limiter;
  // stage: dst = a u0 + b (src + dt / V R(src))
  T a, b;
  int dt_mode;
  T cfl;
  T* dt;                 // per cell
  const T* dt_global;    // DT_GLOBAL
  Control* control;
  int count_step;            // this launch completes a step: control->steps++
  double* partials;          // per block: sum of (R_rho / V)^2, or null
  // implicit prepare: residual, diagonal, lambda A per face, stored in
  // FP32 for the FP32 relaxation whatever the state's precision
  float* res[4];
  float* diag;
  float* lam_i;              // lambda * area per i-face
  float* lam_j;              // lambda * area per j-face
  float* dq[4];
  T omega;
  // residual vector output (TILE_RESIDUAL)
  T* rvec[4];
};
using TileArgs = TileArgsT<float>;


// Block-wide FP64 sum: wave shuffles, then one LDS step across waves.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; w++) s += scratch[w];
  return s;
}

// Which threads of a TX x TY block compute the tile's high-edge faces. Every
// wave computes two faces per lane (its cells' low i- and j-faces); the TX
// j-faces of the top edge take one more wave's worth of work, the TY
// i-faces of the right edge another, each edge on a wave of its own so
// that the two run side by side. The plan counts wave w as running on SIMD
// w % 4 and gives each edge the highest wave on the SIMD with the least
// work so far. The hardware does not keep that order: read from HW_ID, the
// four waves of a 256-thread block sit on four distinct SIMDs in an order
// that changes from block to block, so at 32 x 8 only "two different
// waves" holds. Both edges in one wave (one direction-generic pass of 40
// lanes instead of two) measured no faster: 15.65 against 15.43 us per
// stage at 1280 x 640.
struct EdgePlan {
  int jstart, istart;
};

constexpr EdgePlan edge_plan(int TX, int TY) {
  const int waves = TX * TY / 64;
  int load[4] = {0, 0, 0, 0};
  for (int w = 0; w < waves; w++) load[w % 4] += 2;
  int jw = waves - 1;
  if (TX > 64) {
    // A j-edge wider than a wave takes the last TX / 64 waves.
    jw = waves - (TX + 63) / 64;
    for (int w = jw; w < waves; w++) load[w % 4] += 1;
  } else {
    for (int w = waves - 1; w >= 0; w--)
      if (load[w % 4] < load[jw % 4]) jw = w;
    load[jw % 4] += 1;
  }
  int iw = waves - 1;
  for (int w = waves - 1; w >= 0; w--)
    if (load[w % 4] < load[iw % 4]) iw = w;
  // In one wave, the i-edge lanes follow the j-edge lanes when they fit.
  const int ioff = (iw == jw && TX % 64 + TY <= 64) ? TX % 64 : 0;
  return {jw * 64, iw * 64 + ioff};
}

// Element idx of a device array as the array's base plus a 32-bit byte
// offset: with the base in SGPRs the access takes the scalar-base form,
// one shared offset register for all arrays of the same index instead of a
// 64-bit address formed per array. Grids are held to 2^28 nodes, so every
// offset fits.
template <typename T>
__device__ __forceinline__ T& at(T* base, int idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& at(const T* base, int idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)idx * (unsigned)sizeof(T));
}

// Per-thread inputs of a tile: the metrics of the thread's own i- and
// j-face (and, when asked, of its high-edge face), its cell's volume,
// planar area, stored or global time step and u0, and, on tiles whose whole
// window lies inside the grid, the conserved state of its cell and of its
// halo cell. All are loaded up front, ahead of the staging, so their
// latencies overlap and no later phase waits on a global load.
template <typename T>
struct TileInputs {
  T mi[3], mj[3], me[3];
  T vol, planar, dt, u0[4];
  Vec4<T> own, halo;
};

// Tiles whose window [i0 - 2, i0 + TX + 2) x [j0 - 2, j0 + TY + 2) lies
// inside the grid (all but a ring of them) read no ghost cell and have no
// boundary face.
template <int TX, int TY>
__device__ __forceinline__ bool tile_is_inner(int ni, int nj, int i0, int j0) {
  return i0 >= 2 && j0 >= 2 && i0 + TX + 2 <= ni && j0 + TY + 2 <= nj;
}

// Window position (wi, wj) of halo cell h < 4 (TX + TY): two rows below and
// above the tile, then two columns left and right (corners excluded).
template <int TX, int TY>
__device__ __forceinline__ void halo_cell(int h, int& wi, int& wj) {
  if (h < 4 * TX) {
    const int r = h / TX;
    wi = 2 + h % TX;
    wj = r < 2 ? r : TY + r;
  } else {
    const int r = (h - 4 * TX) / TY;
    wj = 2 + (h - 4 * TX) % TY;
    wi = r < 2 ? r : TX + r;
  }
}

// Loads a tile's TileInputs; EDGE: also the high-edge face metrics
// (otherwise the edge faces load them when they are computed).
template <typename T, int TX, int TY, int MODE, bool EDGE, bool COH = false>
__device__ __forceinline__ void tile_inputs(const TileArgsT<T>& A, int bx, int by, TileInputs<T>& in) {
  constexpr int NT = TX * TY;
  const GridDevT<T>& G = A.G;
  const int ni = G.ni, nj = G.nj;
  const int i0 = bx * TX, j0 = by * TY;
  const int tid = threadIdx.x;
  const int li = tid % TX, lj = tid / TX;
  const int gi = i0 + li, gj = j0 + lj;
  const bool valid = gi < ni && gj < nj;
  const bool has_i = gi <= ni && gj < nj, has_j = gi < ni && gj <= nj;
  // A value is loaded only where its consumer reads it (the face exists,
  // the cell is in the grid, the tile is inner); the others stay unset: a
  // zero written to a register whose load may still be in flight would make
  // the compiler wait there for every outstanding load.
  if (has_i) {
    const int f = gi + gj * (ni + 1);
    in.mi[0] = at(G.inx, f); in.mi[1] = at(G.iny, f); in.mi[2] = at(G.ia, f);
  }
  if (has_j) {
    const int f = gi + gj * ni;
    in.mj[0] = at(G.jnx, f); in.mj[1] = at(G.jny, f); in.mj[2] = at(G.ja, f);
  }
  if constexpr (EDGE) {
    constexpr EdgePlan E = edge_plan(TX, TY);
    static_assert(E.jstart + TX <= E.istart || E.istart + TY <= E.jstart, "edge ranges overlap");
    if (tid >= E.jstart && tid < E.jstart + TX) {
      const int ei = i0 + tid - E.jstart, ej = j0 + TY;
      if (ei < ni && ej <= nj) {
        const int f = ei + ej * ni;
        in.me[0] = at(G.jnx, f); in.me[1] = at(G.jny, f); in.me[2] = at(G.ja, f);
      }
    } else if (tid >= E.istart && tid < E.istart + TY) {
      const int ei = i0 + TX, ej = j0 + tid - E.istart;
      if (ei <= ni && ej < nj) {
        const int f = ei + ej * (ni + 1);
        in.me[0] = at(G.inx, f); in.me[1] = at(G.iny, f); in.me[2] = at(G.ia, f);
      }
    }
  }
  if (valid) {
    const int c = gi + gj * ni;
    in.vol = at(G.vol, c);
    in.planar = at(G.planar, c);
    if (MODE == TILE_STAGE && A.dt_mode == DT_READ) in.dt = at(A.dt, c);
    if (MODE == TILE_STAGE && A.dt_mode == DT_GLOBAL) in.dt = *A.dt_global;
    // u0 of the SSP-RK combination; global stepping also keeps it for a
    // zero time step (the state is left unchanged).
    if (MODE == TILE_STAGE && (A.a != T(0.) || A.dt_mode == DT_GLOBAL)) {
#pragma unroll
      for (int q = 0; q < 4; q++) in.u0[q] = at(A.u0.q[q], c);
    }
  }
  if (tile_is_inner<TX, TY>(ni, nj, i0, j0)) {
    const int c = gi + gj * ni;
    if constexpr (COH) {
      in.own = {ld<true>(A.src.q[0], c), ld<true>(A.src.q[1], c), ld<true>(A.src.q[2], c), ld<true>(A.src.q[3], c)};
    } else {
      in.own = {at(A.src.q[0], c), at(A.src.q[1], c), at(A.src.q[2], c), at(A.src.q[3], c)};
    }
    if (4 * (TX + TY) <= NT && tid < 4 * (TX + TY)) {
      int wi, wj;
      halo_cell<TX, TY>(tid, wi, wj);
      const int h = (i0 + wi - 2) + (j0 + wj - 2) * ni;
      if constexpr (COH) {
        in.halo = {ld<true>(A.src.q[0], h), ld<true>(A.src.q[1], h), ld<true>(A.src.q[2], h),
                   ld<true>(A.src.q[3], h)};
      } else {
        in.halo = {at(A.src.q[0], h), at(A.src.q[1], h), at(A.src.q[2], h), at(A.src.q[3], h)};
      }
    }
  }
}

// The tile (bx, by) of one grid from its inputs; its FP64 partial goes to
// A.partials[slot]. EDGE: the high-edge faces' metrics come with the
// inputs (otherwise those faces load them).
template <typename T, int TX, int TY, int REC, int LIM, int FLUX, int MODE, bool EDGE, bool COH = false>
__device__ __forceinline__ void tile_compute(const TileArgsT<T>& A, int bx, int by, int slot, const TileInputs<T>& in) {
  constexpr int NT = TX * TY;
  constexpr int WX = TX + 4, WY = TY + 4;
  __shared__ T sp[4][WY][WX];        // primitives of the tile and its halo
  __shared__ T fi[4][TY][TX + 1];    // i-face flux times area
  __shared__ T fj[4][TY + 1][TX];    // j-face flux times area
  // The FP64 stage tiles read the time step's face area vectors from the
  // metric arrays instead of staging them: 9 KB less LDS, so four tiles
  // fit on a CU instead of three.
  // The implicit tiles stage lambda A alone: one plane, which keeps the
  // FP64 implicit tile at 40640 B of LDS, four to a CU.
  constexpr bool XLDS = MODE == TILE_IMPLICIT || !std::is_same<T, double>::value;
  constexpr int XP = MODE == TILE_IMPLICIT ? 1 : 2;
  __shared__ T xi[XP][TY][TX + 1];    // i-face area vector (stage) or lambda A (implicit)
  __shared__ T xj[XP][TY + 1][TX];
  __shared__ double red[NT / 64];
  // The implicit tiles with a shared reconstruction form each cell's
  // limiter coefficient once per direction (shared_coeff) instead of once
  // per face side: i-coefficients of window cells x in [1, TX + 3), rows
  // [2, TY + 2), at ci[y - 2][x - 1]; j-coefficients of columns
  // [2, TX + 2), rows [1, TY + 3), at cj[y - 1][x - 2]. At 32 x 8 that is
  // 2,368 B of LDS in FP32 (eight tiles to a CU still) and 4,736 B in FP64
  // (four still: 40,672 B).
  constexpr bool COEF = MODE == TILE_IMPLICIT && REC != PER_VARIABLE;
  __shared__ T ci[COEF ? TY : 1][COEF ? TX + 2 : 1];
  __shared__ T cj[COEF ? TY + 2 : 1][COEF ? TX : 1];

  const GridDevT<T>& G = A.G;
  const int ni = G.ni, nj = G.nj;
  const T g = A.gamma;
  const int i0 = bx * TX, j0 = by * TY;
  const int tid = threadIdx.x;

  // Phase 1: primitives of the window [i0-2, i0+TX+2) x [j0-2, j0+TY+2),
  // corners excluded: each thread's own cell, whose conserved state it
  // keeps in registers for the update, then the 4 (TX + TY) halo cells.
  // Inner tiles take the state from the inputs; the others evaluate ghost
  // cells from the boundary conditions (load_cell).
  const int li = tid % TX, lj = tid / TX;
  const int gi = i0 + li, gj = j0 + lj;
  const bool valid = gi < ni && gj < nj;
  Vec4<T> own = {T(0.), T(0.), T(0.), T(0.)};
  const T* mi = in.mi;
  const T* mj = in.mj;
  const T c_vol = in.vol, c_planar = in.planar, c_dt = in.dt;
  const T* c_u0 = in.u0;
  // The block-uniform test costs no divergence.
  const bool inner = tile_is_inner<TX, TY>(ni, nj, i0, j0);
  // On a tile at the high edge of the grid the thread's position may be a
  // ghost cell (gi in [ni, ni + 1] or gj in [nj, nj + 1]) that the faces
  // near the edge still need.
  if (!(gi >= ni && gj >= nj) && gi <= ni + 1 && gj <= nj + 1) {
    const Vec4<T> q = inner ? in.own : load_cell<T, COH>(A.src, G, *A.B, gi, gj, g);
    if (valid) own = q;
    Prim<T> w = to_prim(q.x, q.y, q.z, q.w, g);
    sp[0][lj + 2][li + 2] = w.rho;
    sp[1][lj + 2][li + 2] = w.u;
    sp[2][lj + 2][li + 2] = w.v;
    sp[3][lj + 2][li + 2] = w.p;
  }
  for (int h = tid; h < 4 * (TX + TY); h += NT) {
    int wi, wj;
    halo_cell<TX, TY>(h, wi, wj);
    const int hi = i0 + wi - 2, hj = j0 + wj - 2;
    const bool oi = hi < 0 || hi >= ni, oj = hj < 0 || hj >= nj;
    if ((oi && oj) || hi > ni + 1 || hj > nj + 1) continue;
    Vec4<T> q;
    if (inner && 4 * (TX + TY) <= NT) q = in.halo;
    else if (inner) {
      const int c = hi + hj * ni;
      q = {ld<COH>(A.src.q[0], c), ld<COH>(A.src.q[1], c), ld<COH>(A.src.q[2], c), ld<COH>(A.src.q[3], c)};
    } else {
      q = load_cell<T, COH>(A.src, G, *A.B, hi, hj, g);
    }
    Prim<T> w = to_prim(q.x, q.y, q.z, q.w, g);
    sp[0][wj][wi] = w.rho;
    sp[1][wj][wi] = w.u;
    sp[2][wj][wi] = w.v;
    sp[3][wj][wi] = w.p;
  }
  __syncthreads();

  auto prim_at = [&](int lj, int li) -> Prim<T> {
    return {sp[0][lj][li], sp[1][lj][li], sp[2][lj][li], sp[3][lj][li]};
  };
  if constexpr (COEF) {
    // Window cells outside what the tile staged give values no face reads:
    // a face reads the coefficients of its two cells, whose stencils are
    // the four cells face_flux reads.
    constexpr int NI = TY * (TX + 2), NJ = (TY + 2) * TX;
    for (int e = tid; e < NI + NJ; e += NT) {
      if (e < NI) {
        const int y = e / (TX + 2), x = e % (TX + 2);
        ci[y][x] = shared_coeff<T, REC>(prim_at(y + 2, x), prim_at(y + 2, x + 1), prim_at(y + 2, x + 2), g);
      } else {
        const int y = (e - NI) / TX, x = (e - NI) % TX;
        cj[y][x] = shared_coeff<T, REC>(prim_at(y, x + 2), prim_at(y + 1, x + 2), prim_at(y + 2, x + 2), g);
      }
    }
    __syncthreads();
  }

  // Phase 2: faces. Face f of direction d at local (fx, fy). Inner tiles
  // (IN) have no boundary face and no face outside the grid, so their
  // instance drops the side tests and the wall mirror.
  // The implicit spectral radii need each cell's sound speed on both of its
  // low faces and in its diagonal: the thread's own cell computes it once.
  // Read only where the own cell lies in the grid (the faces' gi < ni and
  // gj < nj tests, the valid cells).
  T c_own = T(0.);
  if constexpr (MODE == TILE_IMPLICIT) {
    const Prim<T> w = prim_at(lj + 2, li + 2);
    c_own = bs_sqrt(bs_div(g * w.p, w.rho));
  }
  auto do_face = [&](auto inner_tag, int d, int fx, int fy, const T* pre, const T* c_right) {
    constexpr bool IN = decltype(inner_tag)::value;
    const int gi = i0 + fx, gj = j0 + fy;
    T nx, ny, area;
    Vec4<T> f;
    T lam = T(0.);
    if (d == 0) {
      if (!IN && (gi > ni || gj >= nj)) return;
      const int fidx = gi + gj * (ni + 1);
      if (pre) { nx = pre[0]; ny = pre[1]; area = pre[2]; }
      else { nx = G.inx[fidx]; ny = G.iny[fidx]; area = G.ia[fidx]; }
      const int side = IN ? -1 : gi == 0 ? IMIN : gi == ni ? IMAX : -1;
      const Prim<T> l0 = prim_at(fy + 2, fx + 1), r0 = prim_at(fy + 2, fx + 2);
      if (side >= 0 && A.B->face_flux[side < 0 ? 0 : side] != nullptr) {
        const T* b = A.B->face_flux[side] + 4 * gj;
        f = {b[0], b[1], b[2], b[3]};
      } else {
        const int wall = side >= 0 && A.B->kind[side < 0 ? 0 : side] == WALL ? (side == IMIN ? -1 : 1) : 0;
        if constexpr (COEF)
          f = face_flux_coeff<T, REC, FLUX>(prim_at(fy + 2, fx), l0, r0, prim_at(fy + 2, fx + 3), ci[fy][fx],
                                            ci[fy][fx + 1], nx, ny, wall, g);
        else
          f = face_flux<T, REC, FLUX>(LIM, prim_at(fy + 2, fx), l0, r0, prim_at(fy + 2, fx + 3), nx, ny,
                                      wall, g);
      }
      if (MODE == TILE_IMPLICIT) {
        T rl = gi > 0 ? bs_abs(l0.u * nx + l0.v * ny) + bs_sqrt(bs_div(g * l0.p, l0.rho)) : T(0.);
        T rr = gi < ni ? bs_abs(r0.u * nx + r0.v * ny) + (c_right ? *c_right : bs_sqrt(bs_div(g * r0.p, r0.rho)))
                       : T(0.);
        lam = bs_max(rl, rr) * area;
        A.lam_i[fidx] = lam;
      }
      fi[0][fy][fx] = f.x * area;
      fi[1][fy][fx] = f.y * area;
      fi[2][fy][fx] = f.z * area;
      fi[3][fy][fx] = f.w * area;
      if constexpr (MODE == TILE_IMPLICIT) {
        xi[0][fy][fx] = lam;
      } else if constexpr (XLDS) {
        xi[0][fy][fx] = nx * area;
        xi[1][fy][fx] = ny * area;
      }
    } else {
      if (!IN && (gi >= ni || gj > nj)) return;
      const int fidx = gi + gj * ni;
      if (pre) { nx = pre[0]; ny = pre[1]; area = pre[2]; }
      else { nx = G.jnx[fidx]; ny = G.jny[fidx]; area = G.ja[fidx]; }
      const int side = IN ? -1 : gj == 0 ? JMIN : gj == nj ? JMAX : -1;
      const Prim<T> l0 = prim_at(fy + 1, fx + 2), r0 = prim_at(fy + 2, fx + 2);
      if (side >= 0 && A.B->face_flux[side < 0 ? 0 : side] != nullptr) {
        const T* b = A.B->face_flux[side] + 4 * gi;
        f = {b[0], b[1], b[2], b[3]};
      } else {
        const int wall = side >= 0 && A.B->kind[side < 0 ? 0 : side] == WALL ? (side == JMIN ? -1 : 1) : 0;
        if constexpr (COEF)
          f = face_flux_coeff<T, REC, FLUX>(prim_at(fy, fx + 2), l0, r0, prim_at(fy + 3, fx + 2), cj[fy][fx],
                                            cj[fy + 1][fx], nx, ny, wall, g);
        else
          f = face_flux<T, REC, FLUX>(LIM, prim_at(fy, fx + 2), l0, r0, prim_at(fy + 3, fx + 2), nx, ny,
                                      wall, g);
      }
      if (MODE == TILE_IMPLICIT) {
        T rl = gj > 0 ? bs_abs(l0.u * nx + l0.v * ny) + bs_sqrt(bs_div(g * l0.p, l0.rho)) : T(0.);
        T rr = gj < nj ? bs_abs(r0.u * nx + r0.v * ny) + (c_right ? *c_right : bs_sqrt(bs_div(g * r0.p, r0.rho)))
                       : T(0.);
        lam = bs_max(rl, rr) * area;
        A.lam_j[fidx] = lam;
      }
      fj[0][fy][fx] = f.x * area;
      fj[1][fy][fx] = f.y * area;
      fj[2][fy][fx] = f.z * area;
      fj[3][fy][fx] = f.w * area;
      if constexpr (MODE == TILE_IMPLICIT) {
        xj[0][fy][fx] = lam;
      } else if constexpr (XLDS) {
        xj[0][fy][fx] = nx * area;
        xj[1][fy][fx] = ny * area;
      }
    }
  };

  // The tile's high-edge faces (TX j-faces, TY i-faces) go to the waves
  // that leave the SIMDs evenly loaded (edge_plan).
  constexpr EdgePlan E = edge_plan(TX, TY);
  static_assert(E.jstart + TX <= NT && E.istart + TY <= NT, "edge faces outside the block");
  auto faces = [&](auto inner_tag) {
    // The right cell of the thread's own two faces is its own cell.
    const T* own = MODE == TILE_IMPLICIT ? &c_own : nullptr;
    do_face(inner_tag, 0, li, lj, mi, own);
    do_face(inner_tag, 1, li, lj, mj, own);
    const T* me = EDGE ? in.me : nullptr;
    if (tid >= E.jstart && tid < E.jstart + TX) do_face(inner_tag, 1, tid - E.jstart, TY, me, nullptr);
    if (tid >= E.istart && tid < E.istart + TY) do_face(inner_tag, 0, TX, tid - E.istart, me, nullptr);
  };
  if (inner) faces(std::integral_constant<bool, true>());
  else faces(std::integral_constant<bool, false>());
  __syncthreads();

  // Phase 3: cells.
  double r2 = 0.0;
  if (valid) {
    const int c = gi + gj * ni;
    const Prim<T> w = prim_at(lj + 2, li + 2);
    T r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) r[q] = fi[q][lj][li] - fi[q][lj][li + 1] + fj[q][lj][li] - fj[q][lj + 1][li];
    const T planar = c_planar;
    if (A.axisymmetric) r[2] += w.p * planar;
    const T vol = c_vol;
    if (A.partials != nullptr) {
      double rv = (double)r[0] / (double)vol;
      r2 = rv * rv;
    }
    if constexpr (MODE == TILE_STAGE) {
      T dt;
      if (A.dt_mode == DT_COMPUTE_LOCAL) {
        // euler2d.ts:1183-1205 computeTimeSteps: mean area vector of the
        // two faces in each direction.
        const T a = bs_sqrt(bs_div(g * w.p, w.rho));
        T six, siy, sjx, sjy;
        if constexpr (XLDS) {
          six = T(0.5) * (xi[0][lj][li] + xi[0][lj][li + 1]);
          siy = T(0.5) * (xi[1][lj][li] + xi[1][lj][li + 1]);
          sjx = T(0.5) * (xj[0][lj][li] + xj[0][lj + 1][li]);
          sjy = T(0.5) * (xj[1][lj][li] + xj[1][lj + 1][li]);
        } else {
          // Each product rounded on its own (no fused multiply-add), as
          // the staged ones are; the steps differ from the staged form's
          // at round-off (5e-14 after 300 steps at 1280 x 640).
          const int f1 = gi + 1 + gj * (ni + 1), f2 = gi + (gj + 1) * ni;
          six = T(0.5) * (__dmul_rn(mi[0], mi[2]) + __dmul_rn(G.inx[f1], G.ia[f1]));
          siy = T(0.5) * (__dmul_rn(mi[1], mi[2]) + __dmul_rn(G.iny[f1], G.ia[f1]));
          sjx = T(0.5) * (__dmul_rn(mj[0], mj[2]) + __dmul_rn(G.jnx[f2], G.ja[f2]));
          sjy = T(0.5) * (__dmul_rn(mj[1], mj[2]) + __dmul_rn(G.jny[f2], G.ja[f2]));
        }
        const T lam = bs_abs(w.u * six + w.v * siy) + a * bs_sqrt(six * six + siy * siy) +
                          bs_abs(w.u * sjx + w.v * sjy) + a * bs_sqrt(sjx * sjx + sjy * sjy);
        dt = A.cfl * vol * bs_rcp(lam);
        at(A.dt, c) = dt;
      } else {
        dt = c_dt;  // the stored local step, or the global one
      }
      const T k = dt * bs_rcp(vol);
      Vec4<T> out;
      const T srcq[4] = {own.x, own.y, own.z, own.w};
      T o[4];
      // a u0 + b (src + k R) with a = 1 - b (every Shu-Osher stage here),
      // written as u0 + b ((src - u0) + k R): the FP32 coefficients 1/3 and
      // 2/3 do not sum to 1, and the plain form would add (a + b - 1) u0 per
      // step, a source that shifts the steady state by up to 1e-5 in CD.
      if (A.a != T(0.)) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const T u0q = c_u0[q];
          o[q] = u0q + A.b * ((srcq[q] - u0q) + k * r[q]);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) o[q] = srcq[q] + A.b * (k * r[q]);
      }
      if (A.dt_mode == DT_GLOBAL && dt <= T(0.)) {
        // A zero time step (global stepping already at t_end) leaves the
        // state unchanged (euler2d.wgsl update_cells).
#pragma unroll
        for (int q = 0; q < 4; q++) o[q] = c_u0[q];
      }
#pragma unroll
      for (int q = 0; q < 4; q++) at(A.dst.q[q], c) = o[q];
      out = {o[0], o[1], o[2], o[3]};
      flag_non_physical(!is_physical(out, g), A.control);
    } else if (MODE == TILE_IMPLICIT) {
      // euler2d.ts:955-971 / euler2d-implicit.wgsl implicit_prepare: the
      // scalar diagonal and the residual
      // dQ = R / D on the cells with i + j even.
      const double cfl = A.control->cfl;
      const T scale = (T)(0.5 * ((double)A.omega + 1.0 / cfl));
      T d = scale * (xi[0][lj][li] + xi[0][lj][li + 1] + xj[0][lj][li] + xj[0][lj + 1][li]);
      if (A.axisymmetric) d += planar * (bs_abs(w.v) + c_own);
      at(A.diag, c) = d;
      // The first red half-sweep (dQ = R / D on red cells, 0 on black) is
      // formed by the relaxation's first launch from R and D.
#pragma unroll
      for (int q = 0; q < 4; q++) at(A.res[q], c) = r[q];
    } else if (A.rvec[0] != nullptr) {
#pragma unroll
      for (int q = 0; q < 4; q++) at(A.rvec[q], c) = r[q];
    }
  }
  if (A.partials != nullptr) {
    double s = block_sum<NT>(r2, red);
    if (tid == 0) A.partials[slot] = s;
  }
  if (A.count_step && bx == 0 && by == 0 && tid == 0) atomicAdd((unsigned long long*)&A.control->steps, 1ull);
}

template <typename T, int TX, int TY, int REC, int LIM, int FLUX, int MODE, bool COH = false>
__device__ __forceinline__ void tile_body(const TileArgsT<T>& A, int bx, int by, int slot) {
  // The FP32 tiles load the edge faces' metrics with the other inputs; the
  // FP64 tiles, short of registers, when the edge faces are computed.
  constexpr bool EDGE = std::is_same<T, float>::value;
  TileInputs<T> in;
  tile_inputs<T, TX, TY, MODE, EDGE, COH>(A, bx, by, in);
  tile_compute<T, TX, TY, REC, LIM, FLUX, MODE, EDGE, COH>(A, bx, by, slot, in);
}

// Workgroups go to the 8 XCDs round-robin by their linear index, each XCD
// with its own L2. Block b becomes tile t so that every XCD gets one
// contiguous run of tiles in row-major order: vertically adjacent tiles,
// which share two halo rows, then run on the same XCD about one tile row
// apart, and the second read of those rows hits its L2. A bijection of
// [0, n) for any n.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

template <typename T, int TX, int TY, int REC, int LIM, int FLUX, int MODE>
__global__ __launch_bounds__(TX * TY) void tile_kernel(TileArgsT<T> A) {
  // One-dimensional grid of tiles_x * tiles_y blocks.
  const int tiles_x = (A.G.ni + TX - 1) / TX;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  tile_body<T, TX, TY, REC, LIM, FLUX, MODE>(A, t % tiles_x, t / tiles_x, t);
}

// The FP32 stage tiles at eight waves per SIMD, so that a CU holds as many
// tiles as their LDS allows (eight of 32 x 8). The minmod instances need
// 64 VGPRs, eight waves' worth, by themselves. The compiler is asked for
// six waves at least, not eight: at (8, 8) it also held them to 78 SGPRs
// and spilled 39 more to VGPR lanes. At (6, 8) it keeps 94 SGPRs, still
// eight waves, and the 1280 x 640 stage measured 15.33 against 15.62 us.
// (The shared-scaled instances take 67 VGPRs, seven waves, instead of 64
// and four spilled to scratch.)
template <int TX, int TY, int REC, int LIM, int FLUX, int MODE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(6, 8))) void tile_kernel_occ8(
    TileArgsT<float> A) {
  const int tiles_x = (A.G.ni + TX - 1) / TX;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  tile_body<float, TX, TY, REC, LIM, FLUX, MODE>(A, t % tiles_x, t / tiles_x, t);
}

// A batch of grids in one launch: block b takes work item b (the case's
// argument record, the tile and the partial slot); the argument records
// live in device memory, one per case and stage role.
struct WorkItem {
  int args;  // index of the case's TileArgs
  int bx, by;
  int slot;  // FP64 partial slot
};

template <typename T, int TX, int TY, int REC, int LIM, int FLUX, int MODE>
__global__ __launch_bounds__(TX * TY) void batch_tile_kernel(const TileArgsT<T>* __restrict__ args,
                                                             const WorkItem* __restrict__ work) {
  // Work items are ordered by case and tile row; each XCD takes one
  // contiguous run of them (xcd_tile).
  const WorkItem w = work[xcd_tile(blockIdx.x, gridDim.x)];
  tile_body<T, TX, TY, REC, LIM, FLUX, MODE>(args[w.args], w.bx, w.by, w.slot);
}

// The FP32 batched stage tiles, with tile_kernel_occ8's register bounds.
template <int TX, int TY, int REC, int LIM, int FLUX, int MODE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(6, 8))) void batch_tile_kernel_occ8(
    const TileArgsT<float>* __restrict__ args, const WorkItem* __restrict__ work) {
  const WorkItem w = work[xcd_tile(blockIdx.x, gridDim.x)];
  tile_body<float, TX, TY, REC, LIM, FLUX, MODE>(args[w.args], w.bx, w.by, w.slot);
}

// Deterministic FP64 sum of n partials by one block; writes sqrt(sum / cells)
// when `rms` (euler2d.ts:480-491) or the plain sum.
static __global__ __launch_bounds__(256) void finalize_sum(const double* partials, int n, double cells, int rms,
                                                   double* out) {
  __shared__ double red[4];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) v += partials[k];
  double s = block_sum<256>(v, red);
  if (threadIdx.x == 0) *out = rms ? sqrt(s / cells) : s;
}

// Per-cell global time step and block minima (euler2d.wgsl
// compute_time_steps); the finalize takes the minimum and clamps it to the
// FP64 clock's remaining time (finalize_time_step).
template <typename T>
__global__ __launch_bounds__(256) void time_step_kernel(GridDevT<T> G, StateDevT<T> s, T g, T cfl,
                                                        T* partial_min) {
  __shared__ T red[256];
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int ni = G.ni;
  T dt = T(3.0e38);
  if (c < G.ni * G.nj) {
    const int i = c % ni, j = c / ni;
    Prim<T> w = to_prim(s.q[0][c], s.q[1][c], s.q[2][c], s.q[3][c], g);
    T a = bs_sqrt(bs_div(g * w.p, w.rho));
    int f0 = i + j * (ni + 1), f1 = f0 + 1, h0 = c, h1 = c + ni;
    T six = T(0.5) * (G.inx[f0] * G.ia[f0] + G.inx[f1] * G.ia[f1]);
    T siy = T(0.5) * (G.iny[f0] * G.ia[f0] + G.iny[f1] * G.ia[f1]);
    T sjx = T(0.5) * (G.jnx[h0] * G.ja[h0] + G.jnx[h1] * G.ja[h1]);
    T sjy = T(0.5) * (G.jny[h0] * G.ja[h0] + G.jny[h1] * G.ja[h1]);
    T lam = bs_abs(w.u * six + w.v * siy) + a * bs_sqrt(six * six + siy * siy) +
                bs_abs(w.u * sjx + w.v * sjy) + a * bs_sqrt(sjx * sjx + sjy * sjy);
    dt = cfl * G.vol[c] * bs_rcp(lam);
    if (!(dt == dt)) dt = -T(1.);  // NaN marks the block
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = bs_fmin(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial_min[blockIdx.x] = red[0];
}

template <typename T>
__global__ __launch_bounds__(256) void finalize_time_step(const T* partial_min, int n, Control* ctl,
                                                          T* dt_out) {
  __shared__ T red[256];
  T v = T(3.0e38);
  for (int k = threadIdx.x; k < n; k += 256) v = bs_fmin(v, partial_min[k]);
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = bs_fmin(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    T m = red[0];
    double dt;
    if (!(m > T(0.) && m < T(3.0e38))) {
      atomicOr(&ctl->status, STATUS_NON_PHYSICAL);
      dt = 0.0;
    } else {
      double remaining = ctl->t_end - ctl->time;
      double limit = 1e-12 * (fabs(ctl->t_end) > 1.0 ? fabs(ctl->t_end) : 1.0);
      dt = remaining <= limit ? 0.0 : (m < remaining ? (double)m : remaining);
    }
    ctl->dt = dt;
    *dt_out = (T)dt;
  }
}

// After the last stage of a global step: FP64 clock and step count; a step
// of zero length (already at t_end) is not counted.
static __global__ void advance_clock(Control* ctl) {
  if (ctl->dt > 0.0) {
    ctl->time += ctl->dt;
    double limit = 1e-12 * (fabs(ctl->t_end) > 1.0 ? fabs(ctl->t_end) : 1.0);
    if (ctl->t_end - ctl->time < limit) ctl->time = ctl->t_end;
    ctl->steps += 1;
  }
}

// ---------------------------------------------------------------------------
// Implicit red-black relaxation (euler2d.ts:1007-1056, euler2d-implicit.wgsl
// implicit_relax / implicit_update).
// ---------------------------------------------------------------------------

// The relaxation reads the state and the metrics in their own precision
// and relaxes in FP32: an FP64 solver's implicit step is FP64 residual,
// diagonal and update around an FP32 relaxation, as the reference's
// two-float engine solves its implicit system in 32 bits
// (euler2d-two-float.ts:1-14).
template <typename T>
struct RelaxArgsT {
  GridDevT<T> G;
  StateDevT<T> q;
  const float* res[4];
  const float* diag;
  const float* lam_i;
  const float* lam_j;
  float* dq[4];
  float gamma, omega;
};
using RelaxArgs = RelaxArgsT<float>;

template <typename S>
struct UpdateArgsT {
  StateDevT<S> src, dst;
  StateDevT<float> dq;
  int n;
  S gamma;
  Control* control;
  double cfl_max, growth;
};
using UpdateArgs = UpdateArgsT<float>;

// q + alpha dQ, alpha halved (at most 30 times) until density and pressure
// stay within a factor of two (euler2d.ts:816-847).
template <typename S>
__device__ __forceinline__ Vec4<S> update_cell(S rho, S mx, S my, S e, S d0, S d1, S d2, S d3, S g) {
  const S p = (g - S(1)) * (e - S(0.5) * (mx * mx + my * my) * bs_rcp(rho));
  S alpha = S(1);
  for (int tries = 0; tries < 30; tries++) {
    S rho1 = rho + alpha * d0, mx1 = mx + alpha * d1, my1 = my + alpha * d2, e1 = e + alpha * d3;
    S p1 = (g - S(1)) * (e1 - S(0.5) * (mx1 * mx1 + my1 * my1) * bs_rcp(rho1));
    if (rho1 > S(0.5) * rho && rho1 < S(2) * rho && p1 > S(0.5) * p && p1 < S(2) * p) break;
    alpha *= S(0.5);
  }
  return {rho + alpha * d0, mx + alpha * d1, my + alpha * d2, e + alpha * d3};
}

// The CFL ramp and the step count, once per step (transonic-flow.ts:208).
template <typename S>
__device__ __forceinline__ void advance_control(const UpdateArgsT<S>& U) {
  Control* ctl = U.control;
  double next = ctl->cfl * U.growth;
  ctl->cfl = next < U.cfl_max ? next : U.cfl_max;
  ctl->steps += 1;
}


// NH consecutive half-sweeps (colours colour0, colour0 ^ 1, ...) of an
// RTX x RTY tile in one launch. Gauss-Seidel by colours: a cell's update reads
// only cells of the other colour, so half-sweep s is exact on the tile grown
// by NH - 1 - s cells, given q and dQ on the tile grown by NH cells; those
// are staged in LDS once and the tile's dQ written back at the end. The
// cell arithmetic is relax_body's (neighbour_term in the same order), so
// the result is bitwise that of NH single half-sweeps; the grid's own
// faces carry no neighbour term, as there.
// The relaxation tile: 32 x 16 measured fastest (1280 x 640, three
// half-sweeps: 30.9 us; 64 x 8 31.7, 32 x 8 33.4, 64 x 16 37.7, 64 x 4
// 51.4; the population of 500 rockets 1.5 % faster than at 64 x 8): its
// 38 x 22 window stages fewer halo cells per cell than 70 x 14. With the
// residual and diagonal in registers (three 32 x 16 tiles to a CU), 32 x 8
// (five to a CU) measured 28.0 us against 26.6 and the 500-rocket
// population 136.3 against 143.2 solves/s.
#ifndef BS_RTX
#define BS_RTX 32
#endif
#ifndef BS_RTY
#define BS_RTY 16
#endif
constexpr int RTX = BS_RTX, RTY = BS_RTY;
// Threads per relaxation block; every loop of the tile is strided by it.
// The half-sweeps of a 32 x 16 tile update 360, 306 and 256 cells, so 512
// threads leave lanes idle, yet fewer threads are slower (1280 x 640,
// three half-sweeps: 384 threads 49.0 us, 448 34.0, 512 30.8) and more
// gain nothing beyond noise (768 29.9, 1024 30.2; the 500-rocket
// population within 3 %): the tile is latency-bound, and waves in flight
// matter more than busy lanes.
#ifndef BS_RNT
#define BS_RNT (BS_RTX * BS_RTY)
#endif
constexpr int RNT = BS_RNT;
static_assert(RNT % 64 == 0 && RNT <= 1024, "relaxation block size");
// Waves per SIMD the relaxation's registers must allow: three 512-thread
// tiles to a CU (their LDS allows three) need six, 80 VGPRs at most.
#ifndef BS_RELAX_WAVES
#define BS_RELAX_WAVES 6
#endif
constexpr int RELAX_WAVES = BS_RELAX_WAVES;

struct RelaxCtx {
  int i0, j0;  // global index of LDS window cell (0, 0)
};

// INIT: the launch forms the first red half-sweep (a template parameter,
// so that an instance keeps registers for only one of the two stagings).
template <typename T, int NH, bool LIN, bool INIT>
__device__ __forceinline__ void relax_tile_body(const RelaxArgsT<T>& A, int bx, int by, int colour0,
                                                const UpdateArgsT<T>& U, bool fuse_update) {
  static_assert(LIN || std::is_same<T, float>::value, "an FP64 state relaxes with the linearized term");
  const bool fuse = fuse_update;
  constexpr int WX = RTX + 2 * NH, WY = RTY + 2 * NH, NT = RNT;
  // Window cells [0, WX) x [0, WY) hold q and dQ; the cells updated by
  // some half-sweep lie in [1, WX - 1) x [1, WY - 1), and their faces
  // (normal, area, lambda A) are staged too. The residual and diagonal of
  // the cells a thread updates in each half-sweep go to its registers
  // instead of LDS: 16.7 KB less LDS at 32 x 16 and three half-sweeps
  // (50,688 B instead of 67,408), three tiles to a CU instead of two. With
  // INIT the first red half-sweep (dQ = R / D on red cells, 0 on
  // black, as the prepare kernel defines it) is formed here instead of
  // read.
  constexpr int CX = WX - 2, CY = WY - 2;
  __shared__ float sq[4][WY][WX];
  __shared__ float sdq[4][WY][WX];
  __shared__ float sfi[4][CY][CX + 1];  // i-face x (between cells x - 1 and x) of row y: [y - 1][x - 1]
  __shared__ float sfj[4][CY + 1][CX];  // j-face y of column x: [y - 1][x - 1]
  const GridDevT<T>& G = A.G;
  const int ni = G.ni, nj = G.nj;
  const int i0 = bx * RTX - NH, j0 = by * RTY - NH;
  const int tid = threadIdx.x;
  // Every global load of the staging is issued before the first LDS store,
  // so their latencies overlap: a thread's window cells (q, R, D, and dQ
  // unless it is formed here) and faces (normal, area, lambda A) first go
  // to registers.
  constexpr int NC = (WX * WY + NT - 1) / NT, NFI = ((CX + 1) * CY + NT - 1) / NT,
                NFJ = (CX * (CY + 1) + NT - 1) / NT;
  float cq[NC][4], cr[NC][4], cd[NC], cdq[NC][4], vi[NFI][4], vj[NFJ][4];
  bool ok_c[NC], ok_i[NFI], ok_j[NFJ];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    const int e = tid + k * NT;
    const int gi = i0 + e % WX, gj = j0 + e / WX;
    ok_c[k] = e < WX * WY && gi >= 0 && gi < ni && gj >= 0 && gj < nj;
    if (ok_c[k]) {
      const int c = gi + gj * ni;
#pragma unroll
      for (int q = 0; q < 4; q++) cq[k][q] = (float)at(A.q.q[q], c);
      if constexpr (INIT) {
#pragma unroll
        for (int q = 0; q < 4; q++) cr[k][q] = at(A.res[q], c);
        cd[k] = at(A.diag, c);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) cdq[k][q] = at(A.dq[q], c);
      }
    }
  }
  // Half-sweep s updates the cells of one colour in the tile grown by
  // r = NH - 1 - s cells, enumerated row by row, (RTX + 2 r) / 2 to a row;
  // thread tid takes the cells tid, tid + NT, ... of that list (HK at
  // most). hs_cell places entry t: window (wx, wy), grid (gi, gj).
  constexpr int HS0 = (RTX + 2 * (NH - 1)) / 2 * (RTY + 2 * (NH - 1)), HK = (HS0 + NT - 1) / NT;
  auto hs_cell = [&](int s, int t, int& wx, int& wy, int& gi, int& gj) {
    const int colour = (colour0 + s) & 1;
    const int r = NH - 1 - s;
    const int half_w = (RTX + 2 * r) / 2;
    const int row = t / half_w, k = t % half_w;
    wy = NH - r + row;
    gj = j0 + wy;
    const int gx0 = i0 + NH - r;
    gi = gx0 + 2 * k + ((colour - gx0 - gj) & 1);
    wx = gi - i0;
    return t < half_w * (RTY + 2 * r) && gi >= 0 && gi < ni && gj >= 0 && gj < nj;
  };
  float hres[NH][HK][4], hdiag[NH][HK];
#pragma unroll
  for (int s = 0; s < NH; s++) {
#pragma unroll
    for (int k = 0; k < HK; k++) {
      int wx, wy, gi, gj;
      if (hs_cell(s, tid + k * NT, wx, wy, gi, gj)) {
        const int c = gi + gj * ni;
#pragma unroll
        for (int q = 0; q < 4; q++) hres[s][k][q] = at(A.res[q], c);
        hdiag[s][k] = at(A.diag, c);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NFI; k++) {
    const int e = tid + k * NT;
    const int gi = i0 + 1 + e % (CX + 1), gj = j0 + 1 + e / (CX + 1);
    ok_i[k] = e < (CX + 1) * CY && gi >= 0 && gi <= ni && gj >= 0 && gj < nj;
    if (ok_i[k]) {
      const int f = gi + gj * (ni + 1);
      vi[k][0] = (float)at(G.inx, f);
      vi[k][1] = (float)at(G.iny, f);
      vi[k][2] = (float)at(G.ia, f);
      vi[k][3] = at(A.lam_i, f);
    }
  }
#pragma unroll
  for (int k = 0; k < NFJ; k++) {
    const int e = tid + k * NT;
    const int gi = i0 + 1 + e % CX, gj = j0 + 1 + e / CX;
    ok_j[k] = e < CX * (CY + 1) && gi >= 0 && gi < ni && gj >= 0 && gj <= nj;
    if (ok_j[k]) {
      const int f = gi + gj * ni;
      vj[k][0] = (float)at(G.jnx, f);
      vj[k][1] = (float)at(G.jny, f);
      vj[k][2] = (float)at(G.ja, f);
      vj[k][3] = at(A.lam_j, f);
    }
  }
#pragma unroll
  for (int k = 0; k < NC; k++) {
    if (!ok_c[k]) continue;
    const int e = tid + k * NT;
    const int wx = e % WX, wy = e / WX;
    const int gi = i0 + wx, gj = j0 + wy;
#pragma unroll
    for (int q = 0; q < 4; q++) sq[q][wy][wx] = cq[k][q];
    if constexpr (INIT) {
      const bool red = ((gi + gj) & 1) == 0;
#pragma unroll
      for (int q = 0; q < 4; q++) sdq[q][wy][wx] = red ? bs_div(cr[k][q], cd[k]) : 0.f;
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) sdq[q][wy][wx] = cdq[k][q];
    }
  }
#pragma unroll
  for (int k = 0; k < NFI; k++) {
    if (!ok_i[k]) continue;
    const int e = tid + k * NT;
    const int fx = 1 + e % (CX + 1), fy = 1 + e / (CX + 1);
#pragma unroll
    for (int m = 0; m < 4; m++) sfi[m][fy - 1][fx - 1] = vi[k][m];
  }
#pragma unroll
  for (int k = 0; k < NFJ; k++) {
    if (!ok_j[k]) continue;
    const int e = tid + k * NT;
    const int fx = 1 + e % CX, fy = 1 + e / CX;
#pragma unroll
    for (int m = 0; m < 4; m++) sfj[m][fy - 1][fx - 1] = vj[k][m];
  }
  __syncthreads();
  const float g = A.gamma;
  // neighbour_term on LDS values (same arithmetic, same order). With LIN
  // the term is the flux Jacobian applied to dQ instead
  // (euler2d-two-float.wgsl:574-598): linear in dQ, so its FP32 rounding
  // shrinks with dQ, where F(Q + dQ) - F(Q) in FP32 keeps a noise of the
  // size of F that does not (and that would set a converged FP64 state
  // wandering again). The two agree to first order in dQ.
  auto term = [&](int wx, int wy, float nx, float ny, float area, float lam_a, float sum[4]) {
    const float rho = sq[0][wy][wx], mx = sq[1][wy][wx], my = sq[2][wy][wx], e = sq[3][wy][wx];
    const float d0 = sdq[0][wy][wx], d1 = sdq[1][wy][wx], d2 = sdq[2][wy][wx], d3 = sdq[3][wy][wx];
    if constexpr (LIN) {
      const float ir = bs_rcp(rho);
      const float u = mx * ir, v = my * ir;
      const float un = u * nx + v * ny;
      const float p = (g - 1.f) * (e - 0.5f * (mx * mx + my * my) * ir);
      const float dp = (g - 1.f) * (d3 - (u * d1 + v * d2) + 0.5f * (u * u + v * v) * d0);
      const float dun = (d1 * nx + d2 * ny - un * d0) * ir;
      const float half = 0.5f * area, damp = 0.5f * A.omega * lam_a;
      sum[0] += half * (d1 * nx + d2 * ny) - damp * d0;
      sum[1] += half * (d1 * un + mx * dun + dp * nx) - damp * d1;
      sum[2] += half * (d2 * un + my * dun + dp * ny) - damp * d2;
      sum[3] += half * ((d3 + dp) * un + (e + p) * dun) - damp * d3;
      return;
    }
    const float ir0 = bs_rcp(rho);
    const float p0 = (g - 1.f) * (e - 0.5f * (mx * mx + my * my) * ir0);
    const float un0 = (mx * nx + my * ny) * ir0;
    const float rho1 = rho + d0, mx1 = mx + d1, my1 = my + d2, e1 = e + d3;
    const float ir1 = bs_rcp(rho1);
    const float p1 = (g - 1.f) * (e1 - 0.5f * (mx1 * mx1 + my1 * my1) * ir1);
    const float un1 = (mx1 * nx + my1 * ny) * ir1;
    const float half = 0.5f * area, damp = 0.5f * A.omega * lam_a;
    sum[0] += half * (rho1 * un1 - rho * un0) - damp * d0;
    sum[1] += half * (mx1 * un1 + p1 * nx - (mx * un0 + p0 * nx)) - damp * d1;
    sum[2] += half * (my1 * un1 + p1 * ny - (my * un0 + p0 * ny)) - damp * d2;
    sum[3] += half * ((e1 + p1) * un1 - (e + p0) * un0) - damp * d3;
  };
#pragma unroll
  for (int s = 0; s < NH; s++) {
#pragma unroll
    for (int hk = 0; hk < HK; hk++) {
      int wx, wy, gi, gj;
      if (!hs_cell(s, tid + hk * NT, wx, wy, gi, gj)) continue;
      float sum[4] = {0.f, 0.f, 0.f, 0.f};
      const int cx = wx - 1, cy = wy - 1;
      if (gi > 0) term(wx - 1, wy, -sfi[0][cy][cx], -sfi[1][cy][cx], sfi[2][cy][cx], sfi[3][cy][cx], sum);
      if (gi < ni - 1)
        term(wx + 1, wy, sfi[0][cy][cx + 1], sfi[1][cy][cx + 1], sfi[2][cy][cx + 1], sfi[3][cy][cx + 1], sum);
      if (gj > 0) term(wx, wy - 1, -sfj[0][cy][cx], -sfj[1][cy][cx], sfj[2][cy][cx], sfj[3][cy][cx], sum);
      if (gj < nj - 1)
        term(wx, wy + 1, sfj[0][cy + 1][cx], sfj[1][cy + 1][cx], sfj[2][cy + 1][cx], sfj[3][cy + 1][cx], sum);
      const float id = bs_rcp(hdiag[s][hk]);
#pragma unroll
      for (int q = 0; q < 4; q++) sdq[q][wy][wx] = (hres[s][hk][q] - sum[q]) * id;
    }
    __syncthreads();
  }
  // The tile's own cells back to global memory: dQ, or, after a step's
  // last half-sweep (fuse), the updated state itself (update_cell), with
  // the control advanced once.
  for (int e = tid; e < RTX * RTY; e += NT) {
    const int lx = e % RTX, ly = e / RTX;
    const int gi = bx * RTX + lx, gj = by * RTY + ly;
    if (gi >= ni || gj >= nj) continue;
    const int c = gi + gj * ni;
    const int wx = lx + NH, wy = ly + NH;
    if (fuse) {
      // An FP64 state is updated from its own FP64 values (the window
      // holds them rounded to FP32).
      Vec4<T> o;
      if constexpr (std::is_same<T, float>::value)
        o = update_cell<float>(sq[0][wy][wx], sq[1][wy][wx], sq[2][wy][wx], sq[3][wy][wx], sdq[0][wy][wx],
                               sdq[1][wy][wx], sdq[2][wy][wx], sdq[3][wy][wx], U.gamma);
      else
        o = update_cell<T>(at(U.src.q[0], c), at(U.src.q[1], c), at(U.src.q[2], c), at(U.src.q[3], c),
                           (T)sdq[0][wy][wx], (T)sdq[1][wy][wx], (T)sdq[2][wy][wx], (T)sdq[3][wy][wx], U.gamma);
      at(U.dst.q[0], c) = o.x;
      at(U.dst.q[1], c) = o.y;
      at(U.dst.q[2], c) = o.z;
      at(U.dst.q[3], c) = o.w;
      flag_non_physical(!is_physical(o, U.gamma), U.control);
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) at(A.dq[q], c) = sdq[q][wy][wx];
    }
  }
  if (fuse && bx == 0 && by == 0 && tid == 0) advance_control(U);
}

template <typename T, int NH, bool LIN, bool INIT>
__global__ __launch_bounds__(RNT) __attribute__((amdgpu_waves_per_eu(RELAX_WAVES, 8))) void relax_tile_kernel(
    RelaxArgsT<T> A, int colour0, UpdateArgsT<T> U, int fuse_update) {
  // One-dimensional grid, tiles placed per XCD as in tile_kernel.
  const int tiles_x = (A.G.ni + RTX - 1) / RTX;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  relax_tile_body<T, NH, LIN, INIT>(A, t % tiles_x, t / tiles_x, colour0, U, fuse_update != 0);
}

template <typename T, int NH, bool LIN, bool INIT>
__global__ __launch_bounds__(RNT) __attribute__((amdgpu_waves_per_eu(RELAX_WAVES, 8))) void batch_relax_tile_kernel(
    const RelaxArgsT<T>* __restrict__ args, const WorkItem* __restrict__ work, int colour0,
    const UpdateArgsT<T>* upd) {
  const WorkItem w = work[xcd_tile(blockIdx.x, gridDim.x)];
  if (upd) relax_tile_body<T, NH, LIN, INIT>(args[w.args], w.bx, w.by, colour0, upd[w.args], true);
  else relax_tile_body<T, NH, LIN, INIT>(args[w.args], w.bx, w.by, colour0, UpdateArgsT<T>{}, false);
}


// ---------------------------------------------------------------------------
// FP64 boundary fluxes and drag (euler2d.ts:499-525, forces.ts:19-29).
// ---------------------------------------------------------------------------

template <typename S>
struct WallArgsT {
  GridDevT<S> G;
  StateDevT<S> s;
  int side;
  int count;
  int kind;                 // boundary kind of the side
  double state[4];          // primitive outside an inflow / far-field side
  const S* face_cons;       // FACE_STATES side
  const double* face_geo;   // per face: x, y, nx, ny, area (FP64, normals towards increasing index)
  double gamma;             // configured gamma (FP64)
  double gamma32;           // the gamma the state was computed with (FP32 for an FP32 state)
  int rec, limiter, flux;
  double* out;              // 9 per face, or null
  double p_inf;
  double* partials;         // per block: sum of (F_x - p_inf n_x) A, or null
};
using WallArgs = WallArgsT<float>;

template <typename T, typename S>
__device__ __forceinline__ Prim<T> prim64(const StateDevT<S>& s, int c, T g32) {
  T rho = (T)s.q[0][c], mu = (T)s.q[1][c], mv = (T)s.q[2][c], e = (T)s.q[3][c];
  return {rho, mu / rho, mv / rho, (g32 - T(1)) * (e - (T(0.5) * (mu * mu + mv * mv)) / rho)};
}

// Primitive state of cell (i, j) or a ghost cell, in FP64, with the
// boundary conditions of euler2d.ts:1361-1403 applied to primitives.
template <typename S>
__device__ Prim<double> wall_cell(const WallArgsT<S>& A, const int bkind[4], const double bstate[4][4],
                                  const S* const bcons[4], int i, int j) {
  const int ni = A.G.ni, nj = A.G.nj;
  if (i >= 0 && i < ni && j >= 0 && j < nj) return prim64<double>(A.s, i + j * ni, A.gamma32);
  int side, t, mi, mj, wi, wj;
  double nx, ny;
  if (i < 0) { side = IMIN; t = j; mi = -1 - i; mj = j; wi = i + ni; wj = j; }
  else if (i >= ni) { side = IMAX; t = j; mi = 2 * ni - 1 - i; mj = j; wi = i - ni; wj = j; }
  else if (j < 0) { side = JMIN; t = i; mi = i; mj = -1 - j; wi = i; wj = j + nj; }
  else { side = JMAX; t = i; mi = i; mj = 2 * nj - 1 - j; wi = i; wj = j - nj; }
  if (side < 2) {
    int f = (side == IMIN ? 0 : ni) + t * (ni + 1);
    nx = A.G.inx[f]; ny = A.G.iny[f];
  } else {
    int f = t + (side == JMIN ? 0 : nj) * ni;
    nx = A.G.jnx[f]; ny = A.G.jny[f];
  }
  if (side == A.side) {
    // The evaluated side's normals in FP64.
    nx = A.face_geo[5 * t + 2];
    ny = A.face_geo[5 * t + 3];
  }
  const int kind = bkind[side];
  if (kind == INFLOW) return {bstate[side][0], bstate[side][1], bstate[side][2], bstate[side][3]};
  if (kind == FACE_STATES) {
    const S* b = bcons[side] + 4 * t;
    double rho = b[0], mu = b[1], mv = b[2], e = b[3];
    return {rho, mu / rho, mv / rho, (A.gamma32 - 1.0) * (e - (0.5 * (mu * mu + mv * mv)) / rho)};
  }
  if (kind == FAR_FIELD) {
    int ii = side == IMIN ? 0 : side == IMAX ? ni - 1 : i;
    int jj = side == JMIN ? 0 : side == JMAX ? nj - 1 : j;
    Prim<double> in = prim64<double>(A.s, ii + jj * ni, A.gamma32);
    Prim<double> fr = {bstate[side][0], bstate[side][1], bstate[side][2], bstate[side][3]};
    double o = (side == IMIN || side == JMIN) ? -1.0 : 1.0;
    return far_field_state<double>(in, fr, o * nx, o * ny, A.gamma);
  }
  Prim<double> w = prim64<double>(A.s, kind == PERIODIC ? wi + wj * ni : mi + mj * ni, A.gamma32);
  if (kind == WALL) {
    double un = w.u * nx + w.v * ny;
    w.u -= 2.0 * un * nx;
    w.v -= 2.0 * un * ny;
  }
  return w;
}

template <typename S>
struct WallBoundaryT {
  int kind[4];
  double state[4][4];
  const S* cons[4];
};
using WallBoundary = WallBoundaryT<float>;

// Block `block` evaluates faces 256 block ... of the side; its drag partial
// goes to partials[slot].
template <typename S, int REC, int FLUX>
__device__ __forceinline__ void wall_body(const WallArgsT<S>& A, const WallBoundaryT<S>& WB, int block, int slot) {
  __shared__ double red[4];
  const int t = block * 256 + threadIdx.x;
  double contrib = 0.0;
  if (t < A.count) {
    const int ni = A.G.ni, nj = A.G.nj;
    int i, j, di, dj;
    if (A.side < 2) { i = A.side == IMIN ? 0 : ni; j = t; di = 1; dj = 0; }
    else { i = t; j = A.side == JMIN ? 0 : nj; di = 0; dj = 1; }
    Prim<double> ll = wall_cell(A, WB.kind, WB.state, WB.cons, i - 2 * di, j - 2 * dj);
    Prim<double> l0 = wall_cell(A, WB.kind, WB.state, WB.cons, i - di, j - dj);
    Prim<double> r0 = wall_cell(A, WB.kind, WB.state, WB.cons, i, j);
    Prim<double> rr = wall_cell(A, WB.kind, WB.state, WB.cons, i + di, j + dj);
    const double* geo = A.face_geo + 5 * t;
    const double nx = geo[2], ny = geo[3], area = geo[4];
    const int wall = A.kind == WALL ? ((A.side == IMIN || A.side == JMIN) ? -1 : 1) : 0;
    Vec4<double> f = face_flux<double, REC, FLUX>(A.limiter, ll, l0, r0, rr, nx, ny, wall, A.gamma);
    const double o = (A.side == IMIN || A.side == JMIN) ? -1.0 : 1.0;
    if (A.out != nullptr) {
      double* w = A.out + 9 * t;
      w[0] = geo[0]; w[1] = geo[1]; w[2] = o * nx; w[3] = o * ny; w[4] = area;
      w[5] = o * f.x; w[6] = o * f.y; w[7] = o * f.z; w[8] = o * f.w;
    }
    contrib = (o * f.y - A.p_inf * o * nx) * area;
  }
  if (A.partials != nullptr) {
    double s = block_sum<256>(contrib, red);
    if (threadIdx.x == 0) A.partials[slot] = s;
  }
}

template <typename S, int REC, int FLUX>
__global__ __launch_bounds__(256) void wall_flux_kernel(WallArgsT<S> A, WallBoundaryT<S> WB) {
  wall_body<S, REC, FLUX>(A, WB, blockIdx.x, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Fitted shock on jMax (euler2d.wgsl:402-454 compute_shock_boundary;
// shock-fitting.ts:54-113 shockState / compatibleShock, 172-197
// shockBoundary), in FP64 whatever the state's precision.
// ---------------------------------------------------------------------------

struct ShockStateD {
  Prim<double> w;
  double speed, normal_velocity;
};

// The Rankine-Hugoniot state behind a shock of outward normal (nx, ny) that
// the free stream meets at relative normal Mach number `mach`; the shock
// moves at W = mach a + q_n along n.
__device__ inline ShockStateD shock_state(const double free[4], double g, double mach, double nx, double ny) {
  const double a = sqrt(g * free[3] / free[0]);
  const double vn = free[1] * nx + free[2] * ny;
  const double speed = mach * a + vn;
  const double m2 = mach * mach;
  const double density_ratio = ((g + 1.0) * m2) / ((g - 1.0) * m2 + 2.0);
  const double pressure_ratio = 1.0 + ((2.0 * g) / (g + 1.0)) * (m2 - 1.0);
  const double un = speed - (mach * a) / density_ratio;
  return {{free[0] * density_ratio, free[1] + (un - vn) * nx, free[2] + (un - vn) * ny, free[3] * pressure_ratio},
          speed,
          un};
}

template <typename T>
struct ShockArgsT {
  int ni, nj;
  StateDevT<T> s;
  const double* face_geo;  // jMax: x, y, nx, ny, area per face (normal towards increasing j)
  T* cons;                 // jMax face states, conserved (4 per face)
  T* flux;                 // jMax imposed fluxes per unit area (4 per face)
  double* out;             // per face: speed, relative normal Mach number, |u| + a of the cell below
  double free[4];          // primitive free stream ahead of the shock
  double gamma;            // the state's gamma
};

// One thread per jMax face: the shock consistent with the u_n + a
// characteristic from the two cells below the face, extrapolated to it;
// its state fills the ghost cells, its physical flux is the face flux.
template <typename T>
__global__ __launch_bounds__(256) void shock_boundary_kernel(ShockArgsT<T> A) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.ni) return;
  const int ni = A.ni, nj = A.nj;
  const double g = A.gamma;
  const double nx = A.face_geo[5 * i + 2], ny = A.face_geo[5 * i + 3];
  const Prim<double> c1 = prim64<double>(A.s, i + (nj - 1) * ni, g);
  const Prim<double> c2 = prim64<double>(A.s, i + (nj - 2) * ni, g);
  const double a1 = sqrt(g * c1.p / c1.rho);
  const double impedance = c1.rho * a1;
  const double target = (1.5 * c1.p - 0.5 * c2.p) +
                        impedance * (1.5 * (c1.u * nx + c1.v * ny) - 0.5 * (c2.u * nx + c2.v * ny));
  auto mismatch = [&](double m) {
    const ShockStateD st = shock_state(A.free, g, m, nx, ny);
    return st.w.p + impedance * st.normal_velocity - target;
  };
  double mach = 1.0;
  if (mismatch(1.0) < 0.0) {
    double low = 1.0, high = 2.0;
    // Both terms grow with M, so the doubling ends; 1000 doublings would
    // overflow first (a non-physical cell below the face).
    for (int k = 0; k < 1000 && mismatch(high) < 0.0; k++) high *= 2.0;
    for (int k = 0; k < 100 && high - low > 1e-15 * high; k++) {
      const double mid = 0.5 * (low + high);
      if (mismatch(mid) < 0.0) low = mid;
      else high = mid;
    }
    mach = 0.5 * (low + high);
  }
  const ShockStateD sh = shock_state(A.free, g, mach, nx, ny);
  const Prim<double> w = sh.w;
  const double energy = w.p / (g - 1.0) + 0.5 * w.rho * (w.u * w.u + w.v * w.v);
  const double un = w.u * nx + w.v * ny;
  T* c = A.cons + 4 * i;
  c[0] = (T)w.rho;
  c[1] = (T)(w.rho * w.u);
  c[2] = (T)(w.rho * w.v);
  c[3] = (T)energy;
  T* f = A.flux + 4 * i;
  f[0] = (T)(w.rho * un);
  f[1] = (T)(w.rho * w.u * un + w.p * nx);
  f[2] = (T)(w.rho * w.v * un + w.p * ny);
  f[3] = (T)(un * (energy + w.p));
  double* o = A.out + 3 * i;
  o[0] = sh.speed;
  o[1] = mach;
  o[2] = sqrt(c1.u * c1.u + c1.v * c1.v) + a1;
}

}  // namespace bs
