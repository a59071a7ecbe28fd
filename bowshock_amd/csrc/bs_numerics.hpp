// Scalar numerics of the Euler path for gfx950, templated on the real type:
// float for the solver's stencil kernels, double for the wall-flux and drag
// evaluation. The formulas follow src/scheme/{gas,flux,limiters,
// reconstruction}.ts (cited per function); the code is written for the GPU:
// no allocation, structs in registers, the branch-free limiter forms where
// the reference's branches are pure selects.
#pragma once

#include <hip/hip_runtime.h>

namespace bs {

// Boundary kinds (src/engine-cpu/euler2d.ts:34-56), limiters
// (src/scheme/limiters.ts:87-94), reconstructions (reconstruction.ts:143),
// fluxes (the GPU set: euler2d-gpu.ts:57-60).
enum BcKind : int { INFLOW = 0, OUTFLOW = 1, WALL = 2, PERIODIC = 3, FACE_STATES = 4, FAR_FIELD = 5 };
enum Limiter : int { FIRST_ORDER = 0, MINMOD = 1, VAN_LEER = 2, VAN_ALBADA = 3, MONOTONIZED_CENTRAL = 4,
                     UNLIMITED_CENTRAL = 5 };
enum Recon : int { PER_VARIABLE = 0, SHARED = 1, SHARED_SCALED = 2 };
enum Flux : int { HLLC = 0, RUSANOV = 1 };
enum Side : int { IMIN = 0, IMAX = 1, JMIN = 2, JMAX = 3 };

template <typename T> struct Prim { T rho, u, v, p; };
template <typename T> struct FaceState { T rho, un, ut, p; };
template <typename T> struct Vec4 { T x, y, z, w; };

// Reciprocal, square root and division. FP32 (the stencil kernels): the
// hardware v_rcp_f32 / v_rsq_f32 / v_sqrt_f32 (1 ulp), a / b as a * rcp(b);
// the compiler's own division and square root wrap each in range-scaling
// sequences of 6-10 instructions, and a face needs about a dozen.
// FP64 (the FP64 solver's stencils, the wall-flux and drag evaluation): the
// hardware v_rcp_f64 / v_rsq_f64 estimates refined by iteration, instead of
// the compiler's IEEE sequences (div_scale / div_fmas / div_fixup around
// the reciprocal; range scaling and class tests around the square root):
// - reciprocal: two Newton-Raphson steps r += r (1 - x r), each squaring
//   the relative error;
// - quotient: q = a r(b), then one residual correction q += r (a - b q)
//   (Markstein's last step);
// - square root: from y = rsq(x), one step of Goldschmidt's coupled
//   iteration on (g, h) = (x y, y / 2), then two Newton corrections
//   g += h (x - g g); 0 for x = 0;
// - reciprocal square root: two Newton-Raphson steps y += y (1 - x y y) / 2.
// The results are within an ulp or two of the IEEE ones for the finite,
// normal operands a physical state produces, not correctly rounded.
template <typename T> __device__ __forceinline__ T bs_sqrt(T x);
template <> __device__ __forceinline__ float bs_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double bs_sqrt<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
  g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
  return x == 0.0 ? x : g;
}
template <typename T> __device__ __forceinline__ T bs_rcp(T x);
template <> __device__ __forceinline__ float bs_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double bs_rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
template <typename T> __device__ __forceinline__ T bs_rsqrt(T x);
template <> __device__ __forceinline__ float bs_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double bs_rsqrt<double>(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = __builtin_fma(0.5 * y, __builtin_fma(-x * y, y, 1.0), y);
  return __builtin_fma(0.5 * y, __builtin_fma(-x * y, y, 1.0), y);
}
template <typename T> __device__ __forceinline__ T bs_div(T a, T b);
template <> __device__ __forceinline__ float bs_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ __forceinline__ double bs_div<double>(double a, double b) {
  const double r = bs_rcp(b);
  const double q = a * r;
  return __builtin_fma(r, __builtin_fma(-b, q, a), q);
}
template <typename T> __device__ __forceinline__ T bs_abs(T x) { return x < T(0) ? -x : x; }
template <typename T> __device__ __forceinline__ T bs_min(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T bs_max(T a, T b) { return a > b ? a : b; }
// Minimum that ignores a NaN operand (IEEE minNum), for reductions.
__device__ __forceinline__ float bs_fmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double bs_fmin(double a, double b) { return fmin(a, b); }

// gas.ts:43-51 pressureFromConserved2D; gas.wgsl to_primitive_2d.
template <typename T>
__device__ __forceinline__ Prim<T> to_prim(T rho, T mu, T mv, T e, T g) {
  T inv = bs_rcp(rho);
  Prim<T> w;
  w.rho = rho;
  w.u = mu * inv;
  w.v = mv * inv;
  w.p = (g - T(1)) * (e - T(0.5) * (mu * w.u + mv * w.v));
  return w;
}

// euler2d.ts:413-416: conserved from primitive.
template <typename T>
__device__ __forceinline__ Vec4<T> to_cons(Prim<T> w, T g) {
  return {w.rho, w.rho * w.u, w.rho * w.v, bs_div(w.p, g - T(1)) + T(0.5) * w.rho * (w.u * w.u + w.v * w.v)};
}

// minmod(a, b): 0 when the signs differ, else the argument of smaller
// magnitude, which is the median of (a, b, 0); in FP32 one v_med3_f32.
template <typename T> __device__ __forceinline__ T bs_minmod(T a, T b) {
  return a * b <= T(0) ? T(0) : (bs_abs(a) < bs_abs(b) ? a : b);
}
template <> __device__ __forceinline__ float bs_minmod<float>(float a, float b) {
  return __builtin_amdgcn_fmed3f(a, b, 0.f);
}

// limiters.ts:24-94.
template <typename T>
__device__ __forceinline__ T limit(int kind, T a, T b) {
  switch (kind) {
    case FIRST_ORDER: return T(0);
    case MINMOD: return bs_minmod(a, b);
    case VAN_LEER: return a * b <= T(0) ? T(0) : bs_div(T(2) * a * b, a + b);
    case VAN_ALBADA: {
      const T e = T(1e-6);
      return bs_div(a * (b * b + e) + b * (a * a + e), a * a + b * b + T(2) * e);
    }
    case MONOTONIZED_CENTRAL: {
      if (a * b <= T(0)) return T(0);
      T m = bs_min(bs_min(T(2) * bs_abs(a), T(2) * bs_abs(b)), T(0.5) * bs_abs(a + b));
      return a > T(0) ? m : -m;
    }
    default: return T(0.5) * (a + b);
  }
}

// reconstruction.ts:70-103 (shared), 110-151 (shared scaled): the one
// limiter coefficient of the centre cell c between neighbours p and n,
// which its two faces in that direction share (shared_apply).
template <typename T, int REC>
__device__ __forceinline__ T shared_coeff(const Prim<T>& p, const Prim<T>& c, const Prim<T>& n, T g) {
  T ar = c.rho - p.rho, au = c.u - p.u, av = c.v - p.v, ap = c.p - p.p;
  T br = n.rho - c.rho, bu = n.u - c.u, bv = n.v - c.v, bp = n.p - c.p;
  if (REC == SHARED_SCALED) {
    T rho = bs_min(bs_min(p.rho, c.rho), n.rho);
    T rhoa2 = g * bs_min(bs_min(p.p, c.p), n.p);
    T sr = bs_rcp(rho), sp = bs_rcp(rhoa2), su = bs_sqrt(rho * sp);
    ar *= sr; au *= su; av *= su; ap *= sp;
    br *= sr; bu *= su; bv *= su; bp *= sp;
  }
  const T e = T(1e-6);
  T ab = ar * br + au * bu + av * bv + ap * bp;
  T aa = ar * ar + au * au + av * av + ap * ap;
  T bb = br * br + bu * bu + bv * bv + bp * bp;
  T f = bs_div(T(2) * (ab + e), aa + bb + T(2) * e);
  return f < T(0) ? T(0) : f;
}

// The shared reconstructions' state of cell c at its +1 (side > 0) or -1
// face from its coefficient f.
template <typename T>
__device__ __forceinline__ Prim<T> shared_apply(const Prim<T>& p, const Prim<T>& c, const Prim<T>& n, T side, T f) {
  Prim<T> o;
  T k = T(0.25) * side * f;
  o.rho = c.rho + k * (n.rho - p.rho);
  o.u = c.u + k * (n.u - p.u);
  o.v = c.v + k * (n.v - p.v);
  o.p = c.p + k * (n.p - p.p);
  return o;
}

// reconstruction.ts:50-64 (per variable), 70-103 (shared), 110-151 (shared
// scaled): the centre cell's state at its +1 (side > 0) or -1 face.
template <typename T, int REC>
__device__ __forceinline__ Prim<T> reconstruct(int lim, const Prim<T>& p, const Prim<T>& c, const Prim<T>& n,
                                               T side, T g) {
  if (REC == PER_VARIABLE) {
    Prim<T> o;
    T h = T(0.5) * side;
    o.rho = c.rho + h * limit<T>(lim, c.rho - p.rho, n.rho - c.rho);
    o.u = c.u + h * limit<T>(lim, c.u - p.u, n.u - c.u);
    o.v = c.v + h * limit<T>(lim, c.v - p.v, n.v - c.v);
    o.p = c.p + h * limit<T>(lim, c.p - p.p, n.p - c.p);
    return o;
  }
  return shared_apply(p, c, n, side, shared_coeff<T, REC>(p, c, n, g));
}

// flux.ts:34-36 energyOf.
template <typename T>
__device__ __forceinline__ T energy_of(const FaceState<T>& w, T g) {
  return bs_div(w.p, g - T(1)) + T(0.5) * w.rho * (w.un * w.un + w.ut * w.ut);
}

// flux.ts:39-45 physicalFaceFlux.
template <typename T>
__device__ __forceinline__ Vec4<T> physical_flux(const FaceState<T>& w, T e, T g) {
  T mass = w.rho * w.un;
  return {mass, mass * w.un + w.p, mass * w.ut, w.un * (e + w.p)};
}

// flux.ts:164-205 hllcFaceFlux with Einfeldt's wave speeds (flux.ts:87-107).
template <typename T>
__device__ __forceinline__ Vec4<T> hllc(const FaceState<T>& l, const FaceState<T>& r, T g) {
  T el = energy_of(l, g), er = energy_of(r, g);
  // 1 / rho and sqrt(rho) from one reciprocal square root per side.
  T sl_ = bs_rsqrt(l.rho), sr_ = bs_rsqrt(r.rho);
  T il = sl_ * sl_, ir = sr_ * sr_;
  T wl = l.rho * sl_, wr = r.rho * sr_;
  T al = bs_sqrt(g * l.p * il), ar = bs_sqrt(g * r.p * ir);
  T iw = bs_rcp(wl + wr);
  T un = (wl * l.un + wr * r.un) * iw;
  T ut = (wl * l.ut + wr * r.ut) * iw;
  T hl = (el + l.p) * il, hr = (er + r.p) * ir;
  T h = (wl * hl + wr * hr) * iw;
  T a = bs_sqrt(bs_max((g - T(1)) * (h - T(0.5) * (un * un + ut * ut)), T(0)));
  T sl = bs_min(l.un - al, un - a);
  T sr = bs_max(r.un + ar, un + a);
  if (sl >= T(0)) return physical_flux(l, el, g);
  if (sr <= T(0)) return physical_flux(r, er, g);
  T ml = l.rho * (sl - l.un);
  T mr = r.rho * (sr - r.un);
  T ss = bs_div(r.p - l.p + l.un * ml - r.un * mr, ml - mr);
  bool left = ss >= T(0);
  FaceState<T> k = left ? l : r;
  T sk = left ? sl : sr;
  T ek = left ? el : er;
  T ik = left ? il : ir;
  T d = sk - k.un;
  T srho = bs_div(k.rho * d, sk - ss);
  T se = srho * (ek * ik + (ss - k.un) * (ss + bs_div(k.p, k.rho * d)));
  T mass = k.rho * k.un;
  return {mass + sk * (srho - k.rho), mass * k.un + k.p + sk * (srho * ss - mass),
          mass * k.ut + sk * (srho - k.rho) * k.ut, k.un * (ek + k.p) + sk * (se - ek)};
}

// flux.ts:55-78 rusanovFaceFlux.
template <typename T>
__device__ __forceinline__ Vec4<T> rusanov(const FaceState<T>& l, const FaceState<T>& r, T g) {
  T s = bs_max(bs_abs(l.un) + bs_sqrt(bs_div(g * l.p, l.rho)), bs_abs(r.un) + bs_sqrt(bs_div(g * r.p, r.rho)));
  T el = energy_of(l, g), er = energy_of(r, g);
  T ml = l.rho * l.un, mr = r.rho * r.un;
  return {T(0.5) * (ml + mr) - T(0.5) * s * (r.rho - l.rho),
          T(0.5) * (ml * l.un + l.p + mr * r.un + r.p) - T(0.5) * s * (mr - ml),
          T(0.5) * (ml * l.ut + mr * r.ut) - T(0.5) * s * (r.rho * r.ut - l.rho * l.ut),
          T(0.5) * (l.un * (el + l.p) + r.un * (er + r.p)) - T(0.5) * s * (er - el)};
}

template <typename T, int FLUX>
__device__ __forceinline__ Vec4<T> frame_flux(const FaceState<T>& l, const FaceState<T>& r, T g) {
  if (FLUX == RUSANOV) return rusanov(l, r, g);
  return hllc(l, r, g);
}

template <typename T>
__device__ __forceinline__ FaceState<T> rotate_in(const Prim<T>& w, T nx, T ny) {
  return {w.rho, w.u * nx + w.v * ny, -w.u * ny + w.v * nx, w.p};
}

// The flux per unit area through a face of unit normal (nx, ny) between
// reconstructed states: euler2d.ts:1303-1320 faceStates (first-order
// fallback), 1332-1346 wallMirrored, flux.ts:342-366 faceFlux2D.
// `wall` is -1 for a wall face on a minimum side, +1 on a maximum side,
// else 0.
template <typename T, int REC, int FLUX>
__device__ __forceinline__ Vec4<T> face_flux_of(int lim, Prim<T> left, Prim<T> right, const Prim<T>& l0,
                                                const Prim<T>& r0, T nx, T ny, int wall, T g);

template <typename T, int REC, int FLUX>
__device__ __forceinline__ Vec4<T> face_flux(int lim, const Prim<T>& ll, const Prim<T>& l0, const Prim<T>& r0,
                                             const Prim<T>& rr, T nx, T ny, int wall, T g) {
  return face_flux_of<T, REC, FLUX>(lim, reconstruct<T, REC>(lim, ll, l0, r0, T(1), g),
                                    reconstruct<T, REC>(lim, l0, r0, rr, T(-1), g), l0, r0, nx, ny, wall, g);
}

// face_flux with the shared reconstructions' coefficients of the two cells
// (fl of l0, fr of r0) already formed.
template <typename T, int REC, int FLUX>
__device__ __forceinline__ Vec4<T> face_flux_coeff(const Prim<T>& ll, const Prim<T>& l0, const Prim<T>& r0,
                                                   const Prim<T>& rr, T fl, T fr, T nx, T ny, int wall, T g) {
  static_assert(REC != PER_VARIABLE, "per-variable reconstructions have no shared coefficient");
  return face_flux_of<T, REC, FLUX>(0, shared_apply(ll, l0, r0, T(1), fl), shared_apply(l0, r0, rr, T(-1), fr), l0,
                                    r0, nx, ny, wall, g);
}

// The flux from the two reconstructed face states.
template <typename T, int REC, int FLUX>
__device__ __forceinline__ Vec4<T> face_flux_of(int lim, Prim<T> left, Prim<T> right, const Prim<T>& l0,
                                                const Prim<T>& r0, T nx, T ny, int wall, T g) {
  // Per variable with minmod (or first order) the face value of a positive
  // quantity lies between the cell's value and the mean of the cell and its
  // neighbour, so positive density and pressure stay positive and the
  // first-order fallback never triggers: the test is skipped there.
  const bool keeps_positive = REC == PER_VARIABLE && (lim == MINMOD || lim == FIRST_ORDER);
  if (!keeps_positive && !(left.rho > T(0) && left.p > T(0) && right.rho > T(0) && right.p > T(0))) {
    left = l0;
    right = r0;
  }
  FaceState<T> fl = rotate_in(left, nx, ny);
  FaceState<T> fr = rotate_in(right, nx, ny);
  // The mirror of a state about the face is (rho, -un, ut, p) in the face frame.
  if (wall < 0) fl = {fr.rho, -fr.un, fr.ut, fr.p};
  else if (wall > 0) fr = {fl.rho, -fl.un, fl.ut, fl.p};
  Vec4<T> f = frame_flux<T, FLUX>(fl, fr, g);
  return {f.x, f.y * nx - f.z * ny, f.y * ny + f.z * nx, f.w};
}

// euler2d.ts:139-165 farFieldState.
template <typename T>
__device__ __forceinline__ Prim<T> far_field_state(const Prim<T>& in, const Prim<T>& fr, T nx, T ny, T g) {
  T k = T(2) / (g - T(1));
  T ai = bs_sqrt(g * in.p / in.rho), af = bs_sqrt(g * fr.p / fr.rho);
  T uni = in.u * nx + in.v * ny, unf = fr.u * nx + fr.v * ny;
  T plus = uni + ai > T(0) ? uni + k * ai : unf + k * af;
  T minus = uni - ai > T(0) ? uni - k * ai : unf - k * af;
  T un = T(0.5) * (plus + minus);
  T a = (plus - minus) / (T(2) * k);
  Prim<T> up = uni > T(0) ? in : fr;
  T entropy = up.p / pow(up.rho, g);
  T unu = up.u * nx + up.v * ny;
  T rho = pow(a * a / (g * entropy), T(1) / (g - T(1)));
  return {rho, up.u + (un - unu) * nx, up.v + (un - unu) * ny, rho * a * a / g};
}

}  // namespace bs
