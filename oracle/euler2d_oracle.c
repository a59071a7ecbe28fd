/*
 * FP64 CPU oracle of the axisymmetric finite-volume Euler path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path and the CPU
 * baseline of the benchmark, never part of the product. Each function names
 * the reference file and lines whose algorithm it restates; the code is
 * written for this library (plain C, double precision, ghost cells computed
 * on demand instead of stored in padded arrays).
 *
 * Parity pins (tests/test_oracle_*.py): free-stream preservation on
 * distorted grids, closed-form metrics and time steps, flux consistency and
 * invariances, limiter bounds, Taylor-Maccoll cone pressure, the
 * reference's golden cone drags, and the equivalence of every implicit
 * relaxation with the explicit steady state.
 */
#include "euler2d_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char last_error[512];

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(last_error, sizeof last_error, fmt, ap);
  va_end(ap);
  return code;
}

const char *orc_last_error(void) { return last_error; }

typedef struct { double rho, u, v, p; } prim_t;
typedef struct { double rho, un, ut, p; } face_t;
typedef struct { double q[4]; } cons_t;

struct orc_solver {
  orc_config cfg;
  int ni, nj;
  double *x, *y;
  /* metrics */
  double *inx, *iny, *iarea, *jnx, *jny, *jarea, *volume, *planar, *cx, *cy;
  double *source;
  /* state (interleaved conserved) and stage buffers */
  double *u, *stage[2], *res, *dt;
  /* face states / imposed fluxes per side (NULL if none) */
  double *face_states[4];
  double *face_flux[4];
  double time;
  long long steps;
  /* implicit work */
  double *lam_i, *lam_j, *diag, *dq, *pc, *ac, *prev;
};

/* ---------------------------------------------------------------------- */
/* Gas and scalar numerics                                                 */
/* ---------------------------------------------------------------------- */

/* gas.ts:43-51 pressureFromConserved2D */
static double pressure_of(const double *c, double g) {
  return (g - 1.0) * (c[3] - (0.5 * (c[1] * c[1] + c[2] * c[2])) / c[0]);
}

/* euler2d.ts:1348-1358 primitiveAt */
static prim_t to_prim(const double *c, double g) {
  prim_t w;
  w.rho = c[0];
  w.u = c[1] / c[0];
  w.v = c[2] / c[0];
  w.p = pressure_of(c, g);
  return w;
}

/* euler2d.ts:413-416 (conserved from primitive) */
static cons_t to_cons(prim_t w, double g) {
  cons_t c;
  c.q[0] = w.rho;
  c.q[1] = w.rho * w.u;
  c.q[2] = w.rho * w.v;
  c.q[3] = w.p / (g - 1.0) + 0.5 * w.rho * (w.u * w.u + w.v * w.v);
  return c;
}

/* gas.ts:64-66 */
static double sound(double rho, double p, double g) { return sqrt((g * p) / rho); }

/* flux.ts:34-36 */
static double energy_of(face_t w, double g) {
  return w.p / (g - 1.0) + 0.5 * w.rho * (w.un * w.un + w.ut * w.ut);
}

/* flux.ts:39-45 physicalFaceFlux */
static void physical_flux(face_t w, double g, double *out) {
  double mass = w.rho * w.un;
  out[0] = mass;
  out[1] = mass * w.un + w.p;
  out[2] = mass * w.ut;
  out[3] = w.un * (energy_of(w, g) + w.p);
}

/* flux.ts:55-78 rusanovFaceFlux */
static void rusanov(face_t l, face_t r, double g, double *out) {
  double sl = fabs(l.un) + sound(l.rho, l.p, g);
  double sr = fabs(r.un) + sound(r.rho, r.p, g);
  double s = sl > sr ? sl : sr;
  double el = energy_of(l, g), er = energy_of(r, g);
  double ml = l.rho * l.un, mr = r.rho * r.un;
  out[0] = 0.5 * (ml + mr) - 0.5 * s * (r.rho - l.rho);
  out[1] = 0.5 * (ml * l.un + l.p + mr * r.un + r.p) - 0.5 * s * (mr - ml);
  out[2] = 0.5 * (ml * l.ut + mr * r.ut) - 0.5 * s * (r.rho * r.ut - l.rho * l.ut);
  out[3] = 0.5 * (l.un * (el + l.p) + r.un * (er + r.p)) - 0.5 * s * (er - el);
}

/* flux.ts:87-107 einfeldtSpeeds */
static void einfeldt(face_t l, face_t r, double g, double *sl, double *sr) {
  double al = sound(l.rho, l.p, g), ar = sound(r.rho, r.p, g);
  double wl = sqrt(l.rho), wr = sqrt(r.rho);
  double un = (wl * l.un + wr * r.un) / (wl + wr);
  double ut = (wl * l.ut + wr * r.ut) / (wl + wr);
  double hl = (energy_of(l, g) + l.p) / l.rho;
  double hr = (energy_of(r, g) + r.p) / r.rho;
  double h = (wl * hl + wr * hr) / (wl + wr);
  double a2 = (g - 1.0) * (h - 0.5 * (un * un + ut * ut));
  double a = sqrt(a2 > 0.0 ? a2 : 0.0);
  double x = l.un - al, y = un - a;
  *sl = x < y ? x : y;
  x = r.un + ar;
  y = un + a;
  *sr = x > y ? x : y;
}

/* flux.ts:116-147 hllFaceFlux */
static void hll(face_t l, face_t r, double g, double *out) {
  double sl, sr;
  einfeldt(l, r, g, &sl, &sr);
  if (sl >= 0.0) { physical_flux(l, g, out); return; }
  if (sr <= 0.0) { physical_flux(r, g, out); return; }
  double el = energy_of(l, g), er = energy_of(r, g);
  double ml = l.rho * l.un, mr = r.rho * r.un;
  double fl[4] = {ml, ml * l.un + l.p, ml * l.ut, l.un * (el + l.p)};
  double fr[4] = {mr, mr * r.un + r.p, mr * r.ut, r.un * (er + r.p)};
  double ul[4] = {l.rho, ml, l.rho * l.ut, el};
  double ur[4] = {r.rho, mr, r.rho * r.ut, er};
  for (int k = 0; k < 4; k++)
    out[k] = (sr * fl[k] - sl * fr[k] + sl * sr * (ur[k] - ul[k])) / (sr - sl);
}

/* flux.ts:164-205 hllcFaceFlux */
static void hllc(face_t l, face_t r, double g, double *out) {
  double el = energy_of(l, g), er = energy_of(r, g);
  double sl, sr;
  einfeldt(l, r, g, &sl, &sr);
  if (sl >= 0.0) { physical_flux(l, g, out); return; }
  if (sr <= 0.0) { physical_flux(r, g, out); return; }
  double ml = l.rho * (sl - l.un);
  double mr = r.rho * (sr - r.un);
  double ss = (r.p - l.p + l.un * ml - r.un * mr) / (ml - mr);
  int left = ss >= 0.0;
  face_t k = left ? l : r;
  double sk = left ? sl : sr;
  double ek = left ? el : er;
  double srho = (k.rho * (sk - k.un)) / (sk - ss);
  double se = srho * (ek / k.rho + (ss - k.un) * (ss + k.p / (k.rho * (sk - k.un))));
  double mass = k.rho * k.un;
  out[0] = mass + sk * (srho - k.rho);
  out[1] = mass * k.un + k.p + sk * (srho * ss - mass);
  out[2] = mass * k.ut + sk * (srho - k.rho) * k.ut;
  out[3] = k.un * (ek + k.p) + sk * (se - ek);
}

static void frame_flux(int flux, face_t l, face_t r, double g, double *out) {
  if (flux == ORC_RUSANOV) rusanov(l, r, g, out);
  else if (flux == ORC_HLL) hll(l, r, g, out);
  else hllc(l, r, g, out);
}

static face_t rotate_in(prim_t w, double nx, double ny) {
  face_t f;
  f.rho = w.rho;
  f.un = w.u * nx + w.v * ny;
  f.ut = -w.u * ny + w.v * nx;
  f.p = w.p;
  return f;
}

/* flux.ts:342-366 faceFlux2D */
static void flux_2d(int flux, prim_t l, prim_t r, double nx, double ny, double g, double *out) {
  frame_flux(flux, rotate_in(l, nx, ny), rotate_in(r, nx, ny), g, out);
  double n = out[1], t = out[2];
  out[1] = n * nx - t * ny;
  out[2] = n * ny + t * nx;
}

#define VAN_ALBADA_EPSILON 1e-6 /* limiters.ts:51 */

/* limiters.ts:24-94 */
static double limiter(int kind, double a, double b) {
  switch (kind) {
    case ORC_FIRST_ORDER:
      return 0.0;
    case ORC_MINMOD:
      if (a * b <= 0.0) return 0.0;
      return fabs(a) < fabs(b) ? a : b;
    case ORC_VAN_LEER:
      if (a * b <= 0.0) return 0.0;
      return (2.0 * a * b) / (a + b);
    case ORC_VAN_ALBADA: {
      double e = VAN_ALBADA_EPSILON;
      return (a * (b * b + e) + b * (a * a + e)) / (a * a + b * b + 2.0 * e);
    }
    case ORC_MONOTONIZED_CENTRAL: {
      if (a * b <= 0.0) return 0.0;
      double s = a > 0.0 ? 1.0 : -1.0;
      double m = 2.0 * fabs(a);
      if (2.0 * fabs(b) < m) m = 2.0 * fabs(b);
      if (0.5 * fabs(a + b) < m) m = 0.5 * fabs(a + b);
      return s * m;
    }
    default:
      return 0.5 * (a + b);
  }
}

/* reconstruction.ts:50-64 reconstructPerVariable, 70-103 shared,
 * 110-151 sharedScaled. side is +1 or -1. */
static prim_t reconstruct(int rec, int lim, prim_t p, prim_t c, prim_t n, double side,
                          double g) {
  prim_t o;
  if (rec == ORC_PER_VARIABLE) {
    o.rho = c.rho + 0.5 * side * limiter(lim, c.rho - p.rho, n.rho - c.rho);
    o.u = c.u + 0.5 * side * limiter(lim, c.u - p.u, n.u - c.u);
    o.v = c.v + 0.5 * side * limiter(lim, c.v - p.v, n.v - c.v);
    o.p = c.p + 0.5 * side * limiter(lim, c.p - p.p, n.p - c.p);
    return o;
  }
  double sr = 1.0, su = 1.0, sp = 1.0;
  if (rec == ORC_SHARED_SCALED) {
    double rho = p.rho;
    if (c.rho < rho) rho = c.rho;
    if (n.rho < rho) rho = n.rho;
    double pm = p.p;
    if (c.p < pm) pm = c.p;
    if (n.p < pm) pm = n.p;
    double rhoa2 = g * pm;
    sr = 1.0 / rho;
    su = sqrt(rho / rhoa2);
    sp = 1.0 / rhoa2;
  }
  double ar = c.rho - p.rho, au = c.u - p.u, av = c.v - p.v, ap = c.p - p.p;
  double br = n.rho - c.rho, bu = n.u - c.u, bv = n.v - c.v, bp = n.p - c.p;
  if (rec == ORC_SHARED_SCALED) {
    ar *= sr; au *= su; av *= su; ap *= sp;
    br *= sr; bu *= su; bv *= su; bp *= sp;
  }
  double ab = ar * br + au * bu + av * bv + ap * bp;
  double aa = ar * ar + au * au + av * av + ap * ap;
  double bb = br * br + bu * bu + bv * bv + bp * bp;
  double e = VAN_ALBADA_EPSILON;
  double f = (2.0 * (ab + e)) / (aa + bb + 2.0 * e);
  if (f < 0.0) f = 0.0; /* Math.max(0, f), a NaN kept as NaN */
  double k = 0.25 * side * f;
  o.rho = c.rho + k * (n.rho - p.rho);
  o.u = c.u + k * (n.u - p.u);
  o.v = c.v + k * (n.v - p.v);
  o.p = c.p + k * (n.p - p.p);
  return o;
}

/* euler2d.ts:139-165 farFieldState */
static prim_t far_field(prim_t in, prim_t fr, double nx, double ny, double g) {
  double k = 2.0 / (g - 1.0);
  double ai = sqrt((g * in.p) / in.rho);
  double af = sqrt((g * fr.p) / fr.rho);
  double uni = in.u * nx + in.v * ny;
  double unf = fr.u * nx + fr.v * ny;
  double plus = uni + ai > 0.0 ? uni + k * ai : unf + k * af;
  double minus = uni - ai > 0.0 ? uni - k * ai : unf - k * af;
  double un = 0.5 * (plus + minus);
  double a = (plus - minus) / (2.0 * k);
  prim_t up = uni > 0.0 ? in : fr;
  double entropy = up.p / pow(up.rho, g);
  double unu = up.u * nx + up.v * ny;
  double rho = pow((a * a) / (g * entropy), 1.0 / (g - 1.0));
  prim_t o;
  o.rho = rho;
  o.u = up.u + (un - unu) * nx;
  o.v = up.v + (un - unu) * ny;
  o.p = (rho * a * a) / g;
  return o;
}

/* ---------------------------------------------------------------------- */
/* Metrics                                                                 */
/* ---------------------------------------------------------------------- */

/* structured.ts:94-168 computeMetrics */
int orc_compute_metrics(int ni, int nj, const double *x, const double *y, int axisym,
                        double *inx, double *iny, double *iarea, double *jnx, double *jny,
                        double *jarea, double *volume, double *planar, double *cxo,
                        double *cyo) {
  int nn = (ni + 1) * (nj + 1);
  if (axisym)
    for (int k = 0; k < nn; k++)
      if (y[k] < 0.0) return fail(ORC_ERR_RANGE, "axisymmetric grids must have y >= 0 (y is the radius)");
#define NODE(i, j) ((i) + (j) * (ni + 1))
#define FACE(NX, NY, AR, k, a, b)                                         \
  do {                                                                    \
    double dx = x[b] - x[a], dy = y[b] - y[a];                            \
    double len = hypot(dx, dy);                                           \
    if (NX) NX[k] = len > 0.0 ? dy / len : 0.0;                           \
    if (NY) NY[k] = len > 0.0 ? -dx / len : 0.0;                          \
    if (AR) AR[k] = axisym ? len * 0.5 * (y[a] + y[b]) : len;             \
  } while (0)
  for (int j = 0; j < nj; j++)
    for (int i = 0; i <= ni; i++) FACE(inx, iny, iarea, i + j * (ni + 1), NODE(i, j), NODE(i, j + 1));
  for (int j = 0; j <= nj; j++)
    for (int i = 0; i < ni; i++) FACE(jnx, jny, jarea, i + j * ni, NODE(i + 1, j), NODE(i, j));
  for (int j = 0; j < nj; j++) {
    for (int i = 0; i < ni; i++) {
      int corners[4] = {NODE(i, j), NODE(i + 1, j), NODE(i + 1, j + 1), NODE(i, j + 1)};
      double area = 0.0, cx = 0.0, cy = 0.0;
      for (int c = 0; c < 4; c++) {
        int p = corners[c], q = corners[(c + 1) % 4];
        double cross = x[p] * y[q] - x[q] * y[p];
        area += cross;
        cx += (x[p] + x[q]) * cross;
        cy += (y[p] + y[q]) * cross;
      }
      area *= 0.5;
      if (!(area > 0.0)) return fail(ORC_ERR_RANGE, "cell (%d, %d) has non-positive area %g", i, j, area);
      int k = i + j * ni;
      double gx = cx / (6.0 * area), gy = cy / (6.0 * area);
      if (planar) planar[k] = area;
      if (cxo) cxo[k] = gx;
      if (cyo) cyo[k] = gy;
      if (volume) volume[k] = axisym ? area * gy : area;
    }
  }
#undef FACE
#undef NODE
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Solver                                                                  */
/* ---------------------------------------------------------------------- */

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* config.ts:35-55 validateGrid2DConfig */
static int validate(const orc_config *c) {
  if (!(c->cfl > 0.0 && c->cfl <= 1.0)) return fail(ORC_ERR_RANGE, "cfl must be in (0, 1]");
  if (c->ni < 2 || c->nj < 2) return fail(ORC_ERR_RANGE, "the grid needs at least 2 x 2 cells");
  if ((c->reconstruction == ORC_SHARED || c->reconstruction == ORC_SHARED_SCALED) &&
      c->limiter != ORC_VAN_ALBADA)
    return fail(ORC_ERR_RANGE, "shared reconstruction requires the 'vanAlbada' limiter");
  if ((c->bc_kind[0] == ORC_PERIODIC) != (c->bc_kind[1] == ORC_PERIODIC))
    return fail(ORC_ERR_RANGE, "periodic boundaries must be used on both iMin and iMax");
  if ((c->bc_kind[2] == ORC_PERIODIC) != (c->bc_kind[3] == ORC_PERIODIC))
    return fail(ORC_ERR_RANGE, "periodic boundaries must be used on both jMin and jMax");
  return ORC_OK;
}

void orc_destroy(orc_solver *s) {
  if (!s) return;
  double *all[] = {s->x, s->y, s->inx, s->iny, s->iarea, s->jnx, s->jny, s->jarea, s->volume,
                   s->planar, s->cx, s->cy, s->source, s->u, s->stage[0], s->stage[1], s->res,
                   s->dt, s->lam_i, s->lam_j, s->diag, s->dq, s->pc, s->ac, s->prev};
  for (size_t k = 0; k < sizeof all / sizeof all[0]; k++) free(all[k]);
  for (int k = 0; k < 4; k++) { free(s->face_states[k]); free(s->face_flux[k]); }
  free(s);
}

static int along(const orc_solver *s, int side) { return side < 2 ? s->nj : s->ni; }

int orc_create(const orc_config *cfg, orc_solver **out) {
  *out = NULL;
  int rc = validate(cfg);
  if (rc) return rc;
  orc_solver *s = (orc_solver *)calloc(1, sizeof *s);
  s->cfg = *cfg;
  int ni = cfg->ni, nj = cfg->nj;
  s->ni = ni;
  s->nj = nj;
  size_t nn = (size_t)(ni + 1) * (nj + 1), n = (size_t)ni * nj;
  s->x = dalloc(nn);
  s->y = dalloc(nn);
  memcpy(s->x, cfg->x, nn * sizeof(double));
  memcpy(s->y, cfg->y, nn * sizeof(double));
  s->cfg.x = s->x;
  s->cfg.y = s->y;
  size_t fi = (size_t)(ni + 1) * nj, fj = (size_t)ni * (nj + 1);
  s->inx = dalloc(fi); s->iny = dalloc(fi); s->iarea = dalloc(fi);
  s->jnx = dalloc(fj); s->jny = dalloc(fj); s->jarea = dalloc(fj);
  s->volume = dalloc(n); s->planar = dalloc(n); s->cx = dalloc(n); s->cy = dalloc(n);
  rc = orc_compute_metrics(ni, nj, s->x, s->y, cfg->axisymmetric, s->inx, s->iny, s->iarea,
                           s->jnx, s->jny, s->jarea, s->volume, s->planar, s->cx, s->cy);
  if (rc) { orc_destroy(s); return rc; }
  if (cfg->source) {
    s->source = dalloc(4 * n);
    memcpy(s->source, cfg->source, 4 * n * sizeof(double));
  }
  s->cfg.source = s->source;
  s->u = dalloc(4 * n);
  s->stage[0] = dalloc(4 * n);
  s->stage[1] = dalloc(4 * n);
  s->res = dalloc(4 * n);
  s->dt = dalloc(n);
  for (int side = 0; side < 4; side++) {
    if (cfg->bc_kind[side] == ORC_FACE_STATES) {
      int m = along(s, side);
      s->face_states[side] = dalloc(4 * (size_t)m);
      /* Until set, the face states are the side's bc_state. */
      for (int t = 0; t < m; t++) memcpy(s->face_states[side] + 4 * t, cfg->bc_state + 4 * side, 4 * sizeof(double));
    }
  }
  *out = s;
  return ORC_OK;
}

/* euler2d.ts:357-366 setFaceStates */
int orc_set_face_states(orc_solver *s, int side, const double *prim, int count) {
  if (side < 0 || side > 3 || s->cfg.bc_kind[side] != ORC_FACE_STATES)
    return fail(ORC_ERR_RANGE, "the boundary does not take face states");
  if (count != along(s, side)) return fail(ORC_ERR_RANGE, "the boundary needs %d face states", along(s, side));
  memcpy(s->face_states[side], prim, 4 * (size_t)count * sizeof(double));
  return ORC_OK;
}

/* euler2d.ts:376-389 setFaceFluxes */
int orc_set_face_fluxes(orc_solver *s, int side, const double *flux, int count) {
  if (side < 0 || side > 3 || s->cfg.bc_kind[side] != ORC_FACE_STATES)
    return fail(ORC_ERR_RANGE, "the boundary does not take face fluxes");
  if (!flux) { free(s->face_flux[side]); s->face_flux[side] = NULL; return ORC_OK; }
  if (count != 4 * along(s, side)) return fail(ORC_ERR_RANGE, "the boundary needs %d flux values", 4 * along(s, side));
  if (!s->face_flux[side]) s->face_flux[side] = dalloc((size_t)count);
  memcpy(s->face_flux[side], flux, (size_t)count * sizeof(double));
  return ORC_OK;
}

/* euler2d.ts:1471-1483 checkPhysical */
static int check_physical(orc_solver *s, const double *u) {
  int ni = s->ni, nj = s->nj;
  for (int j = 0; j < nj; j++)
    for (int i = 0; i < ni; i++) {
      prim_t w = to_prim(u + 4 * (i + j * ni), s->cfg.gamma);
      if (!(w.rho > 0.0 && w.p > 0.0 && isfinite(w.u) && isfinite(w.v)))
        return fail(ORC_ERR_NONPHYSICAL,
                    "non-physical state in cell (%d, %d) after step %lld: rho=%g, p=%g", i, j,
                    s->steps, w.rho, w.p);
    }
  return ORC_OK;
}

/* Unit normal of boundary face t of a side (pointing to increasing index)
 * and the sign turning it outward: euler2d.ts:1430-1469 boundaryFace. */
static void boundary_normal(const orc_solver *s, int side, int t, double *nx, double *ny,
                            double *area, int *outward) {
  int ni = s->ni, nj = s->nj;
  if (side < 2) {
    int f = (side == ORC_IMIN ? 0 : ni) + t * (ni + 1);
    *nx = s->inx[f]; *ny = s->iny[f]; *area = s->iarea[f];
    *outward = side == ORC_IMIN ? -1 : 1;
  } else {
    int f = t + (side == ORC_JMIN ? 0 : nj) * ni;
    *nx = s->jnx[f]; *ny = s->jny[f]; *area = s->jarea[f];
    *outward = side == ORC_JMIN ? -1 : 1;
  }
}

/* The conserved state of cell (i, j) of u, where (i, j) may be a ghost
 * cell two layers deep on one side (no corners): euler2d.ts:1361-1423
 * fillGhosts and ghostCells, evaluated on demand. */
static cons_t cell_state(const orc_solver *s, const double *u, int i, int j) {
  int ni = s->ni, nj = s->nj;
  double g = s->cfg.gamma;
  cons_t c;
  int side, t, layer;
  if (i >= 0 && i < ni && j >= 0 && j < nj) {
    memcpy(c.q, u + 4 * (i + j * ni), sizeof c.q);
    return c;
  }
  if (i < 0) { side = ORC_IMIN; t = j; layer = -1 - i; }
  else if (i >= ni) { side = ORC_IMAX; t = j; layer = i - ni; }
  else if (j < 0) { side = ORC_JMIN; t = i; layer = -1 - j; }
  else { side = ORC_JMAX; t = i; layer = j - nj; }
  int kind = s->cfg.bc_kind[side];
  double nx, ny, area;
  int outward;
  boundary_normal(s, side, t, &nx, &ny, &area, &outward);
  /* interior cell mirrored by this ghost (layer g mirrors interior layer g),
   * and its periodic image */
  int mi, mj, wi, wj;
  switch (side) {
    case ORC_IMIN: mi = layer; mj = t; wi = ni - 1 - layer; wj = t; break;
    case ORC_IMAX: mi = ni - 1 - layer; mj = t; wi = layer; wj = t; break;
    case ORC_JMIN: mi = t; mj = layer; wi = t; wj = nj - 1 - layer; break;
    default: mi = t; mj = nj - 1 - layer; wi = t; wj = layer; break;
  }
  if (kind == ORC_INFLOW) {
    const double *b = s->cfg.bc_state + 4 * side;
    prim_t w = {b[0], b[1], b[2], b[3]};
    return to_cons(w, g);
  }
  if (kind == ORC_FACE_STATES) {
    const double *b = s->face_states[side] + 4 * t;
    prim_t w = {b[0], b[1], b[2], b[3]};
    return to_cons(w, g);
  }
  if (kind == ORC_FAR_FIELD) {
    int ii = side == ORC_IMIN ? 0 : side == ORC_IMAX ? ni - 1 : t;
    int jj = side == ORC_JMIN ? 0 : side == ORC_JMAX ? nj - 1 : t;
    prim_t in = to_prim(u + 4 * (ii + jj * ni), g);
    const double *b = s->cfg.bc_state + 4 * side;
    prim_t fr = {b[0], b[1], b[2], b[3]};
    return to_cons(far_field(in, fr, outward * nx, outward * ny, g), g);
  }
  const double *src = kind == ORC_PERIODIC ? u + 4 * (wi + wj * ni) : u + 4 * (mi + mj * ni);
  memcpy(c.q, src, sizeof c.q);
  if (kind == ORC_WALL) {
    double mn = src[1] * nx + src[2] * ny;
    c.q[1] = src[1] - 2.0 * mn * nx;
    c.q[2] = src[2] - 2.0 * mn * ny;
  }
  return c;
}

/* euler2d.ts:1303-1320 faceStates and 1332-1346 wallMirrored: the left and
 * right states at face (i, j) of direction dir (0: i, 1: j), `side` the
 * boundary side the face lies on or -1. */
static void face_states(const orc_solver *s, const double *u, int dir, int i, int j, int side,
                        double nx, double ny, prim_t *left, prim_t *right) {
  double g = s->cfg.gamma;
  int di = dir == 0 ? 1 : 0, dj = dir == 0 ? 0 : 1;
  cons_t cll = cell_state(s, u, i - 2 * di, j - 2 * dj);
  cons_t cl = cell_state(s, u, i - di, j - dj);
  cons_t cr = cell_state(s, u, i, j);
  cons_t crr = cell_state(s, u, i + di, j + dj);
  prim_t wll = to_prim(cll.q, g), wl = to_prim(cl.q, g), wr = to_prim(cr.q, g),
         wrr = to_prim(crr.q, g);
  int rec = s->cfg.reconstruction, lim = s->cfg.limiter;
  prim_t l = reconstruct(rec, lim, wll, wl, wr, 1.0, g);
  prim_t r = reconstruct(rec, lim, wl, wr, wrr, -1.0, g);
  if (!(l.rho > 0.0 && l.p > 0.0 && r.rho > 0.0 && r.p > 0.0)) { l = wl; r = wr; }
  if (side >= 0 && s->cfg.bc_kind[side] == ORC_WALL) {
    int at_min = side == ORC_IMIN || side == ORC_JMIN;
    prim_t in = at_min ? r : l;
    double un = in.u * nx + in.v * ny;
    prim_t m = {in.rho, in.u - 2.0 * un * nx, in.v - 2.0 * un * ny, in.p};
    if (at_min) { l = m; r = in; } else { l = in; r = m; }
  }
  *left = l;
  *right = r;
}

/* euler2d.ts:1208-1296 computeResidual: r (interleaved, 4 per cell) for
 * state u; net inflow times area plus source and axisymmetric p A_p. */
static void compute_residual(orc_solver *s, const double *u, double *r) {
  int ni = s->ni, nj = s->nj;
  double g = s->cfg.gamma;
  size_t n = (size_t)ni * nj;
  memset(r, 0, 4 * n * sizeof(double));
  double f[4];
  for (int j = 0; j < nj; j++) {
    for (int i = 0; i <= ni; i++) {
      int fi = i + j * (ni + 1);
      int side = i == 0 ? ORC_IMIN : i == ni ? ORC_IMAX : -1;
      if (side >= 0 && s->face_flux[side]) {
        memcpy(f, s->face_flux[side] + 4 * j, sizeof f);
      } else {
        prim_t l, rr;
        face_states(s, u, 0, i, j, side, s->inx[fi], s->iny[fi], &l, &rr);
        flux_2d(s->cfg.flux, l, rr, s->inx[fi], s->iny[fi], g, f);
      }
      double a = s->iarea[fi];
      if (i > 0) for (int k = 0; k < 4; k++) r[4 * (i - 1 + j * ni) + k] -= f[k] * a;
      if (i < ni) for (int k = 0; k < 4; k++) r[4 * (i + j * ni) + k] += f[k] * a;
    }
  }
  for (int j = 0; j <= nj; j++) {
    for (int i = 0; i < ni; i++) {
      int fj = i + j * ni;
      int side = j == 0 ? ORC_JMIN : j == nj ? ORC_JMAX : -1;
      if (side >= 0 && s->face_flux[side]) {
        memcpy(f, s->face_flux[side] + 4 * i, sizeof f);
      } else {
        prim_t l, rr;
        face_states(s, u, 1, i, j, side, s->jnx[fj], s->jny[fj], &l, &rr);
        flux_2d(s->cfg.flux, l, rr, s->jnx[fj], s->jny[fj], g, f);
      }
      double a = s->jarea[fj];
      if (j > 0) for (int k = 0; k < 4; k++) r[4 * (i + (j - 1) * ni) + k] -= f[k] * a;
      if (j < nj) for (int k = 0; k < 4; k++) r[4 * (i + j * ni) + k] += f[k] * a;
    }
  }
  if (s->source)
    for (size_t c = 0; c < 4 * n; c++) r[c] += s->source[c];
  if (s->cfg.axisymmetric)
    for (size_t c = 0; c < n; c++) r[4 * c + 2] += pressure_of(u + 4 * c, g) * s->planar[c];
}

/* euler2d.ts:1183-1205 computeTimeSteps */
static void compute_time_steps(orc_solver *s, double limit) {
  int ni = s->ni, nj = s->nj;
  double g = s->cfg.gamma, cfl = s->cfg.cfl;
  double smallest = INFINITY;
  for (int j = 0; j < nj; j++) {
    for (int i = 0; i < ni; i++) {
      int c = i + j * ni;
      prim_t w = to_prim(s->u + 4 * c, g);
      double a = sound(w.rho, w.p, g);
      int f0 = i + j * (ni + 1), f1 = f0 + 1;
      double six = 0.5 * (s->inx[f0] * s->iarea[f0] + s->inx[f1] * s->iarea[f1]);
      double siy = 0.5 * (s->iny[f0] * s->iarea[f0] + s->iny[f1] * s->iarea[f1]);
      int g0 = c, g1 = c + ni;
      double sjx = 0.5 * (s->jnx[g0] * s->jarea[g0] + s->jnx[g1] * s->jarea[g1]);
      double sjy = 0.5 * (s->jny[g0] * s->jarea[g0] + s->jny[g1] * s->jarea[g1]);
      double lam = fabs(w.u * six + w.v * siy) + a * hypot(six, siy) +
                   (fabs(w.u * sjx + w.v * sjy) + a * hypot(sjx, sjy));
      double dt = (cfl * s->volume[c]) / lam;
      s->dt[c] = dt;
      if (dt < smallest) smallest = dt;
    }
  }
  if (!s->cfg.local_stepping) {
    double d = smallest < limit ? smallest : limit;
    for (int c = 0; c < ni * nj; c++) s->dt[c] = d;
  }
}

/* runge-kutta.ts:17-28 */
static int rk_stages(int integrator, const double **a, const double **b) {
  static const double ea[] = {0.0}, eb[] = {1.0};
  static const double a2[] = {0.0, 0.5}, b2[] = {1.0, 0.5};
  static const double a3[] = {0.0, 0.75, 1.0 / 3.0}, b3[] = {1.0, 0.25, 2.0 / 3.0};
  if (integrator == ORC_SSPRK2) { *a = a2; *b = b2; return 2; }
  if (integrator == ORC_SSPRK3) { *a = a3; *b = b3; return 3; }
  *a = ea; *b = eb;
  return 1;
}

/* euler2d.ts:1146-1175 step and stage */
static int step(orc_solver *s) {
  const double *a, *b;
  int ns = rk_stages(s->cfg.integrator, &a, &b);
  size_t n = (size_t)s->ni * s->nj;
  const double *prev = s->u;
  for (int k = 0; k < ns; k++) {
    double *out = k == ns - 1 ? s->u : s->stage[k % 2];
    compute_residual(s, prev, s->res);
    for (size_t c = 0; c < n; c++) {
      double factor = s->dt[c] / s->volume[c];
      for (int q = 0; q < 4; q++)
        out[4 * c + q] = a[k] * s->u[4 * c + q] + b[k] * (prev[4 * c + q] + factor * s->res[4 * c + q]);
    }
    prev = out;
  }
  s->steps += 1;
  return check_physical(s, s->u);
}

/* euler2d.ts:554-560 iterate */
int orc_iterate(orc_solver *s, int count) {
  for (int n = 0; n < count; n++) {
    compute_time_steps(s, INFINITY);
    if (!s->cfg.local_stepping) s->time += s->dt[0];
    int rc = step(s);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* euler2d.ts:536-551 advanceTo */
int orc_advance_to(orc_solver *s, double t_end, int max_steps) {
  if (s->cfg.local_stepping)
    return fail(ORC_ERR_RANGE, "advanceTo needs global time stepping; use iterate() for local stepping");
  long long first = s->steps;
  while (s->time < t_end) {
    if (s->steps - first >= max_steps)
      return fail(ORC_ERR_RANGE, "advanceTo(%g) needed more than %d steps", t_end, max_steps);
    compute_time_steps(s, t_end - s->time);
    double dt = s->dt[0];
    int rc = step(s);
    if (rc) return rc;
    s->time += dt;
    double m = fabs(t_end) > 1.0 ? fabs(t_end) : 1.0;
    if (t_end - s->time < 1e-12 * m) s->time = t_end;
  }
  return ORC_OK;
}

/* euler2d.ts:528-533 stableTimeStep */
double orc_stable_time_step(orc_solver *s) {
  int saved = s->cfg.local_stepping;
  s->cfg.local_stepping = 1;
  compute_time_steps(s, INFINITY);
  s->cfg.local_stepping = saved;
  double m = INFINITY;
  for (int c = 0; c < s->ni * s->nj; c++) if (s->dt[c] < m) m = s->dt[c];
  return m;
}

void orc_time_steps(orc_solver *s, double *dt) {
  compute_time_steps(s, INFINITY);
  memcpy(dt, s->dt, (size_t)s->ni * s->nj * sizeof(double));
}

/* euler2d.ts:406-445 setInitialCondition / setCellPrimitives */
int orc_set_primitives(orc_solver *s, const double *prim) {
  size_t n = (size_t)s->ni * s->nj;
  for (size_t c = 0; c < n; c++) {
    prim_t w = {prim[4 * c], prim[4 * c + 1], prim[4 * c + 2], prim[4 * c + 3]};
    cons_t q = to_cons(w, s->cfg.gamma);
    memcpy(s->u + 4 * c, q.q, sizeof q.q);
  }
  s->time = 0.0;
  s->steps = 0;
  return check_physical(s, s->u);
}

/* euler2d.ts:1128-1143 setConserved */
int orc_set_conserved(orc_solver *s, const double *q) {
  memcpy(s->u, q, 4 * (size_t)s->ni * s->nj * sizeof(double));
  s->steps += 1;
  return check_physical(s, s->u);
}

void orc_get_conserved(const orc_solver *s, double *q) {
  memcpy(q, s->u, 4 * (size_t)s->ni * s->nj * sizeof(double));
}

void orc_get_clock(const orc_solver *s, double *time, long long *steps) {
  if (time) *time = s->time;
  if (steps) *steps = s->steps;
}

/* euler2d.ts:462-477 residualVector */
void orc_residual_vector(orc_solver *s, double *r) {
  compute_residual(s, s->u, s->res);
  memcpy(r, s->res, 4 * (size_t)s->ni * s->nj * sizeof(double));
}

/* euler2d.ts:480-491 densityResidualNorm */
double orc_density_residual_norm(orc_solver *s) {
  compute_residual(s, s->u, s->res);
  size_t n = (size_t)s->ni * s->nj;
  double sum = 0.0;
  for (size_t c = 0; c < n; c++) {
    double r = s->res[4 * c] / s->volume[c];
    sum += r * r;
  }
  return sqrt(sum / (double)n);
}

/* euler2d.ts:499-525 boundaryFluxes */
int orc_boundary_fluxes(orc_solver *s, int side, double *out) {
  if (side < 0 || side > 3) return fail(ORC_ERR_RANGE, "no such side");
  int ni = s->ni, nj = s->nj, m = along(s, side);
  double g = s->cfg.gamma;
  for (int t = 0; t < m; t++) {
    double nx, ny, area;
    int outward;
    boundary_normal(s, side, t, &nx, &ny, &area, &outward);
    int i, j, dir, n0, n1;
    if (side < 2) {
      i = side == ORC_IMIN ? 0 : ni; j = t; dir = 0;
      n0 = i + t * (ni + 1); n1 = n0 + ni + 1;
    } else {
      i = t; j = side == ORC_JMIN ? 0 : nj; dir = 1;
      n0 = t + j * (ni + 1); n1 = n0 + 1;
    }
    prim_t l, r;
    face_states(s, s->u, dir, i, j, side, nx, ny, &l, &r);
    double f[4];
    flux_2d(s->cfg.flux, l, r, nx, ny, g, f);
    double *o = out + 9 * t;
    o[0] = 0.5 * (s->x[n0] + s->x[n1]);
    o[1] = 0.5 * (s->y[n0] + s->y[n1]);
    o[2] = outward * nx;
    o[3] = outward * ny;
    o[4] = area;
    for (int k = 0; k < 4; k++) o[5 + k] = outward * f[k];
  }
  return ORC_OK;
}

/* forces.ts:19-29 forebodyPressureDrag */
double orc_forebody_pressure_drag(const double *faces, int count, double p_inf, double q_inf,
                                  double base_radius) {
  double force = 0.0;
  for (int t = 0; t < count; t++) {
    const double *f = faces + 9 * t;
    force += (f[6] - p_inf * f[2]) * f[4];
  }
  return (2.0 * force) / (q_inf * base_radius * base_radius);
}

/* ---------------------------------------------------------------------- */
/* Implicit steps (LU-SGS and its data-parallel relaxations)               */
/* ---------------------------------------------------------------------- */

/* The neighbour term A (1/2 dF_n - 1/2 omega lambda dQ) of cell c for its
 * change d: euler2d.ts:981-1005. */
static void neighbour_term(const orc_solver *s, int c, double nx, double ny, double area,
                           double lam_a, double omega, const double *from, double *sum) {
  double g = s->cfg.gamma;
  const double *q = s->u + 4 * c;
  const double *d = from + 4 * c;
  double rho = q[0], mx = q[1], my = q[2], e = q[3];
  double p0 = s->pc[c];
  double un0 = (mx * nx + my * ny) / rho;
  double rho1 = rho + d[0], mx1 = mx + d[1], my1 = my + d[2], e1 = e + d[3];
  double p1 = (g - 1.0) * (e1 - (0.5 * (mx1 * mx1 + my1 * my1)) / rho1);
  double un1 = (mx1 * nx + my1 * ny) / rho1;
  double half = 0.5 * area, damp = 0.5 * omega * lam_a;
  sum[0] += half * (rho1 * un1 - rho * un0) - damp * d[0];
  sum[1] += half * (mx1 * un1 + p1 * nx - (mx * un0 + p0 * nx)) - damp * d[1];
  sum[2] += half * (my1 * un1 + p1 * ny - (my * un0 + p0 * ny)) - damp * d[2];
  sum[3] += half * ((e1 + p1) * un1 - (e + p0) * un0) - damp * d[3];
}

/* The same term with dF_n the flux Jacobian along n applied to d: linear
 * in d, so its rounding is relative to the term itself (the two-float
 * engine's relaxation, euler2d-two-float.wgsl:574-598). It agrees with
 * neighbour_term to first order in d. */
static void neighbour_term_linear(const double *q, const double *d, double g, double nx, double ny,
                                  double area, double lam_a, double omega, double *sum) {
  double rho = q[0], mx = q[1], my = q[2], e = q[3];
  double u = mx / rho, v = my / rho;
  double un = u * nx + v * ny;
  double p = (g - 1.0) * (e - (0.5 * (mx * mx + my * my)) / rho);
  double dp = (g - 1.0) * (d[3] - (u * d[1] + v * d[2]) + 0.5 * (u * u + v * v) * d[0]);
  double dun = (d[1] * nx + d[2] * ny - un * d[0]) / rho;
  double df[4] = {d[1] * nx + d[2] * ny, d[1] * un + mx * dun + dp * nx, d[2] * un + my * dun + dp * ny,
                  (d[3] + dp) * un + (e + p) * dun};
  double half = 0.5 * area, damp = 0.5 * omega * lam_a;
  for (int k = 0; k < 4; k++) sum[k] += half * df[k] - damp * d[k];
}

static void neighbour_of(const orc_solver *s, int linear, int c, double nx, double ny, double area,
                         double lam_a, double omega, const double *from, double *sum) {
  if (linear)
    neighbour_term_linear(s->u + 4 * c, from + 4 * c, s->cfg.gamma, nx, ny, area, lam_a, omega, sum);
  else
    neighbour_term(s, c, nx, ny, area, lam_a, omega, from, sum);
}

static double radius_of(const orc_solver *s, int c, double nx, double ny) {
  const double *q = s->u + 4 * c;
  return fabs((q[1] * nx + q[2] * ny) / q[0]) + s->ac[c];
}

/* euler2d.ts:878-1105 implicitCorrection */
int orc_implicit_correction(orc_solver *s, double cfl, double omega, int relaxation,
                            int sweeps, double *dq_out) {
  if (!(cfl > 0.0 && omega >= 1.0)) return fail(ORC_ERR_RANGE, "LU-SGS needs cfl > 0 and omega >= 1");
  if (relaxation < ORC_RELAX_SWEEPS || relaxation > ORC_RELAX_RED_BLACK_LINEAR)
    return fail(ORC_ERR_RANGE, "unknown relaxation");
  if (relaxation != ORC_RELAX_SWEEPS && !(sweeps >= 1))
    return fail(ORC_ERR_RANGE, "a relaxation needs a whole number of sweeps, at least one");
  const int linear = relaxation == ORC_RELAX_RED_BLACK_LINEAR;
  int ni = s->ni, nj = s->nj;
  size_t n = (size_t)ni * nj;
  double g = s->cfg.gamma;
  if (!s->lam_i) {
    s->lam_i = dalloc((size_t)(ni + 1) * nj);
    s->lam_j = dalloc((size_t)ni * (nj + 1));
    s->diag = dalloc(n);
    s->dq = dalloc(4 * n);
    s->pc = dalloc(n);
    s->ac = dalloc(n);
    s->prev = dalloc(4 * n);
  }
  compute_residual(s, s->u, s->res);
  for (size_t c = 0; c < n; c++) {
    s->pc[c] = pressure_of(s->u + 4 * c, g);
    s->ac[c] = sqrt((g * s->pc[c]) / s->u[4 * c]);
  }
  for (int j = 0; j < nj; j++)
    for (int i = 0; i <= ni; i++) {
      int f = i + j * (ni + 1), c = i + j * ni;
      double l = i > 0 ? radius_of(s, c - 1, s->inx[f], s->iny[f]) : 0.0;
      double r = i < ni ? radius_of(s, c, s->inx[f], s->iny[f]) : 0.0;
      s->lam_i[f] = (l > r ? l : r) * s->iarea[f];
    }
  for (int j = 0; j <= nj; j++)
    for (int i = 0; i < ni; i++) {
      int f = i + j * ni;
      double b = j > 0 ? radius_of(s, f - ni, s->jnx[f], s->jny[f]) : 0.0;
      double a = j < nj ? radius_of(s, f, s->jnx[f], s->jny[f]) : 0.0;
      s->lam_j[f] = (b > a ? b : a) * s->jarea[f];
    }
  double scale = 0.5 * (omega + 1.0 / cfl);
  for (int j = 0; j < nj; j++)
    for (int i = 0; i < ni; i++) {
      int fi = i + j * (ni + 1), fj = i + j * ni;
      s->diag[fj] = scale * (s->lam_i[fi] + s->lam_i[fi + 1] + s->lam_j[fj] + s->lam_j[fj + ni]) +
                    (s->cfg.axisymmetric ? s->planar[fj] * radius_of(s, fj, 0.0, 1.0) : 0.0);
    }
  double *dq = s->dq, *res = s->res;
  if (relaxation != ORC_RELAX_SWEEPS) {
    memset(dq, 0, 4 * n * sizeof(double));
    for (int sw = 0; sw < sweeps; sw++) {
      const double *from = dq;
      if (relaxation == ORC_RELAX_JACOBI) {
        memcpy(s->prev, dq, 4 * n * sizeof(double));
        from = s->prev;
      }
      for (int colour = 0; colour < (relaxation == ORC_RELAX_JACOBI ? 1 : 2); colour++) {
        for (int j = 0; j < nj; j++) {
          int i0 = relaxation == ORC_RELAX_JACOBI ? 0 : (j + colour) % 2;
          int di = relaxation == ORC_RELAX_JACOBI ? 1 : 2;
          for (int i = i0; i < ni; i += di) {
            int c = i + j * ni;
            double sum[4] = {0, 0, 0, 0};
            if (i > 0) {
              int f = i + j * (ni + 1);
              neighbour_of(s, linear, c - 1, -s->inx[f], -s->iny[f], s->iarea[f], s->lam_i[f], omega, from, sum);
            }
            if (i < ni - 1) {
              int f = i + 1 + j * (ni + 1);
              neighbour_of(s, linear, c + 1, s->inx[f], s->iny[f], s->iarea[f], s->lam_i[f], omega, from, sum);
            }
            if (j > 0)
              neighbour_of(s, linear, c - ni, -s->jnx[c], -s->jny[c], s->jarea[c], s->lam_j[c], omega, from, sum);
            if (j < nj - 1) {
              int f = c + ni;
              neighbour_of(s, linear, c + ni, s->jnx[f], s->jny[f], s->jarea[f], s->lam_j[f], omega, from, sum);
            }
            double d = s->diag[c];
            for (int k = 0; k < 4; k++) dq[4 * c + k] = (res[4 * c + k] - sum[k]) / d;
          }
        }
      }
    }
  } else {
    /* Forward sweep: (D + L) dQ* = R. */
    for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) {
        int c = i + j * ni;
        double sum[4] = {0, 0, 0, 0};
        if (i > 0) {
          int f = i + j * (ni + 1);
          neighbour_term(s, c - 1, -s->inx[f], -s->iny[f], s->iarea[f], s->lam_i[f], omega, dq, sum);
        }
        if (j > 0)
          neighbour_term(s, c - ni, -s->jnx[c], -s->jny[c], s->jarea[c], s->lam_j[c], omega, dq, sum);
        double d = s->diag[c];
        for (int k = 0; k < 4; k++) dq[4 * c + k] = (res[4 * c + k] - sum[k]) / d;
      }
    /* Backward sweep: (D + U) dQ = D dQ*. */
    for (int j = nj - 1; j >= 0; j--)
      for (int i = ni - 1; i >= 0; i--) {
        int c = i + j * ni;
        double sum[4] = {0, 0, 0, 0};
        if (i < ni - 1) {
          int f = i + 1 + j * (ni + 1);
          neighbour_term(s, c + 1, s->inx[f], s->iny[f], s->iarea[f], s->lam_i[f], omega, dq, sum);
        }
        if (j < nj - 1) {
          int f = c + ni;
          neighbour_term(s, c + ni, s->jnx[f], s->jny[f], s->jarea[f], s->lam_j[f], omega, dq, sum);
        }
        double d = s->diag[c];
        for (int k = 0; k < 4; k++) dq[4 * c + k] = dq[4 * c + k] - sum[k] / d;
      }
  }
  if (dq_out) memcpy(dq_out, dq, 4 * n * sizeof(double));
  return ORC_OK;
}

/* euler2d.ts:816-847 applyImplicitUpdate */
static int apply_implicit_update(orc_solver *s, const double *dq) {
  size_t n = (size_t)s->ni * s->nj;
  double g = s->cfg.gamma;
  for (size_t c = 0; c < n; c++) {
    double *q = s->u + 4 * c;
    const double *d = dq + 4 * c;
    double rho = q[0], mx = q[1], my = q[2], e = q[3];
    double p = (g - 1.0) * (e - (0.5 * (mx * mx + my * my)) / rho);
    double alpha = 1.0;
    for (int tries = 0; tries < 30; tries++) {
      double rho1 = rho + alpha * d[0], mx1 = mx + alpha * d[1], my1 = my + alpha * d[2],
             e1 = e + alpha * d[3];
      double p1 = (g - 1.0) * (e1 - (0.5 * (mx1 * mx1 + my1 * my1)) / rho1);
      if (rho1 > 0.5 * rho && rho1 < 2.0 * rho && p1 > 0.5 * p && p1 < 2.0 * p) break;
      alpha *= 0.5;
    }
    q[0] = rho + alpha * d[0];
    q[1] = mx + alpha * d[1];
    q[2] = my + alpha * d[2];
    q[3] = e + alpha * d[3];
  }
  s->steps += 1;
  return check_physical(s, s->u);
}

/* euler2d.ts:581-589 iterateImplicit */
int orc_iterate_implicit(orc_solver *s, int count, double cfl, double omega, int relaxation,
                         int sweeps) {
  for (int k = 0; k < count; k++) {
    int rc = orc_implicit_correction(s, cfl, omega, relaxation, sweeps, NULL);
    if (rc) return rc;
    rc = apply_implicit_update(s, s->dq);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Scalar exports for property tests                                       */
/* ---------------------------------------------------------------------- */

void orc_neighbour_term_linear(const double *q, const double *d, double g, double nx, double ny, double area,
                               double lam_a, double omega, double *out) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  neighbour_term_linear(q, d, g, nx, ny, area, lam_a, omega, out);
}

void orc_face_frame_flux(int flux, const double *l, const double *r, double g, double *out) {
  face_t a = {l[0], l[1], l[2], l[3]}, b = {r[0], r[1], r[2], r[3]};
  frame_flux(flux, a, b, g, out);
}

void orc_face_flux_2d(int flux, const double *l, const double *r, double nx, double ny,
                      double g, double *out) {
  prim_t a = {l[0], l[1], l[2], l[3]}, b = {r[0], r[1], r[2], r[3]};
  flux_2d(flux, a, b, nx, ny, g, out);
}

double orc_limiter(int kind, double a, double b) { return limiter(kind, a, b); }

void orc_reconstruct(int rec, int lim, const double *p, const double *c, const double *n,
                     int side, double g, double *out) {
  prim_t a = {p[0], p[1], p[2], p[3]}, b = {c[0], c[1], c[2], c[3]}, d = {n[0], n[1], n[2], n[3]};
  prim_t o = reconstruct(rec, lim, a, b, d, side > 0 ? 1.0 : -1.0, g);
  out[0] = o.rho; out[1] = o.u; out[2] = o.v; out[3] = o.p;
}

void orc_far_field_state(const double *in, const double *fr, double nx, double ny, double g,
                         double *out) {
  prim_t a = {in[0], in[1], in[2], in[3]}, b = {fr[0], fr[1], fr[2], fr[3]};
  prim_t o = far_field(a, b, nx, ny, g);
  out[0] = o.rho; out[1] = o.u; out[2] = o.v; out[3] = o.p;
}
