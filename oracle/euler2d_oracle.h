/*
 * FP64 CPU oracle of the axisymmetric finite-volume Euler path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain C and double
 * precision, the algorithm of bowshock's CPU reference engine
 * (src/engine-cpu/euler2d.ts) and its numerics (src/scheme/{gas,flux,limiters,reconstruction}.ts,
 * src/grid/structured.ts, src/solver/{config,errors,runge-kutta}.ts, src/analysis/forces.ts). It is
 * the checker that the HIP path is compared against, and the CPU baseline
 * that the benchmark times beside the GPU. Nothing in the product links,
 * loads or calls it.
 *
 * Conventions follow the reference: cell (i, j) at i + j * ni, i-face
 * (i, j) between cells (i - 1, j) and (i, j) at i + j * (ni + 1), j-face
 * (i, j) between cells (i, j - 1) and (i, j) at i + j * ni, nodes at
 * i + j * (ni + 1). Conserved variables (rho, rho u, rho v, E) are stored
 * interleaved, four per cell, as the reference's conserved() returns them.
 */
#ifndef EULER2D_ORACLE_H
#define EULER2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary kinds, sides, limiters, reconstructions, fluxes, integrators:
 * the same names as the reference's string unions. */
enum { ORC_INFLOW = 0, ORC_OUTFLOW = 1, ORC_WALL = 2, ORC_PERIODIC = 3,
       ORC_FACE_STATES = 4, ORC_FAR_FIELD = 5 };
enum { ORC_IMIN = 0, ORC_IMAX = 1, ORC_JMIN = 2, ORC_JMAX = 3 };
enum { ORC_FIRST_ORDER = 0, ORC_MINMOD = 1, ORC_VAN_LEER = 2, ORC_VAN_ALBADA = 3,
       ORC_MONOTONIZED_CENTRAL = 4, ORC_UNLIMITED_CENTRAL = 5 };
enum { ORC_PER_VARIABLE = 0, ORC_SHARED = 1, ORC_SHARED_SCALED = 2 };
enum { ORC_HLLC = 0, ORC_RUSANOV = 1, ORC_HLL = 2 };
enum { ORC_FORWARD_EULER = 0, ORC_SSPRK2 = 1, ORC_SSPRK3 = 2 };
/* RED_BLACK_LINEAR: red-black with the linearized neighbour term of the
 * two-float engine (euler2d-two-float.wgsl:574-598). */
enum { ORC_RELAX_SWEEPS = 0, ORC_RELAX_JACOBI = 1, ORC_RELAX_RED_BLACK = 2, ORC_RELAX_RED_BLACK_LINEAR = 3 };

/* Return codes. */
enum { ORC_OK = 0, ORC_ERR_RANGE = 1, ORC_ERR_NONPHYSICAL = 2 };

typedef struct orc_solver orc_solver;

typedef struct {
  int ni, nj;
  const double *x, *y;      /* (ni + 1) * (nj + 1) node coordinates */
  int axisymmetric;         /* 1: y is the radius */
  double gamma;
  int flux, limiter, reconstruction, integrator;
  double cfl;
  int local_stepping;       /* 1: local time steps, 0: global */
  int bc_kind[4];           /* per side, ORC_IMIN..ORC_JMAX */
  double bc_state[16];      /* per side: rho, u, v, p (inflow / far field) */
  const double *source;     /* 4 per cell or NULL (manufactured solutions) */
} orc_config;

/* Last error message of the calling thread. */
const char *orc_last_error(void);

int orc_create(const orc_config *cfg, orc_solver **out);
void orc_destroy(orc_solver *s);

/* Finite-volume metrics (structured.ts computeMetrics). Any pointer may be
 * NULL. i-faces (ni + 1) * nj, j-faces ni * (nj + 1), cells ni * nj. */
int orc_compute_metrics(int ni, int nj, const double *x, const double *y, int axisymmetric,
                        double *inx, double *iny, double *iarea,
                        double *jnx, double *jny, double *jarea,
                        double *volume, double *planar, double *cx, double *cy);

/* State access (conserved interleaved, 4 per cell). set_primitives takes
 * rho, u, v, p per cell (interleaved) and resets time and steps, as
 * setCellPrimitives. set_conserved counts a step, as setConserved. */
int orc_set_primitives(orc_solver *s, const double *prim);
int orc_set_conserved(orc_solver *s, const double *q);
void orc_get_conserved(const orc_solver *s, double *q);
void orc_get_clock(const orc_solver *s, double *time, long long *steps);

int orc_set_face_states(orc_solver *s, int side, const double *prim, int count);
int orc_set_face_fluxes(orc_solver *s, int side, const double *flux, int count);

int orc_iterate(orc_solver *s, int count);
int orc_advance_to(orc_solver *s, double t_end, int max_steps);
double orc_stable_time_step(orc_solver *s);
/* Per-cell time steps as computeTimeSteps(Infinity) leaves them. */
void orc_time_steps(orc_solver *s, double *dt);

void orc_residual_vector(orc_solver *s, double *r);
double orc_density_residual_norm(orc_solver *s);
/* Per boundary face: x, y, nx, ny, area, f0, f1, f2, f3 (9 doubles). */
int orc_boundary_fluxes(orc_solver *s, int side, double *out);

int orc_implicit_correction(orc_solver *s, double cfl, double omega, int relaxation,
                            int sweeps, double *dq);
int orc_iterate_implicit(orc_solver *s, int count, double cfl, double omega,
                         int relaxation, int sweeps);

/* forces.ts forebodyPressureDrag over faces from orc_boundary_fluxes. */
double orc_forebody_pressure_drag(const double *faces, int count, double p_inf,
                                  double q_inf, double base_radius);

/* Scalar numerics, exported for property tests. A primitive state is
 * (rho, u, v, p); a face state is (rho, un, ut, p); a flux is 4 doubles. */
void orc_face_frame_flux(int flux, const double *left, const double *right, double gamma,
                         double *out);
void orc_face_flux_2d(int flux, const double *left, const double *right, double nx, double ny,
                      double gamma, double *out);
double orc_limiter(int limiter, double a, double b);
/* The linearized neighbour term (4 values) of a cell with conserved state
 * q and change d across a face (normal, area, lambda A). */
void orc_neighbour_term_linear(const double *q, const double *d, double g, double nx, double ny, double area,
                               double lam_a, double omega, double *out);
void orc_reconstruct(int reconstruction, int limiter, const double *prev, const double *centre,
                     const double *next, int side, double gamma, double *out);
void orc_far_field_state(const double *inside, const double *free_state, double nx, double ny,
                         double gamma, double *out);

#ifdef __cplusplus
}
#endif
#endif
