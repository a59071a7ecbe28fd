/*
 * bowshock-mi355x: C ABI of the MI355X (gfx950) axisymmetric Euler engine.
 *
 * One solver holds one structured grid block on one GPU, its state in FP32
 * structure-of-arrays form in HBM, and runs bowshock's explicit SSP-RK and
 * implicit red-black LU-SGS steps with hand-written HIP kernels. Residual
 * norms and wall fluxes / drag are evaluated on the GPU in FP64.
 *
 * Every entry point replaces a method of the reference's GPU engine
 * (src/engine-gpu/euler2d-gpu.ts, class GpuEuler2D) or of the CPU engine
 * it shares its method set with (src/engine-cpu/euler2d.ts, class
 * Euler2D); the citation sits beside each declaration. Plain pointers and
 * sizes only: callers own every host buffer they pass, the solver owns its
 * device memory, and no pointer escapes.
 *
 * Errors: each call returns a bs_status; bs_last_error() gives the message
 * of the calling thread's last failure. A solver is not reentrant: a call
 * made while another call on the same solver is running returns BS_ERR_BUSY
 * (the reference's ExclusiveRunner, src/engine-gpu/gpu-common.ts:122-140);
 * so do bs_solver_destroy, bs_batch_destroy, bs_anderson_destroy and
 * bs_mlp_destroy, which then free nothing. Distinct objects may be driven
 * from distinct threads at once: each object's kernels and copies are
 * ordered on its own stream.
 *
 * Conventions (src/grid/structured.ts:1-8): node (i, j) at i + j (ni + 1);
 * cell (i, j) at i + j ni; sides in the order iMin, iMax, jMin, jMax.
 */
#ifndef BOWSHOCK_H
#define BOWSHOCK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the host facades map each to the reference's error class
 * (src/solver/errors.ts:2-7, src/engine-gpu/gpu-common.ts:7-17). */
typedef enum {
  BS_OK = 0,
  BS_ERR_RANGE = 1,        /* RangeError: invalid configuration or argument */
  BS_ERR_NONPHYSICAL = 2,  /* NonPhysicalStateError */
  BS_ERR_MEMORY = 3,       /* GpuMemoryLimitError */
  BS_ERR_EXECUTION = 4,    /* GpuExecutionError: the device lost or rejected work */
  BS_ERR_PRECISION = 5,    /* GpuPrecisionError */
  BS_ERR_BUSY = 6,         /* "await the previous operation first" */
  BS_ERR_UNSUPPORTED = 7   /* a method of the reference this engine does not provide */
} bs_status;

/* Boundary kinds (src/engine-cpu/euler2d.ts:34-56). */
enum { BS_BC_INFLOW = 0, BS_BC_OUTFLOW = 1, BS_BC_WALL = 2, BS_BC_PERIODIC = 3,
       BS_BC_FACE_STATES = 4, BS_BC_FAR_FIELD = 5 };
enum { BS_SIDE_IMIN = 0, BS_SIDE_IMAX = 1, BS_SIDE_JMIN = 2, BS_SIDE_JMAX = 3 };
/* src/scheme/limiters.ts:87-94 */
enum { BS_LIMITER_FIRST_ORDER = 0, BS_LIMITER_MINMOD = 1, BS_LIMITER_VAN_LEER = 2,
       BS_LIMITER_VAN_ALBADA = 3, BS_LIMITER_MONOTONIZED_CENTRAL = 4,
       BS_LIMITER_UNLIMITED_CENTRAL = 5 };
/* src/scheme/reconstruction.ts:143 */
enum { BS_RECON_PER_VARIABLE = 0, BS_RECON_SHARED = 1, BS_RECON_SHARED_SCALED = 2 };
/* The GPU fluxes, src/engine-gpu/euler2d-gpu.ts:57-60 */
enum { BS_FLUX_HLLC = 0, BS_FLUX_RUSANOV = 1 };
/* src/solver/runge-kutta.ts:17-28 */
enum { BS_RK_FORWARD_EULER = 0, BS_RK_SSPRK2 = 1, BS_RK_SSPRK3 = 2 };
enum { BS_GEOMETRY_PLANAR = 0, BS_GEOMETRY_AXISYMMETRIC = 1 };
enum { BS_STEPPING_GLOBAL = 0, BS_STEPPING_LOCAL = 1 };
/* Precision of a solver's state, metrics and stencil arithmetic. */
enum { BS_PRECISION_FP32 = 0, BS_PRECISION_FP64 = 1 };
/* Status word bits read back after every batch of steps
 * (src/engine-gpu/time-loop.ts:15 STATUS_NON_PHYSICAL). */
enum { BS_STATUS_NON_PHYSICAL = 1u };

typedef struct bs_ctx bs_ctx;
typedef struct bs_solver bs_solver;

/* A boundary condition: kind, and the primitive state (rho, u, v, p)
 * outside an inflow side or the free stream of a far field; for a
 * faceStates side the state given to every face until bs_set_face_states. */
typedef struct {
  int32_t kind;
  double rho, u, v, p;
} bs_boundary;

/* Euler2DConfig / GpuEuler2DConfig (src/engine-cpu/euler2d.ts:58-81,
 * src/engine-gpu/euler2d-gpu.ts:82-92), as plain data. */
typedef struct {
  int32_t ni, nj;
  const double *x, *y;         /* (ni + 1) (nj + 1) node coordinates */
  int32_t geometry;            /* BS_GEOMETRY_* */
  double gamma;
  int32_t flux, limiter, reconstruction, integrator;
  double cfl;
  int32_t time_stepping;       /* BS_STEPPING_* */
  bs_boundary boundaries[4];   /* iMin, iMax, jMin, jMax */
} bs_config;

/* What a batch of steps leaves (src/engine-gpu/time-loop.ts:79-92
 * ControlReading): steps taken since the last reset, status bits, the
 * simulated time (global stepping, FP64 sum of the steps) and the last
 * step's dt. */
typedef struct {
  int64_t steps;
  uint32_t status;
  double time;
  double dt;
} bs_report;

/* Thread-local message of the last failure ("" if none). */
const char *bs_last_error(void);

/* Library version string. */
const char *bs_version(void);

/* A context on one HIP device (the reference's GPUDevice argument,
 * euler2d-gpu.ts:217). */
bs_status bs_ctx_create(int device, bs_ctx **out);
bs_status bs_ctx_destroy(bs_ctx *ctx);

/* GpuEuler2D.create (euler2d-gpu.ts:217-291): validates the configuration
 * (src/solver/config.ts:35-55), computes the metrics in FP64 on the host
 * (src/grid/structured.ts:94-168) and uploads them in FP32. */
bs_status bs_solver_create(bs_ctx *ctx, const bs_config *config, bs_solver **out);
/* The precision extension of the path: the same solver with its state,
 * metrics, time steps and stencil arithmetic in FP64, for convergence
 * below the FP32 residual floor. It replaces the reference's two-float
 * engine (GpuEuler2DTwoFloat.create, src/engine-gpu/euler2d-two-float.ts:151,
 * whose ~48-bit arithmetic WebGPU needs for lack of FP64) with native FP64.
 * Explicit steps, residuals, state input/output, boundary fluxes and drag
 * run as on an FP32 solver. Implicit steps take the FP64 residual,
 * diagonal and update around a relaxation in FP32, as the two-float
 * engine's iterateImplicit (euler2d-two-float.ts:379-448) solves its
 * implicit system in 32 bits: the fixed point is the FP64 residual's.
 * Initial states are formed with the configured (FP64) gamma. */
bs_status bs_solver_create_fp64(bs_ctx *ctx, const bs_config *config, bs_solver **out);
/* BS_PRECISION_FP32 or BS_PRECISION_FP64. */
bs_status bs_solver_precision(bs_solver *s, int32_t *precision);
/* GpuEuler2D.destroy (euler2d-gpu.ts:659); NULL is accepted. A solver
 * that another call is still running on is not freed: BS_ERR_BUSY (the
 * reference's ExclusiveRunner refuses the overlapping call). */
bs_status bs_solver_destroy(bs_solver *s);

/* Grid metrics in FP64 (GridMetrics, src/grid/structured.ts:28-50). Any
 * pointer may be NULL. i-faces (ni + 1) nj, j-faces ni (nj + 1), cells
 * ni nj. Stateless: needs no GPU. */
bs_status bs_compute_metrics(int32_t ni, int32_t nj, const double *x, const double *y,
                             int32_t geometry, double *i_nx, double *i_ny, double *i_area,
                             double *j_nx, double *j_ny, double *j_area, double *volume,
                             double *planar_area, double *centroid_x, double *centroid_y);

/* setInitialCondition / setCellPrimitives (euler2d-gpu.ts:412-459): one
 * primitive state (rho, u, v, p) per cell, interleaved, 4 ni nj doubles;
 * converted to conserved with the FP32 gamma; resets the clock. Returns
 * BS_ERR_NONPHYSICAL before anything reaches the GPU if a state is not
 * physical. */
bs_status bs_set_primitives(bs_solver *s, const double *prim);
/* Euler2D.setConserved without the step count: conserved variables as FP32
 * structure of arrays (rho[n], mu[n], mv[n], energy[n]), `count` = n = ni nj
 * values each (BS_ERR_RANGE otherwise, as the reference's RangeError for a
 * field of the wrong size); resets the clock. */
bs_status bs_set_state(bs_solver *s, const float *rho, const float *mu, const float *mv,
                       const float *energy, int32_t count);
/* GpuEuler2D.readState (euler2d-gpu.ts:549-572, GpuState2D 116-125):
 * conserved FP32 arrays indexed i + j ni. Any pointer may be NULL. Returns
 * BS_ERR_NONPHYSICAL if the status word has the non-physical bit. */
bs_status bs_read_state(bs_solver *s, float *rho, float *mu, float *mv, float *energy,
                        bs_report *report);
/* bs_set_state and bs_read_state with FP64 arrays: exact on an FP64
 * solver; rounded to (or widened from) FP32 on an FP32 solver. The read
 * replaces GpuEuler2DTwoFloat.readConserved (euler2d-two-float.ts:486-501),
 * which returns its ~48-bit state as FP64. */
bs_status bs_set_state_f64(bs_solver *s, const double *rho, const double *mu, const double *mv,
                           const double *energy, int32_t count);
bs_status bs_read_state_f64(bs_solver *s, double *rho, double *mu, double *mv, double *energy,
                            bs_report *report);

/* setFaceStates (euler2d-gpu.ts:575-600): `count` primitive states, one per
 * face along a faceStates side. */
bs_status bs_set_face_states(bs_solver *s, int32_t side, const double *prim, int32_t count);
/* setFaceFluxes (euler2d-gpu.ts:606-630): 4 values per face per unit area
 * along the face normal, or NULL to go back to the Riemann flux. */
bs_status bs_set_face_fluxes(bs_solver *s, int32_t side, const double *flux, int32_t count);

/* iterate (euler2d-gpu.ts:462-473): `count` explicit steps, local or global
 * time steps, in batches between status checks. Returns BS_ERR_NONPHYSICAL
 * when a batch ends with the status bit set, BS_ERR_EXECUTION when the
 * device ran fewer steps than submitted (time-loop.ts:125-150). */
bs_status bs_iterate(bs_solver *s, int32_t count, bs_report *report);
/* advanceTo (euler2d-gpu.ts:511-527, time-loop.ts:158-177): global steps
 * to exactly t_end, at most max_steps. */
bs_status bs_advance_to(bs_solver *s, double t_end, int32_t max_steps, bs_report *report);

/* densityResidualNorm (euler2d-gpu.ts:530-546; euler2d.ts:480-491): RMS of
 * the density residual over volume, reduced in FP64 on the GPU. */
bs_status bs_residual_norm(bs_solver *s, double *norm);
/* residualVector (euler2d.ts:462-477) of the current state, FP32 residual
 * returned as FP64, 4 per cell interleaved. */
bs_status bs_residual_vector(bs_solver *s, double *residual);

/* TransonicGridSolver.checkPrecision (src/analysis/transonic-drag.ts:56-66;
 * gpu-transonic.ts:75-80, 108-131 residualPrecision): the solver's residual
 * vector of its current state against a double-precision one of the same
 * state, here the FP64 stencil kernel's on the state widened to FP64 (with
 * FP64 metrics). BS_ERR_PRECISION when the largest difference exceeds 1e-8
 * (RESIDUAL_PRECISION) of the largest residual, which an FP32 solver does
 * (its 32-bit fluxes lose about 1e-7 of values far larger than the
 * residuals); an FP64 solver's residual is the double one and passes at
 * once. `difference` (may be NULL) receives the relative difference.
 * BS_ERR_NONPHYSICAL if the status word has the non-physical bit. */
bs_status bs_check_precision(bs_solver *s, double *difference);

/* boundaryFluxes (euler2d.ts:499-525) evaluated on the GPU in FP64 from the
 * FP32 state (pressure with the FP32 gamma, gpuPrimitive2D,
 * euler2d-gpu.ts:136-146): 9 doubles per face, x, y, nx, ny (outward),
 * area, flux[4] (outward). `count` must be the faces along the side. */
bs_status bs_wall_fluxes(bs_solver *s, int32_t side, double *faces, int32_t count);
/* forebodyPressureDrag (src/analysis/forces.ts:19-29) over one side's faces,
 * reduced in FP64 on the GPU. */
bs_status bs_forebody_drag(bs_solver *s, int32_t side, double p_inf, double q_inf,
                           double base_radius, double *cd);

/* setImplicitCfl (euler2d-gpu.ts:668-681). */
bs_status bs_set_implicit_cfl(bs_solver *s, double cfl);
/* iterateImplicit (euler2d-gpu.ts:687-723; euler2d.ts:581-589 with the
 * red-black relaxation): `count` steps, each with `sweeps` red-black
 * sweeps; the CFL number starts from the value set, grows by `growth` per
 * step up to cfl_max. */
bs_status bs_iterate_implicit(bs_solver *s, int32_t count, double cfl_max, int32_t sweeps,
                              double omega, double growth, bs_report *report);

/* The fitted shock (GpuEuler2DConfig.fittedShock, euler2d-gpu.ts:85-91):
 * the jMax boundary, which must be a faceStates side, is a shock with this
 * free stream (rho, u, v, p) ahead of it. At the start of every explicit
 * step the GPU sets each jMax face's ghost state and imposed flux to those
 * of the shock consistent with the u_n + a characteristic from the two
 * cells below it (euler2d.wgsl:402-454 compute_shock_boundary;
 * src/solver/shock-fitting.ts:91-113 compatibleShock), in FP64, and records
 * the shock's speed. NULL turns it off. Implicit steps are refused while it
 * is on (euler2d-gpu.ts:705-707). */
bs_status bs_set_fitted_shock(bs_solver *s, const double *free_stream);
/* readShock (euler2d-gpu.ts:630-656): per jMax face (count = ni), the
 * shock's speed along the face's outward normal as set at the start of the
 * last step, its relative normal Mach number, and |u| + a of the cell
 * below; any pointer may be NULL. */
bs_status bs_read_shock(bs_solver *s, double *speed, double *mach, double *signal, int32_t count);

/* The moving grid of a fitted shock (GpuLineGrid, src/engine-gpu/
 * line-grid.ts): the solver's grid lies on ni + 1 straight lines from the
 * body, `lines` holding per line its origin and unit direction (x, y, dx,
 * dy), node (i, j) at the origin plus h_i fractions[i + j (ni + 1)] times
 * the direction (fixed fractions, blunt-grid.ts withFixedFractions), the
 * first node normal to the axis when `first_node_symmetry` (else held).
 * bs_set_heights (setHeights) places the nodes and recomputes the metrics
 * on the GPU; bs_move_grid (moveGrid / moveTo) moves them by less than
 * half the end cells' height and remaps the flow along the lines
 * (remapAlongLines, src/solver/shock-fitting.ts:305-338);
 * bs_iterate_fitted (iterateFitted, euler2d-gpu.ts:481-505) runs `moves`
 * times `steps_per_move` explicit steps with the fitted shock moved after
 * each group by the last step's shock speeds (fitShock,
 * shock-fitting.ts:252-296); bs_read_heights reads the heights back.
 * Heights come and go as `count` = ni + 1 values, one per line (BS_ERR_RANGE
 * otherwise: line-grid.ts:301 throws RangeError for another length). The
 * initial metrics of bs_solver_create stay the solver's until
 * bs_set_heights. */
bs_status bs_set_line_grid(bs_solver *s, const double *lines, const double *fractions,
                           int32_t first_node_symmetry);
bs_status bs_set_heights(bs_solver *s, const double *heights, int32_t count);
bs_status bs_move_grid(bs_solver *s, const double *heights, int32_t count);
bs_status bs_iterate_fitted(bs_solver *s, int32_t moves, int32_t steps_per_move, bs_report *report);
bs_status bs_read_heights(bs_solver *s, double *heights, int32_t count);

/* The neighbour term of the red-black relaxation. BS_RELAX_FLUX_DIFFERENCE
 * is the reference's A (1/2 (F(Q + dQ) - F(Q)) - 1/2 omega lambda dQ)
 * (euler2d.ts:981-1005, euler2d-implicit.wgsl), the default of an FP32
 * solver; BS_RELAX_LINEARIZED applies the flux Jacobian to dQ instead
 * (euler2d-two-float.wgsl:574-598), the form of the two-float engine and
 * the only one of an FP64 solver. The two agree to first order in dQ and
 * share the fixed point (the residual's zero); in FP32 the difference of
 * fluxes keeps a rounding noise of the size of F that does not shrink with
 * dQ, the linearized term's shrinks with it. */
enum { BS_RELAX_FLUX_DIFFERENCE = 0, BS_RELAX_LINEARIZED = 1 };
bs_status bs_set_relaxation(bs_solver *s, int32_t form);

/* Anderson acceleration (src/solver/anderson.ts:20-123 AndersonMixing):
 * the history of the last `depth` (1 to 16) differences of iterates and
 * of their corrections, in FP64 in device memory. bs_anderson_next is
 * AndersonMixing.next on host vectors of `size` values: the next iterate
 * x + f - (dX + dF) gamma, gamma the least-squares fit of f by dF (the
 * oldest columns dropped while a Cholesky pivot of dF^T dF is below 1e-7
 * of its column's norm; the reference's modified Gram-Schmidt drops below
 * 1e-10). bs_anderson_reset is AndersonMixing.reset. */
typedef struct bs_anderson bs_anderson;
bs_status bs_anderson_create(bs_ctx *ctx, int64_t size, int32_t depth, bs_anderson **out);
bs_status bs_anderson_destroy(bs_anderson *a);
bs_status bs_anderson_reset(bs_anderson *a);
bs_status bs_anderson_next(bs_anderson *a, const double *x, const double *f, double *out);
/* Columns held, columns the last solve used, resets by the solver loop,
 * and gamma (depth values, oldest column first); any pointer may be NULL. */
bs_status bs_anderson_state(bs_anderson *a, int32_t *columns, int32_t *used, int32_t *restarts, double *gamma);
/* `count` accelerated implicit steps of solveTransonic's loop
 * (src/analysis/transonic-flow.ts:184-208) on the device: the correction f
 * of an implicit step (bs_iterate_implicit's, before its update), the
 * history reset when |f|^2 exceeds 100 times its smallest since the last
 * reset, the mixed iterate as the new state, or, where a cell of it is not
 * physical, the plain implicit step and a reset; the CFL ramp as
 * bs_iterate_implicit's. The history's size must be 4 ni nj. */
bs_status bs_iterate_anderson(bs_solver *s, bs_anderson *a, int32_t count, double cfl_max, int32_t sweeps,
                              double omega, double growth, bs_report *report);

/* Steps run as captured HIP graphs of 12 steps each (one launch per graph
 * instead of one per kernel), on by default; no counterpart in the
 * reference, whose WebGPU engine encodes command buffers instead. */
bs_status bs_enable_graphs(bs_solver *s, int32_t enable);

/* Mean device time per launch of the solver's kernels since the last call,
 * for throughput reporting: kernel-name index (0 stage, 1 implicit prepare,
 * 2 implicit relaxation, whose last launch of a step also applies the
 * update on an FP32 solver; 3 the FP64 solver's update launch, or the
 * Anderson kernels of an accelerated step), milliseconds per
 * launch and the launches counted. Timing is taken with HIP events on the solver's stream
 * when enabled with bs_enable_timing: enable = 1 records a pair of events
 * around every launch (steps then run without graphs); enable = 2 records
 * one pair around each batch of explicit local-time steps, which launch the
 * stage kernel only, and splits it over the batch's stage launches (graphs
 * stay on; other steps are timed per launch). */
bs_status bs_enable_timing(bs_solver *s, int32_t enable);
bs_status bs_kernel_timing(bs_solver *s, int32_t kernel, double *ms_per_launch, int64_t *launches);

/* ------------------------------------------------------------------------
 * Batched solves. The reference solves its grids one after another
 * (analysis/verified-drag.ts:239-300, each through GpuEuler2D); these entry
 * points advance many cases of one scheme together, every stage one launch
 * over all active cases, with the same per-case kernels, status words and
 * FP64 reductions as the single-solve entry points above. Cases share the
 * scheme fields of bs_config; their grids, boundary kinds and boundary
 * states may differ (H grids with an inflow iMin beside C grids whose iMin
 * is the axis). Cases use local time steps and have no faceStates sides.
 * ------------------------------------------------------------------------ */
typedef struct bs_batch bs_batch;

bs_status bs_batch_create(bs_ctx *ctx, const bs_config *cases, int32_t count, bs_batch **out);
/* The batch of FP64 solvers (bs_solver_create_fp64): state, metrics and
 * time steps in FP64, the implicit steps' residual, diagonal and update in
 * FP64 around an FP32 relaxation with the linearized neighbour term, for
 * convergence below the FP32 residual floor (the reference's two-float
 * engine's role, euler2d-two-float.ts:1-14, 379-448). */
bs_status bs_batch_create_fp64(bs_ctx *ctx, const bs_config *cases, int32_t count, bs_batch **out);
/* BS_PRECISION_FP32 or BS_PRECISION_FP64. */
bs_status bs_batch_precision(bs_batch *b, int32_t *precision);
bs_status bs_batch_destroy(bs_batch *b);
/* Uniform start per case (setInitialCondition, euler2d-gpu.ts:412-437):
 * 4 primitives (rho, u, v, p) per case. */
bs_status bs_batch_set_uniform(bs_batch *b, const double *primitives);
/* setCellPrimitives (euler2d-gpu.ts:439-460) of one case: 4 per cell. */
bs_status bs_batch_set_primitives(bs_batch *b, int32_t index, const double *primitives);
/* The learned warm start's fields for every case at once (no counterpart
 * in the reference: its drivers start from the free stream or a given
 * field, transonic-flow.ts:174): per case 4 x coarse_nj x coarse_ni
 * primitives (rho, u, v, p; row-major, j outer) on a grid uniform in
 * (i / ni, j / nj), multiplied by the case's 4 scale factors, interpolated
 * bilinearly between cell centres (clamped at the ends) to the case's
 * cells, density and pressure held at least `floor_value`. Equal to
 * setting the same interpolated field with bs_batch_set_primitives.
 * BS_ERR_NONPHYSICAL if a field is not finite. */
bs_status bs_batch_set_coarse(bs_batch *b, const double *coarse, int32_t coarse_ni, int32_t coarse_nj,
                              const double *scale, double floor_value);
/* bs_batch_set_coarse for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference. Same arguments and
 * result. Nothing is allocated: the fields pass, as many cases at a time as
 * fit, through a fixed device scratch of BS_COARSE_SCRATCH_BYTES per device
 * that the library holds until the process ends. One case's field (4 x
 * coarse_ni x coarse_nj doubles) may take at most BS_COARSE_SCRATCH_BYTES -
 * 256 bytes; a larger one is refused (BS_ERR_RANGE) before anything runs.
 * Every call on a device shares its scratch: calls that overlap on one
 * device (on any batches) race. */
#define BS_COARSE_SCRATCH_BYTES (8u << 20)
bs_status bs_batch_set_coarse_noalloc(bs_batch *b, const double *coarse, int32_t coarse_ni, int32_t coarse_nj,
                                      const double *scale, double floor_value);
/* The coarse-to-fine start of solveTransonic (transonic-flow.ts:157-168)
 * for every case at once: case c of `b` starts from the current state of
 * case c of `coarse` (same case count and device, grids no finer than b's),
 * each cell taking the state of the coarse cell whose index range holds its
 * centre in (i / ni, j / nj): the cell it lies in, (i / r, j / r), on grids
 * of a family that nest by r. Copies device to device, converting between
 * FP32 and FP64 states; a new start (step counts and status reset, the
 * implicit CFL numbers kept). */
bs_status bs_batch_set_from_coarser(bs_batch *b, bs_batch *coarse);
/* readState (euler2d-gpu.ts:549-573) of one case; null outputs skipped. */
bs_status bs_batch_read_state(bs_batch *b, int32_t index, float *rho, float *mu, float *mv, float *energy);
/* The same with FP64 arrays: exact on an FP64 batch, widened from FP32. */
bs_status bs_batch_read_state_f64(bs_batch *b, int32_t index, double *rho, double *mu, double *mv, double *energy);
/* Cases with active[c] == 0 are skipped by later steps and reductions;
 * `count` must be the batch's case count. */
bs_status bs_batch_set_active(bs_batch *b, const uint8_t *active, int32_t count);
/* Cases of fewer cells than the threshold run on 32 x 4 tiles, the others
 * on 32 x 8 tiles, in either precision (set by the autotuner; 0 = all on
 * the large tile). */
bs_status bs_batch_set_tiling(bs_batch *b, int32_t small_tile_below_cells);
/* No counterpart in the reference (which solves its cases one after
 * another). tile 1 (32 x 4) or 2 (32 x 8): bs_batch_iterate on an FP32
 * batch runs its explicit steps in persistent launches of up to 60 steps,
 * every stage of every active case in one launch, when the active cases'
 * tiles fit on the chip at once (each case's tiles on one XCD); otherwise,
 * and with tile 0 (the default), one launch per stage. The states are
 * bitwise those of the per-stage launches. A launch whose device-side wait
 * does not complete returns BS_ERR_EXECUTION and turns persistent launches
 * off for the batch. FP64 batches refuse a non-zero tile. */
bs_status bs_batch_set_persistent(bs_batch *b, int32_t tile);
/* No counterpart in the reference. At most this many red-black half-sweeps
 * (1, 2 or 3, the default) per relaxation launch of bs_batch_iterate_implicit:
 * fused half-sweeps share one staging of the tile and its halo, and
 * recompute a halo that shrinks by one cell per half-sweep. The states are
 * bitwise the same for every setting (set by the autotuner). */
bs_status bs_batch_set_relax_fusion(bs_batch *b, int32_t half_sweeps);
/* Whether the last bs_batch_iterate ran persistent launches, the blocks a
 * launch over the current active set takes (0: it does not fit, or
 * persistent launches are off) and the blocks the chip holds at once. */
bs_status bs_batch_persistent_info(bs_batch *b, int32_t *used, int32_t *blocks, int32_t *capacity);
/* bs_set_relaxation for every case of the batch (default
 * BS_RELAX_FLUX_DIFFERENCE on an FP32 batch; an FP64 batch takes
 * BS_RELAX_LINEARIZED only). */
bs_status bs_batch_set_relaxation(bs_batch *b, int32_t form);
/* `steps` explicit steps of every active case (iterate). */
bs_status bs_batch_iterate(bs_batch *b, int32_t steps);
/* setImplicitCfl and iterateImplicit (euler2d-gpu.ts:668-723) of every
 * active case. */
bs_status bs_batch_set_implicit_cfl(bs_batch *b, double cfl);
/* Per case (one value each): a case started from a converged or predicted
 * field starts at the final CFL number (transonic-flow.ts:174). */
bs_status bs_batch_set_implicit_cfls(bs_batch *b, const double *cfl);
bs_status bs_batch_iterate_implicit(bs_batch *b, int32_t steps, double cfl_max, int32_t sweeps, double omega,
                                    double growth);
/* densityResidualNorm per case (NaN for inactive cases). */
bs_status bs_batch_residual_norms(bs_batch *b, double *norms);
/* forebodyPressureDrag over each case's jMin side (NaN when inactive). */
bs_status bs_batch_drags(bs_batch *b, const double *p_inf, const double *q_inf, const double *base_radius,
                         double *cd);
/* Status word (bit 0: non-physical) and step count per case. */
bs_status bs_batch_status(bs_batch *b, uint32_t *status, int64_t *steps);
/* bs_enable_timing / bs_kernel_timing for a batch: with timing enabled the
 * steps run without graphs and a pair of HIP events brackets each pass over
 * all active cases (kernel 0: an explicit stage, 1: the implicit prepare,
 * 2: a relaxation launch); bs_batch_kernel_timing reports the mean
 * milliseconds per pass since the last call. */
bs_status bs_batch_enable_timing(bs_batch *b, int32_t enable);
bs_status bs_batch_kernel_timing(bs_batch *b, int32_t kernel, double *ms_per_launch, int64_t *launches);

/* ------------------------------------------------------------------------
 * Learned warm start. No counterpart in the reference, which contains no
 * learned components: a multilayer perceptron, inputs -> hidden (SiLU) x
 * depth -> outputs (linear), trained with BF16 MFMA GEMM kernels (FP32
 * accumulation), FP32 master weights and Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8) on a mean-squared-error loss. hidden, outputs and the batch are
 * multiples of 64. Parameters travel as FP32 in the order: per layer W
 * [in][out] row-major (in = inputs for the first layer, else hidden) then
 * b [out]. A new network has all parameters zero until bs_mlp_set_params.
 * ------------------------------------------------------------------------ */
typedef struct bs_mlp bs_mlp;

bs_status bs_mlp_create(bs_ctx *ctx, int32_t inputs, int32_t hidden, int32_t depth, int32_t outputs,
                        int32_t max_batch, bs_mlp **out);
bs_status bs_mlp_destroy(bs_mlp *p);
int64_t bs_mlp_param_count(const bs_mlp *p);
/* Sets the parameters and resets Adam's moments and step count. */
bs_status bs_mlp_set_params(bs_mlp *p, const float *params, int64_t count);
bs_status bs_mlp_get_params(bs_mlp *p, float *params, int64_t count);
/* Copies a data set to the device: rows x inputs and rows x outputs FP32. */
bs_status bs_mlp_load_data(bs_mlp *p, const float *x, const float *y, int32_t rows);
/* Loss (mean squared error) and gradients (parameter order) over the data
 * rows idx[0 .. batch), without an update. */
bs_status bs_mlp_gradients(bs_mlp *p, const int32_t *idx, int32_t batch, float *grads, double *loss);
/* One Adam step on the data rows idx[0 .. batch); `loss` is the batch's
 * loss before the step. */
bs_status bs_mlp_train_batch(bs_mlp *p, const int32_t *idx, int32_t batch, double lr, double *loss);
/* Outputs for rows x inputs host inputs, in blocks of the maximum batch. */
bs_status bs_mlp_predict(bs_mlp *p, const float *x, int32_t rows, float *y);
/* Device time of one forward GEMM of `layer` at the maximum batch (HIP
 * events over `reps` launches) and its floating-point operations. */
bs_status bs_mlp_time_gemm(bs_mlp *p, int32_t layer, int32_t reps, double *ms_per_gemm, double *flops_per_gemm);

#ifdef __cplusplus
}
#endif
#endif /* BOWSHOCK_H */
