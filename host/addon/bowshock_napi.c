/*
 * N-API addon: the thin JavaScript binding of include/bowshock.h.
 *
 * Every function takes and returns plain JS values and typed arrays and
 * calls one C-ABI entry point; the long-running ones (iterate, advanceTo,
 * iterateImplicit) run on a worker thread through napi_async_work and
 * return a Promise, so Node's event loop is not blocked. A failure rejects
 * or throws an Error whose `code` is the bs_status name; host/index.js maps
 * it to the reference's error classes.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bowshock.h"

#define CHECK_NAPI(call)                                   \
  do {                                                     \
    if ((call) != napi_ok) {                               \
      napi_throw_error(env, NULL, "N-API call failed");    \
      return NULL;                                         \
    }                                                      \
  } while (0)

static const char *status_name(bs_status st) {
  switch (st) {
    case BS_OK: return "BS_OK";
    case BS_ERR_RANGE: return "BS_ERR_RANGE";
    case BS_ERR_NONPHYSICAL: return "BS_ERR_NONPHYSICAL";
    case BS_ERR_MEMORY: return "BS_ERR_MEMORY";
    case BS_ERR_EXECUTION: return "BS_ERR_EXECUTION";
    case BS_ERR_PRECISION: return "BS_ERR_PRECISION";
    case BS_ERR_BUSY: return "BS_ERR_BUSY";
    default: return "BS_ERR_UNSUPPORTED";
  }
}

static napi_value make_error(napi_env env, bs_status st, const char *msg) {
  napi_value code, text, err;
  napi_create_string_utf8(env, status_name(st), NAPI_AUTO_LENGTH, &code);
  napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &text);
  napi_create_error(env, code, text, &err);
  return err;
}

static napi_value throw_status(napi_env env, bs_status st) {
  napi_throw(env, make_error(env, st, bs_last_error()));
  return NULL;
}

static void *get_ptr(napi_env env, napi_value v) {
  void *p = NULL;
  napi_get_value_external(env, v, &p);
  return p;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < want) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return 0;
  }
  return 1;
}

static double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

static int32_t i32(napi_env env, napi_value v) {
  int32_t d = 0;
  napi_get_value_int32(env, v, &d);
  return d;
}

/* Data pointer and element count of a typed array of the given type. */
static void *typed(napi_env env, napi_value v, napi_typedarray_type want, size_t *count) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return NULL;
  napi_typedarray_type type;
  size_t length = 0, offset = 0;
  void *data = NULL;
  napi_value buffer;
  napi_get_typedarray_info(env, v, &type, &length, &data, &buffer, &offset);
  if (type != want) return NULL;
  if (count) *count = length;
  return data;
}

static napi_value report_object(napi_env env, const bs_report *r) {
  napi_value o, v;
  napi_create_object(env, &o);
  napi_create_double(env, (double)r->steps, &v);
  napi_set_named_property(env, o, "steps", v);
  napi_create_uint32(env, r->status, &v);
  napi_set_named_property(env, o, "status", v);
  napi_create_double(env, r->time, &v);
  napi_set_named_property(env, o, "time", v);
  napi_create_double(env, r->dt, &v);
  napi_set_named_property(env, o, "dt", v);
  return o;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value s;
  napi_create_string_utf8(env, bs_version(), NAPI_AUTO_LENGTH, &s);
  return s;
}

static napi_value js_ctx_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bs_ctx *ctx = NULL;
  bs_status st = bs_ctx_create(i32(env, argv[0]), &ctx);
  if (st) return throw_status(env, st);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, ctx, NULL, NULL, &ext));
  return ext;
}

static napi_value js_ctx_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bs_ctx_destroy((bs_ctx *)get_ptr(env, argv[0]));
  return NULL;
}

/* The plain-data configuration {ni, nj, x: Float64Array, y: Float64Array,
 * geometry, gamma, flux, limiter, reconstruction, integrator, cfl,
 * timeStepping, boundaries: [{kind, rho, u, v, p} x 4]} with integer codes
 * as bs_config; x and y stay owned by the JS values. 0 (thrown) on a bad
 * grid. */
static int parse_config(napi_env env, napi_value c, bs_config *cfg) {
  napi_value v;
  memset(cfg, 0, sizeof *cfg);
#define FIELD(name) (napi_get_named_property(env, c, name, &v), v)
  cfg->ni = i32(env, FIELD("ni"));
  cfg->nj = i32(env, FIELD("nj"));
  size_t nx = 0, ny = 0;
  cfg->x = (const double *)typed(env, FIELD("x"), napi_float64_array, &nx);
  cfg->y = (const double *)typed(env, FIELD("y"), napi_float64_array, &ny);
  size_t nodes = (size_t)(cfg->ni + 1) * (size_t)(cfg->nj + 1);
  if (!cfg->x || !cfg->y || nx != nodes || ny != nodes) {
    napi_throw(env, make_error(env, BS_ERR_RANGE, "x and y must be Float64Arrays of (ni + 1) (nj + 1) nodes"));
    return 0;
  }
  cfg->geometry = i32(env, FIELD("geometry"));
  cfg->gamma = num(env, FIELD("gamma"));
  cfg->flux = i32(env, FIELD("flux"));
  cfg->limiter = i32(env, FIELD("limiter"));
  cfg->reconstruction = i32(env, FIELD("reconstruction"));
  cfg->integrator = i32(env, FIELD("integrator"));
  cfg->cfl = num(env, FIELD("cfl"));
  cfg->time_stepping = i32(env, FIELD("timeStepping"));
  napi_value bcs = FIELD("boundaries");
  for (uint32_t k = 0; k < 4; k++) {
    napi_value b;
    napi_get_element(env, bcs, k, &b);
    napi_value f;
    napi_get_named_property(env, b, "kind", &f);
    cfg->boundaries[k].kind = i32(env, f);
    napi_get_named_property(env, b, "rho", &f);
    cfg->boundaries[k].rho = num(env, f);
    napi_get_named_property(env, b, "u", &f);
    cfg->boundaries[k].u = num(env, f);
    napi_get_named_property(env, b, "v", &f);
    cfg->boundaries[k].v = num(env, f);
    napi_get_named_property(env, b, "p", &f);
    cfg->boundaries[k].p = num(env, f);
  }
#undef FIELD
  return 1;
}

/* solverCreate(ctx, cfg[, precision]): precision BS_PRECISION_FP32
 * (default) or _FP64. */
static napi_value js_solver_create(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  size_t argc = 3;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < 2) {
    napi_throw_type_error(env, NULL, "missing arguments");
    return NULL;
  }
  const int precision = argc > 2 ? i32(env, argv[2]) : BS_PRECISION_FP32;
  bs_ctx *ctx = (bs_ctx *)get_ptr(env, argv[0]);
  bs_config cfg;
  if (!parse_config(env, argv[1], &cfg)) return NULL;
  bs_solver *s = NULL;
  bs_status st = precision == BS_PRECISION_FP64 ? bs_solver_create_fp64(ctx, &cfg, &s) : bs_solver_create(ctx, &cfg, &s);
  if (st) return throw_status(env, st);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, s, NULL, NULL, &ext));
  return ext;
}

static napi_value js_solver_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bs_status st = bs_solver_destroy((bs_solver *)get_ptr(env, argv[0]));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value new_f64(napi_env env, size_t n, double **data) {
  napi_value buf, arr;
  napi_create_arraybuffer(env, n * sizeof(double), (void **)data, &buf);
  napi_create_typedarray(env, napi_float64_array, n, buf, 0, &arr);
  return arr;
}

static napi_value new_f32(napi_env env, size_t n, float **data) {
  napi_value buf, arr;
  napi_create_arraybuffer(env, n * sizeof(float), (void **)data, &buf);
  napi_create_typedarray(env, napi_float32_array, n, buf, 0, &arr);
  return arr;
}

static napi_value js_compute_metrics(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  int32_t ni = i32(env, argv[0]), nj = i32(env, argv[1]);
  size_t nx = 0, ny = 0;
  const double *x = (const double *)typed(env, argv[2], napi_float64_array, &nx);
  const double *y = (const double *)typed(env, argv[3], napi_float64_array, &ny);
  if (ni < 1 || nj < 1 || !x || !y || nx != (size_t)(ni + 1) * (nj + 1) || ny != nx) {
    napi_throw(env, make_error(env, BS_ERR_RANGE, "bad grid"));
    return NULL;
  }
  size_t fi = (size_t)(ni + 1) * nj, fj = (size_t)ni * (nj + 1), n = (size_t)ni * nj;
  const char *names[10] = {"iNx", "iNy", "iArea", "jNx", "jNy", "jArea", "volume", "planarArea", "centroidX",
                           "centroidY"};
  size_t sizes[10] = {fi, fi, fi, fj, fj, fj, n, n, n, n};
  double *ptr[10];
  napi_value out, arr[10];
  napi_create_object(env, &out);
  for (int k = 0; k < 10; k++) {
    arr[k] = new_f64(env, sizes[k], &ptr[k]);
    napi_set_named_property(env, out, names[k], arr[k]);
  }
  bs_status st = bs_compute_metrics(ni, nj, x, y, i32(env, argv[4]), ptr[0], ptr[1], ptr[2], ptr[3], ptr[4],
                                    ptr[5], ptr[6], ptr[7], ptr[8], ptr[9]);
  if (st) return throw_status(env, st);
  return out;
}

static napi_value js_set_primitives(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  const double *p = (const double *)typed(env, argv[1], napi_float64_array, NULL);
  if (!p) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_set_primitives((bs_solver *)get_ptr(env, argv[0]), p);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value js_set_state(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  const float *q[4];
  size_t len[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; k++) {
    q[k] = (const float *)typed(env, argv[1 + k], napi_float32_array, &len[k]);
    if (!q[k]) return throw_status(env, BS_ERR_RANGE);
  }
  const size_t n = len[0];
  for (int k = 1; k < 4; k++)
    if (len[k] != n) {
      napi_throw_range_error(env, NULL, "the four state arrays must have one length");
      return NULL;
    }
  bs_status st = bs_set_state((bs_solver *)get_ptr(env, argv[0]), q[0], q[1], q[2], q[3],
                              n > INT32_MAX ? -1 : (int32_t)n);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value js_read_state(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  size_t n = (size_t)num(env, argv[1]);
  float *q[4];
  napi_value out, arr[4];
  napi_create_object(env, &out);
  const char *names[4] = {"rho", "mu", "mv", "energy"};
  for (int k = 0; k < 4; k++) {
    arr[k] = new_f32(env, n, &q[k]);
    napi_set_named_property(env, out, names[k], arr[k]);
  }
  bs_report r;
  memset(&r, 0, sizeof r);
  bs_status st = bs_read_state((bs_solver *)get_ptr(env, argv[0]), q[0], q[1], q[2], q[3], &r);
  if (st) return throw_status(env, st);
  napi_set_named_property(env, out, "report", report_object(env, &r));
  return out;
}

/* readState64(solver, n): the state as Float64Arrays (bs_read_state_f64). */
static napi_value js_read_state_f64(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  size_t n = (size_t)num(env, argv[1]);
  double *q[4];
  napi_value out, arr[4];
  napi_create_object(env, &out);
  const char *names[4] = {"rho", "mu", "mv", "energy"};
  for (int k = 0; k < 4; k++) {
    arr[k] = new_f64(env, n, &q[k]);
    napi_set_named_property(env, out, names[k], arr[k]);
  }
  bs_report r;
  memset(&r, 0, sizeof r);
  bs_status st = bs_read_state_f64((bs_solver *)get_ptr(env, argv[0]), q[0], q[1], q[2], q[3], &r);
  if (st) return throw_status(env, st);
  napi_set_named_property(env, out, "report", report_object(env, &r));
  return out;
}

static napi_value js_set_face_states(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  size_t n = 0;
  const double *p = (const double *)typed(env, argv[2], napi_float64_array, &n);
  if (!p) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_set_face_states((bs_solver *)get_ptr(env, argv[0]), i32(env, argv[1]), p, (int32_t)(n / 4));
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value js_set_face_fluxes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  napi_valuetype t;
  napi_typeof(env, argv[2], &t);
  bs_status st;
  if (t == napi_null || t == napi_undefined) {
    st = bs_set_face_fluxes((bs_solver *)get_ptr(env, argv[0]), i32(env, argv[1]), NULL, 0);
  } else {
    size_t n = 0;
    const double *p = (const double *)typed(env, argv[2], napi_float64_array, &n);
    if (!p) return throw_status(env, BS_ERR_RANGE);
    st = bs_set_face_fluxes((bs_solver *)get_ptr(env, argv[0]), i32(env, argv[1]), p, (int32_t)n);
  }
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value js_residual_norm(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  double norm = 0;
  bs_status st = bs_residual_norm((bs_solver *)get_ptr(env, argv[0]), &norm);
  if (st) return throw_status(env, st);
  napi_value v;
  napi_create_double(env, norm, &v);
  return v;
}

/* checkPrecision(solver): the relative difference; throws
 * GpuPrecisionError (BS_ERR_PRECISION) above the bar. */
static napi_value js_check_precision(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  double diff = 0;
  bs_status st = bs_check_precision((bs_solver *)get_ptr(env, argv[0]), &diff);
  if (st) return throw_status(env, st);
  napi_value v;
  napi_create_double(env, diff, &v);
  return v;
}

static napi_value js_residual_vector(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  size_t n = (size_t)num(env, argv[1]);
  double *r;
  napi_value arr = new_f64(env, 4 * n, &r);
  bs_status st = bs_residual_vector((bs_solver *)get_ptr(env, argv[0]), r);
  if (st) return throw_status(env, st);
  return arr;
}

static napi_value js_wall_fluxes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  int32_t count = i32(env, argv[2]);
  double *f;
  napi_value arr = new_f64(env, 9 * (size_t)(count > 0 ? count : 0), &f);
  bs_status st = bs_wall_fluxes((bs_solver *)get_ptr(env, argv[0]), i32(env, argv[1]), f, count);
  if (st) return throw_status(env, st);
  return arr;
}

static napi_value js_forebody_drag(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  double cd = 0;
  bs_status st = bs_forebody_drag((bs_solver *)get_ptr(env, argv[0]), i32(env, argv[1]), num(env, argv[2]),
                                  num(env, argv[3]), num(env, argv[4]), &cd);
  if (st) return throw_status(env, st);
  napi_value v;
  napi_create_double(env, cd, &v);
  return v;
}

static napi_value js_set_implicit_cfl(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bs_status st = bs_set_implicit_cfl((bs_solver *)get_ptr(env, argv[0]), num(env, argv[1]));
  if (st) return throw_status(env, st);
  return NULL;
}

/* ---- asynchronous steps ------------------------------------------------ */

enum { OP_ITERATE, OP_ADVANCE, OP_IMPLICIT, OP_FITTED };

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  bs_solver *s;
  int op;
  int32_t count, sweeps, max_steps;
  double t_end, cfl, omega, growth;
  bs_status st;
  char message[512];
  bs_report report;
} StepJob;

static void step_execute(napi_env env, void *data) {
  (void)env;
  StepJob *j = (StepJob *)data;
  memset(&j->report, 0, sizeof j->report);
  if (j->op == OP_ITERATE) j->st = bs_iterate(j->s, j->count, &j->report);
  else if (j->op == OP_ADVANCE) j->st = bs_advance_to(j->s, j->t_end, j->max_steps, &j->report);
  else if (j->op == OP_FITTED) j->st = bs_iterate_fitted(j->s, j->count, j->sweeps, &j->report);
  else j->st = bs_iterate_implicit(j->s, j->count, j->cfl, j->sweeps, j->omega, j->growth, &j->report);
  /* The message is thread-local to the worker: copy it here. */
  strncpy(j->message, bs_last_error(), sizeof j->message - 1);
  j->message[sizeof j->message - 1] = 0;
}

static void step_complete(napi_env env, napi_status status, void *data) {
  StepJob *j = (StepJob *)data;
  if (status != napi_ok) {
    napi_reject_deferred(env, j->deferred, make_error(env, BS_ERR_EXECUTION, "the step job was cancelled"));
  } else if (j->st != BS_OK) {
    napi_reject_deferred(env, j->deferred, make_error(env, j->st, j->message));
  } else {
    napi_resolve_deferred(env, j->deferred, report_object(env, &j->report));
  }
  napi_delete_async_work(env, j->work);
  free(j);
}

static napi_value queue_job(napi_env env, StepJob *j) {
  napi_value promise, name;
  napi_create_promise(env, &j->deferred, &promise);
  napi_create_string_utf8(env, "bowshock-steps", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, step_execute, step_complete, j, &j->work);
  napi_queue_async_work(env, j->work);
  return promise;
}

static napi_value js_iterate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  StepJob *j = (StepJob *)calloc(1, sizeof *j);
  j->s = (bs_solver *)get_ptr(env, argv[0]);
  j->op = OP_ITERATE;
  j->count = i32(env, argv[1]);
  return queue_job(env, j);
}

static napi_value js_advance_to(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  StepJob *j = (StepJob *)calloc(1, sizeof *j);
  j->s = (bs_solver *)get_ptr(env, argv[0]);
  j->op = OP_ADVANCE;
  j->t_end = num(env, argv[1]);
  j->max_steps = i32(env, argv[2]);
  return queue_job(env, j);
}

static napi_value js_iterate_implicit(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  StepJob *j = (StepJob *)calloc(1, sizeof *j);
  j->s = (bs_solver *)get_ptr(env, argv[0]);
  j->op = OP_IMPLICIT;
  j->count = i32(env, argv[1]);
  j->cfl = num(env, argv[2]);
  j->sweeps = i32(env, argv[3]);
  j->omega = num(env, argv[4]);
  j->growth = num(env, argv[5]);
  return queue_job(env, j);
}

/* iterateFitted(solver, moves, stepsPerMove): a promise of the report. */
static napi_value js_iterate_fitted(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  StepJob *j = (StepJob *)calloc(1, sizeof *j);
  j->s = (bs_solver *)get_ptr(env, argv[0]);
  j->op = OP_FITTED;
  j->count = i32(env, argv[1]);
  j->sweeps = i32(env, argv[2]);
  return queue_job(env, j);
}

/* setFittedShock(solver, Float64Array(4) | null). */
static napi_value js_set_fitted_shock(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  napi_valuetype t;
  napi_typeof(env, argv[1], &t);
  const double *w = NULL;
  if (t != napi_null && t != napi_undefined) {
    size_t n = 0;
    w = (const double *)typed(env, argv[1], napi_float64_array, &n);
    if (!w || n != 4) return throw_status(env, BS_ERR_RANGE);
  }
  bs_status st = bs_set_fitted_shock((bs_solver *)get_ptr(env, argv[0]), w);
  if (st) return throw_status(env, st);
  return NULL;
}

/* readShock(solver, ni): {speed, mach, signal} as Float64Arrays. */
static napi_value js_read_shock(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  int32_t ni = i32(env, argv[1]);
  double *q[3];
  napi_value out, arr[3];
  napi_create_object(env, &out);
  const char *names[3] = {"speed", "mach", "signal"};
  for (int k = 0; k < 3; k++) {
    arr[k] = new_f64(env, (size_t)(ni > 0 ? ni : 0), &q[k]);
    napi_set_named_property(env, out, names[k], arr[k]);
  }
  bs_status st = bs_read_shock((bs_solver *)get_ptr(env, argv[0]), q[0], q[1], q[2], ni);
  if (st) return throw_status(env, st);
  return out;
}

/* setLineGrid(solver, lines Float64Array, fractions Float64Array, symmetry). */
static napi_value js_set_line_grid(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  size_t nl = 0, nf = 0;
  const double *l = (const double *)typed(env, argv[1], napi_float64_array, &nl);
  const double *f = (const double *)typed(env, argv[2], napi_float64_array, &nf);
  if (!l || !f) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_set_line_grid((bs_solver *)get_ptr(env, argv[0]), l, f, i32(env, argv[3]));
  if (st) return throw_status(env, st);
  return NULL;
}

/* setHeights(solver, heights) / moveGrid(solver, heights). */
static napi_value heights_call(napi_env env, napi_callback_info info, int move) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  size_t n = 0;
  const double *h = (const double *)typed(env, argv[1], napi_float64_array, &n);
  if (!h) return throw_status(env, BS_ERR_RANGE);
  bs_solver *s = (bs_solver *)get_ptr(env, argv[0]);
  const int32_t count = n > INT32_MAX ? -1 : (int32_t)n;
  bs_status st = move ? bs_move_grid(s, h, count) : bs_set_heights(s, h, count);
  if (st) return throw_status(env, st);
  return NULL;
}

static napi_value js_set_heights(napi_env env, napi_callback_info info) { return heights_call(env, info, 0); }
static napi_value js_move_grid(napi_env env, napi_callback_info info) { return heights_call(env, info, 1); }

/* readHeights(solver, ni + 1): Float64Array. */
static napi_value js_read_heights(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  int32_t n = i32(env, argv[1]);
  double *h;
  napi_value arr = new_f64(env, (size_t)(n > 0 ? n : 0), &h);
  bs_status st = bs_read_heights((bs_solver *)get_ptr(env, argv[0]), h, n);
  if (st) return throw_status(env, st);
  return arr;
}

/* ---- batches (bs_batch_*) ------------------------------------------- */

/* batchCreate(ctx, [cfg, ...]): the configurations as for solverCreate. */
static napi_value js_batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bs_ctx *ctx = (bs_ctx *)get_ptr(env, argv[0]);
  uint32_t count = 0;
  napi_get_array_length(env, argv[1], &count);
  if (count == 0) {
    napi_throw(env, make_error(env, BS_ERR_RANGE, "a batch needs at least one case"));
    return NULL;
  }
  bs_config *cfgs = (bs_config *)calloc(count, sizeof(bs_config));
  if (!cfgs) {
    napi_throw(env, make_error(env, BS_ERR_MEMORY, "out of host memory"));
    return NULL;
  }
  for (uint32_t k = 0; k < count; k++) {
    napi_value c;
    napi_get_element(env, argv[1], k, &c);
    if (!parse_config(env, c, &cfgs[k])) {
      free(cfgs);
      return NULL;
    }
  }
  bs_batch *b = NULL;
  bs_status st = bs_batch_create(ctx, cfgs, (int32_t)count, &b);
  free(cfgs);
  if (st) return throw_status(env, st);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, b, NULL, NULL, &ext));
  return ext;
}

static napi_value js_batch_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  bs_status st = bs_batch_destroy((bs_batch *)get_ptr(env, argv[0]));
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchSetUniform(b, Float64Array 4 per case). */
static napi_value js_batch_set_uniform(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  const double *w = (const double *)typed(env, argv[1], napi_float64_array, NULL);
  if (!w) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_batch_set_uniform((bs_batch *)get_ptr(env, argv[0]), w);
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchSetPrimitives(b, index, Float64Array 4 per cell). */
static napi_value js_batch_set_primitives(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  const double *w = (const double *)typed(env, argv[2], napi_float64_array, NULL);
  if (!w) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_batch_set_primitives((bs_batch *)get_ptr(env, argv[0]), i32(env, argv[1]), w);
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchReadState(b, index, cells) -> {rho, mu, mv, energy} Float32Arrays. */
static napi_value js_batch_read_state(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  const int32_t n = i32(env, argv[2]);
  float *q[4];
  napi_value arr[4], o;
  for (int k = 0; k < 4; k++) arr[k] = new_f32(env, (size_t)(n > 0 ? n : 0), &q[k]);
  bs_status st = bs_batch_read_state((bs_batch *)get_ptr(env, argv[0]), i32(env, argv[1]), q[0], q[1], q[2], q[3]);
  if (st) return throw_status(env, st);
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "rho", arr[0]);
  napi_set_named_property(env, o, "mu", arr[1]);
  napi_set_named_property(env, o, "mv", arr[2]);
  napi_set_named_property(env, o, "energy", arr[3]);
  return o;
}

/* batchSetCoarse(b, Float64Array coarse, cni, cnj, Float64Array scale,
 * floor): bs_batch_set_coarse. */
static napi_value js_batch_set_coarse(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  const double *coarse = (const double *)typed(env, argv[1], napi_float64_array, NULL);
  const double *scale = (const double *)typed(env, argv[4], napi_float64_array, NULL);
  if (!coarse || !scale) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_batch_set_coarse((bs_batch *)get_ptr(env, argv[0]), coarse, i32(env, argv[2]), i32(env, argv[3]),
                                     scale, num(env, argv[5]));
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchSetActive(b, Uint8Array one per case). */
static napi_value js_batch_set_active(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  size_t n = 0;
  const uint8_t *a = (const uint8_t *)typed(env, argv[1], napi_uint8_array, &n);
  if (!a) return throw_status(env, BS_ERR_RANGE);
  bs_status st = bs_batch_set_active((bs_batch *)get_ptr(env, argv[0]), a, n > INT32_MAX ? -1 : (int32_t)n);
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchSetInt(b, what, value): 0 tiling threshold, 1 relaxation form,
 * 2 persistent tile, 3 half-sweeps per relaxation launch. */
static napi_value js_batch_set_int(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  bs_batch *b = (bs_batch *)get_ptr(env, argv[0]);
  const int32_t v = i32(env, argv[2]);
  bs_status st;
  switch (i32(env, argv[1])) {
    case 0: st = bs_batch_set_tiling(b, v); break;
    case 1: st = bs_batch_set_relaxation(b, v); break;
    case 2: st = bs_batch_set_persistent(b, v); break;
    case 3: st = bs_batch_set_relax_fusion(b, v); break;
    default: st = BS_ERR_RANGE; break;
  }
  if (st) return throw_status(env, st);
  return NULL;
}

/* batchSetImplicitCfl(b, number | Float64Array one per case). */
static napi_value js_batch_set_implicit_cfl(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bs_batch *b = (bs_batch *)get_ptr(env, argv[0]);
  const double *per = (const double *)typed(env, argv[1], napi_float64_array, NULL);
  bs_status st = per ? bs_batch_set_implicit_cfls(b, per) : bs_batch_set_implicit_cfl(b, num(env, argv[1]));
  if (st) return throw_status(env, st);
  return NULL;
}

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  bs_batch *b;
  int implicit;
  int32_t steps, sweeps;
  double cfl, omega, growth;
  bs_status st;
  char message[512];
} BatchJob;

static void batch_execute(napi_env env, void *data) {
  (void)env;
  BatchJob *j = (BatchJob *)data;
  j->st = j->implicit ? bs_batch_iterate_implicit(j->b, j->steps, j->cfl, j->sweeps, j->omega, j->growth)
                      : bs_batch_iterate(j->b, j->steps);
  strncpy(j->message, bs_last_error(), sizeof j->message - 1);
  j->message[sizeof j->message - 1] = 0;
}

static void batch_complete(napi_env env, napi_status status, void *data) {
  BatchJob *j = (BatchJob *)data;
  if (status != napi_ok) {
    napi_reject_deferred(env, j->deferred, make_error(env, BS_ERR_EXECUTION, "the batch job was cancelled"));
  } else if (j->st != BS_OK) {
    napi_reject_deferred(env, j->deferred, make_error(env, j->st, j->message));
  } else {
    napi_value undefined;
    napi_get_undefined(env, &undefined);
    napi_resolve_deferred(env, j->deferred, undefined);
  }
  napi_delete_async_work(env, j->work);
  free(j);
}

/* batchIterate(b, steps) and batchIterateImplicit(b, steps, cfl, sweeps,
 * omega, growth): Promises, on a worker thread. */
static napi_value batch_queue(napi_env env, napi_callback_info info, int implicit) {
  napi_value argv[6];
  if (!get_args(env, info, implicit ? 6 : 2, argv)) return NULL;
  BatchJob *j = (BatchJob *)calloc(1, sizeof(BatchJob));
  if (!j) {
    napi_throw(env, make_error(env, BS_ERR_MEMORY, "out of host memory"));
    return NULL;
  }
  j->b = (bs_batch *)get_ptr(env, argv[0]);
  j->implicit = implicit;
  j->steps = i32(env, argv[1]);
  if (implicit) {
    j->cfl = num(env, argv[2]);
    j->sweeps = i32(env, argv[3]);
    j->omega = num(env, argv[4]);
    j->growth = num(env, argv[5]);
  }
  napi_value promise, name;
  napi_create_promise(env, &j->deferred, &promise);
  napi_create_string_utf8(env, "bowshock-batch", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, NULL, name, batch_execute, batch_complete, j, &j->work);
  napi_queue_async_work(env, j->work);
  return promise;
}

static napi_value js_batch_iterate(napi_env env, napi_callback_info info) { return batch_queue(env, info, 0); }
static napi_value js_batch_iterate_implicit(napi_env env, napi_callback_info info) {
  return batch_queue(env, info, 1);
}

/* batchResidualNorms(b, count) -> Float64Array. */
static napi_value js_batch_residual_norms(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  const int32_t n = i32(env, argv[1]);
  double *r;
  napi_value arr = new_f64(env, (size_t)(n > 0 ? n : 0), &r);
  bs_status st = bs_batch_residual_norms((bs_batch *)get_ptr(env, argv[0]), r);
  if (st) return throw_status(env, st);
  return arr;
}

/* batchDrags(b, pInf, qInf, baseRadius) (Float64Arrays, one per case) ->
 * Float64Array. */
static napi_value js_batch_drags(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  size_t n = 0, n2 = 0, n3 = 0;
  const double *p = (const double *)typed(env, argv[1], napi_float64_array, &n);
  const double *q = (const double *)typed(env, argv[2], napi_float64_array, &n2);
  const double *R = (const double *)typed(env, argv[3], napi_float64_array, &n3);
  if (!p || !q || !R || n2 != n || n3 != n) return throw_status(env, BS_ERR_RANGE);
  double *cd;
  napi_value arr = new_f64(env, n, &cd);
  bs_status st = bs_batch_drags((bs_batch *)get_ptr(env, argv[0]), p, q, R, cd);
  if (st) return throw_status(env, st);
  return arr;
}

/* batchStatus(b, count) -> {status: Uint32Array, steps: Float64Array}. */
static napi_value js_batch_status(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  const int32_t n = i32(env, argv[1]);
  const size_t m = (size_t)(n > 0 ? n : 0);
  uint32_t *status = (uint32_t *)calloc(m ? m : 1, sizeof(uint32_t));
  int64_t *steps = (int64_t *)calloc(m ? m : 1, sizeof(int64_t));
  if (!status || !steps) {
    free(status);
    free(steps);
    napi_throw(env, make_error(env, BS_ERR_MEMORY, "out of host memory"));
    return NULL;
  }
  bs_status st = bs_batch_status((bs_batch *)get_ptr(env, argv[0]), status, steps);
  if (st) {
    free(status);
    free(steps);
    return throw_status(env, st);
  }
  napi_value o, sbuf, sarr, steps_arr;
  void *sdata = NULL;
  napi_create_arraybuffer(env, m * sizeof(uint32_t), &sdata, &sbuf);
  memcpy(sdata, status, m * sizeof(uint32_t));
  napi_create_typedarray(env, napi_uint32_array, m, sbuf, 0, &sarr);
  double *sd;
  steps_arr = new_f64(env, m, &sd);
  for (size_t k = 0; k < m; k++) sd[k] = (double)steps[k];
  free(status);
  free(steps);
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "status", sarr);
  napi_set_named_property(env, o, "steps", steps_arr);
  return o;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"version", NULL, js_version, NULL, NULL, NULL, napi_enumerable, NULL},
      {"ctxCreate", NULL, js_ctx_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"ctxDestroy", NULL, js_ctx_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"solverCreate", NULL, js_solver_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"solverDestroy", NULL, js_solver_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"computeMetrics", NULL, js_compute_metrics, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setPrimitives", NULL, js_set_primitives, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setState", NULL, js_set_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readState", NULL, js_read_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readState64", NULL, js_read_state_f64, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setFaceStates", NULL, js_set_face_states, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setFaceFluxes", NULL, js_set_face_fluxes, NULL, NULL, NULL, napi_enumerable, NULL},
      {"residualNorm", NULL, js_residual_norm, NULL, NULL, NULL, napi_enumerable, NULL},
      {"checkPrecision", NULL, js_check_precision, NULL, NULL, NULL, napi_enumerable, NULL},
      {"residualVector", NULL, js_residual_vector, NULL, NULL, NULL, napi_enumerable, NULL},
      {"wallFluxes", NULL, js_wall_fluxes, NULL, NULL, NULL, napi_enumerable, NULL},
      {"forebodyDrag", NULL, js_forebody_drag, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setImplicitCfl", NULL, js_set_implicit_cfl, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iterate", NULL, js_iterate, NULL, NULL, NULL, napi_enumerable, NULL},
      {"advanceTo", NULL, js_advance_to, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iterateImplicit", NULL, js_iterate_implicit, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iterateFitted", NULL, js_iterate_fitted, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setFittedShock", NULL, js_set_fitted_shock, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readShock", NULL, js_read_shock, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setLineGrid", NULL, js_set_line_grid, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setHeights", NULL, js_set_heights, NULL, NULL, NULL, napi_enumerable, NULL},
      {"moveGrid", NULL, js_move_grid, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readHeights", NULL, js_read_heights, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchCreate", NULL, js_batch_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchDestroy", NULL, js_batch_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetUniform", NULL, js_batch_set_uniform, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetPrimitives", NULL, js_batch_set_primitives, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchReadState", NULL, js_batch_read_state, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetActive", NULL, js_batch_set_active, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetCoarse", NULL, js_batch_set_coarse, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetInt", NULL, js_batch_set_int, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchSetImplicitCfl", NULL, js_batch_set_implicit_cfl, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchIterate", NULL, js_batch_iterate, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchIterateImplicit", NULL, js_batch_iterate_implicit, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchResidualNorms", NULL, js_batch_residual_norms, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchDrags", NULL, js_batch_drags, NULL, NULL, NULL, napi_enumerable, NULL},
      {"batchStatus", NULL, js_batch_status, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
