'use strict';
/**
 * The HIP engine behind the reference's TypeScript engine interface, for
 * Node: GpuEuler2D with the method set of src/engine-gpu/euler2d-gpu.ts
 * (class GpuEuler2D) and the engine adapters its drivers consume
 * (src/analysis/verified-drag.ts:22-37 EngineFactory / SteadyEngine;
 * src/analysis/transonic-drag.ts:46-87 TransonicEngine).
 *
 * This file is the erasure of index.d.ts (types in index.d.ts, plain
 * CommonJS here) so that it runs on any Node with N-API 8, with or
 * without a TypeScript toolchain. It reaches the GPU only through the
 * N-API addon over include/bowshock.h; there is no other backend.
 */

const path = require('path');

let addon = null;
/** The native addon; throws when it was not built (no fallback). */
function native() {
  if (addon === null) {
    const file = process.env.BOWSHOCK_ADDON || path.join(__dirname, '..', 'bowshock_amd', 'lib', 'bowshock_napi.node');
    addon = require(file);
  }
  return addon;
}

// ---- errors (src/solver/errors.ts:2-7, src/engine-gpu/gpu-common.ts:7-17) ----

class NonPhysicalStateError extends Error {
  constructor(message) {
    super(message);
    this.name = 'NonPhysicalStateError';
  }
}
class GpuMemoryLimitError extends Error {
  constructor(message) {
    super(message);
    this.name = 'GpuMemoryLimitError';
  }
}
class GpuExecutionError extends Error {
  constructor(message) {
    super(message);
    this.name = 'GpuExecutionError';
  }
}
class GpuPrecisionError extends Error {
  constructor(message) {
    super(message);
    this.name = 'GpuPrecisionError';
  }
}

/** Map an addon error (its `code` is the bs_status name) to the reference's class. */
function translate(error) {
  if (!error || typeof error.code !== 'string') return error;
  switch (error.code) {
    case 'BS_ERR_RANGE':
      return new RangeError(error.message);
    case 'BS_ERR_NONPHYSICAL':
      return new NonPhysicalStateError(error.message);
    case 'BS_ERR_MEMORY':
      return new GpuMemoryLimitError(error.message);
    case 'BS_ERR_EXECUTION':
      return new GpuExecutionError(error.message);
    case 'BS_ERR_PRECISION':
      return new GpuPrecisionError(error.message);
    case 'BS_ERR_BUSY':
      return new Error(error.message);
    default:
      return error;
  }
}

function call(fn) {
  try {
    return fn();
  } catch (error) {
    throw translate(error);
  }
}

/**
 * One operation at a time per engine (gpu-common.ts:122-140
 * ExclusiveRunner): a call made while another is running throws.
 */
class ExclusiveRunner {
  constructor(name) {
    this.name = name;
    this.running = false;
  }
  assertIdle() {
    if (this.running) throw new Error(`${this.name}: await the previous operation first`);
  }
  async run(operation) {
    this.assertIdle();
    this.running = true;
    try {
      return await operation();
    } catch (error) {
      throw translate(error);
    } finally {
      this.running = false;
    }
  }
}

// ---- configuration codes (include/bowshock.h) ----

const SIDES = ['iMin', 'iMax', 'jMin', 'jMax'];
const BOUNDARY_KINDS = { inflow: 0, outflow: 1, wall: 2, periodic: 3, faceStates: 4, farField: 5 };
const LIMITERS = { firstOrder: 0, minmod: 1, vanLeer: 2, vanAlbada: 3, monotonizedCentral: 4, unlimitedCentral: 5 };
const RECONSTRUCTIONS = { perVariable: 0, shared: 1, sharedScaled: 2 };
const FLUXES = { hllc: 0, rusanov: 1 };
const INTEGRATORS = { forwardEuler: 0, ssprk2: 1, ssprk3: 2 };

function code(table, value, what) {
  if (!Object.prototype.hasOwnProperty.call(table, value)) throw new RangeError(`unknown ${what} ${value}`);
  return table[value];
}

/** The addon's plain-data form of a GpuEuler2DConfig (euler2d-gpu.ts:82-92). */
function nativeConfig(config) {
  if (config.source !== undefined) {
    throw new RangeError('manufactured-solution sources are only supported by the CPU engine');
  }
  const { grid } = config;
  const boundaries = SIDES.map((side) => {
    const bc = config.boundaries[side];
    const state = bc.state || (bc.states && bc.states[0]) || { rho: 0, u: 0, v: 0, p: 0 };
    return { kind: code(BOUNDARY_KINDS, bc.kind, 'boundary kind'), rho: state.rho, u: state.u, v: state.v, p: state.p };
  });
  return {
    ni: grid.ni,
    nj: grid.nj,
    x: Float64Array.from(grid.x),
    y: Float64Array.from(grid.y),
    geometry: config.geometry === 'axisymmetric' ? 1 : 0,
    gamma: config.gamma,
    flux: code(FLUXES, config.flux, 'GPU flux'),
    limiter: code(LIMITERS, config.limiter, 'limiter'),
    reconstruction: code(RECONSTRUCTIONS, config.reconstruction, 'reconstruction'),
    integrator: code(INTEGRATORS, config.timeIntegrator, 'time integrator'),
    cfl: config.cfl,
    timeStepping: config.timeStepping === 'local' ? 1 : 0,
    boundaries,
  };
}

/** Grid metrics in FP64 (GridMetrics, src/grid/structured.ts:28-50). */
function computeMetrics(grid, geometry) {
  const m = call(() =>
    native().computeMetrics(grid.ni, grid.nj, Float64Array.from(grid.x), Float64Array.from(grid.y), geometry === 'axisymmetric' ? 1 : 0),
  );
  return {
    geometry,
    ni: grid.ni,
    nj: grid.nj,
    iFaces: { nx: m.iNx, ny: m.iNy, area: m.iArea },
    jFaces: { nx: m.jNx, ny: m.jNy, area: m.jArea },
    volume: m.volume,
    planarArea: m.planarArea,
    centroidX: m.centroidX,
    centroidY: m.centroidY,
  };
}

/** A HIP device (the reference's GPUDevice). */
class HipDevice {
  constructor(ordinal) {
    this.ordinal = ordinal || 0;
    this.handle = call(() => native().ctxCreate(this.ordinal));
  }
  destroy() {
    if (this.handle !== null) native().ctxDestroy(this.handle);
    this.handle = null;
  }
}

function packPrimitives(n, at) {
  const packed = new Float64Array(4 * n);
  for (let c = 0; c < n; c++) {
    const w = at(c);
    packed[4 * c] = w.rho;
    packed[4 * c + 1] = w.u;
    packed[4 * c + 2] = w.v;
    packed[4 * c + 3] = w.p;
  }
  return packed;
}

class GpuEuler2D {
  /**
   * euler2d-gpu.ts:217-291: validate, compute the metrics (FP64, host) and
   * upload them. `device` is a HipDevice or a device ordinal.
   */
  static async create(device, config, options = {}) {
    const dev = device instanceof HipDevice ? device : new HipDevice(device);
    return new GpuEuler2D(dev, config, options);
  }

  /**
   * options.precision: 'fp32' (default) or 'fp64', the solver with its
   * state, metrics and residual in FP64 (implicit steps relax in FP32), in place
   * of the two-float engine (euler2d-two-float.ts:151).
   */
  constructor(device, config, options = {}) {
    const precision = options.precision === undefined ? 'fp32' : options.precision;
    if (precision !== 'fp32' && precision !== 'fp64') throw new RangeError(`unknown precision ${precision}`);
    this.config = config;
    this.device = device;
    this.precision = precision;
    this.metrics = computeMetrics(config.grid, config.geometry);
    this.exclusive = new ExclusiveRunner('GpuEuler2D');
    this.handle = call(() =>
      native().solverCreate(device.handle, nativeConfig(config), precision === 'fp64' ? 1 : 0));
    if (config.fittedShock !== undefined) {
      // euler2d-gpu.ts:240-242: the fitted shock is the jMax faceStates side.
      const w = config.fittedShock.freeStream;
      call(() => native().setFittedShock(this.handle, Float64Array.from([w.rho, w.u, w.v, w.p])));
    }
    this.nCells = config.grid.ni * config.grid.nj;
    for (const side of SIDES) {
      const bc = config.boundaries[side];
      if (bc.kind === 'faceStates' && bc.states !== undefined) this.setFaceStates(side, bc.states);
    }
  }

  /** euler2d-gpu.ts:412-432: the field at each cell centroid; resets the clock. */
  setInitialCondition(field) {
    this.exclusive.assertIdle();
    const { centroidX, centroidY } = this.metrics;
    const packed = packPrimitives(this.nCells, (c) => field(centroidX[c], centroidY[c]));
    call(() => native().setPrimitives(this.handle, packed));
  }

  /** euler2d-gpu.ts:439-459: cell(i, j) for every cell; resets the clock. */
  setCellPrimitives(cell) {
    this.exclusive.assertIdle();
    const ni = this.config.grid.ni;
    const packed = packPrimitives(this.nCells, (c) => cell(c % ni, Math.floor(c / ni)));
    call(() => native().setPrimitives(this.handle, packed));
  }

  /** euler2d-gpu.ts:462-473. */
  async iterate(count) {
    await this.exclusive.run(() => native().iterate(this.handle, count));
  }

  /** euler2d-gpu.ts:511-527. */
  async advanceTo(tEnd, maxSteps = 1000000) {
    if (this.config.timeStepping !== 'global') {
      throw new Error('advanceTo needs global time stepping; use iterate() for local stepping');
    }
    await this.exclusive.run(() => native().advanceTo(this.handle, tEnd, maxSteps));
  }

  /** euler2d-gpu.ts:530-546 (the reduction in FP64 on the GPU). */
  async densityResidualNorm() {
    return this.exclusive.run(async () => native().residualNorm(this.handle));
  }

  /** Euler2D.residualVector (euler2d.ts:462-477) of the current state. */
  async residualVector() {
    return this.exclusive.run(async () => native().residualVector(this.handle, this.nCells));
  }

  /** euler2d-gpu.ts:549-572: GpuState2D. */
  async readState() {
    return this.exclusive.run(async () => {
      const s =
        this.precision === 'fp64'
          ? native().readState64(this.handle, this.nCells)
          : native().readState(this.handle, this.nCells);
      return { time: s.report.time, steps: s.report.steps, rho: s.rho, mu: s.mu, mv: s.mv, energy: s.energy };
    });
  }

  /** euler2d-gpu.ts:575-600. */
  setFaceStates(side, states) {
    this.exclusive.assertIdle();
    const packed = packPrimitives(states.length, (t) => states[t]);
    call(() => native().setFaceStates(this.handle, SIDES.indexOf(side), packed));
  }

  /** euler2d-gpu.ts:606-630. */
  setFaceFluxes(side, fluxes) {
    this.exclusive.assertIdle();
    call(() => native().setFaceFluxes(this.handle, SIDES.indexOf(side), fluxes === null ? null : Float64Array.from(fluxes)));
  }

  /**
   * Euler2D.boundaryFluxes (euler2d.ts:499-525) of the current state,
   * evaluated on the GPU in FP64 (BoundaryFaceFlux[]).
   */
  async wallFluxes(side = 'jMin') {
    return this.exclusive.run(async () => {
      const k = SIDES.indexOf(side);
      const count = k < 2 ? this.config.grid.nj : this.config.grid.ni;
      const f = native().wallFluxes(this.handle, k, count);
      const faces = [];
      for (let t = 0; t < count; t++) {
        const o = 9 * t;
        faces.push({ x: f[o], y: f[o + 1], nx: f[o + 2], ny: f[o + 3], area: f[o + 4], flux: [f[o + 5], f[o + 6], f[o + 7], f[o + 8]] });
      }
      return faces;
    });
  }

  /** forces.ts:19-29 over the side's faces, reduced in FP64 on the GPU. */
  async forebodyDrag(freeStream, baseRadius, side = 'jMin') {
    return this.exclusive.run(async () =>
      native().forebodyDrag(this.handle, SIDES.indexOf(side), freeStream.pressure, freeStream.dynamicPressure, baseRadius),
    );
  }

  /** euler2d-gpu.ts:668-681. */
  async setImplicitCfl(cfl) {
    if (!(cfl > 0)) throw new RangeError('the CFL number must be positive');
    await this.exclusive.run(async () => native().setImplicitCfl(this.handle, cfl));
  }

  /** euler2d-gpu.ts:687-723. */
  async iterateImplicit(count, options) {
    const { cfl, sweeps } = options;
    const omega = options.omega === undefined ? 1.5 : options.omega;
    const growth = options.growth === undefined ? 1.2 : options.growth;
    if (!(cfl > 0 && omega >= 1 && growth >= 1)) {
      throw new RangeError('implicit steps need cfl > 0, omega >= 1 and growth >= 1');
    }
    if (!(Number.isInteger(sweeps) && sweeps >= 1)) {
      throw new RangeError('the relaxation needs a whole number of sweeps, at least one');
    }
    await this.exclusive.run(() => native().iterateImplicit(this.handle, count, cfl, sweeps, omega, growth));
  }

  /** GPU state buffers stay inside the library. */
  get stateBuffer() {
    throw new Error('GpuEuler2D (HIP): GPU buffers are not exposed');
  }

  /**
   * euler2d-gpu.ts:376-394: stand-ins for the metric and boundary buffers
   * a GpuLineGrid is made on; they name this solver.
   */
  get geometryBuffer() {
    return { solver: this, part: 'geometry' };
  }
  get boundaryBuffer() {
    return { solver: this, part: 'boundary' };
  }

  /** euler2d-gpu.ts:400-405: the lines to these heights, the flow remapped. */
  moveGrid(lineGrid, heights) {
    this.exclusive.assertIdle();
    lineGrid.checkSolver(this);
    call(() => native().moveGrid(this.handle, lineGrid.heightsOf(heights)));
  }

  /** euler2d-gpu.ts:481-505. */
  async iterateFitted(lineGrid, moves, stepsPerMove) {
    if (this.config.fittedShock === undefined) throw new Error('there is no fitted shock');
    if (!(Number.isInteger(stepsPerMove) && stepsPerMove >= 1 && stepsPerMove <= 256)) {
      throw new RangeError('steps per move must be an integer from 1 to 256');
    }
    lineGrid.checkSolver(this);
    await this.exclusive.run(() => native().iterateFitted(this.handle, moves, stepsPerMove));
  }

  /** euler2d-gpu.ts:630-656 (Float64Arrays here). */
  async readShock() {
    if (this.config.fittedShock === undefined) throw new Error('there is no fitted shock');
    return this.exclusive.run(async () => native().readShock(this.handle, this.config.grid.ni));
  }

  /**
   * TransonicGridSolver.checkPrecision (transonic-drag.ts:56-66): the
   * residual against a double-precision one of the same state; throws
   * GpuPrecisionError above 1e-8 of the largest (an FP32 solver does).
   */
  async checkPrecision() {
    return this.exclusive.run(async () => native().checkPrecision(this.handle));
  }

  /**
   * euler2d-gpu.ts:659. Refused while an operation is pending: its worker
   * thread is still using the solver.
   */
  destroy() {
    this.exclusive.assertIdle();
    if (this.handle !== null) call(() => native().solverDestroy(this.handle));
    this.handle = null;
  }
}

/**
 * euler2d-gpu.ts:136-146 gpuPrimitive2D: primitives of cell c of a state
 * read back, the pressure with the 32-bit gamma the GPU solved with (the
 * gamma itself for the FP64 solver's state).
 */
function gpuPrimitive2D(state, c, gamma) {
  const rho = state.rho[c];
  const mu = state.mu[c];
  const mv = state.mv[c];
  const g = state.rho instanceof Float64Array ? gamma : Math.fround(gamma);
  return { rho, u: mu / rho, v: mv / rho, p: (g - 1) * (state.energy[c] - (0.5 * (mu * mu + mv * mv)) / rho) };
}

/**
 * The HIP engine as a SteadyEngine factory (src/analysis/gpu-engine.ts:14-41,
 * verified-drag.ts:22-37): steps on the GPU in FP32; the wall fluxes in
 * FP64 on the GPU from the FP32 state.
 */
function hipEngine(device) {
  return async (config, freeStream) => {
    const { flux } = config;
    if (!(flux === 'hllc' || flux === 'rusanov')) {
      throw new RangeError(`the GPU engine has no ${flux} flux (only hllc and rusanov)`);
    }
    if (config.source !== undefined) throw new RangeError('the GPU engine has no source terms');
    const gpu = await GpuEuler2D.create(device, config);
    gpu.setInitialCondition(() => freeStream);
    const ni = config.grid.ni;
    return {
      iterate: (steps) => gpu.iterate(steps),
      wallFluxes: () => gpu.wallFluxes('jMin'),
      cells: async () => {
        const state = await gpu.readState();
        return (i, j) => gpuPrimitive2D(state, i + j * ni, config.gamma);
      },
      destroy: () => {
        gpu.destroy();
      },
    };
  };
}

/**
 * The HIP engine as a TransonicEngine (transonic-drag.ts:46-87,
 * gpu-transonic.ts:26-106): implicit red-black steps with a CFL ramp, the
 * drag and residual from the GPU in FP64. Like the reference's, whose
 * engine is the two-float solver, it runs the FP64 solver (relaxing in
 * FP32) unless options.precision says 'fp32'; checkPrecision compares the
 * residual with a double-precision one, so an FP32 solver fails it.
 */
function hipTransonicEngine(device, options) {
  const sweeps = options.sweeps;
  const cfl = options.cfl === undefined ? 1000 : options.cfl;
  const precision = options.precision === undefined ? 'fp64' : options.precision;
  if (!(Number.isInteger(sweeps) && sweeps >= 1)) throw new RangeError('the relaxation needs a whole number of sweeps');
  const solve = async (problem) => {
    const gpu = await GpuEuler2D.create(device, problem.config, { precision });
    const { ni } = problem.config.grid;
    const gamma = problem.config.gamma;
    return {
      start: async (cells, startCfl) => {
        gpu.setCellPrimitives(cells);
        await gpu.setImplicitCfl(startCfl);
      },
      iterate: (steps) => gpu.iterateImplicit(steps, { cfl, sweeps }),
      // The forebody drag of the body's wall (jMin), FP64 on the GPU.
      drag: () => gpu.forebodyDrag(problem.freeStream, problem.body.nose.baseRadius),
      residual: () => gpu.densityResidualNorm(),
      // gpu-transonic.ts:75-80: GpuPrecisionError when the residual is not
      // the double one within RESIDUAL_PRECISION (bs_check_precision).
      checkPrecision: async () => {
        await gpu.checkPrecision();
      },
      cells: async () => {
        const state = await gpu.readState();
        return (i, j) => gpuPrimitive2D(state, i + j * ni, gamma);
      },
      destroy: () => gpu.destroy(),
    };
  };
  return Object.assign(solve, { refuses: () => null });
}

/** verified-drag.ts:225-228 freeStreamAt. */
function freeStreamAt(mach, gamma) {
  const u = Math.fround(mach * Math.sqrt(gamma));
  return { state: { rho: 1, u, v: 0, p: 1 }, pressure: 1, dynamicPressure: 0.5 * u * u };
}

/** forces.ts:19-29 forebodyPressureDrag, on faces from wallFluxes. */
function forebodyPressureDrag(wall, freeStream, baseRadius) {
  let force = 0;
  for (const face of wall) force += (face.flux[1] - freeStream.pressure * face.nx) * face.area;
  return (2 * force) / (freeStream.dynamicPressure * baseRadius * baseRadius);
}

/**
 * GpuLineGrid (src/engine-gpu/line-grid.ts): the grid of a HIP solver on
 * straight lines with fixed fractions (withFixedFractions), its metrics
 * recomputed on the GPU for new heights. Made on the solver's
 * geometryBuffer; the moves themselves run in the solver's moveGrid and
 * iterateFitted.
 */
class GpuLineGrid {
  static async create(device, lines, nj, options) {
    const target = options && options.target;
    if (!target || !(target.solver instanceof GpuEuler2D)) {
      throw new RangeError("the target must be a HIP GpuEuler2D's geometryBuffer");
    }
    return new GpuLineGrid(target.solver, lines, nj, options.firstNode === 'symmetry');
  }

  constructor(solver, lines, nj, symmetry) {
    const ni = lines.length - 1;
    if (ni !== solver.config.grid.ni || nj !== solver.config.grid.nj) {
      throw new RangeError(`the lines make a ${ni} x ${nj} grid, the solver's is another`);
    }
    const packed = new Float64Array(4 * (ni + 1));
    const fractions = new Float64Array((ni + 1) * (nj + 1));
    lines.forEach((l, i) => {
      if (!l.fractions || l.fractions.length !== nj + 1) {
        throw new RangeError(`line ${i} needs ${nj + 1} fixed fractions (withFixedFractions)`);
      }
      packed.set([l.x, l.y, l.dx, l.dy], 4 * i);
      for (let j = 0; j <= nj; j++) fractions[i + j * (ni + 1)] = l.fractions[j];
    });
    this.solver = solver;
    this.ni = ni;
    call(() => native().setLineGrid(solver.handle, packed, fractions, symmetry ? 1 : 0));
  }

  checkSolver(solver) {
    if (solver !== this.solver) throw new RangeError('the line grid was made on another solver');
  }

  /** One height per line (line-grid.ts:301: RangeError otherwise). */
  heightsOf(heights) {
    if (heights.length !== this.ni + 1) {
      throw new RangeError(`expected ${this.ni + 1} heights, got ${heights.length}`);
    }
    return Float64Array.from(heights);
  }

  /** line-grid.ts:195-205: the metrics for these heights. */
  setHeights(heights) {
    this.solver.exclusive.assertIdle();
    call(() => native().setHeights(this.solver.handle, this.heightsOf(heights)));
  }

  /** line-grid.ts:257-270. */
  async readHeights() {
    return Array.from(native().readHeights(this.solver.handle, this.ni + 1));
  }

  destroy() {}
}

/**
 * Many GpuEuler2D cases advanced together in persistent launches
 * (include/bowshock.h bs_batch_*): the cases share the scheme fields of
 * their configs; grids, boundary kinds and states differ. The reference
 * solves its grids one after another (verified-drag.ts:239-300); this is
 * the batched counterpart of that loop's engine.
 */
class GpuBatch2D {
  static async create(device, configs) {
    const dev = device instanceof HipDevice ? device : new HipDevice(device);
    return new GpuBatch2D(dev, configs);
  }

  constructor(device, configs) {
    if (!configs || configs.length === 0) throw new RangeError('a batch needs at least one case');
    for (const c of configs) {
      for (const side of SIDES) {
        if (c.boundaries[side].kind === 'faceStates') throw new RangeError('faceStates boundaries are not batched');
      }
    }
    this.device = device;
    this.configs = configs;
    this.count = configs.length;
    this.cells = configs.map((c) => c.grid.ni * c.grid.nj);
    this.exclusive = new ExclusiveRunner('GpuBatch2D');
    this.handle = call(() => native().batchCreate(device.handle, configs.map(nativeConfig)));
  }

  /** One uniform primitive state per case (or one for all). */
  setUniform(states) {
    this.exclusive.assertIdle();
    const list = Array.isArray(states) ? states : Array(this.count).fill(states);
    if (list.length !== this.count) throw new RangeError('one state per case');
    call(() => native().batchSetUniform(this.handle, packPrimitives(this.count, (k) => list[k])));
  }

  /** setCellPrimitives of case k: cell(i, j) for every cell. */
  setCellPrimitives(k, cell) {
    this.exclusive.assertIdle();
    const ni = this.configs[k].grid.ni;
    const packed = packPrimitives(this.cells[k], (c) => cell(c % ni, Math.floor(c / ni)));
    call(() => native().batchSetPrimitives(this.handle, k, packed));
  }

  /**
   * Every case from a coarse primitive field (rho, u, v, p on a cnj x cni
   * grid uniform in (i / ni, j / nj), rows j-major), times its case's four
   * scale factors, interpolated on the GPU; density and pressure held at
   * least `floor` (bs_batch_set_coarse).
   */
  setCoarse(coarse, scales, shape = [40, 80], floor = 0.05) {
    this.exclusive.assertIdle();
    const [cnj, cni] = shape;
    const per = 4 * cnj * cni;
    if (coarse.length !== this.count * per || scales.length !== this.count) {
      throw new RangeError(`coarse fields of ${per} values and 4 scale factors per case are needed`);
    }
    const s = new Float64Array(4 * this.count);
    scales.forEach((f, k) => s.set(f, 4 * k));
    call(() => native().batchSetCoarse(this.handle, Float64Array.from(coarse), cni, cnj, s, floor));
  }

  /** GpuState2D of case k (Float32Arrays). */
  async readState(k) {
    return this.exclusive.run(async () => native().batchReadState(this.handle, k, this.cells[k]));
  }

  /** Cases with active[k] false are skipped by later steps and reductions. */
  setActive(active) {
    this.exclusive.assertIdle();
    if (active.length !== this.count) throw new RangeError(`expected ${this.count} flags (one per case), got ${active.length}`);
    call(() => native().batchSetActive(this.handle, Uint8Array.from(active, (a) => (a ? 1 : 0))));
  }

  /** Cases of fewer cells than this run on the small tile. */
  setTiling(smallTileBelowCells) {
    call(() => native().batchSetInt(this.handle, 0, smallTileBelowCells));
  }

  /** 'fluxDifference' (default) or 'linearized' neighbour terms. */
  setRelaxation(form) {
    const code = { fluxDifference: 0, linearized: 1 }[form];
    if (code === undefined) throw new RangeError(`unknown relaxation ${form}`);
    call(() => native().batchSetInt(this.handle, 1, code));
  }

  /** Explicit FP32 steps in persistent launches: tile 1 (32 x 4), 2 (32 x 8) or 0 (one launch per stage). */
  setPersistent(tile) {
    this.exclusive.assertIdle();
    call(() => native().batchSetInt(this.handle, 2, tile));
  }

  /** At most this many red-black half-sweeps (1-3) per relaxation launch. */
  setRelaxFusion(halfSweeps) {
    this.exclusive.assertIdle();
    call(() => native().batchSetInt(this.handle, 3, halfSweeps));
  }

  async iterate(steps) {
    await this.exclusive.run(() => native().batchIterate(this.handle, steps));
  }

  /** One CFL number for all cases, or one per case. */
  async setImplicitCfl(cfl) {
    const arg = typeof cfl === 'number' ? cfl : Float64Array.from(cfl);
    await this.exclusive.run(async () => native().batchSetImplicitCfl(this.handle, arg));
  }

  async iterateImplicit(steps, options) {
    const { cfl, sweeps } = options;
    const omega = options.omega === undefined ? 1.5 : options.omega;
    const growth = options.growth === undefined ? 1.2 : options.growth;
    if (!(cfl > 0 && omega >= 1 && growth >= 1 && Number.isInteger(sweeps) && sweeps >= 1)) {
      throw new RangeError('implicit steps need cfl > 0, omega >= 1, growth >= 1 and whole sweeps');
    }
    await this.exclusive.run(() => native().batchIterateImplicit(this.handle, steps, cfl, sweeps, omega, growth));
  }

  /** densityResidualNorm per case, FP64 on the GPU (NaN when inactive). */
  async densityResidualNorms() {
    return this.exclusive.run(async () => native().batchResidualNorms(this.handle, this.count));
  }

  /** forebodyPressureDrag per case over jMin, FP64 on the GPU. */
  async forebodyDrags(freeStreams, baseRadius) {
    const list = Array.isArray(freeStreams) ? freeStreams : Array(this.count).fill(freeStreams);
    const R = Array.isArray(baseRadius) ? baseRadius : Array(this.count).fill(baseRadius);
    return this.exclusive.run(async () =>
      native().batchDrags(this.handle, Float64Array.from(list, (f) => f.pressure),
        Float64Array.from(list, (f) => f.dynamicPressure), Float64Array.from(R)),
    );
  }

  /** Status word (bit 0: non-physical) and step count per case. */
  async status() {
    return this.exclusive.run(async () => native().batchStatus(this.handle, this.count));
  }

  /** Refused while an operation is pending (its worker still runs). */
  destroy() {
    this.exclusive.assertIdle();
    if (this.handle !== null) call(() => native().batchDestroy(this.handle));
    this.handle = null;
  }
}

/** The stopping rule of bowshock_amd/batch.py SolvePlan (same defaults). */
const DEFAULT_PLAN = {
  implicit: false, chunk: 120, maxSteps: 20000, residual: 1e-6, drop: 1e-4, cdTol: 1e-5, settle: 2, stall: 4,
  cfl: 10, cflStart: 1, sweeps: 2, omega: 1.5, growth: 1.2, relaxation: 'fluxDifference',
};

/**
 * Run every case of a batch until it converges or fails (the reference's
 * drivers stop on a residual drop or a settled drag, verified-drag.ts:
 * 258-300, transonic-flow.ts): a case converges when its density residual
 * is at most plan.residual, or has fallen by plan.drop from its first
 * reading, or has stopped falling (no new minimum 10 % lower in plan.stall
 * checks), and its drag changed by at most plan.cdTol over the last
 * plan.settle checks; it fails at plan.maxSteps or on a non-physical
 * state. Finished cases leave the launches at the next check.
 */
async function solveBatch(batch, freeStreams, baseRadius, planIn = {}) {
  const plan = { ...DEFAULT_PLAN, ...planIn };
  const n = batch.count;
  const results = Array.from({ length: n }, () => ({ cd: NaN, residual: NaN, steps: 0, converged: false, reason: 'running' }));
  const track = Array.from({ length: n }, () => ({ first: NaN, best: Infinity, since: 0, cds: [] }));
  const active = Array(n).fill(true);
  batch.setActive(active);
  if (plan.implicit) {
    batch.setRelaxation(plan.relaxation);
    await batch.setImplicitCfl(plan.cflStart);
  }
  let done = 0;
  while (active.some((a) => a)) {
    if (plan.implicit) {
      await batch.iterateImplicit(plan.chunk, { cfl: plan.cfl, sweeps: plan.sweeps, omega: plan.omega, growth: plan.growth });
    } else {
      await batch.iterate(plan.chunk);
    }
    done += plan.chunk;
    const { status, steps } = await batch.status();
    const norms = await batch.densityResidualNorms();
    const cds = await batch.forebodyDrags(freeStreams, baseRadius);
    let changed = false;
    for (let c = 0; c < n; c++) {
      if (!active[c]) continue;
      const r = results[c];
      const t = track[c];
      r.steps = steps[c];
      r.residual = norms[c];
      r.cd = cds[c];
      let finished = false;
      if (status[c] & 1 || !Number.isFinite(r.residual) || !Number.isFinite(r.cd)) {
        r.reason = 'non-physical state';
        finished = true;
      } else {
        if (Number.isNaN(t.first)) t.first = r.residual;
        if (r.residual < 0.9 * t.best) {
          t.best = r.residual;
          t.since = 0;
        } else {
          t.since += 1;
        }
        t.cds.push(r.cd);
        const last = t.cds[t.cds.length - 1];
        const settled =
          t.cds.length > plan.settle &&
          t.cds.slice(-1 - plan.settle, -1).every((v) => Math.abs(last - v) <= plan.cdTol * Math.abs(last));
        const small = r.residual <= plan.residual || r.residual <= plan.drop * t.first;
        const stalled = t.since >= plan.stall;
        if (settled && (small || stalled)) {
          r.converged = true;
          r.reason = small ? 'converged' : 'converged (residual floor)';
          finished = true;
        } else if (done >= plan.maxSteps) {
          r.reason = 'step limit';
          finished = true;
        }
      }
      if (finished) {
        active[c] = false;
        changed = true;
      }
    }
    if (changed && active.some((a) => a)) batch.setActive(active);
  }
  return results;
}

module.exports = {
  GpuEuler2D,
  GpuBatch2D,
  solveBatch,
  GpuLineGrid,
  HipDevice,
  ExclusiveRunner,
  NonPhysicalStateError,
  GpuMemoryLimitError,
  GpuExecutionError,
  GpuPrecisionError,
  computeMetrics,
  gpuPrimitive2D,
  hipEngine,
  hipTransonicEngine,
  freeStreamAt,
  forebodyPressureDrag,
  translate,
  nativeConfig,
  native,
};
