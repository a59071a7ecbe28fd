/**
 * Types of the HIP engine's Node facade (host/index.js): the GpuEuler2D
 * method set of the reference (src/engine-gpu/euler2d-gpu.ts) and the
 * adapters its drivers consume (src/analysis/verified-drag.ts:22-37,
 * src/analysis/transonic-drag.ts:46-87). A TypeScript caller of the
 * reference imports these names from here instead.
 */

export type Side = 'iMin' | 'iMax' | 'jMin' | 'jMax';
export type Geometry = 'planar' | 'axisymmetric';
export type LimiterName =
  | 'firstOrder'
  | 'minmod'
  | 'vanLeer'
  | 'vanAlbada'
  | 'monotonizedCentral'
  | 'unlimitedCentral';
export type Reconstruction = 'perVariable' | 'shared' | 'sharedScaled';
export type GpuFluxName = 'hllc' | 'rusanov';
export type TimeIntegrator = 'forwardEuler' | 'ssprk2' | 'ssprk3';

export interface Primitive2D {
  readonly rho: number;
  readonly u: number;
  readonly v: number;
  readonly p: number;
}

export interface StructuredGrid {
  readonly ni: number;
  readonly nj: number;
  readonly x: Float64Array | readonly number[];
  readonly y: Float64Array | readonly number[];
}

export type BoundaryCondition2D =
  | { readonly kind: 'inflow'; readonly state: Primitive2D }
  | { readonly kind: 'faceStates'; readonly states: readonly Primitive2D[] }
  | { readonly kind: 'outflow' }
  | { readonly kind: 'farField'; readonly state: Primitive2D }
  | { readonly kind: 'wall' }
  | { readonly kind: 'periodic' };

/** euler2d.ts:58-81 / euler2d-gpu.ts:82-92 (no manufactured source). */
export interface GpuEuler2DConfig {
  readonly grid: StructuredGrid;
  readonly geometry: Geometry;
  readonly gamma: number;
  readonly flux: GpuFluxName;
  readonly limiter: LimiterName;
  readonly reconstruction: Reconstruction;
  readonly timeIntegrator: TimeIntegrator;
  readonly cfl: number;
  readonly timeStepping: 'global' | 'local';
  readonly boundaries: Readonly<Record<Side, BoundaryCondition2D>>;
  /** euler2d-gpu.ts:85-91: jMax ('faceStates') is a fitted shock with this free stream ahead of it. */
  readonly fittedShock?: { readonly freeStream: Primitive2D };
}

export interface FaceSet {
  readonly nx: Float64Array;
  readonly ny: Float64Array;
  readonly area: Float64Array;
}

export interface GridMetrics {
  readonly geometry: Geometry;
  readonly ni: number;
  readonly nj: number;
  readonly iFaces: FaceSet;
  readonly jFaces: FaceSet;
  readonly volume: Float64Array;
  readonly planarArea: Float64Array;
  readonly centroidX: Float64Array;
  readonly centroidY: Float64Array;
}

export interface GpuState2D {
  readonly time: number;
  readonly steps: number;
  /** Float64Array from an FP64 solver. */
  readonly rho: Float32Array | Float64Array;
  readonly mu: Float32Array | Float64Array;
  readonly mv: Float32Array | Float64Array;
  readonly energy: Float32Array | Float64Array;
}

export interface BoundaryFaceFlux {
  readonly x: number;
  readonly y: number;
  readonly nx: number;
  readonly ny: number;
  readonly area: number;
  readonly flux: readonly [number, number, number, number];
}

export interface FreeStream {
  readonly state: Primitive2D;
  readonly pressure: number;
  readonly dynamicPressure: number;
}

export class NonPhysicalStateError extends Error {}
export class GpuMemoryLimitError extends Error {}
export class GpuExecutionError extends Error {}
export class GpuPrecisionError extends Error {}

export class ExclusiveRunner {
  constructor(name: string);
  assertIdle(): void;
  run<T>(operation: () => Promise<T> | T): Promise<T>;
}

/** A HIP device ordinal wrapped as the engine's device. */
export class HipDevice {
  constructor(ordinal?: number);
  readonly ordinal: number;
  destroy(): void;
}

/** GpuLineGrid (src/engine-gpu/line-grid.ts) on a HIP solver. */
export class GpuLineGrid {
  static create(
    device: HipDevice | number,
    lines: readonly { x: number; y: number; dx: number; dy: number; fractions?: readonly number[] }[],
    nj: number,
    options: { readonly target: object; readonly boundary?: object; readonly firstNode?: 'symmetry' | 'fixed' },
  ): Promise<GpuLineGrid>;
  setHeights(heights: readonly number[]): void;
  readHeights(): Promise<number[]>;
  destroy(): void;
}

export class GpuEuler2D {
  /** options.precision 'fp64': state, metrics and residual in FP64 (implicit steps relax in FP32). */
  static create(
    device: HipDevice | number,
    config: GpuEuler2DConfig,
    options?: { readonly precision?: 'fp32' | 'fp64' },
  ): Promise<GpuEuler2D>;
  readonly config: GpuEuler2DConfig;
  readonly precision: 'fp32' | 'fp64';
  /** Stand-ins naming this solver, for GpuLineGrid.create. */
  readonly geometryBuffer: object;
  readonly boundaryBuffer: object;
  moveGrid(lineGrid: GpuLineGrid, heights: readonly number[]): void;
  iterateFitted(lineGrid: GpuLineGrid, moves: number, stepsPerMove: number): Promise<void>;
  readShock(): Promise<{ speed: Float64Array; mach: Float64Array; signal: Float64Array }>;
  readonly metrics: GridMetrics;
  setInitialCondition(field: (x: number, y: number) => Primitive2D): void;
  setCellPrimitives(cell: (i: number, j: number) => Primitive2D): void;
  iterate(count: number): Promise<void>;
  advanceTo(tEnd: number, maxSteps?: number): Promise<void>;
  densityResidualNorm(): Promise<number>;
  /**
   * TransonicGridSolver.checkPrecision (transonic-drag.ts:56-66): throws
   * GpuPrecisionError when the residual is not a double-precision one's
   * within 1e-8 of the largest; resolves to the relative difference.
   */
  checkPrecision(): Promise<number>;
  residualVector(): Promise<Float64Array>;
  readState(): Promise<GpuState2D>;
  setFaceStates(side: Side, states: readonly Primitive2D[]): void;
  setFaceFluxes(side: Side, fluxes: Float64Array | null): void;
  wallFluxes(side?: Side): Promise<BoundaryFaceFlux[]>;
  forebodyDrag(freeStream: { readonly pressure: number; readonly dynamicPressure: number }, baseRadius: number,
               side?: Side): Promise<number>;
  setImplicitCfl(cfl: number): Promise<void>;
  iterateImplicit(
    count: number,
    options: { readonly cfl: number; readonly sweeps: number; readonly omega?: number; readonly growth?: number },
  ): Promise<void>;
  destroy(): void;
}

/**
 * Batched solves (include/bowshock.h bs_batch_*): cases of one scheme in
 * persistent launches; grids, boundary kinds and states differ. No
 * faceStates sides; local time steps.
 */
export class GpuBatch2D {
  static create(device: HipDevice | number, configs: readonly GpuEuler2DConfig[]): Promise<GpuBatch2D>;
  readonly count: number;
  readonly configs: readonly GpuEuler2DConfig[];
  setUniform(states: Primitive2D | readonly Primitive2D[]): void;
  setCellPrimitives(index: number, cell: (i: number, j: number) => Primitive2D): void;
  /** Learned warm-start fields for every case, interpolated on the GPU (bs_batch_set_coarse). */
  setCoarse(
    coarse: Float64Array | readonly number[],
    scales: readonly (readonly [number, number, number, number])[],
    shape?: readonly [number, number],
    floor?: number,
  ): void;
  readState(index: number): Promise<{ rho: Float32Array; mu: Float32Array; mv: Float32Array; energy: Float32Array }>;
  setActive(active: readonly boolean[]): void;
  setTiling(smallTileBelowCells: number): void;
  setRelaxation(form: 'fluxDifference' | 'linearized'): void;
  /** Explicit FP32 steps in persistent launches: 1 (32 x 4 tiles), 2 (32 x 8) or 0 (per-stage launches). */
  setPersistent(tile: 0 | 1 | 2): void;
  /** At most this many red-black half-sweeps per relaxation launch (states unchanged). */
  setRelaxFusion(halfSweeps: 1 | 2 | 3): void;
  iterate(steps: number): Promise<void>;
  setImplicitCfl(cfl: number | readonly number[]): Promise<void>;
  iterateImplicit(
    steps: number,
    options: { readonly cfl: number; readonly sweeps: number; readonly omega?: number; readonly growth?: number },
  ): Promise<void>;
  densityResidualNorms(): Promise<Float64Array>;
  forebodyDrags(
    freeStreams: FreeStream | readonly FreeStream[],
    baseRadius: number | readonly number[],
  ): Promise<Float64Array>;
  status(): Promise<{ status: Uint32Array; steps: Float64Array }>;
  destroy(): void;
}

export interface BatchPlan {
  readonly implicit?: boolean;
  readonly chunk?: number;
  readonly maxSteps?: number;
  readonly residual?: number;
  readonly drop?: number;
  readonly cdTol?: number;
  readonly settle?: number;
  readonly stall?: number;
  readonly cfl?: number;
  readonly cflStart?: number;
  readonly sweeps?: number;
  readonly omega?: number;
  readonly growth?: number;
  readonly relaxation?: 'fluxDifference' | 'linearized';
}

export interface BatchResult {
  readonly cd: number;
  readonly residual: number;
  readonly steps: number;
  readonly converged: boolean;
  readonly reason: 'converged' | 'converged (residual floor)' | 'step limit' | 'non-physical state';
}

/** Every case to its stopping rule; finished cases leave the launches. */
export function solveBatch(
  batch: GpuBatch2D,
  freeStreams: FreeStream | readonly FreeStream[],
  baseRadius: number | readonly number[],
  plan?: BatchPlan,
): Promise<BatchResult[]>;

export function computeMetrics(grid: StructuredGrid, geometry: Geometry): GridMetrics;
export function gpuPrimitive2D(state: GpuState2D, c: number, gamma: number): Primitive2D;
export function freeStreamAt(mach: number, gamma: number): FreeStream;
export function forebodyPressureDrag(
  wall: readonly BoundaryFaceFlux[],
  freeStream: { readonly pressure: number; readonly dynamicPressure: number },
  baseRadius: number,
): number;

/** SteadyEngine / EngineFactory of verified-drag.ts:22-37. */
export interface SteadyEngine {
  iterate(steps: number): Promise<void>;
  wallFluxes(): Promise<readonly BoundaryFaceFlux[]>;
  cells(): Promise<(i: number, j: number) => Primitive2D>;
  destroy?(): void;
}
export type EngineFactory = (config: GpuEuler2DConfig, freeStream: Primitive2D) => Promise<SteadyEngine>;
export function hipEngine(device: HipDevice | number): EngineFactory;

/** TransonicGridSolver / TransonicEngine of transonic-drag.ts:46-87. */
export interface TransonicGridSolver {
  start(cells: (i: number, j: number) => Primitive2D, cfl: number): Promise<void>;
  iterate(steps: number): Promise<void>;
  drag(): Promise<number>;
  cells(): Promise<(i: number, j: number) => Primitive2D>;
  residual(): Promise<number>;
  checkPrecision(): Promise<void>;
  destroy(): void;
}
export interface TransonicEngine {
  (problem: {
    readonly config: GpuEuler2DConfig;
    readonly freeStream: FreeStream;
    readonly body: { readonly nose: { readonly baseRadius: number } };
  }): Promise<TransonicGridSolver>;
  readonly refuses?: (cells: { readonly ni: number; readonly nj: number }) => string | null;
}
export function hipTransonicEngine(
  device: HipDevice | number,
  /** precision: the solver ('fp64' by default, as the reference's two-float engine). */
  options: { readonly sweeps: number; readonly cfl?: number; readonly precision?: 'fp32' | 'fp64' },
): TransonicEngine;
