// Types of the MI355X projection addon (fluid_amd/node/index.js).
export interface ProjectionReport {
  flags: number;
  iterations: number;
  outerIterations: number;
  levels: number;
  relativeResidual: number;
  rhsNorm: number;
  maxDivergenceBefore: number;
  maxDivergenceAfter: number;
  deviceMs: number;
  liquidSamples: number;
  samples: number;
  lcpResidual: number;             // CM11a: max projected residual (1/s)
  coarseSweeps: number;            // CM11a: most sweeps of one coarsest LCP solve
  coarseFailures: number;          // CM11a: coarsest solves that reached the 4096-sweep cap
}

export interface ProjectionInputs {
  phi?: Float32Array;              // (nx, ny, nz), z fastest; liquid iff phi < 0
  density?: Float32Array;          // instead of phi (uniform grids): surface density, phi = -(density - 1/2) h
  u: Float32Array;                 // (nx+1, ny, nz)
  v: Float32Array;                 // (nx, ny+1, nz)
  w: Float32Array;                 // (nx, ny, nz+1)
  targetDivergence?: Float32Array; // optional desired divergence (1/s)
  dt: number;
  rho: number;
}

export interface ProjectionOptions {
  relativeTolerance?: number;
  maxIterations?: number;
  maxOuter?: number;
  precision?: 'fp32' | 'fp16' | 'bf16' | 0 | 1 | 2;
  preSweeps?: number;
  postSweeps?: number;
  warmStart?: 0 | 1;
  solver?: 'mgpcg' | 'cm11a' | 0 | 1;  // cm11a: the reference's separating-wall LCP multigrid
  openTop?: 0 | 1;                 // cm11a: air beyond the +y side
  fullCycles?: number;             // cm11a: full cycles (default 3)
  vCycles?: number;                // cm11a: V-cycles after them (default 4)
  lcpSweeps?: number;              // cm11a: projected sweeps per side (default 6)
}

export interface ProjectionResult {
  pressure: Float32Array;
  u: Float32Array;
  v: Float32Array;
  w: Float32Array;
  report: ProjectionReport;
}

export declare class UniformPressureProjection {
  constructor(dims: readonly [number, number, number], spacing: number | readonly number[], context?: unknown);
  readonly levelCount: number;
  readonly levelDimensions: number[][];
  diagnostics: ProjectionReport | undefined;
  project(inputs: ProjectionInputs, options?: ProjectionOptions): Promise<ProjectionResult>;
  /** fp_submit / fp_poll: continuous batching with other queued projections. */
  submit(inputs: ProjectionInputs, options?: ProjectionOptions): Promise<ProjectionResult>;
  /** Velocity extrapolation into the air, in place (fp_extrapolate_velocity); throws while a
   * projection of the grid is in flight. */
  extrapolate(fields: { phi?: Float32Array; density?: Float32Array; u: Float32Array; v: Float32Array;
                        w: Float32Array },
              options?: { openTop?: boolean; maxPasses?: number }): ExtrapolationReport;
  /** Velocity advection and body forces, in place (fp_advect_velocity): traced through `transport`
   * when given, else through the field's own extrapolation; throws while a projection of the grid
   * is in flight. */
  advect(fields: { phi?: Float32Array; density?: Float32Array; u: Float32Array; v: Float32Array;
                   w: Float32Array; transport?: { u: Float32Array; v: Float32Array; w: Float32Array } },
         options: { dt: number; scheme?: 'semi-lagrangian' | 'maccormack'; gravity?: number;
                    viscosity?: number; openTop?: boolean; noSlip?: boolean }): { deviceMs: number };
  /** Frees the grid once no projection of it is in flight (idempotent). */
  destroy(): void;
}

export interface ExtrapolationReport {
  passes: number;        // band passes that ran
  activeLeft: number;    // band faces still active after the last pass (0: converged)
  levels: number;        // hierarchy levels
  bandFaces: number;     // faces known after the band, sources included
  unknownFaces: number;  // open faces left unknown (written 0)
  deviceMs: number;
}

export interface TallCellLayoutDesc {
  nx: number;                      // columns along x (z-fastest arrays)
  fineNy: number;                  // fine rows
  nz: number;
  band: number;                    // cubic band rows B
  h: number;                       // cubic cell size (m)
  columnBases: ArrayLike<number>;  // (nx, nz) tall heights in [2, fineNy]
  maxDelta?: number;               // neighbour base limit (default 4)
}

export declare class TallCellPressureProjection {
  constructor(layout: TallCellLayoutDesc, context?: unknown);
  readonly levelCount: number;
  diagnostics: ProjectionReport | undefined;
  // phi (nx, B+2, nz); u (nx+1, S, nz); v (nx, B+2, nz); w (nx, S, nz+1); S = B+2+maxDelta
  project(inputs: Omit<ProjectionInputs, 'targetDivergence'>, options?: ProjectionOptions): Promise<ProjectionResult>;
  submit(inputs: Omit<ProjectionInputs, 'targetDivergence'>, options?: ProjectionOptions): Promise<ProjectionResult>;
  /** Re-lays the grid out around the surface of `fields.phi` and moves phi, u, v, w to the new
   * layout in place (fp_tall_remesh); the grid projects with the new bases from then on.
   * Throws while a projection of the grid is in flight. */
  remesh(fields: { phi: Float32Array; u: Float32Array; v: Float32Array; w: Float32Array },
         halos?: { liquidHalo?: number; airHalo?: number }): { changedColumns: number; columnBases: Int32Array };
  columnBases: Int32Array;
  destroy(): void;
}

export interface ProjectionReportWithLatency extends ProjectionReport {
  latencyMs?: number;              // submit(): submit-to-completion latency
}

/** Independent projections of several grids in one batched solve. */
export declare function projectBatch(projections: ReadonlyArray<UniformPressureProjection | TallCellPressureProjection>,
                                     inputs: ReadonlyArray<ProjectionInputs>, options?: ProjectionOptions):
  Promise<ProjectionResult[]>;

// ---- the SimulationMethod plugin (lib/core/method-contract.ts:363-602) ----
export type GPUQuality = 'balanced' | 'high' | 'ultra';
export type MethodParamValues = Record<string, number | string | boolean>;

export interface NumberParamSpec {
  kind: 'number'; key: string; label: string; hint?: string; tier: 'coarse' | 'fine'; update?: 'runtime' | 'solver';
  unit: string; min: number; max: number; step: number; digits?: number; default: number;
}
export interface SelectParamSpec {
  kind: 'select'; key: string; label: string; hint?: string; tier: 'coarse' | 'fine'; update?: 'runtime' | 'solver';
  options: ReadonlyArray<{ value: string; label: string }>; default: string;
}

/** The GPUEulerianInfo members this solver publishes (lib/core/webgpu-eulerian.ts:13-38). */
export interface MI355XEulerianInfo {
  nx: number; ny: number; nz: number; storedNy: number; cellCount: number; equivalentUniformCells: number;
  compressionRatio: number; regularLayers: number; maximumNeighborDelta: number; gridKind: 'uniform';
  pressureIterations: number; allocatedBytes: number; quality: GPUQuality; cellSize_m: number;
  pressureSolver: string; pressureIterationsExecuted?: number; pressureRelativeResidual?: number;
  pressureResidual?: number; maxDivergence_s?: number; maxDivergenceBefore_s?: number; maxDivergenceAfter_s?: number;
  pressureSolveConverged?: boolean; pressureIterationCapReached?: boolean;
  pressureConvergenceReason?: 'tolerance' | 'iteration-cap'; pressureCurvatureBreakdown?: boolean;
  pressureFlags?: number; pressureDeviceMs?: number; time_s: number; steps: number; ordinaryGridFallback: boolean;
}

/** GPUSolverInstance (lib/core/method-contract.ts:209-361) without WebGPU:
 * volumeTexture is undefined because the field lives in HIP device memory. */
export interface MI355XSolverInstance {
  readonly info: MI355XEulerianInfo;
  readonly volumeTexture: undefined;
  advanceTo(time_s: number, bodies: unknown[]): boolean;
  applyRuntimeValues(values: MethodParamValues): void;
  readStats(): Promise<MI355XEulerianInfo>;
  destroy(): void;
}

export declare const mi355xProjectionMethod: {
  id: 'mi355x-projection';
  label: string;
  shortLabel: string;
  badge: string;
  description: string;
  detail: string;
  backend: 'webgpu';
  resource: { id: string; lane: 'fluid'; label: string; provides: readonly ['fluid-authority']; blocks: 'transport' };
  qualityLabels: Record<GPUQuality, string>;
  showQualityControl: boolean;
  supportedFieldModes: readonly string[];
  params: ReadonlyArray<NumberParamSpec | SelectParamSpec>;
  pressureMapping: string;
  presetFor(quality: GPUQuality): MethodParamValues;
  normalizeValues(values: MethodParamValues): MethodParamValues;
  runtimeParamKeys: readonly string[];
  capabilities: { volumeRendering: boolean; adoptsRigidRosterShape: boolean };
  /** scene: a SceneDescription (lib/core/model.ts:52-130); closed box containers. */
  createSolverAsync(device: unknown, scene: unknown, quality?: GPUQuality, values?: MethodParamValues,
                    onRigidLoads?: unknown, onProgress?: (p: { phase: string; label: string; completed: number;
                      total: number }) => void, signal?: { aborted: boolean }): Promise<MI355XSolverInstance>;
};
