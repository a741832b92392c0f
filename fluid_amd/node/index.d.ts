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
}

export interface ProjectionInputs {
  phi: Float32Array;               // (nx, ny, nz), z fastest; liquid iff phi < 0
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
  destroy(): void;
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
  destroy(): void;
}

export declare const mi355xProjectionMethod: {
  id: 'mi355x-projection';
  label: string;
  shortLabel: string;
  badge: string;
  params: unknown[];
  presetFor(quality: unknown): Record<string, unknown>;
  createSolverAsync(device: unknown, scene: unknown, quality?: unknown, values?: ProjectionOptions): Promise<{
    info: Record<string, unknown>;
    volumeTexture: undefined;
    advanceTo(time_s: number, bodies?: unknown[]): boolean;
    readStats(): Promise<Record<string, unknown>>;
    destroy(): void;
  }>;
};
