'use strict';
// Host-side drop-in of the MI355X pressure projection for the reference's
// TypeScript host (CommonJS so it runs without a TypeScript toolchain; types in
// index.d.ts).
//
//  * UniformPressureProjection mirrors WebGPUUniformPressureMultigrid
//    (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:228-607):
//    the constructor plans the hierarchy and throws RangeError on an
//    unplannable lattice (:251-256); project() runs divergence -> MGPCG ->
//    gradient subtract; diagnostics holds the decoded report; destroy() is
//    idempotent.
//  * mi355xProjectionMethod has the shape of a SimulationMethod
//    (lib/core/method-contract.ts:363-602) whose createSolverAsync resolves a
//    headless GPUSolverInstance (:209-361): advanceTo() applies gravity and one
//    projection per step; readStats() resolves the last diagnostics.
const path = require('path');

const binding = require(path.join(__dirname, '..', '_lib', 'fluid_projection.node'));

const PRECISION = { fp32: 0, fp16: 1, bf16: 2 };
const FLAGS = {
  NONFINITE: 0x1, NONPOSITIVE_CURVATURE: 0x2, NONPOSITIVE_PRECONDITIONER: 0x4, NONCONVERGENCE: 0x8,
  UNANCHORED_COMPONENT: 0x10, PRECISION_FALLBACK: 0x20, COARSE_SINGULAR: 0x40,
};

let sharedContext = null;
function defaultContext() {
  if (!sharedContext) sharedContext = binding.createContext(0, 0);
  return sharedContext;
}

class UniformPressureProjection {
  constructor(dims, spacing, context) {
    const [nx, ny, nz] = dims;
    const h = Array.isArray(spacing) ? spacing[0] : spacing;
    if (Array.isArray(spacing) && spacing.some((s) => Math.abs(s - h) > 1e-12 * h)) {
      throw new RangeError('MI355X projection needs cubic cells');
    }
    this.context = context || defaultContext();
    this.dims = [nx, ny, nz];
    this.h = h;
    this.grid = binding.createGrid(this.context, { layout: 0, nx, ny, nz, h });
    this.diagnostics = undefined;
  }

  get levelCount() { return binding.gridLevels(this.grid).length; }
  get levelDimensions() { return binding.gridLevels(this.grid); }

  /** One projection; resolves {pressure, u, v, w, report}. */
  async project({ phi, u, v, w, targetDivergence, dt, rho }, options = {}) {
    const [nx, ny, nz] = this.dims;
    const out = {
      pressure: new Float32Array(nx * ny * nz),
      u: new Float32Array((nx + 1) * ny * nz),
      v: new Float32Array(nx * (ny + 1) * nz),
      w: new Float32Array(nx * ny * (nz + 1)),
    };
    const opts = Object.assign({}, options);
    if (typeof opts.precision === 'string') opts.precision = PRECISION[opts.precision];
    const res = await binding.project(this.grid, { phi, u, v, w, targetDivergence, dt, rho }, out, opts);
    this.diagnostics = res.report;
    return res;
  }

  destroy() {
    if (this.grid) {
      binding.destroyGrid(this.grid);
      this.grid = null;
    }
  }
}

/**
 * Restricted tall-cell grid: nx x nz columns of fineNy rows of cubic cells h,
 * each column a tall cell of columnBases[x * nz + z] rows under `band` cubic
 * cells (the reference's createTallCellLayout output, tall-cell-grid.ts:309-436,
 * with x and z named the other way round: this library's arrays are z-fastest).
 * project() takes phi on the samples (nx, band+2, nz) and the face slots
 * u (nx+1, S, nz), v (nx, band+2, nz), w (nx, S, nz+1), S = band+2+maxDelta.
 */
class TallCellPressureProjection {
  constructor({ nx, fineNy, nz, band, h, columnBases, maxDelta = 4 }, context) {
    this.context = context || defaultContext();
    this.nx = nx; this.fineNy = fineNy; this.nz = nz; this.band = band; this.maxDelta = maxDelta;
    this.h = h;
    this.grid = binding.createGrid(this.context, {
      layout: 1, nx, ny: fineNy, nz, band, h, maxDelta, columnBases: Int32Array.from(columnBases),
    });
    this.diagnostics = undefined;
  }

  get levelCount() { return binding.gridLevels(this.grid).length; }

  async project({ phi, u, v, w, dt, rho }, options = {}) {
    const P = this.band + 2, S = this.band + 2 + this.maxDelta;
    const { nx, nz } = this;
    const out = {
      pressure: new Float32Array(nx * P * nz),
      u: new Float32Array((nx + 1) * S * nz),
      v: new Float32Array(nx * P * nz),
      w: new Float32Array(nx * S * (nz + 1)),
    };
    const opts = Object.assign({}, options);
    if (typeof opts.precision === 'string') opts.precision = PRECISION[opts.precision];
    const res = await binding.project(this.grid, { phi, u, v, w, dt, rho }, out, opts);
    this.diagnostics = res.report;
    return res;
  }

  destroy() {
    if (this.grid) {
      binding.destroyGrid(this.grid);
      this.grid = null;
    }
  }
}

// ---- SimulationMethod-shaped plugin (headless projection harness) -----------
function damBreakScene(n, fill, seed) {
  // corner dam-break level set (box signed distance) and zero velocities
  const h = 1 / n;
  const phi = new Float32Array(n * n * n);
  const lo = [0, 0, 0];
  const hi = fill;
  for (let i = 0; i < n; i += 1) for (let j = 0; j < n; j += 1) for (let k = 0; k < n; k += 1) {
    const p = [(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h];
    const q = p.map((x, a) => Math.abs(x - (lo[a] + hi[a]) / 2) - (hi[a] - lo[a]) / 2);
    const outside = Math.hypot(Math.max(q[0], 0), Math.max(q[1], 0), Math.max(q[2], 0));
    phi[(i * n + j) * n + k] = outside + Math.min(Math.max(q[0], q[1], q[2]), 0);
  }
  void seed;
  return { phi, h };
}

const mi355xProjectionMethod = {
  id: 'mi355x-projection',
  label: 'MI355X MGPCG pressure projection',
  shortLabel: 'MI355X',
  badge: 'HIP',
  params: [],
  presetFor: () => ({}),
  async createSolverAsync(_device, scene, _quality, values = {}) {
    const n = (scene && scene.resolution) || 64;
    const fill = (scene && scene.damFill) || [0.4, 0.6, 0.4];
    const rho = (scene && scene.fluid && scene.fluid.density_kg_m3) || 998.2;
    const gravity = (scene && scene.gravity_m_s2) || -9.80665;
    const { phi, h } = damBreakScene(n, fill, 0);
    const proj = new UniformPressureProjection([n, n, n], h);
    let u = new Float32Array((n + 1) * n * n);
    let v = new Float32Array(n * (n + 1) * n);
    let w = new Float32Array(n * n * (n + 1));
    let time = 0;
    let pending = Promise.resolve();
    const info = { method: 'mi355x-projection', grid: [n, n, n], steps: 0, lastReport: undefined };
    return {
      info,
      volumeTexture: undefined,  // no WebGPU texture exists off the browser
      advanceTo(time_s) {
        const dt = time_s - time;
        if (!(dt > 0)) return false;
        time = time_s;
        // gravity impulse on y faces with a liquid neighbour, then project
        for (let i = 0; i < n; i += 1) for (let j = 1; j < n; j += 1) for (let k = 0; k < n; k += 1) {
          if (phi[(i * n + j - 1) * n + k] < 0 || phi[(i * n + j) * n + k] < 0) v[(i * (n + 1) + j) * n + k] += gravity * dt;
        }
        pending = pending.then(() => proj.project({ phi, u, v, w, dt, rho }, values)).then((res) => {
          u = res.u; v = res.v; w = res.w;
          info.steps += 1;
          info.lastReport = res.report;
          return res;
        });
        return true;
      },
      async readStats() {
        await pending;
        return Object.assign({}, info);
      },
      destroy() { proj.destroy(); },
    };
  },
};

module.exports = { binding, UniformPressureProjection, TallCellPressureProjection, mi355xProjectionMethod, PRECISION, FLAGS };
