'use strict';
// Runs one projection through the N-API addon on inputs stored as raw
// float32 files and writes the outputs next to them (used by
// tests/test_node_addon.py).  argv: dir nx ny nz h dt rho precision
const fs = require('fs');
const path = require('path');
const { UniformPressureProjection } = require(path.join(__dirname, '..', '..', 'fluid_amd', 'node'));

function readF32(file) {
  const b = fs.readFileSync(file);
  return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
}

async function main() {
  const [dir, nx, ny, nz, h, dt, rho, precision] = process.argv.slice(2);
  const dims = [Number(nx), Number(ny), Number(nz)];
  const proj = new UniformPressureProjection(dims, Number(h));
  const inputs = {
    phi: readF32(path.join(dir, 'phi.f32')), u: readF32(path.join(dir, 'u.f32')),
    v: readF32(path.join(dir, 'v.f32')), w: readF32(path.join(dir, 'w.f32')),
    dt: Number(dt), rho: Number(rho),
  };
  const res = await proj.project(inputs, { precision, relativeTolerance: 1e-6 });
  for (const k of ['pressure', 'u', 'v', 'w']) fs.writeFileSync(path.join(dir, `${k}.out.f32`), Buffer.from(res[k].buffer));
  console.log(JSON.stringify({ levels: proj.levelCount, report: res.report }));
  proj.destroy();
}

main().catch((e) => { console.error(e.code, e.message); process.exit(1); });
