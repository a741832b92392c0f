'use strict';
// Extrapolates the velocities of a uniform grid through the N-API addon on
// inputs stored as raw float32 files and writes the results next to them
// (used by tests/test_node_addon.py).  argv: dir nx ny nz h
const fs = require('fs');
const path = require('path');
const { UniformPressureProjection } = require(path.join(__dirname, '..', '..', 'fluid_amd', 'node'));

function readF32(file) {
  const b = fs.readFileSync(file);
  return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

const dir = process.argv[2];
const [nx, ny, nz, h] = process.argv.slice(3).map(Number);
const proj = new UniformPressureProjection([nx, ny, nz], h);
const f = {};
for (const k of ['phi', 'u', 'v', 'w']) f[k] = readF32(path.join(dir, `${k}.f32`));
try {
  const rep = proj.extrapolate(f);
  for (const k of ['u', 'v', 'w']) fs.writeFileSync(path.join(dir, `${k}.out.f32`), Buffer.from(f[k].buffer));
  console.log(JSON.stringify(rep));
} catch (e) {
  console.error(e.code, e.message);
  process.exitCode = 1;
}
proj.destroy();
