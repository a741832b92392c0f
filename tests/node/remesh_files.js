'use strict';
// Remeshes a tall-cell grid through the N-API addon on inputs stored as raw
// files and writes the re-laid fields next to them (used by
// tests/test_node_addon.py).  argv: dir nx fineNy nz band h maxDelta
const fs = require('fs');
const path = require('path');
const { TallCellPressureProjection } = require(path.join(__dirname, '..', '..', 'fluid_amd', 'node'));

function read(file, Type) {
  const b = fs.readFileSync(file);
  return new Type(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

async function main() {
  const dir = process.argv[2];
  const [nx, fineNy, nz, band, h, maxDelta] = process.argv.slice(3).map(Number);
  const proj = new TallCellPressureProjection({ nx, fineNy, nz, band, h, maxDelta,
    columnBases: read(path.join(dir, 'bases.i32'), Int32Array) });
  const f = {};
  for (const k of ['phi', 'u', 'v', 'w']) f[k] = read(path.join(dir, `${k}.f32`), Float32Array);
  // refused while a projection of the grid is in flight
  const inflight = proj.project(Object.assign({ dt: 1 / 120, rho: 998.2 }, f));
  let refused = false;
  try { proj.remesh(f); } catch (e) { refused = true; }
  await inflight;
  const first = proj.remesh(f);
  for (const k of ['phi', 'u', 'v', 'w']) fs.writeFileSync(path.join(dir, `${k}.out.f32`), Buffer.from(f[k].buffer));
  fs.writeFileSync(path.join(dir, 'bases.out.i32'), Buffer.from(first.columnBases.buffer));
  const second = proj.remesh(f);
  const res = await proj.project(Object.assign({ dt: 1 / 120, rho: 998.2 }, f));
  console.log(JSON.stringify({ refusedInFlight: refused, changed: first.changedColumns,
    again: second.changedColumns, flags: res.report.flags, basesKept: proj.columnBases === second.columnBases }));
  proj.destroy();
}

main().catch((e) => { console.error(e.code, e.message); process.exit(1); });
