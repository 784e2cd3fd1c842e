'use strict';
// Host-only checks of the N-API binding (no GPU needed): the entry points that run on
// the CPU, called from JS exactly as the reference's host would call them.
//   node napi/test_host.js [brushes.json]   prints one JSON line of results
// Exits non-zero on the first failed check.
const assert = require('assert');
const fs = require('fs');
const clay = require('./index.js');

assert.strictEqual(clay.version(), 1);

// SdfVolume's mip derivation (volume.ts:139-179), with the answers of volume.gpu.ts:28-67.
const g160 = clay.SdfVolume.derive(160, 0);
assert.deepStrictEqual([g160.mipLevels, g160.tileResPerMip], [3, [20, 10, 5]]);
assert.deepStrictEqual(g160.cellOffsets, [0, 8000, 9000]);
assert.strictEqual(g160.totalCells, 9125);
assert.strictEqual(clay.SdfVolume.derive(40, 0).mipLevels, 1);
for (const [res, m] of [[64, 3], [128, 4], [256, 5]]) assert.strictEqual(clay.SdfVolume.derive(res, 0).mipLevels, m);
assert.throws(() => clay.SdfVolume.derive(160, 4), /^Error: SdfVolume: resolution 160 must be a multiple of 64 for 4 mip levels$/);
assert.throws(() => clay.SdfVolume.derive(288, 5), /^Error: SdfVolume: resolution 288 must be a multiple of 128 for 5 mip levels$/);
assert.ok(clay.lastError().includes('resolution 288 must be a multiple of 128'));

// FluidSimOptions defaults (fluid.ts:141-203).
const d = clay.fluidDefaults();
assert.deepStrictEqual([d.capacity, d.bucket, d.spacing, d.substeps], [8192, 64, 0.45, 3]);
assert.deepStrictEqual([d.stiffness, d.viscosity, d.damping, d.maxSpeed], [220, 0.4, 0.995, 15]);
assert.deepStrictEqual([d.bakeResolution, d.bakeBand, d.worldSize], [64, 2, 24]);
assert.deepStrictEqual(d.origin, [-12, -4, -12]);
assert.deepStrictEqual(d.gravity, [0, -9.81, 0]);
assert.ok(d.smoothing < 0 && d.radius < 0 && d.surfaceRadius < 0, 'derived from spacing when unset');

// BrushSet.make (brush.ts:539-564): the records, as hex, for the caller to compare.
assert.throws(() => clay.BrushSet.make({ kind: 'blob' }), /BrushSet: unknown primitive 'blob'/);
let records = [];
if (process.argv[2]) {
  const descs = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  records = descs.map((b) => Buffer.from(clay.BrushSet.make(b).buffer).toString('hex'));
}
console.log(JSON.stringify({ ok: true, records }));
