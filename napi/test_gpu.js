'use strict';
// One FluidSim.step through the N-API binding on device 0, for a GPU test to check.
//   node napi/test_gpu.js case.json out_dir
// case.json: {options, brushes, n, pos: path of n*3 float32 (x[], y[], z[])}.
// Writes out_dir/pos.f32 (x[], y[], z[]), out_dir/codes<m>.u32 for every bake mip and
// out_dir/result.json (particleCount, contacts).
const fs = require('fs');
const path = require('path');
const clay = require('./index.js');

async function main() {
  const c = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  const out = process.argv[3];
  const raw = fs.readFileSync(c.pos);
  const pos = new Float32Array(raw.buffer, raw.byteOffset, raw.byteLength / 4);
  const root = new clay.ClayRoot(0);
  const world = clay.analyticField(root, c.brushes);
  const sim = new clay.FluidSim(root, world, c.options);
  sim.setState(pos);
  sim.step();
  fs.writeFileSync(path.join(out, 'pos.f32'), Buffer.from(sim.getState('pos').buffer));
  for (let m = 0; m < c.options.bakeMipLevels; m++) {
    fs.writeFileSync(path.join(out, `codes${m}.u32`), Buffer.from(sim.surfaceCodes(m).buffer));
  }
  const contacts = await sim.readContacts();
  fs.writeFileSync(path.join(out, 'result.json'), JSON.stringify({ particleCount: sim.particleCount, contacts }));
  sim.destroy();
  world.destroy();
  root.destroy();
  console.log('napi FluidSim step ok');
}

main().catch((e) => { console.error(e); process.exit(1); });
