'use strict';
// The reference's host surface for the fluid-to-field path, over the N-API addon:
// FluidSim (lib/src/sim/fluid.ts:111-582), analyticField (lib/src/trace/field.ts:196-240),
// SdfVolume's geometry rule (lib/src/field/volume.ts:139-179) and BrushSet.make
// (lib/src/field/brush.ts:539-564).  Class and member names are the reference's; a
// root is a device context (the reference's TgpuRoot).
const path = require('path');

const native = require(path.join(__dirname, 'build', 'Release', 'clay_hip.node'));

// ---- brushes (brush.ts:19-48, 150-221, 539-564) ---------------------------------------
const KINDS = {
  sphere: 0, box: 1, roundBox: 2, capsule: 3, torus: 4, cylinder: 5, plane: 6, boxFrame: 7, volume: 8,
  cone: 9, octahedron: 10, tetrahedron: 11, dodecahedron: 12, icosahedron: 13,
};
const SUPPORTED = new Set(['sphere', 'box', 'roundBox', 'capsule', 'torus', 'cylinder', 'plane']);
const OPS = { add: 0, cut: 1, paint: 2 };

function or(x, d) { return x === undefined || x === null ? d : x; }

function eulerToQuat(x, y, z) {
  const cx = Math.cos(x / 2), sx = Math.sin(x / 2);
  const cy = Math.cos(y / 2), sy = Math.sin(y / 2);
  const cz = Math.cos(z / 2), sz = Math.sin(z / 2);
  return [sx * cy * cz + cx * sy * sz, cx * sy * cz - sx * cy * sz, cx * cy * sz + sx * sy * cz,
    cx * cy * cz - sx * sy * sz];
}

function quatMul(a, b) {
  return [
    a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
    a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
    a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3],
    a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2],
  ];
}

function normalizeSize(size) {
  if (size === undefined || size === null) return [1, 1, 1];
  if (typeof size === 'number') return [size, size, size];
  return [size[0], size[1], size[2]];
}

function builtinBound(kind, s, radius) {
  switch (kind) {
    case 'sphere': return s[0];
    case 'box': case 'boxFrame': case 'roundBox': return Math.hypot(s[0], s[1], s[2]) + radius;
    case 'capsule': return s[1] + radius;
    case 'torus': return s[0] + radius;
    case 'cylinder': return Math.hypot(s[0], s[1]);
    case 'plane': return Infinity;
    default: return s[0];
  }
}

const BrushSet = {
  /** One 80-byte Brush record (brush.ts:53-80) into `view` at `offset`. */
  make(desc, view = new DataView(new ArrayBuffer(80)), offset = 0) {
    const kind = desc.kind;
    if (!(kind in KINDS)) {
      throw new Error(`BrushSet: unknown primitive '${kind}'. Register it in \`Game.create({ brushes })\`. ` +
        `Known: ${Object.keys(KINDS).join(', ')}`);
    }
    if (!SUPPORTED.has(kind)) throw new Error(`brush kind '${kind}' is outside the GPU collider slice`);
    const size = normalizeSize(desc.size);
    const scale = or(desc.scale, 1);
    const radius = or(desc.radius, 0);
    const smooth = or(desc.smooth, 0.02);
    const [ex, ey, ez] = or(desc.euler, [0, 0, 0]);
    let q = eulerToQuat(ex, ey, ez);
    if (desc.quat) q = quatMul(desc.quat, q);
    const bound = builtinBound(kind, size, radius) * scale + smooth;
    const pos = or(desc.pos, [0, 0, 0]);
    const f = [pos[0], pos[1], pos[2], scale, q[0], q[1], q[2], q[3], size[0], size[1], size[2], radius,
      or(desc.material, 0), smooth, Number.isFinite(bound) ? bound : 1e9, or(desc.onlyMaterial, -1)];
    f.forEach((x, i) => view.setFloat32(offset + 4 * i, x, true));
    view.setUint32(offset + 64, KINDS[kind], true);
    view.setUint32(offset + 68, OPS[or(desc.op, 'add')], true);
    view.setUint32(offset + 72, or(desc.slot, 0), true);
    view.setUint32(offset + 76, 0, true);
    return view;
  },
};

// ---- device root, fields, volumes ----------------------------------------------------------
class ClayRoot {
  constructor(device = 0) { this.handle = native.ctxCreate(device); }
  destroy() { if (this.handle) native.ctxDestroy(this.handle); this.handle = null; }
}

class TracedField {
  constructor(handle) { this.handle = handle; }
  destroy() { if (this.handle) native.fieldDestroy(this.handle); this.handle = null; }
}

/** analyticField over the fold of `brushes` from (1e9, 0) (field.ts:196-236). */
function analyticField(root, brushes, options = {}) {
  const buf = new ArrayBuffer(80 * brushes.length);
  const view = new DataView(buf);
  brushes.forEach((b, i) => BrushSet.make(b, view, 80 * i));
  return new TracedField(native.fieldAnalytic(root.handle, buf, options));
}

const SdfVolume = {
  /** The geometry rule of `new SdfVolume` (volume.ts:139-179), throwing its messages. */
  derive(resolution, mipLevels = 0) { return native.volumeDerive(resolution, mipLevels); },
};

// ---- FluidSim (fluid.ts:111-582) -------------------------------------------------------------
const STATE = { pos: 0, vel: 1, density: 2, pressure: 3, nvel: 4 };

class FluidSim {
  constructor(root, world, options = {}) {
    this.root = root;
    this.world = world;
    this.handle = native.fluidCreate(root.handle, world ? world.handle : null, options);
  }

  /** FluidSim.step (fluid.ts:536-558): emit, the substeps, the surface bake. */
  step() { native.fluidStep(this.handle); }

  /** FluidSim.spawn (fluid.ts:511-525). */
  spawn(n, pos, vel, spread) { native.fluidSpawn(this.handle, n, pos, vel, spread); }

  /** FluidSim.readContacts (fluid.ts:565-570). */
  async readContacts() { return native.fluidReadContacts(this.handle).points; }

  get particleCount() { return native.fluidParticleCount(this.handle); }

  /** Places particles directly (structure of arrays: x[], y[], z[]). */
  setState(pos, vel = null) { native.fluidSetState(this.handle, pos, vel); }

  getState(what) { return native.fluidGetState(this.handle, STATE[what]); }

  /** The surface bake's u32 codes of one mip (x fastest). */
  surfaceCodes(mip) { return native.fluidSurfaceCodes(this.handle, mip); }

  destroy() { if (this.handle) native.fluidDestroy(this.handle); this.handle = null; }
}

module.exports = {
  native, BrushSet, ClayRoot, TracedField, analyticField, SdfVolume, FluidSim,
  fluidDefaults: native.fluidDefaults, version: native.version, lastError: native.lastError,
};
