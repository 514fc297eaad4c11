'use strict';
// JavaScript host of the MI355X large step (Node >= 12, CommonJS).
//
// Mirrors the reference's host objects for this path over the N-API addon
// (lib/te_napi.node -> libte.so):
//   Replayer          web/replay.js Replayer: plans a recorded command tape
//                     into one word-addressed memory M plus one 64-word
//                     argument slot per dispatch, uploads buffers, runs steps
//                     (optionally hashing every written region), reads back.
//   buildRun(conf, sounding) web/core/driver.js buildRun: a preset / namelist object
//                     and the caller's input_sounding (isnd 7 / 17)
//                     (or flat options) -> step program, FieldArena, t=0 state.
//   presetToConfig    the namelist keys of this path (web/core/config.js).
// Nothing here computes on the CPU: every call goes to the HIP library and
// throws when the addon or a GPU is missing.
const path = require('path');

let native = null;
function addon() {
  if (!native) native = require(path.join(__dirname, '..', 'lib', 'te_napi.node'));
  return native;
}

const ALIGN = 64;  // words: every buffer starts on a 256-byte boundary (replay.js plan)

// ---------------------------------------------------------------- tapes
// plan (web/replay.js:46-121): buffers at 64-word-aligned bases, setBytes
// payloads after them (rounded to 4 words + 4 spare), a 64-word argument slot
// per dispatch (A[0..2] grid, A[4+n] word offset of parameter n), copies as
// "_copy" slots (A[0] words, A[4] source, A[5] destination), and one hash slot
// per (dispatch, written parameter) in kernel-signature order.
function planTape(tape, kernelParams, cmds) {
  cmds = cmds || tape.steps[0];
  const base = [];
  let top = 0;
  for (const b of tape.buffers) { base.push(top); top += Math.ceil(b.length / (4 * ALIGN)) * ALIGN; }
  const buffersEnd = top;
  const bytes = [];
  const byteOff = new Map();
  cmds.forEach((c, i) => {
    if (c.op !== 'bytes') return;
    const raw = Buffer.from(c.data, 'base64');
    const words = Math.ceil(raw.length / 4);
    const w = new Uint32Array(words);
    Buffer.from(w.buffer).set(raw);
    byteOff.set(i, top);
    bytes.push({ off: top, words: w });
    top += Math.ceil(words / 4) * 4 + 4;
  });
  const memWords = Math.ceil(top / ALIGN) * ALIGN;
  const ops = [], slots = [];
  let bind = {}, last = {}, pipe = null;
  const regionLen = off => {
    for (let b = tape.buffers.length - 1; b >= 0; b--) {
      if (off >= base[b]) {
        const end = base[b] + tape.buffers[b].length / 4;
        return tape.buffers[b].arena ? Math.min(tape.fieldStride / 4, end - off) : end - off;
      }
    }
    return 0;
  };
  cmds.forEach((c, i) => {
    switch (c.op) {
      case 'pipe': pipe = c.name; break;
      case 'buf': bind[c.index] = c.buf < 0 ? 0 : base[c.buf] + c.offset / 4; last[c.index] = c.buf; break;
      case 'bufoff': bind[c.index] = base[last[c.index]] + c.offset / 4; break;
      case 'bytes': bind[c.index] = byteOff.get(i); break;
      case 'end': bind = {}; break;
      case 'copy':
        ops.push({ name: '_copy', n: c.size / 4, src: base[c.src] + c.srcOffset / 4, dst: base[c.dst] + c.dstOffset / 4 });
        break;
      case 'dispatch': {
        const o = { name: pipe, grid: c.grid.slice(), bind: Object.assign({}, bind) };
        const sig = kernelParams[pipe];
        if (!sig) throw new Error('kernel ' + pipe + ' has no signature');
        for (const [idx, , write] of sig) {
          if (write) { const off = bind[idx] || 0; slots.push({ op: ops.length, index: idx, off, len: regionLen(off) }); }
        }
        ops.push(o);
        break;
      }
      default: break;
    }
  });
  const names = ops.map(o => o.name);
  const args = new Uint32Array(ops.length * 64);
  ops.forEach((o, n) => {
    const A = args.subarray(n * 64, n * 64 + 64);
    if (o.name === '_copy') { A[0] = o.n; A[1] = 1; A[2] = 1; A[4] = o.src; A[5] = o.dst; return; }
    A[0] = o.grid[0]; A[1] = o.grid[1]; A[2] = o.grid[2];
    for (const k of Object.keys(o.bind)) A[4 + Number(k)] = o.bind[k];
  });
  return { base, buffersEnd, bytes, memWords, ops, slots, names, args };
}

class Replayer {
  // new Replayer(tape, kernelParams): kernelParams maps kernel name ->
  // [[index, kind, write], ...] (the kernels.json signature table).
  constructor(tape, kernelParams) {
    this.tape = tape;
    this.kernelParams = kernelParams;
    this.planned = null;
    this.ctx = null;
  }
  plan() { this.planned = planTape(this.tape, this.kernelParams); return this.planned; }
  build() {
    if (!this.planned) this.plan();
    const p = this.planned;
    this.ctx = addon().programCreate(p.memWords, p.names, p.args);
    for (const b of p.bytes) addon().upload(this.ctx, b.off, b.words);
    return this;
  }
  // upload(id, bytes): t=0 content of buffer `id` (a typed array or Buffer)
  upload(id, data) {
    const u8 = data instanceof Uint8Array ? data : new Uint8Array(data.buffer, data.byteOffset, data.byteLength);
    const words = new Uint32Array(Math.ceil(u8.length / 4));
    new Uint8Array(words.buffer).set(u8);
    addon().upload(this.ctx, this.planned.base[id], words);
  }
  // encodeSteps(n, withHashes, lastOp): runs n steps (ops 0..lastOp); with
  // hashes, one step op by op, hashing every written region after its
  // dispatch into this.hashes (slot order of replay.js).
  encodeSteps(n, withHashes, lastOp) {
    const p = this.planned;
    const last = lastOp == null ? p.names.length - 1 : lastOp;
    if (!withHashes) { addon().run(this.ctx, n, 0, last); addon().sync(this.ctx); return; }
    this.hashes = new Uint32Array(p.slots.length);
    let s = 0;
    for (let i = 0; i <= last; i++) {
      addon().run(this.ctx, 1, i, i);
      while (s < p.slots.length && p.slots[s].op === i) {
        this.hashes[s] = addon().regionHash(this.ctx, p.slots[s].off, p.slots[s].len);
        s++;
      }
    }
  }
  read(words, offWords) { return addon().download(this.ctx, offWords, words); }
  readHashes() { return this.hashes; }
  destroy() { if (this.ctx) addon().destroy(this.ctx); this.ctx = null; }
}

// ---------------------------------------------------------------- presets
// The namelist keys of this path (web/core/config.js reads the same groups):
// grid (param0/param1), dt (dtl), boundary codes (wbc/ebc: 1 periodic,
// 2 open), WENO (advwenos/advwenov = 2), TKE (sgsmodel 1), implicit vertical
// diffusion (doimpl), Rayleigh damping (irdamp, zd, rdalpha), idiss, imove
// (umove, vmove), kdiv, alph, imoist, the explicit nsound (psolver 3) and
// horizontal stretching (stretch_x/y, dx/dy_inner, _outer, nos_/tot_x/y_len).
function presetValue(preset, k) {
  for (const grp of Object.keys(preset || {})) { const v = preset[grp]; if (v && typeof v === 'object' && k in v) return v[k]; }
  return undefined;
}
function presetToConfig(preset, overrides) {
  const g = k => presetValue(preset, k);
  const c = {};
  const set = (key, v) => { if (v !== undefined && v !== null) c[key] = v; };
  set('nx', g('nx')); set('ny', g('ny')); set('nz', g('nz'));
  set('dx', g('dx')); set('dy', g('dy')); set('dz', g('dz'));
  set('dt', g('dtl'));
  const wbc = g('wbc'), sbc = g('sbc');
  if (wbc !== undefined) c.open_x = wbc === 2 ? 1 : 0;
  if (sbc !== undefined) c.open_y = sbc === 2 ? 1 : 0;
  if (g('advwenos') !== undefined) c.weno_scalars = g('advwenos') >= 1 ? 1 : 0;
  if (g('advwenov') !== undefined) c.weno_momentum = g('advwenov') >= 1 ? 1 : 0;
  if (g('sgsmodel') !== undefined) c.tke = g('sgsmodel') === 1 ? 1 : 0;
  if (g('doimpl') !== undefined) c.explicit_vdiff = g('doimpl') === 0 ? 1 : 0;
  if (g('idiss') !== undefined) c.dissipative_heating = g('idiss') === 1 ? 1 : 0;
  if (g('irdamp') !== undefined) c.rayleigh = g('irdamp') === 1 ? 1 : 0;
  set('rayleigh_zd', g('zd')); set('rayleigh_alpha', g('rdalpha'));
  set('alph', g('alph')); set('kdiv', g('kdiv'));
  if (g('imove') === 1) { set('umove', g('umove')); set('vmove', g('vmove')); }
  if (g('imoist') !== undefined) c.moist = g('imoist') === 1 ? 1 : 0;
  if (g('ptype') !== undefined && g('ptype') !== 0) c.microphysics = g('ptype');
  if (g('apmasscon') !== undefined) c.mass_conservation = g('apmasscon');
  if (g('adapt_dt') === 1) {          // config.js:725-733
    c.adaptive_dt = 1;
    set('tapfrq', g('tapfrq')); set('statfrq', g('statfrq')); set('rstfrq', g('rstfrq'));
  }
  set('nsound', g('nsound'));
  for (const a of ['x', 'y']) {              // stretch_x / stretch_y (config.js:562-573)
    if (g('stretch_' + a) === undefined || g('stretch_' + a) === 0) continue;
    c['stretch_' + a] = g('stretch_' + a);
    set('d' + a + '_inner', g('d' + a + '_inner')); set('d' + a + '_outer', g('d' + a + '_outer'));
    set('nos_' + a + '_len', g('nos_' + a + '_len')); set('tot_' + a + '_len', g('tot_' + a + '_len'));
  }
  return Object.assign(c, overrides || {});
}

// ---------------------------------------------------------------- runs
class FieldArena {
  constructor(run) { this.run = run; this.layout = run.layout; this.fieldStride = run.layout.fieldStrideWords * 4; }
  get names() { return Object.keys(this.layout.fields); }
  offsetWords(name) {
    const o = this.layout.fields[name];
    if (o === undefined) throw new Error('no field ' + name);
    return o;
  }
  // field(name): Float32Array copy of strideJ*strideK*(nz+1) values
  field(name) {
    const L = this.layout;
    const w = addon().download(this.run.ctx, this.offsetWords(name), L.strideJ * L.strideK * (L.nz + 1));
    return new Float32Array(w.buffer);
  }
  setField(name, values) {
    const w = new Uint32Array(values.buffer, values.byteOffset, values.length);
    addon().upload(this.run.ctx, this.offsetWords(name), w);
  }
  index(i, j, k) { const L = this.layout; return (i + 3) + L.strideJ * ((j + 3) + L.strideK * k); }
}

// buildRun(configuration, sounding = null, options) as driver.js:23: the
// sounding is the text of a CM1 input_sounding file or an object with its
// `text` (the reference's Sounding keeps it) and `windsFromFile` (isnd 7:
// true, 17: false); options: { encodeOnly, overrides }.
function buildRun(conf, sounding, options) {
  options = options || {};
  const flat = conf && (conf.param0 || conf.param1 || conf.param2) ? presetToConfig(conf, options.overrides) : Object.assign({}, conf || {}, options.overrides || {});
  if (sounding != null) {
    flat.input_sounding = typeof sounding === 'string' ? sounding : sounding.text;
    if (typeof flat.input_sounding !== 'string') throw new TypeError('sounding: input_sounding text expected');
    const isnd = conf && conf.param0 ? presetValue(conf, 'isnd') : undefined;
    const winds = typeof sounding === 'object' && sounding.windsFromFile !== undefined ? sounding.windsFromFile : isnd !== 17;
    flat.sounding_winds = winds ? 1 : 0;
  }
  const ctx = addon().create(flat, !!options.encodeOnly);
  const run = { ctx, config: flat, layout: addon().layout(ctx) };
  run.arena = new FieldArena(run);
  run.program = () => addon().program(ctx);
  run.fill = () => { addon().initState(ctx); return run; };
  run.encodeStep = () => { addon().step(ctx, 1); return run; };
  run.step = n => { addon().step(ctx, n == null ? 1 : n); addon().sync(ctx); return run; };
  run.lastStepMs = () => addon().lastRunMs(ctx);
  // launch shapes chosen by measurement (te_autotune); results are unchanged
  run.autotune = (reps, force) => addon().autotune(ctx, reps == null ? 5 : reps, force ? 1 : 0);
  run.read = (words, offWords) => addon().download(ctx, offWords, words);
  run.fnv1a = (offWords, words) => addon().fnv1a(ctx, offWords, words);
  run.regionHash = (offWords, words) => addon().regionHash(ctx, offWords, words);
  run.memoryWords = () => addon().memoryWords(ctx);
  run.destroy = () => addon().destroy(ctx);
  // run.tape(n) (driver.js:63-73): the header of a tape of n large steps of
  // this configuration as the reference's dispatch list (fuse = 0), with one
  // buffer per profile / parameter region plus the arena and the parameter
  // words as setBytes commands; run.tapeState() gives each buffer's t = 0
  // words (the recordings keep them as state/<id>.bin.gz).
  run.tape = nsteps => exportTape(flat, nsteps == null ? 1 : nsteps).tape;
  run.tapeState = () => exportTape(flat, 0).state;
  // TimeStep.setTimeStep (timestep.js:711-716): the step for a new dt, nsound
  run.setTimeStep = (dt, nsound) => { addon().setTimeStep(ctx, dt, nsound); return run; };
  // adapt_dt = 1 (driver.js:36-50): start() reads the t = 0 stability maxima,
  // step(n) takes n steps of the clock's dt; clock() is its state
  if (flat.adaptive_dt) {
    run.adaptive = {
      start() { addon().adaptiveStart(ctx); return this.clock(); },
      step(n) { addon().adaptiveStep(ctx, n == null ? 1 : n); return this.clock(); },
      clock() { return addon().adaptiveClock(ctx); },
    };
  }
  return run;
}

// ---------------------------------------------------------------- tape export
function exportTape(flat, nsteps) {
  const enc = addon().create(Object.assign({}, flat, { fuse: 0 }), true);
  try {
    addon().initState(enc);
    const L = addon().layout(enc);
    const regions = addon().regions(enc);
    const prog = addon().program(enc);
    const memWords = addon().memoryWords(enc);
    const mem = addon().download(enc, 0, memWords);
    const fieldNames = Object.keys(L.fields);
    const arena = Math.min(...fieldNames.map(n => L.fields[n]));
    const arenaEnd = arena + fieldNames.length * L.fieldStrideWords;
    const buffers = [], starts = [], state = {};
    regions.forEach((r, i) => {
      const end = i + 1 < regions.length ? regions[i + 1].off : arena;
      buffers.push({ id: buffers.length, length: (end - r.off) * 4, label: r.name, arena: false });
      starts.push(r.off);
      state[buffers.length - 1] = mem.slice(r.off, end);
    });
    const arenaId = buffers.length;
    buffers.push({ id: arenaId, length: (arenaEnd - arena) * 4, label: 'arena', arena: true });
    state[arenaId] = mem.slice(arena, arenaEnd);
    const fields = {};
    for (const n of fieldNames) fields[n] = (L.fields[n] - arena) * 4;
    const n = prog.names.length;
    const A = prog.args;
    const payloads = new Set();
    for (let i = 0; i < n; i++) for (let q = 0; q < 48; q++) if (A[i * 64 + 4 + q] >= arenaEnd) payloads.add(A[i * 64 + 4 + q]);
    const sorted = [...payloads].sort((a, b) => a - b);
    const nextPayload = new Map(sorted.map((o, k) => [o, k + 1 < sorted.length ? sorted[k + 1] : memWords]));
    const regionOf = w => {
      for (let b = starts.length - 1; b >= 0; b--) if (w >= starts[b] && buffers[b].length > 0) return b;
      return -1;
    };
    const b64 = (w0, w1) => Buffer.from(mem.buffer, mem.byteOffset + w0 * 4, (w1 - w0) * 4).toString('base64');
    const cmds = [];
    for (let i = 0; i < n; i++) {
      const a = A.subarray(i * 64, i * 64 + 64);
      if (prog.names[i] === '_copy') {
        cmds.push({ op: 'copy', src: arenaId, srcOffset: (a[4] - arena) * 4, dst: arenaId, dstOffset: (a[5] - arena) * 4,
          size: a[0] * 4 });
        continue;
      }
      cmds.push({ op: 'pipe', name: prog.names[i] });
      for (let q = 0; q < 48; q++) {
        const w = a[4 + q];
        if (w === 0 && q !== 0) continue;
        if (w >= arenaEnd) cmds.push({ op: 'bytes', index: q, data: b64(w, nextPayload.get(w)) });
        else if (w >= arena) cmds.push({ op: 'buf', index: q, buf: arenaId, offset: (w - arena) * 4 });
        else {
          const b = regionOf(w);
          cmds.push({ op: 'buf', index: q, buf: b, offset: (w - starts[b]) * 4 });
        }
      }
      cmds.push({ op: 'dispatch', grid: [a[0], a[1], a[2]], kind: 'threads', tg: [1, 1, 1] });
      cmds.push({ op: 'end' });
    }
    const steps = [];
    for (let s = 0; s < nsteps; s++) steps.push(cmds);
    const tape = {
      nx: L.nx, ny: L.ny, nz: L.nz, dt: flat.dt != null ? flat.dt : addon().configDefaults().dt, halo: 3, warmSteps: 0,
      constants: { terrain: false, axisymmetric: false, warmOnly: false, thermoDefault: true, testcase: 0 },
      buffers, fields, fieldStride: L.fieldStrideWords * 4, steps, diag: [],
    };
    return { tape, state };
  } finally {
    addon().destroy(enc);
  }
}

// ---------------------------------------------------------------- clock
// AdaptiveTimeStep (web/core/adaptive.js:15-109): settings { dtl, frequency:
// { tape, restart, stat, domainDiag, azimAvg, hifrq }, ksCheck } and the
// grid's spacings; start / beginStep / endStep as in the reference, the
// state (dt, dbldt, nsound, mtime, cflmax, ksmax) after each call.
const CLOCK_KINDS = ['tape', 'restart', 'stat', 'domainDiag', 'azimAvg', 'hifrq'];
class AdaptiveTimeStep {
  constructor(settings, { minHorizontalSpacing, twoDimensional, minVerticalSpacing = null }) {
    const fr = CLOCK_KINDS.map(k => (settings.frequency && settings.frequency[k] !== undefined ? settings.frequency[k] : NaN));
    this.h = addon().clockCreate(settings.dtl, minHorizontalSpacing, minVerticalSpacing == null ? 0 : minVerticalSpacing,
      twoDimensional ? 1 : 0, settings.ksCheck === false ? 0 : 1, fr);
    this._set(addon().clockCall(this.h, 'get'));
  }
  _set(st) {
    const { fired, ...rest } = st;
    Object.assign(this, rest);
    this.firedKinds = CLOCK_KINDS.filter((k, i) => (fired >> i) & 1);
  }
  fired(kind) { return this.firedKinds.includes(kind); }
  start(cflmax, ksmax) { this._set(addon().clockCall(this.h, 'start', cflmax, ksmax)); }
  beginStep() { this._set(addon().clockCall(this.h, 'begin')); }
  endStep(cflmax, ksmax) { this._set(addon().clockCall(this.h, 'end', cflmax, ksmax)); }
}

// ---------------------------------------------------------------- adjoint
// Autograd-style function over large steps (te_adjoint_*): forward(n) runs n
// checkpointed steps of a run; backward(gradOutputs) takes dJ/d(state after
// the steps) as {field: Float32Array of the field's words} and returns the
// gradient with respect to the state before them.  No counterpart in the
// reference (its step is forward only).
class LargeStepFunction {
  constructor(run, options) {
    options = options || {};
    this.run = run;
    this.layout = run.layout;
    this.adj = addon().adjointCreate(run.ctx, options.every || 10);
  }
  forward(nsteps) {
    addon().fwdCkpt(this.adj, nsteps);
    return this.run;
  }
  backward(gradOutputs) {
    addon().adjZero(this.adj);
    for (const name of Object.keys(gradOutputs)) addon().adjUpload(this.adj, this.run.arena.offsetWords(name), gradOutputs[name]);
    addon().bwd(this.adj);
    return new StateGradient(this);
  }
  lastMs() { const t = addon().adjLastMs(this.adj); return { forward: t[0], backward: t[1] }; }
  destroy() { if (this.adj) addon().adjDestroy(this.adj); this.adj = null; }
}

// Gradient with respect to the state at the first recorded step.
class StateGradient {
  constructor(fn) { this.fn = fn; this.layout = fn.layout; }
  // adjoint of every word of a field (strideJ*strideK*(nz+1) floats)
  field(name) {
    const L = this.layout;
    return addon().adjDownload(this.fn.adj, this.fn.run.arena.offsetWords(name), L.strideJ * L.strideK * (L.nz + 1));
  }
  // gradient with respect to an initial interior field (levels 0..nz-1,
  // i fastest): the t=0 state holds the field and its _a copy, halos filled
  // periodically, so each interior value collects the adjoints of all the
  // words that copy it
  initial(name) {
    const L = this.layout, nx = L.nx, ny = L.ny, nz = L.nz;
    const out = new Float64Array(nx * ny * nz);
    const names = [name].concat(L.fields[name + '_a'] !== undefined ? [name + '_a'] : []);
    for (const f of names) {
      const g = this.field(f);
      for (let k = 0; k < nz; ++k)
        for (let j = -3; j < ny + 4; ++j)
          for (let i = -3; i < nx + 4; ++i) {
            const si = ((i % nx) + nx) % nx, sj = ((j % ny) + ny) % ny;
            out[si + nx * (sj + ny * k)] += g[(i + 3) + L.strideJ * ((j + 3) + L.strideK * k)];
          }
    }
    return out;
  }
}

module.exports = { addon, planTape, Replayer, presetToConfig, buildRun, FieldArena, LargeStepFunction, StateGradient,
  AdaptiveTimeStep };
