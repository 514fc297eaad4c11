'use strict';
// Node-side checks of the JavaScript host, driven by tests/test_node_host.py.
//   node check_host.js plan            -> tape plan summary (CPU)
//   node check_host.js encode '<json>' [sounding file] -> program summary of a config (CPU)
//   node check_host.js ranges          -> bad offsets / counts are refused (CPU)
//   node check_host.js clock '<json>'  -> AdaptiveTimeStep sequence for given maxima (CPU)
//   node check_host.js tape '<json>' <out> [sounding] -> run.tape(2) and run.tapeState() written to <out> (CPU)
//   node check_host.js replay <out>    -> supercell tape replayed on the GPU with hashes
//   node check_host.js run '<json>' n [tune] [sounding] -> field FNV-1a after n steps on the GPU
//        (tune: launch shapes chosen by run.autotune first, kept regardless)
//   node check_host.js adjoint '<json>' n <out> -> gradient of sum of w^2 over a
//        box after n checkpointed steps w.r.t. the initial th, u, v (GPU)
const path = require('path');
const fs = require('fs');
const zlib = require('zlib');
const te = require(path.join(__dirname, '..', '..', 'thunderengine-web_amd', 'js', 'thunderengine.js'));

const SC = path.join(__dirname, '..', 'golden', 'supercell');
const loadTape = () => JSON.parse(zlib.gunzipSync(fs.readFileSync(path.join(SC, 'tape_step1.json.gz'))));
const loadParams = () => JSON.parse(fs.readFileSync(path.join(SC, 'kernel_params.json')));

function fnvWords(w) {
  let h = 0x811c9dc5;
  for (let i = 0; i < w.length; i++) h = Math.imul(h ^ w[i], 0x01000193) >>> 0;
  return h >>> 0;
}

const mode = process.argv[2];
if (mode === 'plan') {
  const p = te.planTape(loadTape(), loadParams());
  console.log(JSON.stringify({ memWords: p.memWords, ops: p.names.length, slots: p.slots.length,
    argsFnv: fnvWords(p.args), slotFnv: fnvWords(Uint32Array.from(p.slots.map(s => s.off))) }));
} else if (mode === 'encode') {
  const r = te.buildRun(JSON.parse(process.argv[3]), process.argv[4] ? fs.readFileSync(process.argv[4], 'utf8') : null, { encodeOnly: true });
  const prog = r.program();
  console.log(JSON.stringify({ ops: prog.names.length, names: prog.names, argsFnv: fnvWords(prog.args),
    fields: r.arena.names.length, memWords: r.memoryWords(), memFnv: fnvWords(r.read(r.memoryWords(), 0)) }));
} else if (mode === 'clock') {
  // AdaptiveTimeStep through the addon: settings JSON, then (cflmax, ksmax) per step
  const a = JSON.parse(process.argv[3]);
  const c = new te.AdaptiveTimeStep({ dtl: a.dtl, frequency: a.frequency }, { minHorizontalSpacing: a.dmin, twoDimensional: false });
  c.start(a.cfl[0], a.ks[0]);
  const seq = [[c.dt, c.nsound, c.mtime, c.firedKinds.join(',')]];
  for (let i = 1; i < a.cfl.length; i++) {
    c.beginStep();
    const fired = c.firedKinds.join(',');
    c.endStep(a.cfl[i], a.ks[i]);
    seq.push([c.dt, c.nsound, c.mtime, fired]);
  }
  console.log(JSON.stringify(seq));
} else if (mode === 'tape') {
  // run.tape(n) / run.tapeState() of a config (+ sounding file) into <out>/tape.json, <out>/state/<id>.bin.gz
  const out = process.argv[4];
  const r = te.buildRun(JSON.parse(process.argv[3]), process.argv[5] ? fs.readFileSync(process.argv[5], 'utf8') : null,
    { encodeOnly: true });
  const tape = r.tape(2);
  const state = r.tapeState();
  fs.mkdirSync(path.join(out, 'state'), { recursive: true });
  fs.writeFileSync(path.join(out, 'tape.json'), JSON.stringify(tape));
  for (const id of Object.keys(state)) {
    const w = state[id];
    fs.writeFileSync(path.join(out, 'state', id + '.bin.gz'), zlib.gzipSync(Buffer.from(w.buffer, w.byteOffset, w.length * 4)));
  }
  console.log(JSON.stringify({ buffers: tape.buffers.length, commands: tape.steps[0].length, steps: tape.steps.length }));
  r.destroy();
} else if (mode === 'ranges') {
  // offsets and counts that are negative, fractional or beyond 2^53 are
  // refused by the addon (RangeError) before they reach the C ABI
  const r = te.buildRun({ nx: 8, ny: 8, nz: 4 }, null, { encodeOnly: true });
  const bad = [[-1, 4], [2 ** 64 - 1, 4], [0.5, 4], [NaN, 4], [0, -1], [0, 2 ** 64 - 1], [r.memoryWords(), 4]];
  const refused = [];
  for (const [off, n] of bad) {
    for (const f of [() => r.read(n, off), () => r.fnv1a(off, n), () => r.regionHash(off, n)]) {
      try { f(); refused.push(false); } catch (e) { refused.push(e instanceof RangeError || /outside memory/.test(e.message)); }
    }
  }
  const ok = r.read(16, 0).length === 16;
  console.log(JSON.stringify({ refused, ok }));
  r.destroy();
} else if (mode === 'replay') {
  const tape = loadTape();
  const rp = new te.Replayer(tape, loadParams());
  const p = rp.plan();
  const have = new Set(te.addon().kernelNames());
  p.names.forEach((n, i) => {   // kernels outside this path become empty copies
    if (!have.has(n)) { p.names[i] = '_copy'; p.args.fill(0, i * 64, i * 64 + 64); }
  });
  rp.build();
  for (const b of tape.buffers) {
    const f = path.join(SC, 'state', b.id + '.bin.gz');
    if (fs.existsSync(f)) rp.upload(b.id, zlib.gunzipSync(fs.readFileSync(f)));
  }
  console.log('replaying', p.names.length, 'dispatches');
  rp.encodeSteps(1, true);
  fs.writeFileSync(process.argv[3], Buffer.from(rp.readHashes().buffer));
  rp.destroy();
  console.log('done');
} else if (mode === 'run') {
  const conf = JSON.parse(process.argv[3]);
  const n = Number(process.argv[4]);
  const r = te.buildRun(conf, process.argv[6] ? fs.readFileSync(process.argv[6], 'utf8') : null);
  r.fill();
  if (process.argv[5] === 'tune' && !(r.autotune(2, true) > 0)) throw new Error('autotune tuned no dispatch');
  r.step(n);
  const out = {};
  const L = r.layout;
  const words = L.strideJ * L.strideK * (L.nz + 1);
  for (const name of r.arena.names) out[name] = r.fnv1a(r.arena.offsetWords(name), words);
  console.log(JSON.stringify(out));
  r.destroy();
} else if (mode === 'adjoint') {
  const conf = JSON.parse(process.argv[3]);
  const n = Number(process.argv[4]);
  const r = te.buildRun(conf);
  r.fill();
  const fn = new te.LargeStepFunction(r, { every: 2 });
  fn.forward(n);
  const L = r.layout;
  const w = r.arena.field('w');
  const seed = new Float32Array(w.length);
  const i0 = L.nx >> 1, j0 = L.ny >> 1;
  for (let k = 1; k < 6; ++k)
    for (let j = j0 - 3; j < j0 + 3; ++j)
      for (let i = i0 - 3; i < i0 + 3; ++i) { const c = r.arena.index(i, j, k); seed[c] = 2 * w[c]; }
  const g = fn.backward({ w: seed });
  const parts = ['th', 'u', 'v'].map(f => g.initial(f));
  const out = new Float64Array(parts.reduce((a, p) => a + p.length, 0));
  let o = 0;
  for (const p of parts) { out.set(p, o); o += p.length; }
  fs.writeFileSync(process.argv[5], Buffer.from(out.buffer));
  fn.destroy();
  r.destroy();
  console.log('done');
} else {
  console.error('unknown mode ' + mode);
  process.exit(2);
}
