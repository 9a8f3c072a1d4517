'use strict';
// MI355X GpuSolver for Node.js: the interface of the reference's WebGPU solver (web/js/solver_gpu.js:17-523) over
// the C ABI (include/weave.h) through the N-API addon (weave_napi.node).  Plain CommonJS for Node 12+.
//
//   const W = require('./solver_mi355x.js');
//   const device = await W.getDevice();                 // null without a HIP device (solver_gpu.js:17-29)
//   const m = W.Model.hexDome({ span: 3000, spacing: 300, segment: 50, contraction: 0.15, noise: 1, seed: 1 },
//                             { E: 10000, G: 600, density: 500, width: 40, height: 4, mu: 0.3, tol: 1e-6 });
//   const g = new W.GpuSolver(device, [m], { perSubmit: 64 });
//   const ctl = await g.step(600);                      // [{ iter, done, rollback, ke, resid, ... }]
//   await g.download(0);                                // m.st.X, m.st.Q, ... overwritten in place
//   g.dispose();
//
// Device calls that wait for the GPU return Promises; calls on one solver are serialised (one host thread per
// batch handle), as the reference's app serialises them (app.js:164-166).

const path = require('path');

let addon = null;
function native() {
  if (!addon) addon = require(path.join(__dirname, '..', 'lib', 'weave_napi.node'));
  return addon;
}

const FP32_FP64 = 0;
const FP64 = 1;

/** getDevice() (solver_gpu.js:17-29): the HIP device, or null when there is none */
async function getDevice(index = 0) {
  const ctx = native().deviceCreate(index);
  if (!ctx) return null;
  return { ctx, info: native().deviceInfo(ctx) };
}

/** A host model: one gridshell built by the C++ model builder; `st` holds typed arrays over its state (the arrays a
 *  download overwrites, named as the reference's CpuSolver state: X, X0, V, Q, Q0, Wr, u, nPrev, slip, stuck, lam) */
class Model {
  constructor(handle) {
    this.h = handle;
    this.info = native().modelInfo(handle);
    this.st = native().modelState(handle);
    this.st.history = { kinetic: [], strain: [], bending: [], torsion: [], stretch: [], shear: [], contact: [], residual: [] };
  }
  static hexDome(spec, scene = {}) { return new Model(native().modelHexDome(spec, scene)); }
  /** a model on a triangle mesh (buildTopology … CpuSolver._buildGoals): V Float64Array [3nV], F Int32Array [3nF],
   *  scene fields as the addon names them, anchors [{ type, s, k, fix }], point loads [{ s, k, F, local }], flipped
   *  knots [[s, k], ...]; duplicate vertices are welded unless weld is false (scene_mi355x.js reads scene.json) */
  /** opts: anchors, loads, flips, weld, and for a hybrid weave groups (lath groups 2, 3, … with the scene's group
   *  keys) and assign ([strand, group] pairs; strands left out are in group 1, the scene's own) */
  static fromMesh(V, F, scene = {}, opts = {}) {
    return new Model(native().modelFromMesh(Float64Array.from(V), Int32Array.from(F), scene, opts.anchors || [],
      opts.loads || [], opts.flips || [], opts.weld !== false, opts.groups || null, opts.assign || []));
  }
  /** md.protocol of the reference (solver_gpu.js:256-318, calibrate.js:345-365): { mode: 'force' | 'displacement' |
   *  null, probe, axis, sched, settle, avg }; the schedule is used as given (flatten the load cycles and drop
   *  non-positive displacements first, as CalibrationSet's jobs do) */
  setProtocol(p) { native().modelSetProtocol(this.h, p || {}); this.protocol = p || null; return this; }
  clone() { return new Model(native().modelClone(this.h)); }
  get n() { return this.info.n; }
  get m() { return this.info.m; }
  get K() { return this.info.K; }
  get span() { return this.info.span; }
}

class GpuSolver {
  /** new GpuSolver(device, models, { perSubmit, tile, precision }) (solver_gpu.js:65-123).  Instances may differ in
   *  topology; precision 'fp32' (FP32 forces, FP64 positions / frames / energy / residual) or 'fp64' */
  constructor(device, models, opts = {}) {
    if (!device || !device.ctx) throw new Error('GpuSolver: no device');
    this.dev = device;
    this.models = models;
    this.B = models.length;
    const precision = opts.precision === 'fp64' ? FP64 : FP32_FP64;
    this.h = native().batchCreate(device.ctx, models.map((m) => m.h),
      { precision, perSubmit: opts.perSubmit || 0, tile: opts.tile === 'autotune' ? -1 : (opts.tile || 0), history: opts.history || 0,
        crossLanes: opts.crossLanes || 0 });
    this.perSubmit = native().batchOpts(this.h).perSubmit;
    this.ctlHost = new Array(this.B).fill(null);
    this._chain = Promise.resolve();
    this.sizes = models.map((m) => ({ n: m.n, m: m.m, K: m.K }));
    // fixed-point scales of the reference (solver_gpu.js:132-159) have no use here: positions and frames are FP64
    this.PB = 0; this.FB = 0; this.TB = 0; this.KB = 0;
  }

  /** fixed-point scales (solver_gpu.js:132): none on this device */
  static computeScales() { return { PB: 0, FB: 0, TB: 0, KB: 0 }; }

  /** compileInfo() (solver_gpu.js:126): the kernels are compiled ahead of time */
  async compileInfo() { return []; }

  _serial(fn) {
    const p = this._chain.then(fn);
    this._chain = p.catch(() => {});
    return p;
  }

  /** upload(b) (solver_gpu.js:191): the model's state and goals to the device */
  upload(b) { native().batchUpload(this.h, b, this.models[b].h); }

  /** updateParams(b) (solver_gpu.js:323-333): new goals and parameters; throws if the rotation clamps changed */
  updateParams(b) { native().batchUpdateParams(this.h, b, this.models[b].h); }
  _uploadParams(b) { this.updateParams(b); }

  /** step(nIter, { instances, force }) (solver_gpu.js:338-361) → control records */
  step(nIter, opts = {}) {
    const inst = opts.instances || null;
    const force = opts.force !== false;
    return this._serial(() => native().batchStep(this.h, nIter, inst, force).then((c) => { this.ctlHost = c; return c; }));
  }

  /** readControl() (solver_gpu.js:374-386) */
  readControl() {
    return this._serial(() => native().batchReadControl(this.h).then((c) => { this.ctlHost = c; return c; }));
  }

  /** readHistory(b) (solver_gpu.js:389-405): the rows written since the last call, appended to model.st.history */
  readHistory(b) {
    return this._serial(() => native().batchReadHistory(this.h, b, 4096).then((rows) => {
      const H = this.models[b].st.history;
      const n = rows.length / 9;
      for (let k = 0; k < n; k++) {
        const o = 9 * k;
        H.kinetic.push(rows[o]); H.strain.push(rows[o + 1]); H.bending.push(rows[o + 2]); H.torsion.push(rows[o + 3]);
        H.stretch.push(rows[o + 4]); H.shear.push(rows[o + 5]); H.contact.push(rows[o + 6]); H.residual.push(rows[o + 7]);
      }
      return n;
    }));
  }

  /** download(b) (solver_gpu.js:431-503): the device state of instance b into its model's arrays, in place */
  download(b) {
    const m = this.models[b];
    return this._serial(() => native().batchDownload(this.h, b, m.h).then((s) => {
      Object.assign(m.st, { iteration: s.iteration, residual: s.residual });
      m.converged = !!s.converged; m.kePrev = s.kePrev;
      m.residualParts = { force: s.resForce, torque: s.resTorque, fref: s.fref };
      m.energies = { stretch: s.eStretch, shear: s.eShear, bending: s.eBend, torsion: s.eTwist, contact: s.eContact };
      return m;
    }));
  }

  /** downloadAll(instances) (solver_gpu.js:418-428) */
  async downloadAll(instances = null) {
    const out = [];
    for (const b of instances || [...Array(this.B).keys()]) out.push(await this.download(b));
    return out;
  }

  /** internalize(b) (solver_gpu.js:507-520): the current shape becomes the displacement origin, slip lengths reset */
  internalize(b) { native().batchInternalize(this.h, b); }

  /** readOutputs(b) (solver_gpu.js:408-411): the replay outputs recorded so far, a Float64Array of rows (travel,
   *  reaction, value, iteration) */
  readOutputs(b) { return this._serial(() => native().batchReadOutputs(this.h, b)); }

  /** readAllOutputs() (solver_gpu.js:412-416): every instance's outputs */
  async readAllOutputs() {
    const out = [];
    for (let b = 0; b < this.B; b++) out.push(await this.readOutputs(b));
    return out;
  }

  /** dispose() (solver_gpu.js:522) */
  dispose() { if (this.h) { native().batchDestroy(this.h); this.h = null; } }
}

/** buildResults (solver_cpu.js:878-922) over a downloaded model: per-segment Mx My T N Vx Vy, per-crossing contact
 *  layers, per-node displacement, reactions of the anchored nodes keyed "strand,node", energies.  Knots are indexed
 *  by crossing; the ground's per-node force is not part of a download. */
function buildResults(model) {
  const st = model.st, info = model.info, n = info.n, m = info.m, K = info.K;
  const mom = native().modelSegmentMoments(model.h);
  const V2 = st.V2 || new Float64Array(2 * m);
  const segment = { Mx: mom.Mx, My: mom.My, T: mom.T, N: Float64Array.from(st.N || new Float64Array(m)),
    Vx: Float64Array.from({ length: m }, (_, j) => V2[2 * j]), Vy: Float64Array.from({ length: m }, (_, j) => V2[2 * j + 1]) };
  const disp = new Float64Array(3 * n);
  for (let i = 0; i < 3 * n; i++) disp[i] = st.X[i] - st.X0[i];
  const knot = {};
  if (info.contacts && K) {
    const pick = (a) => Float64Array.from(a || new Float64Array(K));
    Object.assign(knot, { Fn: pick(st.Fn), Ft: pick(st.Ft), slip: pick(st.slip), state: pick(st.state), gap: pick(st.gap),
      ctype: pick(st.ctype), pressure: pick(st.pressure), util: pick(st.util), theta: pick(st.theta),
      phiA: Float64Array.from({ length: K }, (_, k) => st.phi[2 * k]), phiB: Float64Array.from({ length: K }, (_, k) => st.phi[2 * k + 1]) });
    if (info.augmented) knot.lambda = pick(st.lam);
  } else for (const k of ['Fn', 'Ft', 'slip', 'state']) knot[k] = new Float64Array(K);
  const reactions = new Map();
  for (let s = 0; s < info.S; s++) {
    for (let i = info.nodeOffset[s]; i < info.nodeOffset[s + 1]; i++) {
      if (info.nodeFree[3 * i] && info.nodeFree[3 * i + 1] && info.nodeFree[3 * i + 2]) continue;
      reactions.set(`${s},${i - info.nodeOffset[s]}`, [st.reactions[3 * i], st.reactions[3 * i + 1], st.reactions[3 * i + 2], 0, 0, 0]);
    }
  }
  const e = model.energies || {};
  const beam = (e.bending || 0) + (e.torsion || 0) + (e.stretch || 0) + (e.shear || 0);
  return { iterations: st.iteration, converged: !!model.converged, residual: st.residual, segment, knot, disp, reactions,
    energies: { kinetic: model.kePrev || 0, bending: e.bending || 0, torsion: e.torsion || 0, stretch: e.stretch || 0,
      shear: e.shear || 0, contact: e.contact || 0, beam, strain: beam + (e.contact || 0) }, rod: 'cosserat' };
}

/** resultsJSON (app.js:360-390): per-strand segment arrays, knot values by crossing index, per-strand displacements,
 *  reactions */
function resultsJSON(model, r) {
  const info = model.info;
  const perStrand = (a) => { const out = []; for (let s = 0; s < info.S; s++) out.push(Array.from(a.slice(info.segOffset[s], info.segOffset[s + 1]))); return out; };
  const segment = {}; for (const k of ['Mx', 'My', 'T', 'N', 'Vx', 'Vy']) segment[k] = perStrand(r.segment[k]);
  const knot = {}; for (const [k, v] of Object.entries(r.knot)) { const o = {}; v.forEach((x, c) => { o[String(c)] = x; }); knot[k] = o; }
  const disp = [];
  for (let s = 0; s < info.S; s++) { const rows = []; for (let i = info.nodeOffset[s]; i < info.nodeOffset[s + 1]; i++) rows.push([r.disp[3 * i], r.disp[3 * i + 1], r.disp[3 * i + 2]]); disp.push(rows); }
  const reactions = {}; for (const [k, v] of r.reactions) reactions[k] = v;
  return { iterations: r.iterations, converged: r.converged, residual: r.residual, segment, knot, disp, reactions, energies: r.energies };
}

/** Continuous-batching request server (include/weave.h, serving): submit solves as they arrive, pump, poll */
class Server {
  constructor(device, opts = {}) {
    this.h = native().serverCreate(device.ctx, {
      precision: opts.precision === 'fp64' ? FP64 : FP32_FP64, chunk: opts.chunk || 0, tile: opts.tile || 0,
      maxSlots: opts.maxSlots || 0, hbmBudget: opts.hbmBudget || 0,
    });
    this.keep = new Map();
  }
  submit(model, maxIter = 0) {
    const id = native().serverSubmit(this.h, model.h, maxIter);
    this.keep.set(id, model);
    return id;
  }
  cancel(id) { native().serverCancel(this.h, id); }
  pump() { return native().serverPump(this.h); }
  poll() {
    return native().serverPoll(this.h).map((r) => { r.model = this.keep.get(r.id); this.keep.delete(r.id); return r; });
  }
  clockMs() { return native().serverClockMs(this.h); }
  dispose() { if (this.h) { native().serverDestroy(this.h); this.h = null; } }
}

module.exports = { getDevice, Model, GpuSolver, Server, buildResults, resultsJSON, FP32_FP64, FP64 };
