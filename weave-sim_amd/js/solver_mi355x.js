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
// Every call on a solver or server returns a Promise and runs after the calls made before it on the same handle (the
// C ABI serialises calls on a handle too), as the reference's app serialises them (app.js:164-166).  The models may be
// the reference's own CpuSolver objects (their typed arrays are bound without a copy) or this package's.

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
      opts.loads || [], opts.flips || [], opts.weld !== false, opts.groups || null, opts.assign || [],
      opts.joints || [], opts.pieces || []));
  }
  /** md.protocol of the reference (solver_gpu.js:256-318, calibrate.js:345-365): { mode: 'force' | 'displacement' |
   *  null, probe, axis, sched, settle, avg }; the schedule is used as given (flatten the load cycles and drop
   *  non-positive displacements first, as CalibrationSet's jobs do) */
  setProtocol(p) { native().modelSetProtocol(this.h, p || {}); this.protocol = p || null; return this; }
  clone() { return new Model(native().modelClone(this.h)); }
  /** copies of the model's ws_model / ws_state arrays and scalars under their include/weave.h names */
  fields() { return native().modelFields(this.h); }
  /** per crossing: mesh edge id, strand_a, k_a, strand_b, k_b, over strand (Int32Array [6K]) */
  knots() { return native().modelKnots(this.h); }
  /** the frames (strand, k, node) in frame-key order (Int32Array [3·frames]) */
  frames() { return native().modelFrames(this.h); }
  /** CpuSolver.lockContacts(edges) / unlockContacts(edges) (solver_cpu.js:869-870): the crossings on the given mesh
   *  edges, all when edges is null; GpuSolver.updateParams takes the locks to the device */
  lockContacts(edges = null) { native().modelLockContacts(this.h, edges ? Array.from(edges) : null, true); return this; }
  unlockContacts(edges = null) { native().modelLockContacts(this.h, edges ? Array.from(edges) : null, false); return this; }
  get n() { return this.info.n; }
  get m() { return this.info.m; }
  get K() { return this.info.K; }
  get span() { return this.info.span; }
}

// ------------------------------------------------------------------------------------------ the reference's models
/** the reference's CpuSolver (web/js/solver_cpu.js:662-720): its state `st` (model.js:10-75), rod `beam`
 *  (rod_cosserat.js:47-100), goals `mass` / `kRotStatic` / `lRef` / `loads` / `anchors`, contact goal `contact`
 *  (solver_cpu.js:59-98), ground `floor`, scene parameters and replay `protocol` */
function isCpuSolver(md) { return !!md && !md.h && md.st && md.beam && md.scene; }

const asType = (x, T) => (x instanceof T ? x : T.from(x || []));

/** the fields of ws_model / ws_state (include/weave.h) over a CpuSolver's arrays: typed arrays of the C type are
 *  passed as they are (no copy; a download overwrites them in place, as GpuSolver.download does,
 *  solver_gpu.js:431-503), the rest converted; `back` lists the state arrays a download must copy back */
function cpuSolverFields(md) {
  const st = md.st, rod = md.beam, ct = md.contact || null, sc = md.scene, sp = sc.solver, cp = sc.contact;
  const fl = md.floor || null, pr = md.protocol || null;
  const S = st.S, n = st.n, m = st.m, p = st.p, K = ct ? st.K : 0, C = rod.cSeg.length;
  // strands: a closed lath has as many segments as nodes, its joints wrap (buildState, model.js:36-37)
  const closed = new Uint8Array(S), jOff = new Int32Array(S + 1);
  for (let s = 0; s < S; s++) {
    const nn = st.nodeOffset[s + 1] - st.nodeOffset[s], ns = st.segOffset[s + 1] - st.segOffset[s];
    closed[s] = nn > 0 && ns === nn ? 1 : 0;
    jOff[s + 1] = jOff[s] + (closed[s] ? ns : Math.max(ns - 1, 0));
  }
  const cB = new Float64Array(3 * C);
  rod.cB.forEach((b, c) => { for (let k = 0; k < 3; k++) cB[3 * c + k] = b[k]; });
  const pl = Array.from(md.loads.pointLocal || []);
  const plNode = Int32Array.from(pl, ([nid]) => nid), plF = new Float64Array(3 * pl.length);
  pl.forEach(([, Fv], q) => { for (let k = 0; k < 3; k++) plF[3 * q + k] = +Fv[k]; });
  const back = [];
  const stateArr = (x, T) => { if (x instanceof T) return x; const y = T.from(x); back.push([y, x]); return y; };
  const f = {
    n, m, p, K, C, S, n_point_local: pl.length, l_ref: md.lRef,
    par: {
      mass_scale: +sp.mass_scale, damping: +sp.damping, tol: +sp.tol, kinetic_damping: sp.kinetic_damping ? 1 : 0,
      warmup_iter: sp.warmup_iter | 0, max_iter: sp.max_iter | 0, contacts: ct ? 1 : 0,
      ramp_iter: Math.max(cp.ramp_iter | 0, 0), augmented: cp.augmented ? 1 : 0, al_every: cp.al_every | 0, mu: +cp.mu,
      strength: +cp.strength, stick_strength: +cp.stick_strength, torsion_factor: +cp.torsion_factor,
      epsilon: +cp.epsilon, slip_rate: +cp.slip_rate, rot_friction: cp.rot_friction ? 1 : 0, rot_factor: +cp.rot_factor,
      seat_law: cp.seat_law && cp.seat_law !== 'off' ? 1 : 0, seat_gap: cp.seat_gap ? 1 : 0, dent_rate: +cp.dent_rate,
      dent_after: cp.dent_after_iter >= 0 ? cp.dent_after_iter : (ct ? ct.dentAfterDefault : 0),
      edge_threshold: +cp.edge_threshold, floor_on: fl ? 1 : 0, floor_z: fl ? +fl.p.z : 0,
      floor_mu: fl ? Math.max(+fl.p.mu || 0, 0) : 0,
    },
    node_offset: asType(st.nodeOffset, Int32Array), seg_offset: asType(st.segOffset, Int32Array), joint_offset: jOff,
    strand_closed: closed, seg_i0: asType(st.segI0, Int32Array), seg_i1: asType(st.segI1, Int32Array),
    j_seg0: asType(st.jSeg0, Int32Array), j_seg1: asType(st.jSeg1, Int32Array), j_node: asType(st.jNode, Int32Array),
    l0: asType(st.l0, Float64Array), j_len: asType(st.jLen, Float64Array),
    EA: asType(rod.segEA, Float64Array), GA: asType(rod.segGA, Float64Array), EIx: asType(rod.segEIx, Float64Array),
    EIy: asType(rod.segEIy, Float64Array), GJ: asType(rod.segGJ, Float64Array), omega0: asType(st.Omega0, Float64Array),
    c_seg: asType(rod.cSeg, Int32Array), c_node: asType(rod.cNode, Int32Array), c_len: asType(rod.cLen, Float64Array),
    c_B: cB, c_Q: asType(rod.cQ, Float64Array),
    mass: asType(md.mass, Float64Array), krot_static: asType(md.kRotStatic, Float64Array),
    load_const: asType(md.loads.const, Float64Array), point_local_node: plNode, point_local_F: plF,
    node_free: asType(md.anchors.nodeFree, Uint8Array),
    proto_mode: pr ? pr.mode : 0,
    // the state, read on upload and written by a download
    X: stateArr(st.X, Float64Array), X0: stateArr(st.X0, Float64Array), V: stateArr(st.V, Float64Array),
    Q: stateArr(st.Q, Float64Array), Q0: stateArr(st.Q0, Float64Array), W: stateArr(st.Wr, Float64Array),
    N: stateArr(rod.N, Float64Array), Vsh: stateArr(rod.V, Float64Array), M: stateArr(rod.M, Float64Array),
    cM: stateArr(rod.cM, Float64Array), reactions: stateArr(md.anchors.reactions, Float64Array),
    iteration: st.iteration | 0, residual: st.residual, ke_prev: md.kePrev || 0, pen_n: ct ? ct.penN | 0 : 0,
  };
  if (pr) Object.assign(f, { proto_probe: pr.probe, proto_n: pr.sched.length, proto_settle: pr.settle, proto_avg: pr.avg,
    proto_axis: Array.from(pr.axis || [0, 0, -1]), proto_sched: Float64Array.from(pr.sched) });
  if (ct) {
    Object.assign(f, {
      k_strand: asType(ct.strand, Int32Array), k_seg_off: asType(ct.segOff, Int32Array), k_nseg: asType(ct.nseg, Int32Array),
      k_wk: asType(ct.wk, Float64Array), k_hk: asType(ct.hk, Float64Array), k_fc90: asType(ct.fc90, Float64Array),
      k_E90: asType(ct.E90, Float64Array), k_Hs: asType(ct.Hs, Float64Array), k_w: asType(ct.w, Float64Array),
      k_h: asType(ct.h, Float64Array), k_enabled: asType(ct.enabled, Uint8Array), k_locked: asType(ct.locked, Uint8Array),
      u: stateArr(st.u, Float64Array), n_prev: stateArr(st.nPrev, Float64Array), slip: stateArr(st.slip, Float64Array),
      stuck: stateArr(st.stuck, Uint8Array), lam: stateArr(st.lam, Float64Array), pen_acc: stateArr(ct.penAcc, Float64Array),
      theta0: stateArr(ct.theta0, Float64Array), rot_slip: stateArr(ct.rotSlip, Float64Array),
      dent: stateArr(ct.dent, Float64Array), climb: stateArr(ct.climb, Float64Array), tanpsi: stateArr(ct.tanpsi, Float64Array),
      dent_e: stateArr(ct.dentE, Float64Array), Fn: stateArr(ct.Fn, Float64Array), Ft: stateArr(ct.Ft, Float64Array),
      gap: stateArr(ct.gap, Float64Array), kstate: stateArr(ct.state, Float64Array), theta: stateArr(ct.theta, Float64Array),
      Anom: stateArr(ct.Anom, Float64Array), pressure: stateArr(ct.pressure, Float64Array), util: stateArr(ct.util, Float64Array),
      Mrot: stateArr(ct.Mrot, Float64Array), phi: stateArr(ct.phi, Float64Array), mid: stateArr(ct.mid, Float64Array),
      ctype: stateArr(ct.ctype, Int32Array), face_side: stateArr(ct.faceSide, Int32Array),
    });
  }
  let floorA = null;
  if (fl) {
    floorA = new Float64Array(2 * n);
    for (let i = 0; i < n; i++) { floorA[2 * i] = fl.ax[i]; floorA[2 * i + 1] = fl.ay[i]; }
    Object.assign(f, { node_w: asType(fl.w, Float64Array), node_h: asType(fl.h, Float64Array), floor_a: floorA,
      floor_F: stateArr(fl.F || new Float64Array(3 * n), Float64Array),
      floor_state: stateArr(fl.state || new Uint8Array(n), Uint8Array),
      // the rigid freedoms the Floor carries (uploaded as _uploadFloor writes them, solver_gpu.js:236-237) and its Rb
      rb_v: Array.from(fl.rv || [0, 0, 0]), rb_w: Array.from(fl.rw || [0, 0, 0]), rb_radius: +fl.Rb || 0 });
  }
  return { fields: f, back, floorA };
}

/** one CpuSolver as a device instance: a native model over its arrays, rebuilt when the CpuSolver's goals were
 *  rebuilt (upload, updateParams) */
class CpuSolverInstance {
  constructor(md) { this.cpu = md; this.bind(); }
  bind() {
    const { fields, back, floorA } = cpuSolverFields(this.cpu);
    this.back = back; this.floorA = floorA;
    this.h = native().modelFromFields(fields);
  }
  /** after a download: converted arrays copied back, the scalars and derived fields the reference's download sets
   *  (solver_gpu.js:474-497) */
  afterDownload(s) {
    const md = this.cpu, st = md.st, ct = md.contact, rod = md.beam;
    for (const [from, to] of this.back) for (let i = 0; i < to.length; i++) to[i] = from[i];
    if (this.floorA && md.floor) for (let i = 0; i < st.n; i++) { md.floor.ax[i] = this.floorA[2 * i]; md.floor.ay[i] = this.floorA[2 * i + 1]; }
    st.iteration = s.iteration; st.residual = s.residual; md.kePrev = s.kePrev;
    md.converged = !!s.converged;
    md.residualParts = { force: s.resForce, torque: s.resTorque, rigid: s.resRigid };
    rod.energyParts = { bending: s.eBend, torsion: s.eTwist, stretch: s.eStretch, shear: s.eShear };
    rod.energy = s.eStretch + s.eShear + s.eBend + s.eTwist;
    if (ct) { ct.energy = s.eContact; ct.penN = s.penN; }
    if (md.floor) {
      // the ground's energy, largest Fn and the ramp of the iteration that evaluated the state (solver_gpu.js:460-462)
      const fl = md.floor, rampN = Math.max(md.scene.contact.ramp_iter | 0, 0), it = Math.max(s.iteration - 1, 0);
      fl.energy = s.eFloor; fl.FnMax = s.fnFloor;
      fl.ramp = rampN <= 0 ? 1.0 : Math.min(1.0, 0.01 + 0.99 * it / rampN);
      // its rigid velocities and reference point (solver_gpu.js:461)
      for (let k = 0; k < 3; k++) {
        if (fl.rv) fl.rv[k] = s.rbV[k];
        if (fl.rw) fl.rw[k] = s.rbW[k];
        if (fl.c) fl.c[k] = s.rbC[k];
      }
    }
    // the frame inertia of the state (CpuSolver.step), when the object carries the reference's methods
    if (md.inertia3 && typeof rod.updateBendCoupling === 'function') {
      const M = st.m, kTw = new Float64Array(M);
      if (ct && typeof ct.addTwistStiffness === 'function') ct.addTwistStiffness(st, kTw);
      const ms = md.scene.solver.mass_scale, kD = rod.updateBendCoupling(st);
      for (let j = 0; j < M; j++) for (let k = 0; k < 3; k++) md.inertia3[3 * j + k] = ms * Math.max(md.kRotStatic[3 * j + k] + kTw[j] + kD[3 * j + k], 1e-12);
    }
    return md;
  }
}

/** a per-handle chain: every call on a solver or server runs after the previous one has finished, so a call made
 *  while a step or a pump is on a worker thread waits for it (the reference's app serialises with A.busy,
 *  app.js:164-166) */
class Chain {
  constructor() { this.tail = Promise.resolve(); }
  run(fn) {
    const p = this.tail.then(fn);
    this.tail = p.catch(() => {});
    return p;
  }
}

class GpuSolver {
  /** new GpuSolver(device, models, { perSubmit, tile, precision }) (solver_gpu.js:65-123).  models: the reference's
   *  CpuSolver objects, or models of this package's builder; instances may differ in topology; precision 'fp32'
   *  (FP32 forces, FP64 positions / frames / energy / residual) or 'fp64'; also crossLanes, ctrlMode and streams
   *  (include/weave.h ws_batch_opts) */
  constructor(device, models, opts = {}) {
    if (!device || !device.ctx) throw new Error('GpuSolver: no device');
    this.dev = device;
    this.models = models;
    this.B = models.length;
    this.inst = models.map((md) => (isCpuSolver(md) ? new CpuSolverInstance(md) : md));
    const precision = opts.precision === 'fp64' ? FP64 : FP32_FP64;
    this.h = native().batchCreate(device.ctx, this.inst.map((m) => m.h),
      { precision, perSubmit: opts.perSubmit || 0, tile: opts.tile === 'autotune' ? -1 : (opts.tile || 0), history: opts.history || 0,
        crossLanes: opts.crossLanes || 0, ctrlMode: opts.ctrlMode || 0, streams: opts.streams || 0 });
    this.perSubmit = native().batchOpts(this.h).perSubmit;
    this.ctlHost = new Array(this.B).fill(null);
    this._chain = new Chain();
    this.sizes = models.map((m) => (isCpuSolver(m) ? { n: m.st.n, m: m.st.m, K: m.contact ? m.st.K : 0 } : { n: m.n, m: m.m, K: m.K }));
    // fixed-point scales of the reference (solver_gpu.js:132-159) have no use here: positions and frames are FP64
    this.PB = 0; this.FB = 0; this.TB = 0; this.KB = 0;
  }

  /** fixed-point scales (solver_gpu.js:132): none on this device */
  static computeScales() { return { PB: 0, FB: 0, TB: 0, KB: 0 }; }

  /** compileInfo() (solver_gpu.js:126): the kernels are compiled ahead of time */
  async compileInfo() { return []; }

  _serial(fn) {
    if (!this.h) return Promise.reject(new Error('GpuSolver: disposed'));
    return this._chain.run(() => { if (!this.h) throw new Error('GpuSolver: disposed'); return fn(); });
  }

  /** upload(b) (solver_gpu.js:191): the model's state and goals to the device (after the calls queued before it) */
  upload(b) {
    return this._serial(() => {
      const it = this.inst[b];
      if (it instanceof CpuSolverInstance) it.bind();
      native().batchUpload(this.h, b, it.h);
    });
  }

  /** updateParams(b) (solver_gpu.js:323-333): new goals and parameters; rejects if the rotation clamps changed */
  updateParams(b) {
    return this._serial(() => {
      const it = this.inst[b];
      if (it instanceof CpuSolverInstance) it.bind();    // a rebuild replaced the CpuSolver's goal arrays
      native().batchUpdateParams(this.h, b, it.h);
    });
  }
  _uploadParams(b) { return this.updateParams(b); }

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
  readHistory(b) { return this._serial(() => this._readHistory(b)); }
  _readHistory(b) {
    return native().batchReadHistory(this.h, b, 4096).then((rows) => {
      const st = this.models[b].st;
      const H = st.history || (st.history = { kinetic: [], strain: [], bending: [], torsion: [], stretch: [], shear: [], contact: [], residual: [] });
      const n = rows.length / 9;
      for (let k = 0; k < n; k++) {
        const o = 9 * k;
        H.kinetic.push(rows[o]); H.strain.push(rows[o + 1]); H.bending.push(rows[o + 2]); H.torsion.push(rows[o + 3]);
        H.stretch.push(rows[o + 4]); H.shear.push(rows[o + 5]); H.contact.push(rows[o + 6]); H.residual.push(rows[o + 7]);
      }
      return n;
    });
  }

  /** download(b, { history }) (solver_gpu.js:431-503): the device state of instance b into its model's arrays, in
   *  place, then (history, default true) the history rows written since the last read */
  download(b, opts = {}) {
    const it = this.inst[b], md = this.models[b];
    return this._serial(() => native().batchDownload(this.h, b, it.h).then((s) => {
      if (it instanceof CpuSolverInstance) it.afterDownload(s);
      else {
        Object.assign(md.st, { iteration: s.iteration, residual: s.residual });
        md.converged = !!s.converged; md.kePrev = s.kePrev;
        md.residualParts = { force: s.resForce, torque: s.resTorque, fref: s.fref };
        md.energies = { stretch: s.eStretch, shear: s.eShear, bending: s.eBend, torsion: s.eTwist, contact: s.eContact,
          floor: s.eFloor };
        md.floorFnMax = s.fnFloor;
      }
      return opts.history === false ? md : this._readHistory(b).then(() => md);
    }));
  }

  /** downloadAll(instances, opts) (solver_gpu.js:418-428) */
  async downloadAll(instances = null, opts = {}) {
    const out = [];
    for (const b of instances || [...Array(this.B).keys()]) out.push(await this.download(b, opts));
    return out;
  }

  /** internalize(b) (solver_gpu.js:507-520): the current shape becomes the displacement origin, slip lengths reset */
  internalize(b) { return this._serial(() => { native().batchInternalize(this.h, b); }); }

  /** readOutputs(b) (solver_gpu.js:408-411): the replay outputs recorded so far, a Float64Array of rows (travel,
   *  reaction, value, iteration) */
  readOutputs(b) { return this._serial(() => native().batchReadOutputs(this.h, b)); }

  /** readAllOutputs() (solver_gpu.js:412-416): every instance's outputs */
  async readAllOutputs() {
    const out = [];
    for (let b = 0; b < this.B; b++) out.push(await this.readOutputs(b));
    return out;
  }

  /** dispose() (solver_gpu.js:522): after the calls queued before it */
  dispose() {
    if (!this.h) return Promise.resolve();
    return this._chain.run(() => { if (this.h) { native().batchDestroy(this.h); this.h = null; } });
  }
}

/** buildResults (solver_cpu.js:878-922) over a downloaded model: per-segment Mx My T N Vx Vy, per-crossing contact
 *  layers, per-node displacement, reactions of the anchored frames keyed "strand,k" (a fixed node that is no frame:
 *  "node,strand,node"), the ground (its force on every node, state, summary; null without one), energies */
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
  const fixed = (i) => !(info.nodeFree[3 * i] && info.nodeFree[3 * i + 1] && info.nodeFree[3 * i + 2]);
  const fr = model.frames(), framed = new Set();
  for (let r = 0; r < fr.length / 3; r++) {
    const s = fr[3 * r], k = fr[3 * r + 1], i = fr[3 * r + 2];
    if (!fixed(i)) continue;
    reactions.set(`${s},${k}`, [st.reactions[3 * i], st.reactions[3 * i + 1], st.reactions[3 * i + 2], 0, 0, 0]);
    framed.add(i);
  }
  for (let s = 0; s < info.S; s++) {
    for (let i = info.nodeOffset[s]; i < info.nodeOffset[s + 1]; i++) {
      if (!fixed(i) || framed.has(i)) continue;
      reactions.set(`node,${s},${i - info.nodeOffset[s]}`, [st.reactions[3 * i], st.reactions[3 * i + 1], st.reactions[3 * i + 2], 0, 0, 0]);
    }
  }
  const e = model.energies || {};
  const beam = (e.bending || 0) + (e.torsion || 0) + (e.stretch || 0) + (e.shear || 0);
  let floor = null;
  if (info.floorOn) {
    // Floor.summary (solver_cpu.js:629-633): nodes on the ground, the slipping ones, the resultant of their forces
    const F = Float64Array.from(st.floorF), state = Uint8Array.from(st.floorState), sum = [0, 0, 0];
    let nodes = 0, slipping = 0;
    for (let i = 0; i < n; i++) if (state[i]) { nodes++; if (state[i] === 2) slipping++; for (let k = 0; k < 3; k++) sum[k] += F[3 * i + k]; }
    floor = { z: info.floorZ, mu: info.floorMu, F, state, nodes, slipping, sum, FnMax: model.floorFnMax || 0 };
  }
  const ef = info.floorOn ? (e.floor || 0) : 0;
  return { iterations: st.iteration, converged: !!model.converged, residual: st.residual, segment, knot, disp, reactions, floor,
    energies: { kinetic: model.kePrev || 0, bending: e.bending || 0, torsion: e.torsion || 0, stretch: e.stretch || 0,
      shear: e.shear || 0, contact: e.contact || 0, floor: ef, beam, strain: beam + (e.contact || 0) + ef }, rod: 'cosserat' };
}

/** resultsJSON (app.js:360-390): per-strand segment arrays, knot values by the knot's mesh edge id, per-strand
 *  displacements, reactions, and with a ground its plane and the nodes resting on it */
function resultsJSON(model, r) {
  const info = model.info;
  const perStrand = (a) => { const out = []; for (let s = 0; s < info.S; s++) out.push(Array.from(a.slice(info.segOffset[s], info.segOffset[s + 1]))); return out; };
  const segment = {}; for (const k of ['Mx', 'My', 'T', 'N', 'Vx', 'Vy']) segment[k] = perStrand(r.segment[k]);
  const kn = info.K ? model.knots() : new Int32Array(0);
  const knot = {}; for (const [k, v] of Object.entries(r.knot)) { const o = {}; v.forEach((x, c) => { o[String(kn[6 * c])] = x; }); knot[k] = o; }
  const disp = [];
  for (let s = 0; s < info.S; s++) { const rows = []; for (let i = info.nodeOffset[s]; i < info.nodeOffset[s + 1]; i++) rows.push([r.disp[3 * i], r.disp[3 * i + 1], r.disp[3 * i + 2]]); disp.push(rows); }
  const reactions = {}; for (const [k, v] of r.reactions) reactions[k] = v;
  const doc = { iterations: r.iterations, converged: r.converged, residual: r.residual, energies: r.energies, segment, knot,
    node: { disp }, reactions };
  if (r.floor) {
    // the plane and its force on every node resting on it (app.js:377-382)
    const contacts = [], X = model.st.X;
    for (let s = 0; s < info.S; s++) {
      for (let i = info.nodeOffset[s]; i < info.nodeOffset[s + 1]; i++) {
        if (!r.floor.state[i]) continue;
        contacts.push({ strand: s, node: i - info.nodeOffset[s], point: [X[3 * i], X[3 * i + 1], X[3 * i + 2]].map(v => +v.toFixed(4)),
          F: [r.floor.F[3 * i], r.floor.F[3 * i + 1], r.floor.F[3 * i + 2]], sliding: r.floor.state[i] === 2 });
      }
    }
    doc.floor = { z: r.floor.z, mu: r.floor.mu, nodes: r.floor.nodes, sliding: r.floor.slipping, resultant: r.floor.sum, contacts };
  }
  return doc;
}

/** a float as Python's repr writes it (the reports' tables were written by Python): the shortest round-trip digits,
 *  positional for decimal exponents −4 … 15, else d.ddde±XX */
function fmt(x) {
  const v = +x;
  if (!Number.isFinite(v)) return Number.isNaN(v) ? 'nan' : (v > 0 ? 'inf' : '-inf');
  if (v === 0) return Object.is(v, -0) ? '-0.0' : '0.0';
  const parts = v.toExponential().split('e'), exp = parseInt(parts[1], 10);
  const neg = parts[0][0] === '-', digits = parts[0].replace('-', '').replace('.', '');
  let body;
  if (exp < -4 || exp >= 16) {
    body = (digits.length > 1 ? digits[0] + '.' + digits.slice(1) : digits) + 'e' + (exp < 0 ? '-' : '+') + String(Math.abs(exp)).padStart(2, '0');
  } else if (exp >= 0) {
    body = digits.length <= exp + 1 ? digits + '0'.repeat(exp + 1 - digits.length) + '.0' : digits.slice(0, exp + 1) + '.' + digits.slice(exp + 1);
  } else body = '0.' + '0'.repeat(-exp - 1) + digits;
  return (neg ? '-' : '') + body;
}

/** the knots table of the reference's reports (examples/scenes/report-DATE/knots.csv): one row per crossing in knot-edge order */
function knotsCSV(model, r) {
  const kn = model.knots(), K = kn.length / 6, out = ['knot_edge,strand_a,k_a,strand_b,k_b,over,Fn,Ft,gap,slip,state'];
  for (let c = 0; c < K; c++) {
    const ids = Array.from(kn.slice(6 * c, 6 * c + 6)).join(',');
    out.push(ids + ',' + ['Fn', 'Ft', 'gap', 'slip', 'state'].map((k) => fmt(r.knot[k][c])).join(','));
  }
  return out.join('\n') + '\n';
}

/** the segments table of the reference's reports (examples/scenes/report-DATE/segments.csv): strand, segment, Mx, My, T, N */
function segmentsCSV(model, r) {
  const info = model.info, out = ['strand,segment,Mx,My,T,N'];
  for (let s = 0; s < info.S; s++) for (let j = info.segOffset[s]; j < info.segOffset[s + 1]; j++)
    out.push(`${s},${j - info.segOffset[s]},` + ['Mx', 'My', 'T', 'N'].map((k) => fmt(r.segment[k][j])).join(','));
  return out.join('\n') + '\n';
}

/** Continuous-batching request server (include/weave.h, serving): submit solves as they arrive, pump, poll.  Every
 *  call is queued on the server's chain, so a request submitted while a pump runs on a worker joins after it */
class Server {
  constructor(device, opts = {}) {
    this.h = native().serverCreate(device.ctx, {
      precision: opts.precision === 'fp64' ? FP64 : FP32_FP64, chunk: opts.chunk || 0, tile: opts.tile || 0,
      maxSlots: opts.maxSlots || 0, hbmBudget: opts.hbmBudget || 0,
    });
    this.keep = new Map();
    this._chain = new Chain();
  }
  _serial(fn) {
    if (!this.h) return Promise.reject(new Error('Server: disposed'));
    return this._chain.run(() => { if (!this.h) throw new Error('Server: disposed'); return fn(); });
  }
  /** → Promise<request id>; the model (a package model or a CpuSolver) stays referenced until its result is polled */
  submit(model, maxIter = 0) {
    return this._serial(() => {
      const it = isCpuSolver(model) ? new CpuSolverInstance(model) : model;
      const id = native().serverSubmit(this.h, it.h, maxIter);
      this.keep.set(id, { model, it });
      return id;
    });
  }
  cancel(id) { return this._serial(() => { native().serverCancel(this.h, id); }); }
  /** → Promise<{ active, queued }>: admit, relax one chunk, retire */
  pump() { return this._serial(() => native().serverPump(this.h)); }
  /** → Promise<results>: each with its model, whose arrays hold the final state */
  poll() {
    return this._serial(() => native().serverPoll(this.h).map((r) => {
      const k = this.keep.get(r.id);
      this.keep.delete(r.id);
      r.model = k ? k.model : undefined;
      if (k && k.it instanceof CpuSolverInstance) {
        const md = k.model;
        for (const [from, to] of k.it.back) for (let i = 0; i < to.length; i++) to[i] = from[i];
        md.st.iteration = r.iterations; md.st.residual = r.residual;
      }
      return r;
    }));
  }
  clockMs() { return native().serverClockMs(this.h); }
  dispose() {
    if (!this.h) return Promise.resolve();
    return this._chain.run(() => { if (this.h) { native().serverDestroy(this.h); this.h = null; } });
  }
}

module.exports = { getDevice, Model, GpuSolver, Server, buildResults, resultsJSON, knotsCSV, segmentsCSV, FP32_FP64, FP64,
  // the CpuSolver binding, for its round-trip test: the fields it builds and the native model over them
  _internal: { cpuSolverFields, CpuSolverInstance, modelFromFields: (f) => native().modelFromFields(f),
    modelFields: (h) => native().modelFields(h) } };
