'use strict';
// Scenes as the reference stores them (scene.json version 1 and the mesh it names) read into a host model, for the
// Node.js host: Scene.fromJSON (web/js/scene.js:136-187) — every section from the params.js defaults plus the keys
// the file holds, E90 / fc90 / H of older files from their N/mm² defaults in the scene's stress unit — and
// Mesh.fromBytes (web/js/mesh.js:24-31): STL (binary when the size says so, else ASCII) or OBJ (fan-triangulated),
// scaled, then welded by the model builder.  A hybrid weave's lath groups and strand assignments (scene.js:50-62) go
// to the builder with the scene, and so do groups assigned to pieces of a discretisation (hybrid.pieces, with the
// scene's joints) and anchors by every selector, the 'region' one included.
//
//   const S = require('./scene_mi355x.js');
//   const m = S.loadScene('examples/scenes/dome_basic.json');        // a W.Model, ready for new W.GpuSolver(...)

const fs = require('fs');
const path = require('path');
const { Model } = require('./solver_mi355x.js');

const SCENE_VERSION = 1;
const LENGTH_TO_M = { mm: 1e-3, cm: 1e-2, m: 1.0 };
const FORCE_TO_N = { N: 1.0, kN: 1e3 };
const FLOOR_MU = 0.4;
const MATERIAL_MM_N = { E90: 800.0, fc90: 8.0, H: 30.0 };
const SELECTOR = { frame: 0, all_start: 1, all_end: 2, all_ends: 3, strand: 4, boundary: 5, region: 6 };

/** triangle soup { V, F } of an STL file's bytes (Buffer / Uint8Array) */
function parseSTL(buf) {
  const u8 = buf instanceof Uint8Array ? buf : new Uint8Array(buf);
  const dv = new DataView(u8.buffer, u8.byteOffset, u8.byteLength);
  if (u8.byteLength >= 84) {
    const t = dv.getUint32(80, true);
    if (84 + 50 * t === u8.byteLength) {
      const V = new Float64Array(9 * t), F = new Int32Array(3 * t);
      for (let q = 0; q < t; q++) {
        const o = 84 + 50 * q + 12;
        for (let k = 0; k < 9; k++) V[9 * q + k] = dv.getFloat32(o + 4 * k, true);
      }
      for (let i = 0; i < 3 * t; i++) F[i] = i;
      return { V, F };
    }
  }
  const text = Buffer.from(u8).toString('utf8');
  const re = /vertex\s+([-+0-9.eE]+)\s+([-+0-9.eE]+)\s+([-+0-9.eE]+)/g;
  const vals = [];
  let m;
  while ((m = re.exec(text))) vals.push(+m[1], +m[2], +m[3]);
  const t = Math.floor(vals.length / 9);
  if (!t) throw new Error('no triangles found (binary or ASCII STL expected)');
  const F = new Int32Array(3 * t);
  for (let i = 0; i < 3 * t; i++) F[i] = i;
  return { V: Float64Array.from(vals.slice(0, 9 * t)), F };
}

/** vertices and fan-triangulated faces of an OBJ file's text */
function parseOBJ(text) {
  const V = [], F = [];
  for (const raw of text.split(/\r?\n/)) {
    const line = raw.trim();
    if (line.startsWith('v ')) { const p = line.split(/\s+/); V.push(+p[1], +p[2], +p[3]); }
    else if (line.startsWith('f ')) {
      const nv = V.length / 3;
      const idx = line.split(/\s+/).slice(1).map((tok) => { const i = parseInt(tok.split('/')[0], 10); return i < 0 ? nv + i : i - 1; });
      for (let k = 1; k + 1 < idx.length; k++) F.push(idx[0], idx[k], idx[k + 1]);
    }
  }
  if (!F.length) throw new Error('no faces in the OBJ file');
  return { V: Float64Array.from(V), F: Int32Array.from(F) };
}

/** { V, F } of a mesh file's bytes, scaled before welding */
function readMesh(bytes, name, scale = 1.0) {
  const r = String(name).toLowerCase().endsWith('.obj') ? parseOBJ(Buffer.from(bytes).toString('utf8')) : parseSTL(bytes);
  for (let i = 0; i < r.V.length; i++) r.V[i] *= scale;
  return r;
}

const has = (o, k) => o != null && Object.prototype.hasOwnProperty.call(o, k) && o[k] !== null && o[k] !== undefined;
const flag = (v) => (v ? 1 : 0);

// params.js:6-19: what a section of a scene or of a lath group holds when the file leaves a key out
const DEFAULTS = {
  material: { E: 10000.0, G: 625.0, density: 700.0, E90: 800.0, fc90: 8.0, H: 30.0 },
  section: { width: 60.0, height: 30.0, Ix: null, Iy: null, J: null },
  beam: { mode: 'simple', activation: 1.0, stiffness: 1.0, activation_flat: 1.0, activation_edge: 1.0, activation_twist: 1.0,
    stiffness_flat: 1.0, stiffness_edge: 1.0, stiffness_twist: 1.0, stiffness_axial: 1.0, stiffness_shear: 1.0 },
};

/** the builder's keys of a material / section / beam, over `into` */
function fillGroup(into, mat, sec, beam) {
  for (const k of ['E', 'G', 'density', 'E90', 'fc90', 'H']) if (has(mat, k)) into[k] = +mat[k];
  for (const k of ['width', 'height']) if (has(sec, k)) into[k] = +sec[k];
  for (const k of ['Ix', 'Iy', 'J']) into[k] = has(sec, k) ? +sec[k] : 0;
  into.beam_advanced = flag(beam.mode === 'advanced');
  const bmap = { activation: 'activation', stiffness: 'stiffness', activation_flat: 'act_flat', activation_edge: 'act_edge',
    activation_twist: 'act_twist', stiffness_axial: 'k_axial', stiffness_flat: 'k_flat', stiffness_edge: 'k_edge',
    stiffness_twist: 'k_twist', stiffness_shear: 'k_shear' };
  for (const k of Object.keys(bmap)) if (has(beam, k)) into[bmap[k]] = +beam[k];
  return into;
}

/** groups 2, 3, … and [strand, group] pairs of an enabled hybrid weave, or null: each group from the params.js
 *  defaults (fromDict, scene.js:167), its E90 / fc90 / H as given (not rescaled) */
function hybridGroups(d) {
  const h = d.hybrid || {};
  if (!h.enabled) return null;
  const groups = (h.groups || []).map((g) => fillGroup(fillGroup({}, DEFAULTS.material, DEFAULTS.section, DEFAULTS.beam),
    g.material || {}, g.section || {}, g.beam || {}));
  const assign = Object.entries(h.assign || {}).map(([k, v]) => [+k | 0, +v | 0]);
  // groups of pieces of a discretisation (scene.js:57-90): a piece is looked up by its key "s:p" (pieceKey,
  // scene.js:10), so only keys in that exact form apply; the joints are Scene.joints, arc lengths along each strand
  const pieces = [];
  for (const [k, v] of Object.entries(h.pieces || {})) {
    const m = /^(\d+):(\d+)$/.exec(String(k));
    if (m && String(k) === `${+m[1]}:${+m[2]}`) pieces.push([+m[1], +m[2], +v | 0]);
  }
  const joints = [];
  for (const [k, vs] of Object.entries(d.joints || {})) if (/^\d+$/.test(String(k)) && vs) for (const x of vs) joints.push([+k, +x]);
  return { groups, assign, joints, pieces };
}

/** { scene, anchors, loads, flips, mesh, hybrid } of a scene.json object: the scene as the addon's builder reads it */
function sceneFromJSON(d) {
  d = d || {};
  const v = has(d, 'version') ? d.version : SCENE_VERSION;
  if (v > SCENE_VERSION) throw new Error(`scene version ${v} newer than supported ${SCENE_VERSION}`);
  const units = d.units || {};
  const lu = units.length || 'mm', fu = units.force || 'N';
  if (!(lu in LENGTH_TO_M) || !(fu in FORCE_TO_N)) throw new Error(`unknown units ${lu}, ${fu}`);
  const L = LENGTH_TO_M[lu], Fn = FORCE_TO_N[fu];
  const stress = Fn / (L * L) / 1e6;
  const sc = { length_to_m: L, force_to_n: Fn };
  const mat = d.material || {};
  fillGroup(sc, mat, d.section || {}, d.beam || {});
  for (const k of Object.keys(MATERIAL_MM_N)) if (!(k in mat)) sc[k] = MATERIAL_MM_N[k] / stress;
  const sol = d.solver || {};
  for (const k of ['mass_scale', 'damping', 'tol', 'warmup_iter', 'max_iter']) if (has(sol, k)) sc[k] = +sol[k];
  if (has(sol, 'kinetic_damping')) sc.kinetic_damping = flag(sol.kinetic_damping);
  sc.prestress_asbuilt = flag(sol.prestress === 'asbuilt');
  const con = d.contact || {};
  for (const k of ['mu', 'strength', 'stick_strength', 'torsion_factor', 'epsilon', 'slip_rate', 'rot_factor', 'dent_rate',
    'edge_threshold', 'ramp_iter', 'al_every']) if (has(con, k)) sc[k] = +con[k];
  for (const k of ['rot_friction', 'seat_gap', 'augmented']) if (has(con, k)) sc[k] = flag(con[k]);
  if (has(con, 'seat_law')) sc.seat_law = flag(con.seat_law !== 'off');
  if (has(con, 'dent_after_iter')) sc.dent_after = +con.dent_after_iter;
  sc.locked = flag(con.locked);
  const fl = d.floor || {};
  const fz = has(fl, 'z') ? +fl.z : null;
  const fmu = has(fl, 'mu') && +fl.mu >= 0 ? +fl.mu : FLOOR_MU;
  sc.floor_on = flag(fl.enabled && fz !== null && Number.isFinite(fz));
  sc.floor_z = fz !== null && Number.isFinite(fz) ? fz : 0;
  sc.floor_mu = fmu;
  const wv = d.weave || {};
  if (has(wv, 'subdivisions')) sc.subdivisions = +wv.subdivisions;
  if (has(wv, 'overhang')) sc.overhang = +wv.overhang;
  sc.offset_half_height = flag((wv.offset_mode || 'half_height') === 'half_height');
  if (has(wv, 'seed_strand_phase')) sc.seed_strand_phase = +wv.seed_strand_phase;
  const flips = (wv.flip_knots || []).map((f) => [f[0] | 0, f[1] | 0]);
  const lo = d.loads || {};
  const gr = lo.gravity || {};
  sc.gravity = flag(has(gr, 'enabled') ? gr.enabled : true);
  sc.gravity_dir = (gr.direction || [0, 0, -1]).map(Number);
  const udl = lo.udl || {};
  sc.udl = !udl.enabled ? 0 : ((udl.per || 'length') === 'length' ? 1 : 2);
  sc.udl_q = (udl.q || [0, 0, 0]).map(Number);
  const loads = (lo.point || []).map((p) => ({ s: p.frame[0] | 0, k: p.frame[1] | 0, F: p.F.map(Number), local: (p.space || 'world') === 'local' ? 1 : 0 }));
  const anchors = [];
  for (const a of d.anchors || []) {
    const sel = a.select || {}, t = sel.type || 'frame', fix = (a.fix || [1, 1, 1, 0, 0, 0]).map((f) => (f ? 1 : 0));
    if (t === 'frame') anchors.push({ type: SELECTOR.frame, s: sel.id[0] | 0, k: sel.id[1] | 0, fix });
    else if (t === 'knots') for (const i of sel.ids || []) anchors.push({ type: SELECTOR.frame, s: i[0] | 0, k: i[1] | 0, fix });
    else if (t === 'strand') anchors.push({ type: SELECTOR.strand, s: sel.strand | 0, k: 0, fix });
    // regionFrames (model.js:127-145): the knots within `radius` of the seed frame's knot along the mesh edges
    else if (t === 'region') anchors.push({ type: SELECTOR.region, s: sel.id[0] | 0, k: sel.id[1] | 0, fix, radius: +(sel.radius || 0) });
    else if (t in SELECTOR) anchors.push({ type: SELECTOR[t], s: 0, k: 0, fix });
    else throw new Error(`unknown selector type ${t}`);
  }
  const mesh = d.mesh || {};
  return { scene: sc, anchors, loads, flips, mesh: { path: mesh.path || '', scale: has(mesh, 'scale') ? +mesh.scale : 1.0 },
    hybrid: hybridGroups(d) };
}

/** the host model of a scene.json object and its mesh file's bytes */
function modelFromScene(d, meshBytes, meshName) {
  const s = sceneFromJSON(d);
  const { V, F } = readMesh(meshBytes, meshName || s.mesh.path, s.mesh.scale);
  return Model.fromMesh(V, F, s.scene, { anchors: s.anchors, loads: s.loads, flips: s.flips, weld: true,
    groups: s.hybrid ? s.hybrid.groups : null, assign: s.hybrid ? s.hybrid.assign : [],
    joints: s.hybrid ? s.hybrid.joints : [], pieces: s.hybrid ? s.hybrid.pieces : [] });
}

/** the host model of a scene.json file; the mesh path inside it is relative to the file's own folder */
function loadScene(file) {
  const d = JSON.parse(fs.readFileSync(file, 'utf8'));
  const s = sceneFromJSON(d);
  if (!s.mesh.path) throw new Error('the scene names no mesh');
  const meshPath = path.resolve(path.dirname(file), s.mesh.path);
  return modelFromScene(d, fs.readFileSync(meshPath), meshPath);
}

module.exports = { parseSTL, parseOBJ, readMesh, sceneFromJSON, modelFromScene, loadScene };
