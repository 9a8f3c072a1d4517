'use strict';
// An object with the fields of the reference's CpuSolver (web/js/solver_cpu.js:662-720; its state buildState,
// model.js:10-75; rod CosseratRod, rod_cosserat.js:47-100; contact goal PolylineContact, solver_cpu.js:59-98),
// assembled from copies of a built model's arrays (Model.fields()).  The GpuSolver binding reads exactly these names.

function asCpuSolver(f) {
  const C = f.C, K = f.K, P = f.par;
  const st = {
    S: f.S, n: f.n, m: f.m, p: f.p, K: f.K, nodeOffset: f.node_offset, segOffset: f.seg_offset,
    segI0: f.seg_i0, segI1: f.seg_i1, jSeg0: f.j_seg0, jSeg1: f.j_seg1, jNode: f.j_node, jLen: f.j_len, l0: f.l0,
    Omega0: f.omega0, X: f.X, X0: f.X0, V: f.V, Q: f.Q, Q0: f.Q0, Wr: f.W, u: f.u, nPrev: f.n_prev, slip: f.slip,
    stuck: f.stuck, lam: f.lam, iteration: f.iteration, residual: f.residual,
    history: { kinetic: [], strain: [], bending: [], torsion: [], stretch: [], shear: [], contact: [], residual: [] },
  };
  const beam = {
    segEA: f.EA, segGA: f.GA, segEIx: f.EIx, segEIy: f.EIy, segGJ: f.GJ,
    cSeg: Array.from(f.c_seg), cNode: Array.from(f.c_node), cLen: Array.from(f.c_len),
    cB: Array.from({ length: C }, (_, c) => [f.c_B[3 * c], f.c_B[3 * c + 1], f.c_B[3 * c + 2]]), cQ: f.c_Q,
    N: new Float64Array(f.m), V: new Float64Array(2 * f.m), M: new Float64Array(3 * f.p), cM: new Float64Array(2 * C),
  };
  const contact = K ? {
    strand: f.k_strand, segOff: f.k_seg_off, nseg: f.k_nseg, wk: f.k_wk, hk: f.k_hk, fc90: f.k_fc90, E90: f.k_E90,
    Hs: f.k_Hs, w: f.k_w, h: f.k_h, enabled: f.k_enabled, locked: f.k_locked, theta0: f.theta0, rotSlip: f.rot_slip,
    dent: f.dent, dentE: f.dent_e, climb: f.climb, tanpsi: f.tanpsi, Fn: f.Fn, Ft: f.Ft, state: f.kstate, gap: f.gap,
    ctype: f.ctype, faceSide: f.face_side, phi: f.phi, theta: f.theta, Anom: f.Anom, pressure: f.pressure,
    util: f.util, Mrot: f.Mrot, mid: f.mid, penAcc: f.pen_acc, penN: f.pen_n, dentAfterDefault: P.dent_after,
  } : null;
  const scene = {
    solver: { mass_scale: P.mass_scale, damping: P.damping, tol: P.tol, kinetic_damping: !!P.kinetic_damping,
      warmup_iter: P.warmup_iter, max_iter: P.max_iter },
    contact: { ramp_iter: P.ramp_iter, augmented: !!P.augmented, al_every: P.al_every, mu: P.mu, strength: P.strength,
      stick_strength: P.stick_strength, torsion_factor: P.torsion_factor, epsilon: P.epsilon, slip_rate: P.slip_rate,
      rot_friction: !!P.rot_friction, rot_factor: P.rot_factor, seat_law: P.seat_law ? 'L1p' : 'off', seat_gap: !!P.seat_gap,
      dent_rate: P.dent_rate, dent_after_iter: P.dent_after, edge_threshold: P.edge_threshold, locked: false },
  };
  // loads as the reference's Loads holds them: world loads summed per node, frame-local point loads as [node, F]
  const pointLocal = Array.from({ length: f.n_point_local }, (_, q) =>
    [f.point_local_node[q], [f.point_local_F[3 * q], f.point_local_F[3 * q + 1], f.point_local_F[3 * q + 2]]]);
  return {
    st, beam, contact, scene, mass: f.mass, kRotStatic: f.krot_static, lRef: f.l_ref,
    loads: { const: f.load_const, pointLocal }, anchors: { nodeFree: f.node_free, reactions: new Float64Array(3 * f.n) },
    kePrev: f.ke_prev, converged: false, floor: null, protocol: null,
  };
}

module.exports = { asCpuSolver };
