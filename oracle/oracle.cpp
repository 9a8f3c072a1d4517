// ORACLE — test infrastructure, not product code.
//
// A CPU f64 restatement of the reference's dynamic-relaxation solver `CpuSolver` (web/js/solver_cpu.js) with the
// Cosserat rod (web/js/rod_cosserat.js), the crossing contact `PolylineContact` (solver_cpu.js:59-507), the ground
// `Floor` (solver_cpu.js:535-634) and the quaternion helpers (web/js/quat.js).  It is the checker the device solver is
// compared against: only tests, the smoke check and the benchmark's CPU-baseline leg load it.  The product library
// never links or calls it.
//
// Each function names the reference lines it follows.  The arithmetic follows the reference's formulas operation by
// operation in f64 (built without fast-math and without FP contraction), including numpy's pairwise summation order for
// the kinetic energy (solver_cpu.js:42-54).
//
// Pinning (what checks this restatement; see tests/test_oracle.py):
//   CosseratRod.apply        finite-difference gradient of the rod energy (rod_check.mjs method) on a random bent,
//                            twisted, sheared strand with clamps; Timoshenko / Haringx / Kirchhoff closed forms
//                            (rod_bench.mjs) through full relaxations
//   PolylineContact.apply    single-crossing invariants (cases.mjs:14-28): stick → friction equals the in-plane load
//                            share and u is constant; slip → |Ft| = μ Fn on the cone; locked → u constant; force balance
//   Floor.apply              floor_check.mjs invariants: ΣF = weight, cone ratio μ, μ = 0 frictionless
//   step / KD / residual     Σ reactions + Σ loads = 0 at convergence; kinetic-energy monotonicity between resets;
//                            pairwiseSum against numpy.sum order; bit-identical reruns
//   _classify / _support     constructed crossings (tests/test_oracle_contact_laws.py): face / edge / edge-on-edge
//                            types from the section rolls, the strip overlap area wA·wB/sinθ against a polygon
//                            clipping, the edge-on-face Winkler wedge (Fn = ½ κ φ b² ℓ), the support half-extent and
//                            point against the section's corners
//   W3 _seatFriction/_walls  the inclined-wall friction law μ_eff = tan(atan μ + ψ) on a constructed wall angle
//   W2 _rotationalFriction   stick below and return mapping onto the Coulomb moment μ Fn r_eq above it; r_eq against
//                            the friction radius of a uniformly pressed disk (⅔√(A/π)) and of a line contact (ℓ/4)
//   AL updateMultipliers     a pressed crossing's penetration driven to zero with λ = the load, and λ held at the
//                            crushing cap fc90·wA·wB/sinθ with the penetration carrying the rest
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../include/weave.h"

namespace {

constexpr double SUPPORT_SMOOTH = 0.03;                // solver_cpu.js:16
constexpr double TWIST_RATE_MAX = 0.5 * SUPPORT_SMOOTH;  // solver_cpu.js:21
constexpr double ROLL_DAMP = 0.5;                      // solver_cpu.js:28
constexpr double AL_FREE_MIN = 0.25;                   // solver_cpu.js:33
constexpr double RB_MASS = 2.0, RB_VMAX = 0.05, RB_LEVER = 0.05;  // solver_cpu.js:37-39

// numpy-style pairwise summation (solver_cpu.js:42-54)
double pairwiseSum(const double* a, size_t n) {
  if (n < 8) { double r = 0; for (size_t i = 0; i < n; i++) r += a[i]; return r; }
  if (n <= 128) {
    double r[8]; for (int k = 0; k < 8; k++) r[k] = a[k];
    size_t i = 8;
    for (; i < n - (n % 8); i += 8) for (int k = 0; k < 8; k++) r[k] += a[i + k];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  }
  size_t n2 = n >> 1; n2 -= n2 % 8;
  return pairwiseSum(a, n2) + pairwiseSum(a + n2, n - n2);
}

inline double jsSign(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : v); }

// quat.js
void toFrame(const double* q, double* R) {
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double n = std::sqrt(w * w + x * x + y * y + z * z); if (n == 0) n = 1.0;
  w /= n; x /= n; y /= n; z /= n;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y + z * w); R[2] = 2 * (x * z - y * w);
  R[3] = 2 * (x * y - z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z + x * w);
  R[6] = 2 * (x * z + y * w); R[7] = 2 * (y * z - x * w); R[8] = 1 - 2 * (x * x + y * y);
}
void qnormalize(double* q) {
  double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]); if (n == 0) n = 1.0;
  q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}
void qmul(const double* a, const double* b, double* out) {
  double aw = a[0], ax = a[1], ay = a[2], az = a[3], bw = b[0], bx = b[1], by = b[2], bz = b[3];
  out[0] = aw * bw - ax * bx - ay * by - az * bz;
  out[1] = aw * bx + ax * bw + ay * bz - az * by;
  out[2] = aw * by - ax * bz + ay * bw + az * bx;
  out[3] = aw * bz + ax * by - ay * bx + az * bw;
}
void qexp(double tx, double ty, double tz, double* out) {
  double a = std::sqrt(tx * tx + ty * ty + tz * tz), half = 0.5 * a;
  double s = a > 1e-12 ? std::sin(half) / a : 0.5;
  out[0] = std::cos(half); out[1] = tx * s; out[2] = ty * s; out[3] = tz * s;
}
// rod_cosserat.js:33-45
double relRotation(const double* qa, const double* qb, double* out) {
  double aw = qa[0], ax = -qa[1], ay = -qa[2], az = -qa[3];
  double bw = qb[0], bx = qb[1], by = qb[2], bz = qb[3];
  double w = aw * bw - ax * bx - ay * by - az * bz;
  double x = aw * bx + ax * bw + ay * bz - az * by;
  double y = aw * by - ax * bz + ay * bw + az * bx;
  double z = aw * bz + ax * by - ay * bx + az * bw;
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  double s = std::sqrt(x * x + y * y + z * z), th = 2 * std::atan2(s, w);
  double f = s > 1e-12 ? th / s : 2 / std::max(w, 1e-12);
  out[0] = f * x; out[1] = f * y; out[2] = f * z;
  return th;
}

template <class T> std::vector<T> cp(const T* p, size_t n) { return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T{}); }

struct Oracle {
  // ---- model (immutable)
  int n = 0, m = 0, p = 0, K = 0, C = 0, S = 0, npl = 0;
  ws_params par{};
  double lRef = 1.0;
  std::vector<int32_t> segI0, segI1, jSeg0, jSeg1, jNode, cSeg, cNode, kSegOff, kNseg, plNode;
  std::vector<double> l0, jLen, EA, GA, EIxS, EIyS, GJS, EIx, EIy, GJ, omega0, cLen, cB, cQ, mass, krot, loadConst, plF;
  std::vector<double> wk, hk, fc90, E90, Hs, kw, kh, nodeW, nodeH;
  std::vector<uint8_t> nodeFree, fixedAny, nodeFixed, enabled, locked;
  // ---- state
  std::vector<double> X, X0, V, Q, Q0, W, u, nPrev, slip, lam, penAcc, theta0, rotSlip, dent, climb, tanpsi, dentE;
  std::vector<uint8_t> stuck;
  int iteration = 0; int penN = 0; double kePrev = 0.0, residual = INFINITY; bool converged = false;
  // ---- rod outputs (CosseratRod fields)
  std::vector<double> N, Vsh, M, cM, kDyn, Rseg;
  double eStretch = 0, eShear = 0, eBend = 0, eTwist = 0;
  // ---- contact outputs (PolylineContact fields)
  double ramp = 1.0, cEnergy = 0, rotEnergy = 0;
  std::vector<double> Fn, Ft, state, gap, theta, Anom, pressure, util, Mrot, phi, mid;
  std::vector<int32_t> ctype, faceSide;
  // ---- floor
  bool floorOn = false;
  std::vector<double> fax, fay, fF; std::vector<uint8_t> fState;
  double fEnergy = 0, fFnMax = 0, fRamp = 1.0, fc[3] = {0, 0, 0}, fRb = 1e-9, rv[3] = {0, 0, 0}, rw[3] = {0, 0, 0};
  double Qt[3] = {0, 0, 0}, Qr[3] = {0, 0, 0}, d2[3] = {0, 0, 0}; int fnc = 0; double vmax = 0, wmax = 0;
  bool anyAnchor = false;
  // ---- solver
  std::vector<double> F, T, Ltotal, reactions, inertia3;
  double resForce = 0, resTorque = 0, resRigid = 0, fref = 1e-12, ke = 0;
  std::vector<double> histKE, histStrain, histResid;
  bool keepHistory = false;
  // scratch
  struct Ev { std::vector<double> P, t, R, L, tp; std::vector<int> seg, a, b; };
  Ev ev[2];
  std::vector<double> Tn, Qn;

  void init(const ws_model& md, const ws_state& st) {
    n = md.n; m = md.m; p = md.p; K = md.K; C = md.C; S = md.S; npl = md.n_point_local; par = md.par; lRef = md.l_ref;
    segI0 = cp(md.seg_i0, m); segI1 = cp(md.seg_i1, m); jSeg0 = cp(md.j_seg0, p); jSeg1 = cp(md.j_seg1, p); jNode = cp(md.j_node, p);
    l0 = cp(md.l0, m); jLen = cp(md.j_len, p);
    EA = cp(md.EA, m); GA = cp(md.GA, m); EIxS = cp(md.EIx, m); EIyS = cp(md.EIy, m); GJS = cp(md.GJ, m);
    EIx.resize(p); EIy.resize(p); GJ.resize(p);
    for (int j = 0; j < p; j++) { int s0 = jSeg0[j]; EIx[j] = EIxS[s0]; EIy[j] = EIyS[s0]; GJ[j] = GJS[s0]; }  // rod_cosserat.js:56
    omega0 = cp(md.omega0, 3 * p);
    cSeg = cp(md.c_seg, C); cNode = cp(md.c_node, C); cLen = cp(md.c_len, C); cB = cp(md.c_B, 3 * C); cQ = cp(md.c_Q, 4 * C);
    mass = cp(md.mass, n); krot = cp(md.krot_static, 3 * m); loadConst = cp(md.load_const, 3 * n);
    plNode = cp(md.point_local_node, npl); plF = cp(md.point_local_F, 3 * npl);
    nodeFree = cp(md.node_free, 3 * n);
    fixedAny.assign(n, 0); nodeFixed.assign(n, 0);
    for (int i = 0; i < n; i++) {
      bool all = nodeFree[3 * i] && nodeFree[3 * i + 1] && nodeFree[3 * i + 2];
      fixedAny[i] = all ? 0 : 1;
      nodeFixed[i] = (nodeFree[3 * i] || nodeFree[3 * i + 1] || nodeFree[3 * i + 2]) ? 0 : 1;  // setFixed (:481)
      if (!all) anyAnchor = true;
    }
    if (C > 0) anyAnchor = true;
    kSegOff = cp(md.k_seg_off, 2 * K); kNseg = cp(md.k_nseg, 2 * K);
    wk = cp(md.k_wk, 2 * K); hk = cp(md.k_hk, 2 * K); fc90 = cp(md.k_fc90, 2 * K); E90 = cp(md.k_E90, 2 * K); Hs = cp(md.k_Hs, 2 * K);
    kw = cp(md.k_w, K); kh = cp(md.k_h, K); enabled = cp(md.k_enabled, K); locked = cp(md.k_locked, K);
    // state
    X = cp(st.X, 3 * n); X0 = st.X0 ? cp(st.X0, 3 * n) : X; V = cp(st.V, 3 * n);
    Q = cp(st.Q, 4 * m); Q0 = st.Q0 ? cp(st.Q0, 4 * m) : Q; W = cp(st.W, 3 * m);
    u = cp(st.u, 2 * K); nPrev = cp(st.n_prev, 3 * K); slip = cp(st.slip, K); lam = cp(st.lam, K); penAcc = cp(st.pen_acc, K);
    stuck = st.stuck ? cp(st.stuck, K) : std::vector<uint8_t>(K, 1);
    theta0 = st.theta0 ? cp(st.theta0, K) : std::vector<double>(K, NAN); rotSlip = cp(st.rot_slip, K);
    dent = cp(st.dent, 2 * K); climb = cp(st.climb, 2 * K); tanpsi = st.tanpsi ? cp(st.tanpsi, 2 * K) : std::vector<double>(2 * K, 1.0);
    dentE = cp(st.dent_e, K);
    iteration = st.iteration; penN = st.pen_n; kePrev = st.ke_prev; residual = st.residual;
    N.assign(m, 0); Vsh.assign(2 * m, 0); M.assign(3 * p, 0); cM.assign(2 * C, 0); kDyn.assign(3 * m, 0); Rseg.assign(9 * m, 0);
    Fn = st.Fn ? cp(st.Fn, K) : std::vector<double>(K, 0.0);
    Ft.assign(K, 0); state.assign(K, 0); gap.assign(K, 0); theta.assign(K, 0); Anom.assign(K, 0); pressure.assign(K, 0);
    util.assign(K, 0); Mrot.assign(K, 0); phi.assign(2 * K, 0); mid.assign(3 * K, 0); ctype.assign(K, 0); faceSide.assign(K, -1);
    F.assign(3 * n, 0); T.assign(3 * m, 0); Ltotal.assign(3 * n, 0); reactions.assign(3 * n, 0); inertia3.assign(3 * m, 0);
    // Floor constructor (solver_cpu.js:536-563)
    floorOn = par.floor_on != 0;
    if (floorOn) {
      nodeW = cp(md.node_w, n); nodeH = cp(md.node_h, n);
      fax.resize(n); fay.resize(n); fF.assign(3 * n, 0); fState.assign(n, 0);
      for (int i = 0; i < n; i++) { fax[i] = st.floor_a ? st.floor_a[2 * i] : X[3 * i]; fay[i] = st.floor_a ? st.floor_a[2 * i + 1] : X[3 * i + 1]; }
      double c[3] = {0, 0, 0};
      for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) c[k] += X[3 * i + k] / std::max(n, 1);
      double R2 = 0;
      for (int i = 0; i < n; i++) R2 = std::max(R2, (X[3 * i] - c[0]) * (X[3 * i] - c[0]) + (X[3 * i + 1] - c[1]) * (X[3 * i + 1] - c[1]) + (X[3 * i + 2] - c[2]) * (X[3 * i + 2] - c[2]));
      for (int k = 0; k < 3; k++) fc[k] = c[k];
      fRb = std::max(std::sqrt(R2), 1e-9);
      // a state that carries the rigid freedoms (a download of them: Floor.rv / rw, and Rb as the Floor kept it)
      if (st.rb_radius > 0) fRb = st.rb_radius;
      for (int k = 0; k < 3; k++) { rv[k] = st.rb_v[k]; rw[k] = st.rb_w[k]; }
      double hmin = INFINITY;
      for (int i = 0; i < n; i++) hmin = std::min(hmin, std::min(nodeH[i], nodeW[i]));
      vmax = RB_VMAX * (std::isfinite(hmin) && hmin > 0 ? hmin : 1.0); wmax = vmax / fRb;
    }
    for (int s = 0; s < 2; s++) {
      ev[s].P.assign(3 * K, 0); ev[s].t.assign(3 * K, 0); ev[s].R.assign(9 * K, 0); ev[s].L.assign(K, 0); ev[s].tp.assign(K, 0);
      ev[s].seg.assign(K, 0); ev[s].a.assign(K, 0); ev[s].b.assign(K, 0);
    }
  }

  // ============================================================================================ CosseratRod.apply
  // rod_cosserat.js:114-183
  void rodApply(std::vector<double>& Fo, std::vector<double>& To) {
    double* R = Rseg.data();
    for (int j = 0; j < m; j++) toFrame(&Q[4 * j], R + 9 * j);
    double eS = 0, eSh = 0, eB = 0, eT = 0;
    for (int j = 0; j < m; j++) {
      int a = segI0[j], b = segI1[j], o = 9 * j; double L0 = l0[j];
      double ex = X[3 * b] - X[3 * a], ey = X[3 * b + 1] - X[3 * a + 1], ez = X[3 * b + 2] - X[3 * a + 2];
      double s1 = (R[o] * ex + R[o + 1] * ey + R[o + 2] * ez) / L0;
      double s2 = (R[o + 3] * ex + R[o + 4] * ey + R[o + 5] * ez) / L0;
      double s3 = (R[o + 6] * ex + R[o + 7] * ey + R[o + 8] * ez) / L0 - 1.0;
      double n1 = GA[j] * s1, n2 = GA[j] * s2, n3 = EA[j] * s3;
      double nx = n1 * R[o] + n2 * R[o + 3] + n3 * R[o + 6];
      double ny = n1 * R[o + 1] + n2 * R[o + 4] + n3 * R[o + 7];
      double nz = n1 * R[o + 2] + n2 * R[o + 5] + n3 * R[o + 8];
      Fo[3 * a] += nx; Fo[3 * a + 1] += ny; Fo[3 * a + 2] += nz;
      Fo[3 * b] -= nx; Fo[3 * b + 1] -= ny; Fo[3 * b + 2] -= nz;
      To[3 * j] += ey * nz - ez * ny; To[3 * j + 1] += ez * nx - ex * nz; To[3 * j + 2] += ex * ny - ey * nx;
      N[j] = n3; Vsh[2 * j] = n1; Vsh[2 * j + 1] = n2;
      eS += 0.5 * L0 * n3 * s3; eSh += 0.5 * L0 * (n1 * s1 + n2 * s2);
    }
    double psi[3];
    std::fill(kDyn.begin(), kDyn.end(), 0.0);
    for (int j = 0; j < p; j++) {
      int sa = jSeg0[j], sb = jSeg1[j], o = 9 * sa; double lb = jLen[j];
      double th = relRotation(&Q[4 * sa], &Q[4 * sb], psi);
      double d0 = psi[0] / lb - omega0[3 * j], d1 = psi[1] / lb - omega0[3 * j + 1], d2_ = psi[2] / lb - omega0[3 * j + 2];
      double m0 = EIx[j] * d0, m1 = EIy[j] * d1, m2 = GJ[j] * d2_;
      {
        double q0 = psi[0] * psi[0], q1 = psi[1] * psi[1], q2 = psi[2] * psi[2], f = 0.25 / lb;
        double k0 = f * (EIy[j] * q2 + GJ[j] * q1), k1 = f * (EIx[j] * q2 + GJ[j] * q0), k2 = f * (EIx[j] * q1 + EIy[j] * q0);
        kDyn[3 * sa] += k0; kDyn[3 * sa + 1] += k1; kDyn[3 * sa + 2] += k2;
        kDyn[3 * sb] += k0; kDyn[3 * sb + 1] += k1; kDyn[3 * sb + 2] += k2;
      }
      double beta = th < 1e-3 ? 1.0 / 12 + th * th / 720 : 1 / (th * th) - (1 + std::cos(th)) / (2 * th * std::sin(th));
      double c0 = psi[1] * m2 - psi[2] * m1, c1 = psi[2] * m0 - psi[0] * m2, c2 = psi[0] * m1 - psi[1] * m0;
      double e0 = psi[1] * c2 - psi[2] * c1, e1 = psi[2] * c0 - psi[0] * c2, e2 = psi[0] * c1 - psi[1] * c0;
      double j0 = m0 + 0.5 * c0 + beta * e0, j1 = m1 + 0.5 * c1 + beta * e1, j2 = m2 + 0.5 * c2 + beta * e2;
      double tx = j0 * R[o] + j1 * R[o + 3] + j2 * R[o + 6];
      double ty = j0 * R[o + 1] + j1 * R[o + 4] + j2 * R[o + 7];
      double tz = j0 * R[o + 2] + j1 * R[o + 5] + j2 * R[o + 8];
      To[3 * sa] += tx; To[3 * sa + 1] += ty; To[3 * sa + 2] += tz;
      To[3 * sb] -= tx; To[3 * sb + 1] -= ty; To[3 * sb + 2] -= tz;
      M[3 * j] = m0; M[3 * j + 1] = m1; M[3 * j + 2] = m2;
      eB += 0.5 * lb * (m0 * d0 + m1 * d1); eT += 0.5 * lb * m2 * d2_;
    }
    if (C) {
      double Rc[9], QQ[8];
      for (int c = 0; c < C; c++) {
        int j = cSeg[c]; double lc = cLen[c]; const double* B = &cB[3 * c];
        for (int k = 0; k < 4; k++) { QQ[k] = cQ[4 * c + k]; QQ[4 + k] = Q[4 * j + k]; }
        double th = relRotation(QQ, QQ + 4, psi);
        toFrame(QQ, Rc);
        double d0 = psi[0] / lc, d1 = psi[1] / lc, d2_ = psi[2] / lc;
        double m0 = B[0] * d0, m1 = B[1] * d1, m2 = B[2] * d2_;
        double beta = th < 1e-3 ? 1.0 / 12 + th * th / 720 : 1 / (th * th) - (1 + std::cos(th)) / (2 * th * std::sin(th));
        double c0 = psi[1] * m2 - psi[2] * m1, c1 = psi[2] * m0 - psi[0] * m2, c2 = psi[0] * m1 - psi[1] * m0;
        double e0 = psi[1] * c2 - psi[2] * c1, e1 = psi[2] * c0 - psi[0] * c2, e2 = psi[0] * c1 - psi[1] * c0;
        double j0 = m0 + 0.5 * c0 + beta * e0, j1 = m1 + 0.5 * c1 + beta * e1, j2 = m2 + 0.5 * c2 + beta * e2;
        To[3 * j] -= j0 * Rc[0] + j1 * Rc[3] + j2 * Rc[6];
        To[3 * j + 1] -= j0 * Rc[1] + j1 * Rc[4] + j2 * Rc[7];
        To[3 * j + 2] -= j0 * Rc[2] + j1 * Rc[5] + j2 * Rc[8];
        cM[2 * c] = m0; cM[2 * c + 1] = m1;
        eB += 0.5 * lc * (m0 * d0 + m1 * d1); eT += 0.5 * lc * m2 * d2_;
      }
    }
    eStretch = eS; eShear = eSh; eBend = eB; eTwist = eT;
  }

  // ============================================================================================ node caches
  // nodeQuats (model.js:89-105)
  void nodeQuats(std::vector<double>& out) {
    std::vector<int> ref(n, -1);
    for (int j = 0; j < m; j++) ref[segI1[j]] = j;
    for (int j = 0; j < m; j++) ref[segI0[j]] = j;
    out.assign(4 * n, 0.0);
    for (const auto* idx : {&segI0, &segI1}) {
      for (int j = 0; j < m; j++) {
        int i = (*idx)[j], r = ref[i];
        double d = Q[4 * j] * Q[4 * r] + Q[4 * j + 1] * Q[4 * r + 1] + Q[4 * j + 2] * Q[4 * r + 2] + Q[4 * j + 3] * Q[4 * r + 3];
        double sg = d < 0 ? -1 : 1;
        for (int c = 0; c < 4; c++) out[4 * i + c] += sg * Q[4 * j + c];
      }
    }
    for (int i = 0; i < n; i++) qnormalize(&out[4 * i]);
  }
  // PolylineContact._prepare (solver_cpu.js:179-193)
  void prepare() {
    Tn.assign(3 * n, 0.0);
    std::vector<double> eh(3 * m);
    for (int j = 0; j < m; j++) {
      int a = segI0[j], b = segI1[j];
      double ex = X[3 * b] - X[3 * a], ey = X[3 * b + 1] - X[3 * a + 1], ez = X[3 * b + 2] - X[3 * a + 2];
      double l = std::max(std::sqrt(ex * ex + ey * ey + ez * ez), 1e-12);
      eh[3 * j] = ex / l; eh[3 * j + 1] = ey / l; eh[3 * j + 2] = ez / l;
    }
    for (int j = 0; j < m; j++) { int a = segI0[j]; for (int k = 0; k < 3; k++) Tn[3 * a + k] += eh[3 * j + k]; }
    for (int j = 0; j < m; j++) { int b = segI1[j]; for (int k = 0; k < 3; k++) Tn[3 * b + k] += eh[3 * j + k]; }
    for (int i = 0; i < n; i++) {
      double l = std::max(std::sqrt(Tn[3 * i] * Tn[3 * i] + Tn[3 * i + 1] * Tn[3 * i + 1] + Tn[3 * i + 2] * Tn[3 * i + 2]), 1e-12);
      for (int k = 0; k < 3; k++) Tn[3 * i + k] /= l;
    }
    nodeQuats(Qn);
  }
  // _segment (solver_cpu.js:144-149)
  void segmentOf(int side, int c, int& seg, double& t) const {
    double uu = u[2 * c + side]; int ns = kNseg[2 * c + side];
    int i = int(std::min(std::max(std::floor(uu), 0.0), double(std::max(ns - 1, 0))));
    t = std::min(std::max(uu - i, 0.0), 1.0);
    seg = kSegOff[2 * c + side] + i;
  }
  // _eval (solver_cpu.js:197-228)
  void evalSide(int side) {
    Ev& e = ev[side];
    double Rn[9], q[4];
    for (int c = 0; c < K; c++) {
      int seg; double t; segmentOf(side, c, seg, t);
      int a = segI0[seg], b = segI1[seg];
      double dx = X[3 * b] - X[3 * a], dy = X[3 * b + 1] - X[3 * a + 1], dz = X[3 * b + 2] - X[3 * a + 2];
      double L = std::sqrt(dx * dx + dy * dy + dz * dz);
      double t2 = t * t, t3 = t * t * t;
      double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
      double d00 = 6 * t2 - 6 * t, d10 = 3 * t2 - 4 * t + 1, d01 = -6 * t2 + 6 * t, d11 = 3 * t2 - 2 * t;
      double tg[3];
      for (int k = 0; k < 3; k++) {
        e.P[3 * c + k] = h00 * X[3 * a + k] + (h10 * L) * Tn[3 * a + k] + h01 * X[3 * b + k] + (h11 * L) * Tn[3 * b + k];
        tg[k] = d00 * X[3 * a + k] + (d10 * L) * Tn[3 * a + k] + d01 * X[3 * b + k] + (d11 * L) * Tn[3 * b + k];
      }
      double tl = std::max(std::sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]), 1e-12);
      tg[0] /= tl; tg[1] /= tl; tg[2] /= tl;
      const double* qa = &Qn[4 * a]; const double* qb = &Qn[4 * b];
      double sgn = (qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2] + qa[3] * qb[3]) < 0 ? -1 : 1;
      for (int k = 0; k < 4; k++) q[k] = (1 - t) * qa[k] + t * sgn * qb[k];
      qnormalize(q);
      toFrame(q, Rn);
      double y[3] = {Rn[3], Rn[4], Rn[5]};
      double yt = y[0] * tg[0] + y[1] * tg[1] + y[2] * tg[2];
      y[0] -= yt * tg[0]; y[1] -= yt * tg[1]; y[2] -= yt * tg[2];
      double yl = std::max(std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]), 1e-12);
      y[0] /= yl; y[1] /= yl; y[2] /= yl;
      double x[3] = {y[1] * tg[2] - y[2] * tg[1], y[2] * tg[0] - y[0] * tg[2], y[0] * tg[1] - y[1] * tg[0]};
      double* R = &e.R[9 * c];
      R[0] = x[0]; R[1] = x[1]; R[2] = x[2]; R[3] = y[0]; R[4] = y[1]; R[5] = y[2]; R[6] = tg[0]; R[7] = tg[1]; R[8] = tg[2];
      e.t[3 * c] = tg[0]; e.t[3 * c + 1] = tg[1]; e.t[3 * c + 2] = tg[2];
      e.L[c] = L; e.seg[c] = seg; e.tp[c] = t; e.a[c] = a; e.b[c] = b;
    }
  }

  // ============================================================================================ PolylineContact
  // _classify (solver_cpu.js:231-271)
  void classify(const std::vector<double>& nrm, std::vector<double>& cFace) {
    cFace.assign(2 * K, 0); std::vector<double> Z0(2 * K); std::vector<uint8_t> edge(2 * K);
    const Ev& A = ev[0]; const Ev& B = ev[1];
    for (int c = 0; c < K; c++) {
      for (int side = 0; side < 2; side++) {
        const double* R = side == 0 ? &A.R[9 * c] : &B.R[9 * c];
        double sg = side == 0 ? 1.0 : -1.0; int i = 2 * c + side;
        double nx = (nrm[3 * c] * R[0] + nrm[3 * c + 1] * R[1] + nrm[3 * c + 2] * R[2]) * sg;
        double ny = (nrm[3 * c] * R[3] + nrm[3 * c + 1] * R[4] + nrm[3 * c + 2] * R[5]) * sg;
        double sx = std::fabs(std::tanh(nx / SUPPORT_SMOOTH)), sy = std::fabs(std::tanh(ny / SUPPORT_SMOOTH));
        edge[i] = sx * sy > par.edge_threshold ? 1 : 0;
        bool yFace = std::fabs(ny) >= std::fabs(nx);
        cFace[i] = yFace ? wk[i] : hk[i];
        Z0[i] = yFace ? hk[i] : wk[i];
        phi[i] = std::atan(std::min(std::fabs(nx), std::fabs(ny)) / std::max(std::max(std::fabs(nx), std::fabs(ny)), 1e-12));
      }
      const double* tA = &A.t[3 * c]; const double* tB = &B.t[3 * c];
      double cr[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
      double sinT = std::min(std::max(std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]), 1e-6), 1.0);
      theta[c] = std::asin(sinT);
      int nEdge = edge[2 * c] + edge[2 * c + 1];
      ctype[c] = nEdge;
      faceSide[c] = nEdge == 1 ? (edge[2 * c + 1] ? 0 : 1) : -1;
      double Fa = std::fabs(Fn[c]);
      double Aa = cFace[2 * c] * cFace[2 * c + 1] / sinT, pres = Fa / std::max(Aa, 1e-12);
      double ut = pres / std::max(std::min(Hs[2 * c], Hs[2 * c + 1]), 1e-12);
      if (nEdge == 1) {
        int fs = faceSide[c], es = 1 - fs;
        double ell = cFace[2 * c + fs] / sinT, kap = E90[2 * c + fs] / std::max(Z0[2 * c + fs], 1e-12);
        double phiE = std::max(phi[2 * c + es], 1e-4);
        double b = std::sqrt(2.0 * Fa / std::max(ell * kap * phiE, 1e-12));
        bool full = b >= cFace[2 * c + es];
        double Afe = full ? cFace[2 * c] * cFace[2 * c + 1] / sinT : b * ell;
        double pfe = full ? Fa / std::max(Afe, 1e-12) : kap * phiE * b;
        Aa = Afe; pres = pfe; ut = pfe / std::max(fc90[2 * c + es], 1e-12);
      } else if (nEdge == 2) {
        double fc = std::min(fc90[2 * c], fc90[2 * c + 1]);
        Aa = Fa / std::max(fc, 1e-12); pres = fc * (Fa > 0 ? 1 : 0); ut = Fa > 0 ? 1.0 : 0.0;
      }
      Anom[c] = Aa; pressure[c] = pres; util[c] = ut;
    }
  }
  // _support (solver_cpu.js:274-282)
  double support(const double* R, const double* nv, int c, double sign, int side, double* r) const {
    double w = wk[2 * c + side], h_ = hk[2 * c + side];
    double nx = (nv[0] * R[0] + nv[1] * R[1] + nv[2] * R[2]) * sign, ny = (nv[0] * R[3] + nv[1] * R[4] + nv[2] * R[5]) * sign;
    double h = 0.5 * w * std::fabs(nx) + 0.5 * h_ * std::fabs(ny);
    double ax = 0.5 * w * std::tanh(nx / SUPPORT_SMOOTH), ay = 0.5 * h_ * std::tanh(ny / SUPPORT_SMOOTH);
    for (int k = 0; k < 3; k++) r[k] = ax * R[k] + ay * R[3 + k];
    return h;
  }
  // _walls (solver_cpu.js:285-293)
  double walls(int sideCut, const double* R, const double* nv, const double* mm) const {
    double sg = sideCut == 0 ? 1.0 : -1.0;
    double nx = (nv[0] * R[0] + nv[1] * R[1] + nv[2] * R[2]) * sg, ny = (nv[0] * R[3] + nv[1] * R[4] + nv[2] * R[5]) * sg;
    double fx[3], fy[3]; for (int k = 0; k < 3; k++) { fx[k] = jsSign(nx) * R[k]; fy[k] = jsSign(ny) * R[3 + k]; }
    const double* f = (fx[0] * mm[0] + fx[1] * mm[1] + fx[2] * mm[2]) > 0 ? fx : fy;
    double cosPsi = std::min(std::max(std::fabs(f[0] * nv[0] + f[1] * nv[1] + f[2] * nv[2]), 1e-6), 1.0);
    return std::tan(std::acos(cosPsi));
  }
  // _seatFriction (solver_cpu.js:295-315)
  double seatFriction(int c, const double* St, const double* nv, const double* tA, const double* tB) {
    double mu = par.mu;
    double norm = std::sqrt(St[0] * St[0] + St[1] * St[1] + St[2] * St[2]);
    double dn = std::max(norm, 1e-12), dhat[3] = {St[0] / dn, St[1] / dn, St[2] / dn}; bool has = norm > 1e-12;
    for (int j = 0; j < 2; j++) {
      bool act = dent[2 * c + j] > 0 && has;
      if (!act) continue;
      const double* tCut = j == 0 ? tB : tA; const double* Rcut = j == 0 ? &ev[1].R[9 * c] : &ev[0].R[9 * c];
      double mm[3] = {nv[1] * tCut[2] - nv[2] * tCut[1], nv[2] * tCut[0] - nv[0] * tCut[2], nv[0] * tCut[1] - nv[1] * tCut[0]};
      double ml = std::max(std::sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), 1e-12);
      for (double& v : mm) v /= ml;
      double sgn = j == 0 ? 1.0 : -1.0;
      double dj = (dhat[0] * mm[0] + dhat[1] * mm[1] + dhat[2] * mm[2]) * sgn;
      double wallDir[3]; for (int k = 0; k < 3; k++) wallDir[k] = jsSign(dj) * mm[k] * sgn;
      double tp = walls(1 - j, Rcut, nv, wallDir);
      tanpsi[2 * c + j] = tp;
      double add = std::fabs(dj) * (par.mu * tp < 1.0 ? (par.mu + tp) / std::max(1.0 - par.mu * tp, 1e-6) - par.mu : 9.0 * par.mu);
      mu += std::min(add, 9.0 * par.mu);
    }
    return mu;
  }
  // _seatUpdate (solver_cpu.js:317-344)
  void seatUpdate(int c, bool slipping, double da, double db, const double* tA, const double* tB, const double* nv, double fn,
                  const std::vector<double>& cFace, bool active) {
    double inc[3] = {da * tA[0] - db * tB[0], da * tA[1] - db * tB[1], da * tA[2] - db * tB[2]};
    for (int j = 0; j < 2; j++) {
      const double* tCut = j == 0 ? tB : tA;
      double mm[3] = {nv[1] * tCut[2] - nv[2] * tCut[1], nv[2] * tCut[0] - nv[0] * tCut[2], nv[0] * tCut[1] - nv[1] * tCut[0]};
      double ml = std::max(std::sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), 1e-12);
      for (double& v : mm) v /= ml;
      double stp = (inc[0] * mm[0] + inc[1] * mm[1] + inc[2] * mm[2]) * (j == 0 ? 1.0 : -1.0);
      int i = 2 * c + j; bool has = dent[i] > 0;
      if (has && slipping) climb[i] += stp;
      if (has && std::fabs(climb[i]) >= dent[i] / std::max(tanpsi[i], 1e-6)) { dent[i] = 0; climb[i] = 0; }
    }
    int after = par.dent_after;
    if (iteration < after) return;
    double sinp0 = std::sin(phi[2 * c]) * std::cos(phi[2 * c]), sinp1 = std::sin(phi[2 * c + 1]) * std::cos(phi[2 * c + 1]);
    double sinT = std::sin(theta[c]);
    if (ctype[c] == 2 && active && !locked[c]) {
      double fc = std::min(fc90[2 * c], fc90[2 * c + 1]);
      double dEst = std::sqrt(std::fabs(fn) / std::max(fc, 1e-12) * sinT * sinp0 * sinp1);
      for (int j = 0; j < 2; j++) { double cur = dent[2 * c + j]; dent[2 * c + j] = std::max(cur, std::min(cur + par.dent_rate * dEst, dEst)); }
    }
    if (ctype[c] == 1 && active && !locked[c]) {
      int fs = faceSide[c], es = 1 - fs;
      double ell = cFace[2 * c + fs] / std::max(sinT, 1e-6);
      double dE = std::fabs(fn) / std::max(fc90[2 * c + es] * ell, 1e-12) * (es == 0 ? sinp0 : sinp1);
      dentE[c] = std::max(dentE[c], dE);
    }
  }
  // _rotationalFriction (solver_cpu.js:347-379)
  void rotationalFriction(std::vector<double>& To, const std::vector<double>& nrm, const std::vector<uint8_t>& active, double kt) {
    const Ev& A = ev[0]; const Ev& B = ev[1];
    std::vector<double> Mv(K), dthA(K), krA(K);
    for (int c = 0; c < K; c++) {
      const double* tA = &A.t[3 * c]; const double* tB = &B.t[3 * c];
      double cr[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
      double th = std::atan2(cr[0] * nrm[3 * c] + cr[1] * nrm[3 * c + 1] + cr[2] * nrm[3 * c + 2], tA[0] * tB[0] + tA[1] * tB[1] + tA[2] * tB[2]);
      if (std::isnan(theta0[c]) && active[c]) theta0[c] = th;
      bool ok = active[c] && !std::isnan(theta0[c]);
      double sinT = std::max(std::sin(theta[c]), 1e-6);
      double req = 0.38 * std::sqrt(wk[2 * c] * wk[2 * c + 1] / sinT);
      if (ctype[c] == 1) { double cw = faceSide[c] == 0 ? wk[2 * c] : wk[2 * c + 1]; req = 0.25 * cw / sinT; }
      else if (ctype[c] == 2) req = 0.38 * std::sqrt(std::max(Anom[c], 1e-12));
      double kr = kt * req * req;
      double dth = ok ? th - theta0[c] : 0.0;
      dth = std::fmod(std::fmod(dth + M_PI, 2 * M_PI) + 2 * M_PI, 2 * M_PI) - M_PI;
      double Mc = kr * dth;
      double Mlim = locked[c] ? INFINITY : par.rot_factor * par.mu * std::fabs(Fn[c]) * req;
      if (ok && std::fabs(Mc) > Mlim) {
        double rate = std::min(std::max(par.slip_rate, 1e-3), 1.0);
        double adv = rate * (1.0 - Mlim / std::max(std::fabs(Mc), 1e-12)) * dth;
        adv = std::min(std::max(adv, -0.05), 0.05);
        theta0[c] += adv; rotSlip[c] += std::fabs(adv); dth -= adv;
        Mc = kr * dth;
      }
      Mv[c] = ok ? Mc : 0.0; dthA[c] = dth; krA[c] = kr;
    }
    Mrot = Mv;
    for (int c = 0; c < K; c++) { int s = A.seg[c]; for (int k = 0; k < 3; k++) To[3 * s + k] += Mv[c] * nrm[3 * c + k]; }
    for (int c = 0; c < K; c++) { int s = B.seg[c]; for (int k = 0; k < 3; k++) To[3 * s + k] -= Mv[c] * nrm[3 * c + k]; }
    std::vector<double> e(K); for (int c = 0; c < K; c++) e[c] = krA[c] * dthA[c] * dthA[c];
    rotEnergy = 0.5 * pairwiseSum(e.data(), K);
  }
  // apply (solver_cpu.js:381-478)
  void contactApply(std::vector<double>& Fo, std::vector<double>& To) {
    if (!K) return;
    prepare(); evalSide(0); evalSide(1);
    const Ev& A = ev[0]; const Ev& B = ev[1];
    double kn = ramp * par.strength, kt = ramp * par.stick_strength;
    bool seat = par.seat_law != 0, aug = par.augmented != 0;
    bool accum = iteration >= std::max(par.ramp_iter, 0);
    std::vector<double> nrm(3 * K), FA(3 * K), FB(3 * K), rAx(3 * K), rBx(3 * K), daA(K), dbA(K), ePen(K), eTan(K);
    std::vector<uint8_t> slipA(K), activeA(K);
    double rA[3], rB[3];
    for (int c = 0; c < K; c++) {
      const double* tA = &A.t[3 * c]; const double* tB = &B.t[3 * c];
      double nx = tA[1] * tB[2] - tA[2] * tB[1], ny = tA[2] * tB[0] - tA[0] * tB[2], nz = tA[0] * tB[1] - tA[1] * tB[0];
      double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (nn < 1e-6) { nx = nPrev[3 * c]; ny = nPrev[3 * c + 1]; nz = nPrev[3 * c + 2]; }
      else { double d = std::max(nn, 1e-12); nx /= d; ny /= d; nz /= d; }
      if (nx * nPrev[3 * c] + ny * nPrev[3 * c + 1] + nz * nPrev[3 * c + 2] < 0) { nx = -nx; ny = -ny; nz = -nz; }
      nPrev[3 * c] = nx; nPrev[3 * c + 1] = ny; nPrev[3 * c + 2] = nz;
      nrm[3 * c] = nx; nrm[3 * c + 1] = ny; nrm[3 * c + 2] = nz;
      double nv[3] = {nx, ny, nz};
      double Sv[3] = {B.P[3 * c] - A.P[3 * c], B.P[3 * c + 1] - A.P[3 * c + 1], B.P[3 * c + 2] - A.P[3 * c + 2]};
      for (int k = 0; k < 3; k++) mid[3 * c + k] = 0.5 * (A.P[3 * c + k] + B.P[3 * c + k]);
      double g = Sv[0] * nx + Sv[1] * ny + Sv[2] * nz;
      double hA = support(&A.R[9 * c], nv, c, +1.0, 0, rA);
      double hB = support(&B.R[9 * c], nv, c, -1.0, 1, rB);
      double D = hA + hB;
      if (seat && par.seat_gap) D -= dent[2 * c] + dent[2 * c + 1] + dentE[c];
      double pen = D - g;
      if (accum) penAcc[c] += pen;
      double fn = 0; bool active;
      if (aug) {
        double trial = lam[c] + kn * pen;
        active = (trial > 0 || locked[c]) && enabled[c];
        fn = active ? (locked[c] ? trial : std::max(trial, 0.0)) : 0;
      } else {
        active = (pen > 0 || locked[c]) && enabled[c];
        fn = active ? kn * pen : 0.0;
      }
      activeA[c] = active ? 1 : 0;
      double St[3] = {Sv[0] - g * nx, Sv[1] - g * ny, Sv[2] - g * nz};
      double taa = tA[0] * tA[0] + tA[1] * tA[1] + tA[2] * tA[2], tbb = tB[0] * tB[0] + tB[1] * tB[1] + tB[2] * tB[2], tab = tA[0] * tB[0] + tA[1] * tB[1] + tA[2] * tB[2];
      double sa = St[0] * tA[0] + St[1] * tA[1] + St[2] * tA[2], sb = St[0] * tB[0] + St[1] * tB[1] + St[2] * tB[2];
      double det = taa * tbb - tab * tab; bool ok = det > 1e-8;
      double ca = ok ? (sa * tbb - sb * tab) / det : 0.0, cb = ok ? (sb * taa - sa * tab) / det : 0.0;
      double Ftv = std::sqrt((kt * St[0]) * (kt * St[0]) + (kt * St[1]) * (kt * St[1]) + (kt * St[2]) * (kt * St[2]));
      double muDir = par.mu;
      if (seat) muDir = seatFriction(c, St, nv, tA, tB);
      double limit = locked[c] ? INFINITY : muDir * std::fabs(fn);
      bool inContact = active && ok;
      bool slip_ = inContact && Ftv > limit;
      bool free_ = !active && ok && enabled[c];
      double scale = 0;
      if (slip_) scale = std::min(std::max(par.slip_rate, 1e-3), 1.0) * (1.0 - limit / std::max(Ftv, 1e-12));
      if (free_) scale = 1.0;
      double eps = std::max(par.epsilon, 1e-9);
      double da = std::min(std::max(ca * scale, -eps), eps), db = std::min(std::max(cb * scale, -eps), eps);
      bool moved = slip_ || free_;
      if (moved) { u[2 * c] += da / std::max(A.L[c], 1e-12); u[2 * c + 1] -= db / std::max(B.L[c], 1e-12); }
      u[2 * c] = std::min(std::max(u[2 * c], 0.0), double(kNseg[2 * c]));
      u[2 * c + 1] = std::min(std::max(u[2 * c + 1], 0.0), double(kNseg[2 * c + 1]));
      if (slip_) {
        double q0 = da * tA[0] + db * tB[0], q1 = da * tA[1] + db * tB[1], q2 = da * tA[2] + db * tB[2];
        slip[c] += std::sqrt(q0 * q0 + q1 * q1 + q2 * q2);
      }
      stuck[c] = moved ? 0 : 1;
      double Se[3] = {0, 0, 0};
      if (active) { Se[0] = St[0] - da * tA[0] - db * tB[0]; Se[1] = St[1] - da * tA[1] - db * tB[1]; Se[2] = St[2] - da * tA[2] - db * tB[2]; }
      double fT[3] = {kt * Se[0], kt * Se[1], kt * Se[2]};
      for (int k = 0; k < 3; k++) { FA[3 * c + k] = -fn * nv[k] + fT[k]; FB[3 * c + k] = fn * nv[k] - fT[k]; }
      Fn[c] = fn;
      Ft[c] = std::sqrt(fT[0] * fT[0] + fT[1] * fT[1] + fT[2] * fT[2]);
      state[c] = slip_ ? 1 : 0; gap[c] = g;
      slipA[c] = slip_ ? 1 : 0; daA[c] = da; dbA[c] = db;
      for (int k = 0; k < 3; k++) { rAx[3 * c + k] = rA[k]; rBx[3 * c + k] = rB[k]; }
      ePen[c] = active ? (aug ? fn * fn / std::max(kn, 1e-300) : kn * pen * pen) : 0.0;
      eTan[c] = Se[0] * Se[0] + Se[1] * Se[1] + Se[2] * Se[2];
    }
    if (accum) penN += 1;
    std::vector<double> cFace;
    classify(nrm, cFace);
    if (seat) for (int c = 0; c < K; c++) seatUpdate(c, slipA[c], daA[c], dbA[c], &A.t[3 * c], &B.t[3 * c], &nrm[3 * c], Fn[c], cFace, activeA[c]);
    for (int c = 0; c < K; c++) { int i = A.a[c]; double w = 1 - A.tp[c]; for (int k = 0; k < 3; k++) Fo[3 * i + k] += w * FA[3 * c + k]; }
    for (int c = 0; c < K; c++) { int i = A.b[c]; double w = A.tp[c]; for (int k = 0; k < 3; k++) Fo[3 * i + k] += w * FA[3 * c + k]; }
    for (int c = 0; c < K; c++) { int i = B.a[c]; double w = 1 - B.tp[c]; for (int k = 0; k < 3; k++) Fo[3 * i + k] += w * FB[3 * c + k]; }
    for (int c = 0; c < K; c++) { int i = B.b[c]; double w = B.tp[c]; for (int k = 0; k < 3; k++) Fo[3 * i + k] += w * FB[3 * c + k]; }
    double tf = par.torsion_factor;
    for (int c = 0; c < K; c++) {
      int s = A.seg[c]; const double* r = &rAx[3 * c]; const double* f = &FA[3 * c];
      To[3 * s] += tf * (r[1] * f[2] - r[2] * f[1]); To[3 * s + 1] += tf * (r[2] * f[0] - r[0] * f[2]); To[3 * s + 2] += tf * (r[0] * f[1] - r[1] * f[0]);
    }
    for (int c = 0; c < K; c++) {
      int s = B.seg[c]; const double* r = &rBx[3 * c]; const double* f = &FB[3 * c];
      To[3 * s] += tf * (r[1] * f[2] - r[2] * f[1]); To[3 * s + 1] += tf * (r[2] * f[0] - r[0] * f[2]); To[3 * s + 2] += tf * (r[0] * f[1] - r[1] * f[0]);
    }
    if (par.rot_friction) rotationalFriction(To, nrm, activeA, kt);
    cEnergy = 0.5 * pairwiseSum(ePen.data(), K) + 0.5 * kt * pairwiseSum(eTan.data(), K) + (par.rot_friction ? rotEnergy : 0.0);
  }
  // addTwistStiffness (solver_cpu.js:170-176)
  void addTwistStiffness(std::vector<double>& kSeg) {
    double kn = ramp * par.strength, kt = ramp * par.stick_strength, tf = std::max(par.torsion_factor, 1e-3);
    for (int side = 0; side < 2; side++) for (int c = 0; c < K; c++) {
      double kth = tf * ((Fn[c] + kn * 0.25 * kh[c]) * (0.5 * kw[c]) / SUPPORT_SMOOTH + kt * (0.5 * kh[c]) * (0.5 * kh[c]));
      int seg; double t; segmentOf(side, c, seg, t);
      kSeg[seg] += kth;
    }
  }
  double freeWeight(int side, int c) const {  // solver_cpu.js:483-487
    int seg; double t; segmentOf(side, c, seg, t);
    int a = segI0[seg], b = segI1[seg];
    return (nodeFixed[a] ? 0 : (1 - t) * (1 - t)) + (nodeFixed[b] ? 0 : t * t);
  }
  double capForce(int c) const { return std::min(fc90[2 * c], fc90[2 * c + 1]) * wk[2 * c] * wk[2 * c + 1] / std::max(std::sin(theta[c]), 1e-6); }
  static double alValue(double l, double penMean, double kn, bool lk, bool en, bool pinned, double cap) {  // :492-497
    if (!en || pinned) return 0;
    double v = l + kn * penMean;
    if (!lk) v = std::max(v, 0.0);
    return std::min(std::max(v, -cap), cap);
  }
  void updateMultipliers() {  // :499-506
    double kn = ramp * par.strength; int nn = std::max(penN, 1);
    for (int c = 0; c < K; c++) {
      bool pinned = freeWeight(0, c) < AL_FREE_MIN || freeWeight(1, c) < AL_FREE_MIN;
      lam[c] = alValue(lam[c], penAcc[c] / nn, kn, locked[c], enabled[c], pinned, capForce(c));
    }
    std::fill(penAcc.begin(), penAcc.end(), 0.0); penN = 0;
  }

  // ============================================================================================ Floor
  void rigidMasses(double ms, double* out) const {  // :565-572
    double kn = fRamp * par.strength, kt = fRamp * par.stick_strength, f = ms * RB_MASS;
    double nc = std::max(fnc, 1), rmin = (RB_LEVER * fRb) * (RB_LEVER * fRb) * kn;
    out[0] = std::max(f * nc * kt, 1e-30); out[1] = out[0]; out[2] = std::max(f * nc * kn, 1e-30);
    out[3] = std::max(f * std::max(kn * d2[1] + kt * d2[2], rmin), 1e-30); out[4] = std::max(f * std::max(kn * d2[0] + kt * d2[2], rmin), 1e-30);
    out[5] = std::max(f * std::max(kt * (d2[0] + d2[1]), rmin), 1e-30);
  }
  double rigidResidual(double fr) const {
    return std::max(std::hypot(Qt[0], Qt[1], Qt[2]) / fr, std::hypot(Qr[0], Qr[1], Qr[2]) / (fr * fRb));
  }
  void floorApply(std::vector<double>& Fo, double rmp, const std::vector<double>* loads) {  // :583-627
    double q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, dxx = 0, dyy = 0, dzz = 0; int nc = 0;
    double z0 = par.floor_z, mu = std::max(par.floor_mu, 0.0);
    double kn = rmp * par.strength, kt = rmp * par.stick_strength;
    if (n) { double cx = 0, cy = 0, cz = 0; for (int i = 0; i < n; i++) { cx += X[3 * i]; cy += X[3 * i + 1]; cz += X[3 * i + 2]; } fc[0] = cx / n; fc[1] = cy / n; fc[2] = cz / n; }
    std::vector<double> qn; nodeQuats(qn);
    double R[9], e = 0, fmax = 0;
    fRamp = rmp;
    for (int i = 0; i < n; i++) {
      toFrame(&qn[4 * i], R);
      double pen = (z0 - X[3 * i + 2]) + (0.5 * nodeW[i] * std::fabs(R[2]) + 0.5 * nodeH[i] * std::fabs(R[5]));
      double* out = &fF[3 * i];
      if (pen > 0) {
        double fn = kn * pen;
        double sx = X[3 * i] - fax[i], sy = X[3 * i + 1] - fay[i];
        double Ftv = kt * std::sqrt(sx * sx + sy * sy), limit = mu * fn;
        double scale = Ftv > limit ? 1.0 - limit / std::max(Ftv, 1e-12) : 0.0;
        double dax = sx * scale, day = sy * scale;
        if (scale > 0) { fax[i] += dax; fay[i] += day; }
        double ex = sx - dax, ey = sy - day;
        out[0] = -kt * ex; out[1] = -kt * ey; out[2] = fn;
        fState[i] = scale > 0 ? 2 : 1;
        e += 0.5 * kn * pen * pen + 0.5 * kt * (ex * ex + ey * ey);
        if (fn > fmax) fmax = fn;
        double dx = X[3 * i] - fc[0], dy = X[3 * i + 1] - fc[1], dz = X[3 * i + 2] - fc[2];
        dxx += dx * dx; dyy += dy * dy; dzz += dz * dz; nc++;
      } else {
        fax[i] = X[3 * i]; fay[i] = X[3 * i + 1];
        out[0] = 0; out[1] = 0; out[2] = 0; fState[i] = 0;
      }
      Fo[3 * i] += out[0]; Fo[3 * i + 1] += out[1]; Fo[3 * i + 2] += out[2];
      if (loads) {
        const auto& L = *loads;
        double gx = L[3 * i] + out[0], gy = L[3 * i + 1] + out[1], gz = L[3 * i + 2] + out[2];
        double dx = X[3 * i] - fc[0], dy = X[3 * i + 1] - fc[1], dz = X[3 * i + 2] - fc[2];
        q0 += gx; q1 += gy; q2 += gz; q3 += dy * gz - dz * gy; q4 += dz * gx - dx * gz; q5 += dx * gy - dy * gx;
      }
    }
    fEnergy = e; fFnMax = fmax;
    if (loads) { Qt[0] = q0; Qt[1] = q1; Qt[2] = q2; Qr[0] = q3; Qr[1] = q4; Qr[2] = q5; d2[0] = dxx; d2[1] = dyy; d2[2] = dzz; fnc = nc; }
  }

  // ============================================================================================ CpuSolver.step
  double forceRef() const {  // :843-853
    double a = 0; for (double v : Ltotal) a = std::max(a, std::fabs(v));
    double b = 0; for (double v : N) b = std::max(b, std::fabs(v));
    double c = 0; for (double v : M) c = std::max(c, std::fabs(v));
    double lmin = INFINITY; for (double v : jLen) lmin = std::min(lmin, v);
    c = p ? c / std::max(lmin, 1e-12) : 0;
    double d = 0; if (par.contacts) { d = -INFINITY; for (double v : Fn) d = std::max(d, v); }
    double e = floorOn ? fFnMax : 0;
    return std::max(std::max(std::max(a, b), std::max(c, d)), std::max(e, 1e-12));
  }

  // UTM load-test replays (calibrate.js:154-215): force control (replayForce: the probe's constant load is axis·F
  // for each schedule value) or displacement control (replayDisplacement: the probe pinned, placed at P0 + d·axis at
  // rest).  Per value: settle (runAsync, calibrate.js:119-130: at most `settle` iterations, ending at convergence),
  // then `avg` single iterations averaging the probe's travel (X − P0)·axis and its reaction·axis.  One row per value:
  // travel, reaction, value, iteration.  The schedule is used as given (the caller flattens the load cycles,
  // forceSequence :147-151, and drops non-positive displacements as the reference's job setup does).
  void replay(int mode, int probe, const double* ax, const double* sched, int nsched, int settle, int avg, double* out) {
    const double P0[3] = {X[3 * probe], X[3 * probe + 1], X[3 * probe + 2]};
    double applied[3] = {0, 0, 0};
    if (mode == 2) {   // pin the probe (calibrate.js:190-193): reactions reported, the rigid freedoms off
      for (int k = 0; k < 3; k++) nodeFree[3 * probe + k] = 0;
      fixedAny[probe] = 1; anyAnchor = true;
    }
    for (int q = 0; q < nsched; q++) {
      const double v = sched[q];
      if (mode == 1) {
        for (int k = 0; k < 3; k++) loadConst[3 * probe + k] -= applied[k];
        for (int k = 0; k < 3; k++) { applied[k] = ax[k] * v; loadConst[3 * probe + k] += applied[k]; }
      } else {
        for (int k = 0; k < 3; k++) { X[3 * probe + k] = P0[k] + ax[k] * v; V[3 * probe + k] = 0.0; }
      }
      converged = false;
      step(std::max(settle, 1));
      double accD = 0, accR = 0;
      const int na = std::max(avg, 1);
      for (int i = 0; i < na; i++) {
        step(1);
        accD += (X[3 * probe] - P0[0]) * ax[0] + (X[3 * probe + 1] - P0[1]) * ax[1] + (X[3 * probe + 2] - P0[2]) * ax[2];
        accR += reactions[3 * probe] * ax[0] + reactions[3 * probe + 1] * ax[1] + reactions[3 * probe + 2] * ax[2];
      }
      out[4 * q] = accD / na; out[4 * q + 1] = accR / na; out[4 * q + 2] = v; out[4 * q + 3] = iteration;
    }
  }

  // The evaluator of the stability check (stability.js:27-52) at positions Xe and frames Qe: node forces before the
  // anchors (rod, constant and frame-local loads, contacts, ground), body-axis torques Rᵀ T, and the frame inertia of
  // that state (static + contact righting + bending coupling).  The contact and ground passes run as in an iteration
  // (ramp at the current iteration, as the device's crossing kernel has it) and every piece of state is restored.
  void evaluate(const double* Xe, const double* Qe, double* Fo, double* To, double* Io) {
    const Oracle saved = *this;
    const int rampN = std::max(par.ramp_iter, 0);
    const bool contact = par.contacts != 0 && K > 0;
    std::copy(Xe, Xe + 3 * n, X.begin());
    std::copy(Qe, Qe + 4 * m, Q.begin());
    if (contact) ramp = rampN == 0 ? 1.0 : std::min(1.0, 0.01 + 0.99 * iteration / rampN);
    std::fill(F.begin(), F.end(), 0.0); std::fill(T.begin(), T.end(), 0.0);
    rodApply(F, T);
    Ltotal = loadConst;
    if (npl) {
      std::vector<double> nq; nodeQuats(nq); double Rl[9];
      for (int q = 0; q < npl; q++) {
        int nid = plNode[q]; const double* Fv = &plF[3 * q];
        toFrame(&nq[4 * nid], Rl);
        for (int k = 0; k < 3; k++) Ltotal[3 * nid + k] += Rl[k] * Fv[0] + Rl[3 + k] * Fv[1] + Rl[6 + k] * Fv[2];
      }
    }
    for (int i = 0; i < 3 * n; i++) F[i] += Ltotal[i];
    if (contact) contactApply(F, T);
    if (floorOn) floorApply(F, rampN <= 0 ? 1.0 : std::min(1.0, 0.01 + 0.99 * iteration / rampN), nullptr);
    std::copy(F.begin(), F.end(), Fo);
    double R[9];
    std::vector<double> kTw(m, 0.0);
    if (contact) addTwistStiffness(kTw);
    for (int j = 0; j < m; j++) {
      toFrame(&Q[4 * j], R);
      for (int k = 0; k < 3; k++) {
        To[3 * j + k] = R[3 * k] * T[3 * j] + R[3 * k + 1] * T[3 * j + 1] + R[3 * k + 2] * T[3 * j + 2];
        Io[3 * j + k] = par.mass_scale * std::max(krot[3 * j + k] + kTw[j] + kDyn[3 * j + k], 1e-12);
      }
    }
    *this = saved;
  }

  // one or more iterations (solver_cpu.js:737-841); returns the number of iterations run
  int step(int nIter) {
    double ms = par.mass_scale;
    int rampN = std::max(par.ramp_iter, 0);
    bool contact = par.contacts != 0 && K > 0;
    std::vector<double> R(9), tm(3 * m), kTw(m), Xprev(3 * n), Qprev(4 * m), dq(4), qq(4), q0(4), kin(3 * n), kinR(3 * m), qw(4, 0.0);
    bool rb = floorOn && !anyAnchor;
    double Mrb[6] = {0, 0, 0, 0, 0, 0};
    if (floorOn && !rb) { for (int k = 0; k < 3; k++) { rv[k] = 0; rw[k] = 0; } }
    bool broke = false; int done = 0;
    for (int it = 0; it < nIter; it++) {
      if (contact) ramp = rampN == 0 ? 1.0 : std::min(1.0, 0.01 + 0.99 * iteration / rampN);
      std::fill(F.begin(), F.end(), 0.0); std::fill(T.begin(), T.end(), 0.0);
      rodApply(F, T);
      Ltotal = loadConst;
      if (npl) {
        std::vector<double> nq; nodeQuats(nq); double Rl[9];
        for (int q = 0; q < npl; q++) {
          int nid = plNode[q]; const double* Fv = &plF[3 * q];
          toFrame(&nq[4 * nid], Rl);
          for (int k = 0; k < 3; k++) Ltotal[3 * nid + k] += Rl[k] * Fv[0] + Rl[3 + k] * Fv[1] + Rl[6 + k] * Fv[2];
        }
      }
      for (int i = 0; i < 3 * n; i++) F[i] += Ltotal[i];
      if (contact) contactApply(F, T);
      if (floorOn) floorApply(F, rampN <= 0 ? 1.0 : std::min(1.0, 0.01 + 0.99 * iteration / rampN), &Ltotal);
      std::fill(reactions.begin(), reactions.end(), 0.0);
      for (int i = 0; i < n; i++) if (fixedAny[i]) for (int k = 0; k < 3; k++) reactions[3 * i + k] = -F[3 * i + k];
      for (int i = 0; i < 3 * n; i++) if (!nodeFree[i]) F[i] = 0;
      for (int j = 0; j < m; j++) {
        toFrame(&Q[4 * j], R.data());
        for (int k = 0; k < 3; k++) tm[3 * j + k] = R[3 * k] * T[3 * j] + R[3 * k + 1] * T[3 * j + 1] + R[3 * k + 2] * T[3 * j + 2];
      }
      std::fill(kTw.begin(), kTw.end(), 0.0);
      if (contact) addTwistStiffness(kTw);
      for (int j = 0; j < m; j++) for (int k = 0; k < 3; k++) inertia3[3 * j + k] = ms * std::max(krot[3 * j + k] + kTw[j] + kDyn[3 * j + k], 1e-12);
      Xprev = X; Qprev = Q;
      bool warm = iteration < par.warmup_iter && contact;
      if (warm) { std::fill(F.begin(), F.end(), 0.0); std::fill(V.begin(), V.end(), 0.0); if (rb) for (int k = 0; k < 3; k++) { rv[k] = 0; rw[k] = 0; } }
      for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) V[3 * i + k] += F[3 * i + k] / mass[i];
      for (int j = 0; j < 3 * m; j++) W[j] += tm[j] / inertia3[j];
      for (int j = 0; j < m; j++) { W[3 * j + 2] = std::min(std::max(W[3 * j + 2], -TWIST_RATE_MAX), TWIST_RATE_MAX); if (kTw[j] > 0) W[3 * j + 2] *= ROLL_DAMP; }
      if (par.damping != 1.0) { for (int i = 0; i < 3 * n; i++) V[i] *= par.damping; for (int j = 0; j < 3 * m; j++) W[j] *= par.damping; }
      for (int i = 0; i < 3 * n; i++) if (!nodeFree[i]) V[i] = 0;
      for (int i = 0; i < 3 * n; i++) X[i] += V[i];
      if (rb) {
        rigidMasses(ms, Mrb);
        if (!warm) for (int k = 0; k < 3; k++) {
          rv[k] = std::min(std::max((rv[k] + Qt[k] / Mrb[k]) * par.damping, -vmax), vmax);
          rw[k] = std::min(std::max((rw[k] + Qr[k] / Mrb[3 + k]) * par.damping, -wmax), wmax);
        }
      }
      bool rbMove = rb && (rv[0] != 0 || rv[1] != 0 || rv[2] != 0 || rw[0] != 0 || rw[1] != 0 || rw[2] != 0);
      if (rbMove) {
        for (int i = 0; i < n; i++) {
          double dx = Xprev[3 * i] - fc[0], dy = Xprev[3 * i + 1] - fc[1], dz = Xprev[3 * i + 2] - fc[2];
          X[3 * i] += rv[0] + rw[1] * dz - rw[2] * dy; X[3 * i + 1] += rv[1] + rw[2] * dx - rw[0] * dz; X[3 * i + 2] += rv[2] + rw[0] * dy - rw[1] * dx;
        }
        qexp(rw[0], rw[1], rw[2], qw.data());
      }
      for (int j = 0; j < m; j++) {
        qexp(W[3 * j], W[3 * j + 1], W[3 * j + 2], dq.data());
        for (int k = 0; k < 4; k++) q0[k] = Q[4 * j + k];
        qmul(q0.data(), dq.data(), qq.data());
        if (rbMove) { for (int k = 0; k < 4; k++) q0[k] = qq[k]; qmul(qw.data(), q0.data(), qq.data()); }
        qnormalize(qq.data());
        for (int k = 0; k < 4; k++) Q[4 * j + k] = qq[k];
      }
      for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) kin[3 * i + k] = mass[i] * V[3 * i + k] * V[3 * i + k];
      for (int j = 0; j < 3 * m; j++) kinR[j] = inertia3[j] * W[j] * W[j];
      double kev = 0.5 * (pairwiseSum(kin.data(), kin.size()) + pairwiseSum(kinR.data(), kinR.size()));
      if (rb) for (int k = 0; k < 3; k++) kev += 0.5 * (Mrb[k] * rv[k] * rv[k] + Mrb[3 + k] * rw[k] * rw[k]);
      if (par.kinetic_damping && kev < kePrev) {
        X = Xprev; Q = Qprev; std::fill(V.begin(), V.end(), 0.0); std::fill(W.begin(), W.end(), 0.0); kev = 0.0;
        if (rb) for (int k = 0; k < 3; k++) { rv[k] = 0; rw[k] = 0; }
      }
      kePrev = kev; ke = kev;
      iteration += 1;
      if (contact && par.augmented && par.al_every > 0 && iteration > rampN && (iteration - rampN) % par.al_every == 0) updateMultipliers();
      double fmax = 0, tmax = 0;
      for (int i = 0; i < n; i++) { double f = std::sqrt(F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] + F[3 * i + 2] * F[3 * i + 2]); if (f > fmax) fmax = f; }
      for (int j = 0; j < m; j++) { double t = std::sqrt(tm[3 * j] * tm[3 * j] + tm[3 * j + 1] * tm[3 * j + 1] + tm[3 * j + 2] * tm[3 * j + 2]); if (t > tmax) tmax = t; }
      fref = forceRef();
      resForce = n ? fmax / fref : 0; resTorque = m ? tmax / (fref * lRef) : 0; resRigid = rb ? rigidResidual(fref) : 0;
      double resid = std::max(std::max(resForce, resTorque), resRigid);
      residual = iteration > par.warmup_iter ? resid : INFINITY;
      if (keepHistory) {
        histKE.push_back(kev);
        histStrain.push_back(eStretch + eShear + eBend + eTwist + (contact ? cEnergy : 0) + (floorOn ? fEnergy : 0));
        histResid.push_back(residual);
      }
      done++;
      if (residual <= par.tol) { converged = true; broke = true; break; }
    }
    if (!broke) converged = residual <= par.tol;
    return done;
  }

  void exportState(ws_state& o) const {
    auto put = [](double* dst, const std::vector<double>& src) { if (dst) std::copy(src.begin(), src.end(), dst); };
    o.iteration = iteration; o.converged = converged; o.pen_n = penN; o.ke_prev = kePrev; o.residual = residual;
    o.res_force = resForce; o.res_torque = resTorque; o.fref = fref; o.ke = ke;
    o.e_stretch = eStretch; o.e_shear = eShear; o.e_bend = eBend; o.e_twist = eTwist; o.e_contact = cEnergy;
    put(o.X, X); put(o.X0, X0); put(o.V, V); put(o.Q, Q); put(o.Q0, Q0); put(o.W, W);
    put(o.u, u); put(o.n_prev, nPrev); put(o.slip, slip); put(o.lam, lam); put(o.pen_acc, penAcc);
    put(o.theta0, theta0); put(o.rot_slip, rotSlip); put(o.dent, dent); put(o.climb, climb); put(o.tanpsi, tanpsi); put(o.dent_e, dentE);
    if (o.stuck) std::copy(stuck.begin(), stuck.end(), o.stuck);
    if (o.floor_a && floorOn) for (int i = 0; i < n; i++) { o.floor_a[2 * i] = fax[i]; o.floor_a[2 * i + 1] = fay[i]; }
    put(o.N, N); put(o.Vsh, Vsh); put(o.M, M); put(o.cM, cM); put(o.reactions, reactions);
    put(o.Fn, Fn); put(o.Ft, Ft); put(o.gap, gap); put(o.kstate, state); put(o.theta, theta); put(o.Anom, Anom);
    put(o.pressure, pressure); put(o.util, util); put(o.Mrot, Mrot); put(o.phi, phi); put(o.mid, mid);
    if (o.ctype) std::copy(ctype.begin(), ctype.end(), o.ctype);
    if (o.face_side) std::copy(faceSide.begin(), faceSide.end(), o.face_side);
    if (floorOn) {                 // Floor.F / state / energy / FnMax of the last apply
      put(o.floor_F, fF);
      if (o.floor_state) std::copy(fState.begin(), fState.end(), o.floor_state);
      o.e_floor = fEnergy; o.fn_floor = fFnMax;
    } else {
      o.e_floor = 0; o.fn_floor = 0;
    }
    // the rigid freedoms (Floor.rv / rw, Rb) and the reference point the next evaluation takes: the centroid of the
    // current positions (Floor.apply, solver_cpu.js:590)
    const bool rb = floorOn && !anyAnchor;
    double cs[3] = {0, 0, 0};
    if (rb && n) {
      for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) cs[k] += X[3 * i + k];
      for (int k = 0; k < 3; k++) cs[k] /= n;
    }
    for (int k = 0; k < 3; k++) { o.rb_v[k] = rb ? rv[k] : 0.0; o.rb_w[k] = rb ? rw[k] : 0.0; o.rb_c[k] = cs[k]; }
    o.rb_radius = rb ? fRb : 0.0;
  }
};

thread_local std::string g_err;

}  // namespace

extern "C" {

struct oracle_t { Oracle o; };

WS_API const char* oracle_last_error(void) { return g_err.c_str(); }

WS_API oracle_t* oracle_create(const ws_model* md, const ws_state* st, int32_t keep_history) {
  if (!md || !st) { g_err = "oracle_create: null model"; return nullptr; }
  auto* h = new oracle_t();
  h->o.init(*md, *st);
  h->o.keepHistory = keep_history != 0;
  return h;
}
WS_API void oracle_free(oracle_t* h) { delete h; }
WS_API int32_t oracle_step(oracle_t* h, int32_t n_iter) { return h->o.step(n_iter); }
WS_API void oracle_evaluate(oracle_t* h, const double* X, const double* Q, double* F, double* T, double* I3) {
  h->o.evaluate(X, Q, F, T, I3);
}
WS_API void oracle_replay(oracle_t* h, int32_t mode, int32_t probe, const double* axis, const double* sched, int32_t n,
                          int32_t settle, int32_t avg, double* out) {
  h->o.replay(mode, probe, axis, sched, n, settle, avg, out);
}
WS_API int32_t oracle_converged(oracle_t* h) { return h->o.converged ? 1 : 0; }
WS_API int32_t oracle_iteration(oracle_t* h) { return h->o.iteration; }
WS_API double oracle_residual(oracle_t* h) { return h->o.residual; }
WS_API double oracle_ke(oracle_t* h) { return h->o.ke; }
WS_API void oracle_get(oracle_t* h, ws_state* out) { h->o.exportState(*out); }
WS_API void oracle_inertia(oracle_t* h, double* out) { std::copy(h->o.inertia3.begin(), h->o.inertia3.end(), out); }
// the ground's per-node forces of the last iteration and state (0 off, 1 stick, 2 slide) (Floor.F / Floor.state)
WS_API int32_t oracle_floor(oracle_t* h, double* F, uint8_t* state, double* rigid_residual) {
  if (!h->o.floorOn) return 0;
  std::copy(h->o.fF.begin(), h->o.fF.end(), F);
  std::copy(h->o.fState.begin(), h->o.fState.end(), state);
  if (rigid_residual) *rigid_residual = h->o.resRigid;
  return 1;
}
WS_API int32_t oracle_history(oracle_t* h, double* ke, double* strain, double* resid, int32_t cap) {
  int32_t nn = int32_t(std::min<size_t>(h->o.histKE.size(), size_t(cap)));
  for (int i = 0; i < nn; i++) { if (ke) ke[i] = h->o.histKE[i]; if (strain) strain[i] = h->o.histStrain[i]; if (resid) resid[i] = h->o.histResid[i]; }
  return nn;
}
// rod forces, world torques and energy at the oracle's current X, Q (CosseratRod.apply), for gradient checks
WS_API double oracle_rod_apply(oracle_t* h, const double* X, const double* Q, double* Fo, double* To) {
  Oracle& o = h->o;
  if (X) std::copy(X, X + 3 * o.n, o.X.begin());
  if (Q) std::copy(Q, Q + 4 * o.m, o.Q.begin());
  std::vector<double> Fv(3 * o.n, 0.0), Tv(3 * o.m, 0.0);
  o.rodApply(Fv, Tv);
  if (Fo) std::copy(Fv.begin(), Fv.end(), Fo);
  if (To) std::copy(Tv.begin(), Tv.end(), To);
  return o.eStretch + o.eShear + o.eBend + o.eTwist;
}
// numpy-order pairwise sum, exposed for its known-answer test
WS_API double oracle_pairwise_sum(const double* a, int64_t n) { return pairwiseSum(a, size_t(n)); }

// ---- probes of single contact laws on constructed geometry (the closed-form tests of _classify, _support,
// _seatFriction / _walls, _rotationalFriction and alValue).  RA, RB: the evaluated frames of both sides of crossing c,
// rows x (width axis), y (height axis), t (tangent), as _eval leaves them (solver_cpu.js:219-226); nrm: the contact
// normal of crossing c.
static void probe_sides(Oracle& o, int c, const double* RA, const double* RB) {
  for (int s = 0; s < 2; s++) {
    const double* R = s == 0 ? RA : RB;
    std::copy(R, R + 9, o.ev[s].R.begin() + 9 * c);
    std::copy(R + 6, R + 9, o.ev[s].t.begin() + 3 * c);
  }
}
static std::vector<double> probe_normal(const Oracle& o, int c, const double* nrm) {
  std::vector<double> nv(3 * o.K, 0.0);
  std::copy(nrm, nrm + 3, nv.begin() + 3 * c);
  return nv;
}
// _classify (solver_cpu.js:231-271) with Fn[c] = fn: out = ctype, faceSide, theta, phi A, phi B, Anom, pressure,
// util, the face extent cFace of A and B
WS_API void oracle_probe_classify(oracle_t* h, int32_t c, const double* RA, const double* RB, const double* nrm, double fn,
                                  double* out) {
  Oracle& o = h->o;
  probe_sides(o, c, RA, RB);
  o.Fn[c] = fn;
  std::vector<double> cFace;
  o.classify(probe_normal(o, c, nrm), cFace);
  const double v[10] = {double(o.ctype[c]), double(o.faceSide[c]), o.theta[c], o.phi[2 * c], o.phi[2 * c + 1], o.Anom[c],
                        o.pressure[c], o.util[c], cFace[2 * c], cFace[2 * c + 1]};
  std::copy(v, v + 10, out);
}
// _support (solver_cpu.js:274-282): the half-extent of the section of `side` towards sign·nrm, its support point r
WS_API double oracle_probe_support(oracle_t* h, int32_t c, const double* R, const double* nrm, double sign, int32_t side,
                                   double* r) {
  return h->o.support(R, nrm, c, sign, side, r);
}
// _seatFriction (solver_cpu.js:295-315) with dents dent2 on both sides of crossing c and tangential offset St:
// returns the directional friction coefficient, tan psi of both sides in tanpsi2
WS_API double oracle_probe_seat(oracle_t* h, int32_t c, const double* RA, const double* RB, const double* nrm,
                                const double* St, const double* dent2, double* tanpsi2) {
  Oracle& o = h->o;
  probe_sides(o, c, RA, RB);
  o.dent[2 * c] = dent2[0]; o.dent[2 * c + 1] = dent2[1];
  o.tanpsi[2 * c] = 1.0; o.tanpsi[2 * c + 1] = 1.0;
  const double mu = o.seatFriction(c, St, nrm, RA + 6, RB + 6);
  tanpsi2[0] = o.tanpsi[2 * c]; tanpsi2[1] = o.tanpsi[2 * c + 1];
  return mu;
}
// _rotationalFriction (solver_cpu.js:347-379) of crossing c after _classify, with Fn[c] = fn, the reference angle
// theta0 (NaN: unset) and the stick spring kt: out = M about n, theta0 after, rotational slip of the call
WS_API void oracle_probe_rot(oracle_t* h, int32_t c, const double* RA, const double* RB, const double* nrm, double fn,
                             double theta0, int32_t active, double kt, double* out) {
  Oracle& o = h->o;
  probe_sides(o, c, RA, RB);
  o.Fn[c] = fn;
  const std::vector<double> nv = probe_normal(o, c, nrm);
  std::vector<double> cFace;
  o.classify(nv, cFace);
  o.theta0[c] = theta0; o.rotSlip[c] = 0.0;
  std::vector<uint8_t> act(o.K, 0);
  act[c] = active ? 1 : 0;
  std::vector<double> Tv(3 * o.m, 0.0);
  o.rotationalFriction(Tv, nv, act, kt);
  out[0] = o.Mrot[c]; out[1] = o.theta0[c]; out[2] = o.rotSlip[c];
}
// PolylineContact.alValue (solver_cpu.js:492-497)
WS_API double oracle_probe_al_value(double lam, double pen_mean, double kn, int32_t locked, int32_t enabled, int32_t pinned,
                                    double cap) {
  return Oracle::alValue(lam, pen_mean, kn, locked != 0, enabled != 0, pinned != 0, cap);
}

}  // extern "C"
