// HIP kernels (gfx950) of the dynamic-relaxation iteration.  One iteration is two launches:
//
//   k_cross   one thread per crossing: node tangents and mean node frames around both sides' contact points,
//             C¹ Hermite point / tangent / frame on each side, contact normal with side persistence, penalty normal
//             force, Coulomb stick / slip with the u advance, W1 classification, W2 / W3 / AL branches; writes the
//             crossing state and one record per crossing side (force, lever torque, righting stiffness) — no atomics.
//             Follows PolylineContact.apply (web/js/solver_cpu.js:381-478) and the crossing branch of k_force
//             (web/js/wgsl.js:571-777).
//   k_strand  one workgroup per strand tile (a run of consecutive nodes of one lath): positions and frames of the
//             tile plus a one-node halo staged in LDS; segment stretch / shear forces (rod_cosserat.js:118-133), joint
//             bending / twist torques with J_l⁻ᵀ and the bending-coupling inertia (:136-159), clamp joints (:161-180);
//             the crossing-side records of the tile gathered in a fixed order; node and frame integration
//             (solver_cpu.js:751-808); FP64 partial sums; the last tile of an instance reduces the partials in index
//             order and runs kinetic damping, the residual and the stop test (solver_cpu.js:809-837).
//
// The kinetic-damping roll-back needs no copy: positions and frames are double-buffered and `cur` simply does not
// flip on a roll-back, while the stale velocities are read as zero by the next iteration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "device.hpp"

namespace wsd {

constexpr double SUPPORT_SMOOTH = 0.03;              // solver_cpu.js:16
constexpr double TWIST_RATE_MAX = 0.5 * SUPPORT_SMOOTH;
constexpr double ROLL_DAMP = 0.5;
constexpr double AL_FREE_MIN = 0.25;

template <typename A, typename B> __device__ __forceinline__ typename std::common_type<A, B>::type dmax(A a, B b) {
  typedef typename std::common_type<A, B>::type T; return T(a) > T(b) ? T(a) : T(b);
}
template <typename A, typename B> __device__ __forceinline__ typename std::common_type<A, B>::type dmin(A a, B b) {
  typedef typename std::common_type<A, B>::type T; return T(a) < T(b) ? T(a) : T(b);
}
template <typename R> __device__ __forceinline__ R clampv(R v, R lo, R hi) { return dmin(dmax(v, lo), hi); }

__device__ __forceinline__ float fsqrt_(float x) { return sqrtf(x); }
__device__ __forceinline__ double fsqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ float ftanh_(float x) { return tanhf(x); }
__device__ __forceinline__ double ftanh_(double x) { return tanh(x); }
__device__ __forceinline__ float fatan_(float x) { return atanf(x); }
__device__ __forceinline__ double fatan_(double x) { return atan(x); }
__device__ __forceinline__ float fasin_(float x) { return asinf(x); }
__device__ __forceinline__ double fasin_(double x) { return asin(x); }
__device__ __forceinline__ float fsin_(float x) { return sinf(x); }
__device__ __forceinline__ double fsin_(double x) { return sin(x); }
__device__ __forceinline__ float fcos_(float x) { return cosf(x); }
__device__ __forceinline__ double fcos_(double x) { return cos(x); }
__device__ __forceinline__ float ftan_(float x) { return tanf(x); }
__device__ __forceinline__ double ftan_(double x) { return tan(x); }
__device__ __forceinline__ float facos_(float x) { return acosf(x); }
__device__ __forceinline__ double facos_(double x) { return acos(x); }
__device__ __forceinline__ float fatan2_(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double fatan2_(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float ffmod_(float a, float b) { return fmodf(a, b); }
__device__ __forceinline__ double ffmod_(double a, double b) { return fmod(a, b); }

template <typename R> __device__ __forceinline__ R jsSign(R v) { return v > R(0) ? R(1) : (v < R(0) ? R(-1) : v); }

// ------------------------------------------------------------------------------------------------ quaternions
// frame of a quaternion, normalised first (quat.js:122-130); columns X | Y | Z
template <typename R>
__device__ __forceinline__ void qframe(double w, double x, double y, double z, R* F) {
  double n = sqrt(w * w + x * x + y * y + z * z);
  if (n == 0) n = 1.0;
  w /= n; x /= n; y /= n; z /= n;
  F[0] = R(1 - 2 * (y * y + z * z)); F[1] = R(2 * (x * y + z * w)); F[2] = R(2 * (x * z - y * w));
  F[3] = R(2 * (x * y - z * w)); F[4] = R(1 - 2 * (x * x + z * z)); F[5] = R(2 * (y * z + x * w));
  F[6] = R(2 * (x * z + y * w)); F[7] = R(2 * (y * z - x * w)); F[8] = R(1 - 2 * (x * x + y * y));
}
template <typename R>
__device__ __forceinline__ void qframeR(R w, R x, R y, R z, R* F) {
  R n = fsqrt_(w * w + x * x + y * y + z * z);
  if (n == R(0)) n = R(1);
  w /= n; x /= n; y /= n; z /= n;
  F[0] = 1 - 2 * (y * y + z * z); F[1] = 2 * (x * y + z * w); F[2] = 2 * (x * z - y * w);
  F[3] = 2 * (x * y - z * w); F[4] = 1 - 2 * (x * x + z * z); F[5] = 2 * (y * z + x * w);
  F[6] = 2 * (x * z + y * w); F[7] = 2 * (y * z - x * w); F[8] = 1 - 2 * (x * x + y * y);
}

// Ψ = log(q_a* ⊗ q_b) in f64 (rod_cosserat.js:33-45); also returns the relative quaternion d = ±(q_a* ⊗ q_b)
__device__ __forceinline__ double relrot(const double* qa, const double* qb, double* psi, double* d) {
  double aw = qa[0], ax = -qa[1], ay = -qa[2], az = -qa[3];
  double bw = qb[0], bx = qb[1], by = qb[2], bz = qb[3];
  double w = aw * bw - ax * bx - ay * by - az * bz;
  double x = aw * bx + ax * bw + ay * bz - az * by;
  double y = aw * by - ax * bz + ay * bw + az * bx;
  double z = aw * bz + ax * by - ay * bx + az * bw;
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  double s = sqrt(x * x + y * y + z * z), th = 2 * atan2(s, w);
  double f = s > 1e-12 ? th / s : 2 / fmax(w, 1e-12);
  psi[0] = f * x; psi[1] = f * y; psi[2] = f * z;
  d[0] = w; d[1] = x; d[2] = y; d[3] = z;
  return th;
}

// J_l⁻ᵀ(Ψ) m = m + ½ Ψ × m + β Ψ × (Ψ × m), β = 1/θ² − (1 + cos θ)/(2θ sin θ) (rod_cosserat.js:147-151).
// f64: the reference's branch point 1e-3; f32: the series below θ = 1 (the closed form cancels in f32).
template <typename R>
__device__ __forceinline__ void jlt(const R* psi, R th, const R* m, R* J) {
  R beta;
  if constexpr (std::is_same<R, double>::value) {
    beta = th < 1e-3 ? 1.0 / 12 + th * th / 720 : 1 / (th * th) - (1 + cos(th)) / (2 * th * sin(th));
  } else {
    if (th < 1.0f) {
      float t2 = th * th;
      beta = 0.08333333333333333f + t2 * (1.388888888888889e-3f + t2 * (3.306878306878307e-5f + t2 * (8.267195767195767e-7f + t2 * 2.087675698786810e-8f)));
    } else {
      beta = 1 / (th * th) - (1 + cosf(th)) / (2 * th * sinf(th));
    }
  }
  R c0 = psi[1] * m[2] - psi[2] * m[1], c1 = psi[2] * m[0] - psi[0] * m[2], c2 = psi[0] * m[1] - psi[1] * m[0];
  R e0 = psi[1] * c2 - psi[2] * c1, e1 = psi[2] * c0 - psi[0] * c2, e2 = psi[0] * c1 - psi[1] * c0;
  J[0] = m[0] + R(0.5) * c0 + beta * e0; J[1] = m[1] + R(0.5) * c1 + beta * e1; J[2] = m[2] + R(0.5) * c2 + beta * e2;
}

// R(d)ᵀ v for the relative rotation d = q_a* ⊗ q_b: the body axes of b from those of a
template <typename R>
__device__ __forceinline__ void rotT(const double* d, const R* v, R* out) {
  R F[9];
  qframe<R>(d[0], d[1], d[2], d[3], F);
  out[0] = F[0] * v[0] + F[1] * v[1] + F[2] * v[2];
  out[1] = F[3] * v[0] + F[4] * v[1] + F[5] * v[2];
  out[2] = F[6] * v[0] + F[7] * v[1] + F[8] * v[2];
}

__device__ __forceinline__ bool inst_live(const Ctl* c) { return c->done == 0 && c->iter < c->stop_at; }

// ================================================================================================= k_begin
// step(nIter, {instances, force}) (solver_gpu.js:338-345): stop_at = iter + nIter, clear the convergence flag
template <typename Real>
__global__ void k_begin(const Inst<Real>* insts, const int32_t* sel, int32_t nsel, int32_t n_iter, int32_t force) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsel) return;
  Ctl* c = insts[sel[t]].ctl;
  c->stop_at = c->iter + n_iter;
  if (force) c->done = 0;
}

// ================================================================================================= k_cross
template <typename Real>
struct SideEv {
  int seg, a, b, loc;  // segment (instance index), its nodes, strand-local segment index
  Real t, L;
  Real P[3], tg[3], R[9];
};

// PolylineContact._segment + _eval (solver_cpu.js:144-149, 197-228) on one side, positions relative to (ox, oy, oz).
// A strand's nodes, segments and joints are contiguous, so the neighbours of segment i are implicit: nodes
// i-1 .. i+2 and segments i-1 .. i+1 (modulo the strand for closed laths).  All loads issue together once u is known.
// Node tangents: normalised sum of the adjacent unit chords, the starting segment first (_prepare); node frames:
// hemisphere-consistent mean of the adjacent segment quaternions referenced to the starting segment (nodeQuats).
template <typename Real, typename G>
__device__ __forceinline__ void eval_side(const Inst<Real>& I, int cur, int c, int side, double uu, double ox, double oy, double oz,
                                          SideEv<G>& E) {
  const int ns = I.k_nseg[2 * c + side];
  const int noff = I.k_noff[2 * c + side];
  const int nnr = I.k_nn[2 * c + side];
  const bool closed = (nnr >> 30) & 1;
  const int nn = nnr & 0x3fffffff;
  const int soff = I.k_seg_off[2 * c + side];
  int i = int(fmin(fmax(floor(uu), 0.0), double(max(ns - 1, 0))));
  double td = fmin(fmax(uu - i, 0.0), 1.0);
  G t = G(td);
  E.loc = i;
  E.seg = soff + i;
  const int ia = i, ib = closed ? (i + 1) % nn : i + 1;
  E.a = noff + ia; E.b = noff + ib;
  // neighbours: node before a, node after b; segment before i, segment after i
  const bool hasPrevSeg = closed || i >= 1;          // segment ending at a
  const bool hasNextSeg = closed || i + 1 < ns;      // segment starting at b
  const int iam = closed ? (i - 1 + nn) % nn : i - 1;
  const int ibp = closed ? (ib + 1) % nn : ib + 1;
  const int sm = closed ? (i - 1 + ns) % ns : i - 1;
  const int sp = closed ? (i + 1) % ns : i + 1;
  const double* X = I.x[cur]; const double* Y = I.y[cur]; const double* Z = I.z[cur];
  const double xa = X[E.a], ya = Y[E.a], za = Z[E.a];
  const double xb = X[E.b], yb = Y[E.b], zb = Z[E.b];
  double xm = 0, ym = 0, zm = 0, xp = 0, yp = 0, zp = 0;
  if (hasPrevSeg) { xm = X[noff + iam]; ym = Y[noff + iam]; zm = Z[noff + iam]; }
  if (hasNextSeg) { xp = X[noff + ibp]; yp = Y[noff + ibp]; zp = Z[noff + ibp]; }
  const double* QW = I.qw[cur]; const double* QX = I.qx[cur]; const double* QY = I.qy[cur]; const double* QZ = I.qz[cur];
  G q0[4] = {G(QW[E.seg]), G(QX[E.seg]), G(QY[E.seg]), G(QZ[E.seg])};
  G qm[4] = {0, 0, 0, 0}, qp[4] = {0, 0, 0, 0};
  if (hasPrevSeg) { int j = soff + sm; qm[0] = G(QW[j]); qm[1] = G(QX[j]); qm[2] = G(QY[j]); qm[3] = G(QZ[j]); }
  if (hasNextSeg) { int j = soff + sp; qp[0] = G(QW[j]); qp[1] = G(QX[j]); qp[2] = G(QY[j]); qp[3] = G(QZ[j]); }
  auto unitAdd = [](G* T, double dx, double dy, double dz) {
    G ex = G(dx), ey = G(dy), ez = G(dz);
    G l = dmax(fsqrt_(ex * ex + ey * ey + ez * ez), G(1e-12));
    T[0] += ex / l; T[1] += ey / l; T[2] += ez / l;
  };
  auto normT = [](G* T) {
    G l = dmax(fsqrt_(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]), G(1e-12));
    T[0] /= l; T[1] /= l; T[2] /= l;
  };
  auto meanQ = [](G* out, const G* qs, bool hasStart, const G* qe, bool hasEnd) {
    const G* r = hasStart ? qs : qe;
    out[0] = out[1] = out[2] = out[3] = G(0);
    if (hasStart) for (int k = 0; k < 4; k++) out[k] += qs[k];   // sign of qs · qs is +1
    if (hasEnd) {
      G d = qe[0] * r[0] + qe[1] * r[1] + qe[2] * r[2] + qe[3] * r[3];
      G sg = d < G(0) ? G(-1) : G(1);
      for (int k = 0; k < 4; k++) out[k] += sg * qe[k];
    }
    G nq = fsqrt_(out[0] * out[0] + out[1] * out[1] + out[2] * out[2] + out[3] * out[3]);
    if (nq == G(0)) nq = G(1);
    for (int k = 0; k < 4; k++) out[k] /= nq;
  };
  G Ta[3] = {0, 0, 0}, Tb[3] = {0, 0, 0}, qa[4], qb[4];
  unitAdd(Ta, xb - xa, yb - ya, zb - za);                           // segment i starts at a
  if (hasPrevSeg) unitAdd(Ta, xa - xm, ya - ym, za - zm);           // segment i-1 ends at a
  normT(Ta);
  if (hasNextSeg) unitAdd(Tb, xp - xb, yp - yb, zp - zb);           // segment i+1 starts at b
  unitAdd(Tb, xb - xa, yb - ya, zb - za);                           // segment i ends at b
  normT(Tb);
  meanQ(qa, q0, true, qm, hasPrevSeg);
  meanQ(qb, qp, hasNextSeg, q0, true);
  G pa[3] = {G(xa - ox), G(ya - oy), G(za - oz)};
  G pb[3] = {G(xb - ox), G(yb - oy), G(zb - oz)};
  G dx = G(xb - xa), dy = G(yb - ya), dz = G(zb - za);
  G L = fsqrt_(dx * dx + dy * dy + dz * dz);
  G t2 = t * t, t3 = t * t * t;
  G h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
  G d00 = 6 * t2 - 6 * t, d10 = 3 * t2 - 4 * t + 1, d01 = -6 * t2 + 6 * t, d11 = 3 * t2 - 2 * t;
  G tg[3];
  for (int k = 0; k < 3; k++) {
    E.P[k] = h00 * pa[k] + (h10 * L) * Ta[k] + h01 * pb[k] + (h11 * L) * Tb[k];
    tg[k] = d00 * pa[k] + (d10 * L) * Ta[k] + d01 * pb[k] + (d11 * L) * Tb[k];
  }
  normT(tg);
  G sgn = (qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2] + qa[3] * qb[3]) < G(0) ? G(-1) : G(1);
  G q[4];
  for (int k = 0; k < 4; k++) q[k] = (1 - t) * qa[k] + t * sgn * qb[k];
  G Rn[9];
  qframeR(q[0], q[1], q[2], q[3], Rn);   // normalises q first (qnormalize then toFrame)
  G y[3] = {Rn[3], Rn[4], Rn[5]};
  G yt = y[0] * tg[0] + y[1] * tg[1] + y[2] * tg[2];
  y[0] -= yt * tg[0]; y[1] -= yt * tg[1]; y[2] -= yt * tg[2];
  normT(y);
  E.R[0] = y[1] * tg[2] - y[2] * tg[1]; E.R[1] = y[2] * tg[0] - y[0] * tg[2]; E.R[2] = y[0] * tg[1] - y[1] * tg[0];
  E.R[3] = y[0]; E.R[4] = y[1]; E.R[5] = y[2]; E.R[6] = tg[0]; E.R[7] = tg[1]; E.R[8] = tg[2];
  E.tg[0] = tg[0]; E.tg[1] = tg[1]; E.tg[2] = tg[2];
  E.L = L; E.t = t;
}

// half-extent h(n) and smoothed support point r of a side's section facing sign·n (solver_cpu.js:274-282)
template <typename Real>
__device__ __forceinline__ Real support(const Real* R, const Real* n, Real sign, Real w, Real h_, Real* r) {
  Real nx = (n[0] * R[0] + n[1] * R[1] + n[2] * R[2]) * sign, ny = (n[0] * R[3] + n[1] * R[4] + n[2] * R[5]) * sign;
  Real h = Real(0.5) * w * fabs(nx) + Real(0.5) * h_ * fabs(ny);
  Real ax = Real(0.5) * w * ftanh_(nx / Real(SUPPORT_SMOOTH)), ay = Real(0.5) * h_ * ftanh_(ny / Real(SUPPORT_SMOOTH));
  for (int k = 0; k < 3; k++) r[k] = ax * R[k] + ay * R[3 + k];
  return h;
}

// W3 wall slope tan ψ (solver_cpu.js:285-293)
template <typename Real>
__device__ __forceinline__ Real walls(int sideCut, const Real* R, const Real* n, const Real* mm) {
  Real sg = sideCut == 0 ? Real(1) : Real(-1);
  Real nx = (n[0] * R[0] + n[1] * R[1] + n[2] * R[2]) * sg, ny = (n[0] * R[3] + n[1] * R[4] + n[2] * R[5]) * sg;
  Real fx[3], fy[3];
  for (int k = 0; k < 3; k++) { fx[k] = jsSign(nx) * R[k]; fy[k] = jsSign(ny) * R[3 + k]; }
  const Real* f = (fx[0] * mm[0] + fx[1] * mm[1] + fx[2] * mm[2]) > Real(0) ? fx : fy;
  Real cp = clampv(Real(fabs(f[0] * n[0] + f[1] * n[1] + f[2] * n[2])), Real(1e-6), Real(1));
  return ftan_(facos_(cp));
}

template <typename Real, typename G>
__device__ __forceinline__ G free_weight(const Inst<Real>& I, const SideEv<G>& E) {
  bool fa = (I.fix[E.a] & 16) != 0, fb = (I.fix[E.b] & 16) != 0;
  return (fa ? G(0) : (1 - E.t) * (1 - E.t)) + (fb ? G(0) : E.t * E.t);
}

template <typename Real>
__global__ __launch_bounds__(CROSS_BLOCK) void k_cross(const Inst<Real>* __restrict__ insts, const CrossItem* __restrict__ items,
                                                      int32_t nitems) {
  int bid = blockIdx.x;
  if (bid >= nitems) return;
  CrossItem it = items[bid];
  if (it.inst < 0) return;
  const Inst<Real>& I = insts[it.inst];
  const Ctl* ctl = I.ctl;
  if (!inst_live(ctl)) return;
  const int cur = ctl->cur, iter = ctl->iter;
  const bool al_update = ctl->al_update != 0;
  const int pen_n = ctl->pen_n;
  const int tid = threadIdx.x;
  double eC = 0.0, mxFn = -INFINITY;
  if (tid < it.count) {
    using G = double;   // contact kinematics in FP64: the penetration D − g cancels mm-scale lengths

    const int c = it.c0 + tid;
    const int rampN = max(I.ramp_iter, 0);
    const double rampD = rampN == 0 ? 1.0 : fmin(1.0, 0.01 + 0.99 * iter / rampN);
    const G ramp = G(rampD);
    const G kn = ramp * G(I.strength), kt = ramp * G(I.stick);
    const bool enabled = I.k_en[c] != 0;
    const bool locked = I.k_lock[c] != 0;
    const bool seat = I.seat != 0, aug = I.aug != 0;
    const G mu0 = G(I.mu);
    // ---- both sides, relative to the first node of A's segment
    const double uA = I.u[2 * c], uB = I.u[2 * c + 1];
    const int o = I.k_noff[2 * c] + int(fmin(fmax(floor(uA), 0.0), double(max(I.k_nseg[2 * c] - 1, 0))));
    const double ox = I.x[cur][o], oy = I.y[cur][o], oz = I.z[cur][o];
    SideEv<G> A, B;
    eval_side(I, cur, c, 0, uA, ox, oy, oz, A);
    eval_side(I, cur, c, 1, uB, ox, oy, oz, B);
    const G wA = I.wk[2 * c], wB = I.wk[2 * c + 1], hA = I.hk[2 * c], hB = I.hk[2 * c + 1];
    // ---- AL multiplier update deferred from the end of the previous iteration (updateMultipliers, :499-506)
    if (aug && al_update) {
      G kn1 = G(I.strength);   // the ramp is complete whenever the multipliers update
      bool pinned = free_weight(I, A) < G(AL_FREE_MIN) || free_weight(I, B) < G(AL_FREE_MIN);
      G cap = dmin(I.fc90[2 * c], I.fc90[2 * c + 1]) * wA * wB / dmax(fsin_(I.theta[c]), G(1e-6));
      G v = I.lam[c] + kn1 * (I.penacc[c] / G(max(pen_n, 1)));
      if (!locked) v = dmax(v, G(0));
      v = clampv(v, -cap, cap);
      I.lam[c] = (!enabled || pinned) ? G(0) : v;
      I.penacc[c] = G(0);
    }
    const G* tA = A.tg; const G* tB = B.tg;
    G np[3] = {I.nprev[3 * c], I.nprev[3 * c + 1], I.nprev[3 * c + 2]};
    G n[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
    G nn = fsqrt_(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn < G(1e-6)) { n[0] = np[0]; n[1] = np[1]; n[2] = np[2]; }
    else { G d = dmax(nn, G(1e-12)); n[0] /= d; n[1] /= d; n[2] /= d; }
    if (n[0] * np[0] + n[1] * np[1] + n[2] * np[2] < G(0)) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
    I.nprev[3 * c] = n[0]; I.nprev[3 * c + 1] = n[1]; I.nprev[3 * c + 2] = n[2];
    G S[3] = {B.P[0] - A.P[0], B.P[1] - A.P[1], B.P[2] - A.P[2]};
    G g = S[0] * n[0] + S[1] * n[1] + S[2] * n[2];
    G rA[3], rB[3];
    G supA = support(A.R, n, G(1), wA, hA, rA);
    G supB = support(B.R, n, G(-1), wB, hB, rB);
    G D = supA + supB;
    if (seat && I.seat_gap) D -= I.dent[2 * c] + I.dent[2 * c + 1] + I.dente[c];
    G pen = D - g;
    if (aug && iter >= rampN) I.penacc[c] += pen;
    G Fn = 0; bool act;
    if (aug) {
      G trial = I.lam[c] + kn * pen;
      act = (trial > G(0) || locked) && enabled;
      Fn = act ? (locked ? trial : dmax(trial, G(0))) : G(0);
    } else {
      act = (pen > G(0) || locked) && enabled;
      Fn = act ? kn * pen : G(0);
    }
    G St[3] = {S[0] - g * n[0], S[1] - g * n[1], S[2] - g * n[2]};
    G taa = tA[0] * tA[0] + tA[1] * tA[1] + tA[2] * tA[2], tbb = tB[0] * tB[0] + tB[1] * tB[1] + tB[2] * tB[2];
    G tab = tA[0] * tB[0] + tA[1] * tB[1] + tA[2] * tB[2];
    G sa = St[0] * tA[0] + St[1] * tA[1] + St[2] * tA[2], sb = St[0] * tB[0] + St[1] * tB[1] + St[2] * tB[2];
    G det = taa * tbb - tab * tab;
    bool ok = det > G(1e-8);
    G ca = ok ? (sa * tbb - sb * tab) / det : G(0), cb = ok ? (sb * taa - sa * tab) / det : G(0);
    G Ftv = fsqrt_((kt * St[0]) * (kt * St[0]) + (kt * St[1]) * (kt * St[1]) + (kt * St[2]) * (kt * St[2]));
    // ---- W3 directional seat friction (_seatFriction, :295-315)
    G mu = mu0;
    if (seat) {
      G norm = fsqrt_(St[0] * St[0] + St[1] * St[1] + St[2] * St[2]);
      G dn = dmax(norm, G(1e-12));
      G dh[3] = {St[0] / dn, St[1] / dn, St[2] / dn};
      bool has = norm > G(1e-12);
      for (int j = 0; j < 2; j++) {
        if (!(I.dent[2 * c + j] > G(0) && has)) continue;
        const G* tCut = j == 0 ? tB : tA;
        const G* Rcut = j == 0 ? B.R : A.R;
        G mm[3] = {n[1] * tCut[2] - n[2] * tCut[1], n[2] * tCut[0] - n[0] * tCut[2], n[0] * tCut[1] - n[1] * tCut[0]};
        G ml = dmax(fsqrt_(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), G(1e-12));
        mm[0] /= ml; mm[1] /= ml; mm[2] /= ml;
        G sgn = j == 0 ? G(1) : G(-1);
        G dj = (dh[0] * mm[0] + dh[1] * mm[1] + dh[2] * mm[2]) * sgn;
        G wd[3] = {jsSign(dj) * mm[0] * sgn, jsSign(dj) * mm[1] * sgn, jsSign(dj) * mm[2] * sgn};
        G tp = walls(1 - j, Rcut, n, wd);
        I.tanpsi[2 * c + j] = tp;
        G add = fabs(dj) * (mu0 * tp < G(1) ? (mu0 + tp) / dmax(G(1) - mu0 * tp, G(1e-6)) - mu0 : G(9) * mu0);
        mu += dmin(add, G(9) * mu0);
      }
    }
    G limit = locked ? G(INFINITY) : mu * fabs(Fn);
    bool slip = act && ok && Ftv > limit;
    bool fre = !act && ok && enabled;
    G sc = 0;
    if (slip) sc = clampv(G(I.slip_rate), G(1e-3), G(1)) * (G(1) - limit / dmax(Ftv, G(1e-12)));
    if (fre) sc = G(1);
    G eps = dmax(G(I.epsilon), G(1e-9));
    G da = clampv(ca * sc, -eps, eps), db = clampv(cb * sc, -eps, eps);
    bool moved = slip || fre;
    if (moved) {
      double u0 = I.u[2 * c] + double(da) / double(dmax(A.L, G(1e-12)));
      double u1 = I.u[2 * c + 1] - double(db) / double(dmax(B.L, G(1e-12)));
      I.u[2 * c] = fmin(fmax(u0, 0.0), double(I.k_nseg[2 * c]));
      I.u[2 * c + 1] = fmin(fmax(u1, 0.0), double(I.k_nseg[2 * c + 1]));
    }
    if (slip) {
      G q0 = da * tA[0] + db * tB[0], q1 = da * tA[1] + db * tB[1], q2 = da * tA[2] + db * tB[2];
      I.slip[c] += double(fsqrt_(q0 * q0 + q1 * q1 + q2 * q2));
    }
    G Se[3] = {0, 0, 0};
    if (act) { Se[0] = St[0] - da * tA[0] - db * tB[0]; Se[1] = St[1] - da * tA[1] - db * tB[1]; Se[2] = St[2] - da * tA[2] - db * tB[2]; }
    G fT[3] = {kt * Se[0], kt * Se[1], kt * Se[2]};
    G FA[3], FB[3];
    for (int k = 0; k < 3; k++) { FA[k] = -Fn * n[k] + fT[k]; FB[k] = Fn * n[k] - fT[k]; }
    I.fn[c] = Fn;
    I.ft[c] = fsqrt_(fT[0] * fT[0] + fT[1] * fT[1] + fT[2] * fT[2]);
    I.gap[c] = g;
    I.kflags[c] = uint8_t((slip ? 1 : 0) | (moved ? 0 : 2));
    for (int k = 0; k < 3; k++) I.mid[3 * c + k] = G(0.5) * (A.P[k] + B.P[k]) + G(k == 0 ? ox : (k == 1 ? oy : oz));
    // ---- W1 classification (_classify, :231-271)
    G cface[2], z0[2], phiv[2]; int edge[2];
    for (int sd = 0; sd < 2; sd++) {
      const G* R = sd == 0 ? A.R : B.R;
      G sg = sd == 0 ? G(1) : G(-1);
      G wv = sd == 0 ? wA : wB, hv = sd == 0 ? hA : hB;
      G nx = (n[0] * R[0] + n[1] * R[1] + n[2] * R[2]) * sg, ny = (n[0] * R[3] + n[1] * R[4] + n[2] * R[5]) * sg;
      G sx = fabs(ftanh_(nx / G(SUPPORT_SMOOTH))), sy = fabs(ftanh_(ny / G(SUPPORT_SMOOTH)));
      edge[sd] = sx * sy > G(I.edge_thr) ? 1 : 0;
      bool yf = fabs(ny) >= fabs(nx);
      cface[sd] = yf ? wv : hv; z0[sd] = yf ? hv : wv;
      phiv[sd] = fatan_(dmin(G(fabs(nx)), G(fabs(ny))) / dmax(dmax(G(fabs(nx)), G(fabs(ny))), G(1e-12)));
    }
    G cr[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
    G sinT = clampv(fsqrt_(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]), G(1e-6), G(1));
    G theta = fasin_(sinT);
    int nEdge = edge[0] + edge[1];
    int faceSide = nEdge == 1 ? (edge[1] ? 0 : 1) : -1;
    G Fa = fabs(Fn);
    G Aa = cface[0] * cface[1] / sinT, pres = Fa / dmax(Aa, G(1e-12));
    G ut = pres / dmax(dmin(I.Hs[2 * c], I.Hs[2 * c + 1]), G(1e-12));
    if (nEdge == 1) {
      int fs = faceSide, es = 1 - fs;
      G ell = cface[fs] / sinT, kap = I.E90[2 * c + fs] / dmax(z0[fs], G(1e-12));
      G phiE = dmax(phiv[es], G(1e-4));
      G bw = fsqrt_(G(2) * Fa / dmax(ell * kap * phiE, G(1e-12)));
      bool full = bw >= cface[es];
      G Afe = full ? cface[0] * cface[1] / sinT : bw * ell;
      G pfe = full ? Fa / dmax(Afe, G(1e-12)) : kap * phiE * bw;
      Aa = Afe; pres = pfe; ut = pfe / dmax(I.fc90[2 * c + es], G(1e-12));
    } else if (nEdge == 2) {
      G fc = dmin(I.fc90[2 * c], I.fc90[2 * c + 1]);
      Aa = Fa / dmax(fc, G(1e-12)); pres = Fa > G(0) ? fc : G(0); ut = Fa > G(0) ? G(1) : G(0);
    }
    I.theta[c] = theta; I.anom[c] = Aa; I.pres[c] = pres; I.util[c] = ut;
    I.phi[2 * c] = phiv[0]; I.phi[2 * c + 1] = phiv[1];
    I.ctype[c] = int8_t(nEdge); I.face[c] = int8_t(faceSide);
    // ---- W3 seat update (_seatUpdate, :317-344)
    if (seat) {
      G inc[3] = {da * tA[0] - db * tB[0], da * tA[1] - db * tB[1], da * tA[2] - db * tB[2]};
      for (int j = 0; j < 2; j++) {
        const G* tCut = j == 0 ? tB : tA;
        G mm[3] = {n[1] * tCut[2] - n[2] * tCut[1], n[2] * tCut[0] - n[0] * tCut[2], n[0] * tCut[1] - n[1] * tCut[0]};
        G ml = dmax(fsqrt_(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), G(1e-12));
        mm[0] /= ml; mm[1] /= ml; mm[2] /= ml;
        G stp = (inc[0] * mm[0] + inc[1] * mm[1] + inc[2] * mm[2]) * (j == 0 ? G(1) : G(-1));
        int i2 = 2 * c + j;
        bool has = I.dent[i2] > G(0);
        if (has && slip) I.climb[i2] += stp;
        if (has && fabs(I.climb[i2]) >= I.dent[i2] / dmax(I.tanpsi[i2], G(1e-6))) { I.dent[i2] = G(0); I.climb[i2] = G(0); }
      }
      if (iter >= I.dent_after) {
        G sp0 = fsin_(phiv[0]) * fcos_(phiv[0]), sp1 = fsin_(phiv[1]) * fcos_(phiv[1]);
        G sT = fsin_(theta);
        if (nEdge == 2 && act && !locked) {
          G fc = dmin(I.fc90[2 * c], I.fc90[2 * c + 1]);
          G dEst = fsqrt_(Fa / dmax(fc, G(1e-12)) * sT * sp0 * sp1);
          for (int j = 0; j < 2; j++) { G cu = I.dent[2 * c + j]; I.dent[2 * c + j] = dmax(cu, dmin(cu + G(I.dent_rate) * dEst, dEst)); }
        }
        if (nEdge == 1 && act && !locked) {
          int fs = faceSide, es = 1 - fs;
          G ell = cface[fs] / dmax(sT, G(1e-6));
          G dE = Fa / dmax(I.fc90[2 * c + es] * ell, G(1e-12)) * (es == 0 ? sp0 : sp1);
          I.dente[c] = dmax(I.dente[c], dE);
        }
      }
    }
    // ---- lever torques, W2 rotational friction (_rotationalFriction, :347-379)
    G tf = G(I.torsion_factor);
    G TA[3] = {tf * (rA[1] * FA[2] - rA[2] * FA[1]), tf * (rA[2] * FA[0] - rA[0] * FA[2]), tf * (rA[0] * FA[1] - rA[1] * FA[0])};
    G TB[3] = {tf * (rB[1] * FB[2] - rB[2] * FB[1]), tf * (rB[2] * FB[0] - rB[0] * FB[2]), tf * (rB[0] * FB[1] - rB[1] * FB[0])};
    G eRot = 0;
    if (I.rot_fric) {
      G thc = fatan2_(cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2], tab);
      G th0 = I.theta0[c];
      if (isnan(th0) && act) th0 = thc;
      bool okr = act && !isnan(th0);
      G sT2 = dmax(fsin_(theta), G(1e-6));
      G req = G(0.38) * fsqrt_(wA * wB / sT2);
      if (nEdge == 1) req = G(0.25) * (faceSide == 0 ? wA : wB) / sT2;
      else if (nEdge == 2) req = G(0.38) * fsqrt_(dmax(Aa, G(1e-12)));
      G kr = kt * req * req;
      G dth = okr ? thc - th0 : G(0);
      const G PI = G(3.141592653589793);
      dth = ffmod_(ffmod_(dth + PI, 2 * PI) + 2 * PI, 2 * PI) - PI;
      G Mc = kr * dth;
      G Mlim = locked ? G(INFINITY) : G(I.rot_factor) * mu0 * Fa * req;
      if (okr && fabs(Mc) > Mlim) {
        G adv = clampv(clampv(G(I.slip_rate), G(1e-3), G(1)) * (G(1) - Mlim / dmax(G(fabs(Mc)), G(1e-12))) * dth, G(-0.05), G(0.05));
        th0 += adv; I.rotslip[c] += fabs(adv); dth -= adv;
        Mc = kr * dth;
      }
      I.theta0[c] = th0;
      G Mr = okr ? Mc : G(0);
      I.mrot[c] = Mr;
      for (int k = 0; k < 3; k++) { TA[k] += Mr * n[k]; TB[k] -= Mr * n[k]; }
      eRot = G(0.5) * kr * dth * dth;
    }
    // ---- righting stiffness (addTwistStiffness, :170-176) on the segments after the slip
    G tfk = dmax(G(I.torsion_factor), G(1e-3));
    G kth = tfk * ((Fn + kn * G(0.25) * I.kh[c]) * (G(0.5) * I.kw[c]) / G(SUPPORT_SMOOTH) + kt * (G(0.5) * I.kh[c]) * (G(0.5) * I.kh[c]));
    int postA = int(fmin(fmax(floor(I.u[2 * c]), 0.0), double(max(I.k_nseg[2 * c] - 1, 0))));
    int postB = int(fmin(fmax(floor(I.u[2 * c + 1]), 0.0), double(max(I.k_nseg[2 * c + 1] - 1, 0))));
    SideRec<Real> ra, rb;
    ra.seg_pre = A.loc; ra.seg_post = postA; ra.t = A.t; ra.kth = kth;
    rb.seg_pre = B.loc; rb.seg_post = postB; rb.t = B.t; rb.kth = kth;
    for (int k = 0; k < 3; k++) { ra.F[k] = FA[k]; ra.T[k] = TA[k]; rb.F[k] = FB[k]; rb.T[k] = TB[k]; }
    I.side[I.side_pos[2 * c]] = ra;
    I.side[I.side_pos[2 * c + 1]] = rb;
    G ePen = act ? (aug ? Fn * Fn / dmax(kn, G(1e-30)) : kn * pen * pen) : G(0);
    eC = 0.5 * double(ePen) + 0.5 * double(kt) * double(Se[0] * Se[0] + Se[1] * Se[1] + Se[2] * Se[2]) + double(eRot);
    mxFn = double(Fn);
  }
  // ---- block partial (fixed tree order)
  __shared__ double rs[CROSS_BLOCK], rm[CROSS_BLOCK];
  rs[tid] = eC; rm[tid] = mxFn;
  __syncthreads();
  for (int k = CROSS_BLOCK / 2; k > 0; k >>= 1) {
    if (tid < k) { rs[tid] += rs[tid + k]; rm[tid] = fmax(rm[tid], rm[tid + k]); }
    __syncthreads();
  }
  if (tid == 0) { I.cpart[it.blk * CPART] = rs[0]; I.cpart[it.blk * CPART + 1] = rm[0]; }
}

// ================================================================================================= k_strand
template <typename Real, int TILE>
struct StrandLDS {
  double X[TILE + 2][3];      // nodes k0-1 .. k1
  double Q[TILE + 2][4];      // segments k0-1 .. k1
  Real SF[TILE + 1][3];       // world force of segments k0-1 .. k1-1
  Real JA[TILE + 1][3];       // joint k0-1+t torque on its first segment (body axes)
  Real JB[TILE + 1][3];       // ... on its second segment (body axes of that segment)
  Real KD[TILE + 1][3];       // bending-coupling inertia of joint k0-1+t (both segments)
  Real ST[TILE][3];           // stretch torque (Rᵀe) × (Rᵀn) of owned segment k0+t (body axes)
  SideRec<Real> rec[TILE];    // gathered crossing sides (one chunk)
  int wcnt[TILE / 64];
  double red[13][TILE / 64];
};

// fixed-order block reduction: v[0, nsum) are sums, v[nsum, NV) maxima; every thread receives the result
template <int TILE, int NV>
__device__ __forceinline__ void block_reduce(double (*red)[TILE / 64], double* v, int nsum) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < NV; k++) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) {
      double y = __shfl_down(x, off, 64);
      x = k < nsum ? x + y : fmax(x, y);
    }
    if (lane == 0) red[k][wv] = x;
  }
  __syncthreads();
  for (int k = 0; k < NV; k++) {
    double x = red[k][0];
    for (int w = 1; w < TILE / 64; w++) x = k < nsum ? x + red[k][w] : fmax(x, red[k][w]);
    v[k] = x;
  }
  __syncthreads();
}

template <typename Real, int TILE, bool OUT>
__device__ void strand_tile(const Inst<Real>& I, const TileItem tl, StrandLDS<Real, TILE>& L, int tile_local,
                            double* outN, double* outV, double* outM, double* outCM, double* outR) {
  Ctl* ctl = I.ctl;
  const int tid = threadIdx.x;
  int cur = ctl->cur;
  const bool rbk = ctl->rollback != 0;
  if (OUT) cur = rbk ? cur : (cur ^ 1);            // the state the last iteration evaluated
  const int iter = ctl->iter;
  const int s = tl.strand;
  const int noff = I.node_off[s], nn = I.node_off[s + 1] - noff;
  const int soff = I.seg_off[s], nseg = I.seg_off[s + 1] - soff;
  const int joff = I.joint_off[s], njs = I.joint_off[s + 1] - joff;
  const bool closed = I.closed[s] != 0;
  const int k0 = tl.k0, k1 = tl.k1, cnt = k1 - k0;
  auto node_ok = [&](int k, int& kk) { if (closed) { kk = ((k % nn) + nn) % nn; return true; } kk = k; return k >= 0 && k < nn; };
  auto seg_ok = [&](int j, int& jj) { if (closed) { if (nseg <= 0) { jj = 0; return false; } jj = ((j % nseg) + nseg) % nseg; return true; } jj = j; return j >= 0 && j < nseg; };
  auto joint_ok = [&](int j, int& jj) { if (closed) { if (njs <= 0) { jj = 0; return false; } jj = ((j % njs) + njs) % njs; return true; } jj = j; return j >= 0 && j < njs; };

  // ---- stage positions and frames of the tile and its halo
  for (int h = tid; h < cnt + 2; h += TILE) {
    int kk;
    if (node_ok(k0 - 1 + h, kk)) { int i = noff + kk; L.X[h][0] = I.x[cur][i]; L.X[h][1] = I.y[cur][i]; L.X[h][2] = I.z[cur][i]; }
    else { L.X[h][0] = L.X[h][1] = L.X[h][2] = 0.0; }
    int jj;
    if (seg_ok(k0 - 1 + h, jj)) { int j = soff + jj; L.Q[h][0] = I.qw[cur][j]; L.Q[h][1] = I.qx[cur][j]; L.Q[h][2] = I.qy[cur][j]; L.Q[h][3] = I.qz[cur][j]; }
    else { L.Q[h][0] = 1.0; L.Q[h][1] = L.Q[h][2] = L.Q[h][3] = 0.0; }
  }
  __syncthreads();

  double eS = 0, eSh = 0, eB = 0, eT = 0, mxN = 0, mxM = 0;
  // ---- segment stretch / shear (CosseratRod.apply :118-133) and joints (:136-159), k0-1+t for t in [0, cnt]
  for (int t = tid; t < cnt + 1; t += TILE) {
    int jj;
    const int j = k0 - 1 + t;
    if (seg_ok(j, jj)) {
      const int sg = soff + jj;
      double Rd[9];
      {
        double w = L.Q[t][0], x = L.Q[t][1], y = L.Q[t][2], z = L.Q[t][3];
        double nq = sqrt(w * w + x * x + y * y + z * z); if (nq == 0) nq = 1.0;
        w /= nq; x /= nq; y /= nq; z /= nq;
        Rd[0] = 1 - 2 * (y * y + z * z); Rd[1] = 2 * (x * y + z * w); Rd[2] = 2 * (x * z - y * w);
        Rd[3] = 2 * (x * y - z * w); Rd[4] = 1 - 2 * (x * x + z * z); Rd[5] = 2 * (y * z + x * w);
        Rd[6] = 2 * (x * z + y * w); Rd[7] = 2 * (y * z - x * w); Rd[8] = 1 - 2 * (x * x + y * y);
      }
      double ex = L.X[t + 1][0] - L.X[t][0], ey = L.X[t + 1][1] - L.X[t][1], ez = L.X[t + 1][2] - L.X[t][2];
      double l0 = I.l0[sg];
      double b0 = Rd[0] * ex + Rd[1] * ey + Rd[2] * ez, b1 = Rd[3] * ex + Rd[4] * ey + Rd[5] * ez, b2 = Rd[6] * ex + Rd[7] * ey + Rd[8] * ez;
      double s1 = b0 / l0, s2 = b1 / l0, s3 = b2 / l0 - 1.0;           // σ in f64: the strain cancels in f32
      Real GA = I.GA[sg], EA = I.EA[sg];
      Real n1 = GA * Real(s1), n2 = GA * Real(s2), n3 = EA * Real(s3);
      L.SF[t][0] = n1 * Real(Rd[0]) + n2 * Real(Rd[3]) + n3 * Real(Rd[6]);
      L.SF[t][1] = n1 * Real(Rd[1]) + n2 * Real(Rd[4]) + n3 * Real(Rd[7]);
      L.SF[t][2] = n1 * Real(Rd[2]) + n2 * Real(Rd[5]) + n3 * Real(Rd[8]);
      if (t >= 1) {   // owned segment k0 + (t-1)
        Real e0 = Real(b0), e1 = Real(b1), e2 = Real(b2);
        L.ST[t - 1][0] = e1 * n3 - e2 * n2; L.ST[t - 1][1] = e2 * n1 - e0 * n3; L.ST[t - 1][2] = e0 * n2 - e1 * n1;
        eS += 0.5 * l0 * double(n3) * s3; eSh += 0.5 * l0 * (double(n1) * s1 + double(n2) * s2);
        mxN = fmax(mxN, fabs(double(n3)));
        if (OUT) { outN[sg] = double(n3); outV[2 * sg] = double(n1); outV[2 * sg + 1] = double(n2); }
      }
    } else {
      L.SF[t][0] = L.SF[t][1] = L.SF[t][2] = Real(0);
      if (t >= 1) L.ST[t - 1][0] = L.ST[t - 1][1] = L.ST[t - 1][2] = Real(0);
    }
    int ji;
    if (joint_ok(j, ji)) {
      const int jg = joff + ji;
      const int sa = soff + (closed ? ((j % nseg) + nseg) % nseg : j);
      double psi[3], dq[4];
      double th = relrot(L.Q[t], L.Q[t + 1], psi, dq);
      Real lb = I.jlen[jg];
      Real ps[3] = {Real(psi[0]), Real(psi[1]), Real(psi[2])};
      Real d0 = ps[0] / lb - I.om0[3 * jg], d1 = ps[1] / lb - I.om0[3 * jg + 1], d2 = ps[2] / lb - I.om0[3 * jg + 2];
      Real Bx = I.EIx[sa], By = I.EIy[sa], Bz = I.GJ[sa];
      Real mm[3] = {Bx * d0, By * d1, Bz * d2};
      Real q0 = ps[0] * ps[0], q1 = ps[1] * ps[1], q2 = ps[2] * ps[2], f = Real(0.25) / lb;
      L.KD[t][0] = f * (By * q2 + Bz * q1); L.KD[t][1] = f * (Bx * q2 + Bz * q0); L.KD[t][2] = f * (Bx * q1 + By * q0);
      Real J[3];
      jlt(ps, Real(th), mm, J);
      L.JA[t][0] = J[0]; L.JA[t][1] = J[1]; L.JA[t][2] = J[2];
      Real Jb[3];
      rotT(dq, J, Jb);
      L.JB[t][0] = -Jb[0]; L.JB[t][1] = -Jb[1]; L.JB[t][2] = -Jb[2];
      if (t >= 1) {   // owned joint (its first segment is owned)
        eB += 0.5 * double(lb) * (double(mm[0]) * double(d0) + double(mm[1]) * double(d1));
        eT += 0.5 * double(lb) * double(mm[2]) * double(d2);
        mxM = fmax(mxM, fmax(fabs(double(mm[0])), fmax(fabs(double(mm[1])), fabs(double(mm[2])))));
        if (OUT) { outM[3 * jg] = double(mm[0]); outM[3 * jg + 1] = double(mm[1]); outM[3 * jg + 2] = double(mm[2]); }
      }
    } else {
      for (int k = 0; k < 3; k++) { L.JA[t][k] = Real(0); L.JB[t][k] = Real(0); L.KD[t][k] = Real(0); }
    }
  }
  __syncthreads();

  // ---- gather the crossing sides of this strand that touch the tile, in the strand's fixed side order
  const int t = tid;
  const int k = k0 + t;
  const bool ownNode = t < cnt;
  int jself;
  const bool ownSeg = ownNode && seg_ok(k, jself) && (closed || k < nseg);
  Real Fc[3] = {0, 0, 0}, Tc[3] = {0, 0, 0}, kTw = 0;
  const int so = I.side_off[s], ns = I.side_off[s + 1] - so;
  const int lane = tid & 63, wv = tid >> 6;
  auto modn = [&](int v) { return (closed && nseg > 0) ? ((v % nseg) + nseg) % nseg : v; };
  for (int base = 0; base < ns; base += TILE) {
    const int i = base + tid;
    bool rel = false;
    SideRec<Real> r;
    if (i < ns) {
      r = I.side[so + i];
      if (closed) {
        // membership by cyclic distance from k0-1 (the tile covers at most the whole loop)
        const int nz = nseg > 0 ? nseg : 1;
        int dp = ((r.seg_pre - (k0 - 1)) % nz + nz) % nz, dq = ((r.seg_post - k0) % nz + nz) % nz;
        rel = dp < cnt + 1 || dq < cnt;
      } else {
        rel = (r.seg_pre >= k0 - 1 && r.seg_pre < k1) || (r.seg_post >= k0 && r.seg_post < k1);
      }
    }
    unsigned long long mask = __ballot(rel);
    int pos = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) L.wcnt[wv] = __popcll(mask);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < TILE / 64; w++) { if (w < wv) off += L.wcnt[w]; tot += L.wcnt[w]; }
    if (rel) L.rec[off + pos] = r;
    __syncthreads();
    if (ownNode) {
      const int km1 = modn(k - 1);
      for (int q = 0; q < tot; q++) {
        const SideRec<Real>& g = L.rec[q];
        if (g.seg_pre == k && (closed || k < nseg)) { Real w = Real(1) - g.t; Fc[0] += w * g.F[0]; Fc[1] += w * g.F[1]; Fc[2] += w * g.F[2]; }
        if ((closed || k >= 1) && g.seg_pre == km1) { Real w = g.t; Fc[0] += w * g.F[0]; Fc[1] += w * g.F[1]; Fc[2] += w * g.F[2]; }
        if (ownSeg) {
          if (g.seg_pre == k) { Tc[0] += g.T[0]; Tc[1] += g.T[1]; Tc[2] += g.T[2]; }
          if (g.seg_post == k) kTw += g.kth;
        }
      }
    }
    __syncthreads();
  }

  // ---- nodes: loads, reactions, anchors, v += F/m, x += v (solver_cpu.js:753-782)
  double keN = 0, keR = 0, mxF = 0, mxT = 0, mxL = 0;
  const bool warm = iter < I.warmup && I.contacts;
  const Real damping = Real(I.damping);
  if (ownNode) {
    int kk; node_ok(k, kk);
    const int i = noff + kk;
    const bool hasEnd = closed || k >= 1, hasStart = closed || k < nseg;
    Real F[3];
    for (int c = 0; c < 3; c++) {
      if (hasEnd && hasStart) F[c] = L.SF[t + 1][c] - L.SF[t][c];
      else if (hasStart) F[c] = L.SF[t + 1][c];
      else if (hasEnd) F[c] = -L.SF[t][c];
      else F[c] = Real(0);
    }
    Real ld[3] = {I.lx[i], I.ly[i], I.lz[i]};
    if (I.npl) {
      Real lx = I.llx[i], ly = I.lly[i], lz = I.llz[i];
      if (lx != Real(0) || ly != Real(0) || lz != Real(0)) {
        // mean node frame from the adjacent segments (nodeQuats: the starting segment is the reference)
        const double* qs = hasStart ? L.Q[t + 1] : L.Q[t];
        double acc[4] = {0, 0, 0, 0};
        for (int pass = 0; pass < 2; pass++) {
          if (pass == 0 && !hasStart) continue;
          if (pass == 1 && !hasEnd) continue;
          const double* q = pass == 0 ? L.Q[t + 1] : L.Q[t];
          double d = q[0] * qs[0] + q[1] * qs[1] + q[2] * qs[2] + q[3] * qs[3];
          double sg = d < 0 ? -1.0 : 1.0;
          for (int c = 0; c < 4; c++) acc[c] += sg * q[c];
        }
        Real Rn[9];
        qframe<Real>(acc[0], acc[1], acc[2], acc[3], Rn);
        for (int c = 0; c < 3; c++) ld[c] += Rn[c] * lx + Rn[3 + c] * ly + Rn[6 + c] * lz;
      }
    }
    mxL = fmax(fabs(double(ld[0])), fmax(fabs(double(ld[1])), fabs(double(ld[2]))));
    for (int c = 0; c < 3; c++) F[c] += ld[c];
    for (int c = 0; c < 3; c++) F[c] += Fc[c];
    const uint8_t fx = I.fix[i];
    if (OUT) {
      if (fx & 8) { outR[3 * i] = -double(F[0]); outR[3 * i + 1] = -double(F[1]); outR[3 * i + 2] = -double(F[2]); }
      else { outR[3 * i] = outR[3 * i + 1] = outR[3 * i + 2] = 0.0; }
    }
    if (fx & 1) F[0] = Real(0);
    if (fx & 2) F[1] = Real(0);
    if (fx & 4) F[2] = Real(0);
    mxF = sqrt(double(F[0]) * double(F[0]) + double(F[1]) * double(F[1]) + double(F[2]) * double(F[2]));
    if (!OUT) {
      Real V[3];
      if (rbk || warm) { V[0] = V[1] = V[2] = Real(0); } else { V[0] = I.vx[i]; V[1] = I.vy[i]; V[2] = I.vz[i]; }
      if (warm) { F[0] = F[1] = F[2] = Real(0); }
      const Real ms = I.mass[i];
      for (int c = 0; c < 3; c++) V[c] += F[c] / ms;
      if (damping != Real(1)) for (int c = 0; c < 3; c++) V[c] *= damping;
      if (fx & 1) V[0] = Real(0);
      if (fx & 2) V[1] = Real(0);
      if (fx & 4) V[2] = Real(0);
      const int nx = cur ^ 1;
      I.x[nx][i] = L.X[t + 1][0] + double(V[0]);
      I.y[nx][i] = L.X[t + 1][1] + double(V[1]);
      I.z[nx][i] = L.X[t + 1][2] + double(V[2]);
      I.vx[i] = V[0]; I.vy[i] = V[1]; I.vz[i] = V[2];
      keN = double(ms) * (double(V[0]) * double(V[0]) + double(V[1]) * double(V[1]) + double(V[2]) * double(V[2]));
    }
  }
  // ---- frames: body torque, inertia, ω += tm / I, roll limit and damping, q ← q ⊗ exp(ω) (:766-808)
  if (ownSeg) {
    const int j = soff + jself;
    Real tm[3] = {L.ST[t][0], L.ST[t][1], L.ST[t][2]};
    int jx;
    if (joint_ok(k, jx)) for (int c = 0; c < 3; c++) tm[c] += L.JA[t + 1][c];
    if (joint_ok(k - 1, jx)) for (int c = 0; c < 3; c++) tm[c] += L.JB[t][c];
    if (I.C > 0) {
      for (int e = I.clamp_off[j]; e < I.clamp_off[j + 1]; e++) {
        const int cc = I.clamp_idx[e];
        double cq[4] = {I.c_Q[4 * cc], I.c_Q[4 * cc + 1], I.c_Q[4 * cc + 2], I.c_Q[4 * cc + 3]};
        double psi[3], dq[4];
        double th = relrot(cq, L.Q[t + 1], psi, dq);
        Real lc = I.c_len[cc];
        Real ps[3] = {Real(psi[0]), Real(psi[1]), Real(psi[2])};
        Real d[3] = {ps[0] / lc, ps[1] / lc, ps[2] / lc};
        Real mm[3] = {I.c_B[3 * cc] * d[0], I.c_B[3 * cc + 1] * d[1], I.c_B[3 * cc + 2] * d[2]};
        Real J[3], Jb[3];
        jlt(ps, Real(th), mm, J);
        rotT(dq, J, Jb);
        for (int c = 0; c < 3; c++) tm[c] -= Jb[c];
        eB += 0.5 * double(lc) * (double(mm[0]) * double(d[0]) + double(mm[1]) * double(d[1]));
        eT += 0.5 * double(lc) * double(mm[2]) * double(d[2]);
        if (OUT) { outCM[2 * cc] = double(mm[0]); outCM[2 * cc + 1] = double(mm[1]); }
      }
    }
    Real Rs[9];
    qframe<Real>(L.Q[t + 1][0], L.Q[t + 1][1], L.Q[t + 1][2], L.Q[t + 1][3], Rs);
    tm[0] += Rs[0] * Tc[0] + Rs[1] * Tc[1] + Rs[2] * Tc[2];
    tm[1] += Rs[3] * Tc[0] + Rs[4] * Tc[1] + Rs[5] * Tc[2];
    tm[2] += Rs[6] * Tc[0] + Rs[7] * Tc[1] + Rs[8] * Tc[2];
    mxT = sqrt(double(tm[0]) * double(tm[0]) + double(tm[1]) * double(tm[1]) + double(tm[2]) * double(tm[2]));
    if (!OUT) {
      const Real msc = Real(I.mass_scale);
      Real kd[3] = {L.KD[t][0] + L.KD[t + 1][0], L.KD[t][1] + L.KD[t + 1][1], L.KD[t][2] + L.KD[t + 1][2]};
      Real I3[3] = {msc * dmax(I.krx[j] + kTw + kd[0], Real(1e-12)), msc * dmax(I.kry[j] + kTw + kd[1], Real(1e-12)),
                    msc * dmax(I.krz[j] + kTw + kd[2], Real(1e-12))};
      Real W[3];
      if (rbk) { W[0] = W[1] = W[2] = Real(0); } else { W[0] = I.wx[j]; W[1] = I.wy[j]; W[2] = I.wz[j]; }
      for (int c = 0; c < 3; c++) W[c] += tm[c] / I3[c];
      W[2] = clampv(W[2], Real(-TWIST_RATE_MAX), Real(TWIST_RATE_MAX));
      if (kTw > Real(0)) W[2] *= Real(ROLL_DAMP);
      if (damping != Real(1)) for (int c = 0; c < 3; c++) W[c] *= damping;
      I.wx[j] = W[0]; I.wy[j] = W[1]; I.wz[j] = W[2];
      // q ⊗ exp(ω) in f64 (quat.js qexp / qmul), normalised
      double wxd = W[0], wyd = W[1], wzd = W[2];
      double a = sqrt(wxd * wxd + wyd * wyd + wzd * wzd), half = 0.5 * a;
      double sc = a > 1e-12 ? sin(half) / a : 0.5;
      double dw = cos(half), dx = wxd * sc, dy = wyd * sc, dz = wzd * sc;
      const double* q = L.Q[t + 1];
      double rw = q[0] * dw - q[1] * dx - q[2] * dy - q[3] * dz;
      double rx = q[0] * dx + q[1] * dw + q[2] * dz - q[3] * dy;
      double ry = q[0] * dy - q[1] * dz + q[2] * dw + q[3] * dx;
      double rz = q[0] * dz + q[1] * dy - q[2] * dx + q[3] * dw;
      double nq = sqrt(rw * rw + rx * rx + ry * ry + rz * rz); if (nq == 0) nq = 1.0;
      const int nx = cur ^ 1;
      I.qw[nx][j] = rw / nq; I.qx[nx][j] = rx / nq; I.qy[nx][j] = ry / nq; I.qz[nx][j] = rz / nq;
      keR = double(I3[0]) * double(W[0]) * double(W[0]) + double(I3[1]) * double(W[1]) * double(W[1]) + double(I3[2]) * double(W[2]) * double(W[2]);
    }
  }
  if (OUT) return;

  // ---- tile partial: sums keN, keR, eS, eSh, eB, eT | maxima |F|, |tm|, |load|, |N|, |M| (reduced by k_ctrl)
  double v[11] = {keN, keR, eS, eSh, eB, eT, mxF, mxT, mxL, mxN, mxM};
  block_reduce<TILE, 11>(L.red, v, 6);
  if (tid < 11) I.spart[size_t(tile_local) * SPART + tid] = v[tid];
}

// ================================================================================================= k_ctrl
// one workgroup per instance: fixed-order reduction of the tile and crossing-block partials, then kinetic damping,
// the residual, the stop test and the history ring (solver_cpu.js:809-837; wgsl.js:990-1047)
template <typename Real>
__global__ __launch_bounds__(256) void k_ctrl(const Inst<Real>* __restrict__ insts, int32_t B) {
  __shared__ double red[13][4];
  const int b = blockIdx.x;
  if (b >= B) return;
  const Inst<Real>& I = insts[b];
  Ctl* ctl = I.ctl;
  if (!inst_live(ctl)) return;
  const int tid = threadIdx.x;
  double s12[13];
  for (int q = 0; q < 7; q++) s12[q] = 0.0;
  for (int q = 7; q < 13; q++) s12[q] = q == 12 ? -INFINITY : 0.0;
  for (int w = tid; w < I.ntiles; w += 256) {
    const double* sp = I.spart + size_t(w) * SPART;
    double t0 = sp[0], t1 = sp[1], t2 = sp[2], t3 = sp[3], t4 = sp[4], t5 = sp[5];
    double m0 = sp[6], m1 = sp[7], m2 = sp[8], m3 = sp[9], m4 = sp[10];
    s12[0] += t0; s12[1] += t1; s12[2] += t2; s12[3] += t3; s12[4] += t4; s12[5] += t5;
    s12[7] = fmax(s12[7], m0); s12[8] = fmax(s12[8], m1); s12[9] = fmax(s12[9], m2); s12[10] = fmax(s12[10], m3); s12[11] = fmax(s12[11], m4);
  }
  for (int w = tid; w < I.ncblk; w += 256) {
    const double* cp = I.cpart + size_t(w) * CPART;
    s12[6] += cp[0];
    s12[12] = fmax(s12[12], cp[1]);
  }
  block_reduce<256, 13>(red, s12, 7);
  if (tid != 0) return;
  const double eCont = s12[6];
  Ctl c = *ctl;
  double ke = 0.5 * (s12[0] + s12[1]);
  int rollback = 0;
  if (I.kin_damp && ke < c.ke_prev) { rollback = 1; ke = 0.0; }
  const int iterPre = c.iter;
  c.rollback = rollback;
  c.ke_prev = ke; c.ke = ke;
  c.iter = iterPre + 1;
  if (I.aug && I.contacts) {
    if (c.al_update) c.pen_n = 0;
    if (iterPre >= max(I.ramp_iter, 0)) c.pen_n += 1;
    const int rampN = max(I.ramp_iter, 0);
    c.al_update = (I.al_every > 0 && c.iter > rampN && (c.iter - rampN) % I.al_every == 0) ? 1 : 0;
  } else {
    c.al_update = 0;
  }
  double fref = fmax(s12[9], s12[10]);
  if (I.p > 0) fref = fmax(fref, s12[11] / fmax(I.jlen_min, 1e-12));
  if (I.contacts) fref = fmax(fref, s12[12]);
  fref = fmax(fref, 1e-12);
  const double rf = I.n ? s12[7] / fref : 0.0, rt = I.m ? s12[8] / (fref * I.l_ref) : 0.0;
  double resid = fmax(rf, rt);
  if (!(c.iter > I.warmup)) resid = INFINITY;
  c.resid = resid; c.fref = fref; c.res_f = rf; c.res_t = rt;
  c.e_stretch = s12[2]; c.e_shear = s12[3]; c.e_bend = s12[4]; c.e_tors = s12[5]; c.e_cont = eCont;
  if (resid <= I.tol) c.done = 1;
  if (!rollback) c.cur = c.cur ^ 1;
  c.live_flag = 1;   // an iteration has been evaluated
  if (I.hist_len > 0) {
    double* h = I.hist + size_t((c.iter - 1) % I.hist_len) * HIST_STRIDE;
    h[0] = ke; h[1] = s12[2] + s12[3] + s12[4] + s12[5] + eCont; h[2] = s12[4]; h[3] = s12[5];
    h[4] = s12[2]; h[5] = s12[3]; h[6] = eCont; h[7] = resid; h[8] = double(c.iter);
  }
  c.hist_n += 1;
  *ctl = c;
}

// XCD-aware block mapping: consecutive logical tiles (neighbouring tiles of one lath share halo nodes) run on the
// same XCD and its L2; the hardware deals workgroups round-robin over the 8 XCDs
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int nx = 8;
  const int per = (nblk + nx - 1) / nx;
  const int x = bid % nx, r = bid / nx;
  const int full = nblk % nx == 0 ? nx : nblk % nx;   // XCDs that receive `per` blocks
  int base = x < full ? x * per : full * per + (x - full) * (per - 1);
  return base + r;
}

template <typename Real, int TILE>
__global__ __launch_bounds__(TILE) void k_strand(const Inst<Real>* __restrict__ insts, const TileItem* __restrict__ tiles, int32_t ntiles) {
  __shared__ StrandLDS<Real, TILE> L;
  int lt = xcd_tile(blockIdx.x, gridDim.x);
  if (lt >= ntiles) return;
  TileItem tl = tiles[lt];
  if (tl.inst < 0) return;
  const Inst<Real>& I = insts[tl.inst];
  if (!inst_live(I.ctl)) return;
  strand_tile<Real, TILE, false>(I, tl, L, lt - I.tile0, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// outputs of the last evaluated state of one instance: N, V, M, clamp moments, reactions (download)
template <typename Real, int TILE>
__global__ __launch_bounds__(TILE) void k_outputs(const Inst<Real>* __restrict__ insts, const TileItem* __restrict__ tiles, int32_t inst,
                                                  double* outN, double* outV, double* outM, double* outCM, double* outR) {
  __shared__ StrandLDS<Real, TILE> L;
  const Inst<Real>& I = insts[inst];
  int lt = I.tile0 + blockIdx.x;
  if (blockIdx.x >= I.ntiles) return;
  TileItem tl = tiles[lt];
  strand_tile<Real, TILE, true>(I, tl, L, blockIdx.x, outN, outV, outM, outCM, outR);
}

// explicit instantiations
#define WSD_INST(R, T) \
  template __global__ void k_strand<R, T>(const Inst<R>*, const TileItem*, int32_t); \
  template __global__ void k_outputs<R, T>(const Inst<R>*, const TileItem*, int32_t, double*, double*, double*, double*, double*);
WSD_INST(float, 64) WSD_INST(float, 128) WSD_INST(float, 256)
WSD_INST(double, 64) WSD_INST(double, 128) WSD_INST(double, 256)
template __global__ void k_cross<float>(const Inst<float>*, const CrossItem*, int32_t);
template __global__ void k_cross<double>(const Inst<double>*, const CrossItem*, int32_t);
template __global__ void k_ctrl<float>(const Inst<float>*, int32_t);
template __global__ void k_ctrl<double>(const Inst<double>*, int32_t);
template __global__ void k_begin<float>(const Inst<float>*, const int32_t*, int32_t, int32_t, int32_t);
template __global__ void k_begin<double>(const Inst<double>*, const int32_t*, int32_t, int32_t, int32_t);

}  // namespace wsd
