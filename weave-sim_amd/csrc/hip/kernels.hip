// HIP kernels (gfx950) of the dynamic-relaxation iteration.  One iteration is three launches (two when a batch is
// built with ctrl_mode 1, the control step then run by the strand launch; a rigid-step launch after them when a batch
// holds a structure only the ground holds):
//
//   k_cross   one or two lanes per crossing: node tangents and mean node frames around both sides' contact points,
//             C¹ Hermite point / tangent / frame on each side, contact normal with side persistence, penalty normal
//             force, Coulomb stick / slip with the u advance, W1 classification, W2 / W3 / AL branches; writes the
//             crossing state and one record per crossing side (force, lever torque, righting stiffness) and a partial
//             (contact energy, max Fn) per block — no atomics.  Follows PolylineContact.apply
//             (web/js/solver_cpu.js:381-478) and the crossing branch of k_force (web/js/wgsl.js:571-777).
//   k_strand  one workgroup per strand tile (a run of consecutive nodes of one lath): positions and frames of the
//             tile plus a one-node halo staged in LDS; segment stretch / shear forces (rod_cosserat.js:118-133), joint
//             bending / twist torques with J_l⁻ᵀ and the bending-coupling inertia (:136-159), clamp joints (:161-180);
//             the crossing-side records of the tile gathered in a fixed order; node and frame integration
//             (solver_cpu.js:751-808); one FP64 partial per tile (with ctrl_mode 1 the last tiles of an instance to
//             finish reduce the partials and run the control step: strand_control_tail).
//   k_ctrl    per instance, one or more workgroups: the partials reduced in a fixed order, then kinetic damping,
//             the residual, the stop test and the history ring (solver_cpu.js:809-837; control_step).
//
// The kinetic-damping roll-back needs no copy: positions and frames are double-buffered and `cur` simply does not
// flip on a roll-back, while the stale velocities are read as zero by the next iteration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <type_traits>

#include "device.hpp"

namespace wsd {

constexpr double SUPPORT_SMOOTH = 0.03;              // solver_cpu.js:16
constexpr double INV_SMOOTH = 1.0 / SUPPORT_SMOOTH;
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


// ------------------------------------------------------------------------------------------------ fast arithmetic
// Reciprocals: FP32 uses the hardware reciprocal (1 ulp; the FP32 form's forces carry 1e-7 anyway); FP64 refines
// the hardware estimate with two Newton steps to full double precision.
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0); r = fma(r, e, r);
  e = fma(-x, r, 1.0); return fma(r, e, r);
}
// 1/√x in f64: hardware estimate and two Newton steps
__device__ __forceinline__ double rsqrt_d(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// Wavefront reductions with DPP row shifts and row broadcasts (no LDS traffic).  Inclusive scan within each row of
// 16 lanes (row_shr 1, 2, 4, 8), then row 0 → row 1 and row 2 → row 3 (row_bcast:15), rows 0–1 → rows 2–3
// (row_bcast:31): lane 63 holds the wavefront's result.  Lanes whose source lies outside the row read 0, the neutral
// value of both the sums and the maxima taken here (all maxima are of non-negative quantities).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWS, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <bool SUM>
__device__ __forceinline__ double op_(double a, double b) { return SUM ? a + b : fmax(a, b); }
template <bool SUM>
__device__ __forceinline__ double wave_reduce63(double x) {
  x = op_<SUM>(x, dpp_<0x111, 0xF>(x));
  x = op_<SUM>(x, dpp_<0x112, 0xF>(x));
  x = op_<SUM>(x, dpp_<0x114, 0xF>(x));
  x = op_<SUM>(x, dpp_<0x118, 0xF>(x));
  x = op_<SUM>(x, dpp_<0x142, 0xA>(x));
  x = op_<SUM>(x, dpp_<0x143, 0xC>(x));
  return x;
}

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
  // the frame of q/|q| as I + (2/|q|²)(…): normalisation folded into one reciprocal, no square root
  const R n2 = w * w + x * x + y * y + z * z;
  const R s = n2 > R(0) ? R(2) * rcp_(n2) : R(2);
  F[0] = 1 - s * (y * y + z * z); F[1] = s * (x * y + z * w); F[2] = s * (x * z - y * w);
  F[3] = s * (x * y - z * w); F[4] = 1 - s * (x * x + z * z); F[5] = s * (y * z + x * w);
  F[6] = s * (x * z + y * w); F[7] = s * (y * z - x * w); F[8] = 1 - s * (x * x + y * y);
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

// the first eight words of a control record (iter, done, rollback, al_update, stop_at, hist_n, cur, counter) in one
// scalar load: separate narrow loads of the same record can be split over two memory latencies
struct CtlHead { int32_t iter, done, rollback, al_update, stop_at, hist_n, cur, counter; };
__device__ __forceinline__ CtlHead ctl_head(const Ctl* c) {
  typedef int32_t v8 __attribute__((ext_vector_type(8))) __attribute__((aligned(8)));
  const v8 v = *reinterpret_cast<const v8*>(c);
  return CtlHead{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
}
static_assert(offsetof(Ctl, iter) == 0 && offsetof(Ctl, done) == 4 && offsetof(Ctl, rollback) == 8 &&
              offsetof(Ctl, al_update) == 12 && offsetof(Ctl, stop_at) == 16 && offsetof(Ctl, cur) == 24,
              "control record head layout");

// one of a descriptor's two state buffers: both pointers are read and one is selected, so the read does not wait for
// the control record that names the buffer (a read indexed by it would)
template <typename T> __device__ __forceinline__ T* buf(T* const (&p)[2], int k) {
  T* const a = p[0];
  T* const b = p[1];
  return k ? b : a;
}

// An array pointer read from an instance descriptor, in the global address space: its loads and stores become
// global_* instead of flat_* (a flat access also counts against the LDS wait counter, so every LDS wait would wait
// for it too).  The descriptor's pointers all come from hipMalloc.
template <typename T> __device__ __forceinline__ __attribute__((address_space(1))) T* gl(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// ================================================================================================= k_begin
// step(nIter, {instances, force}) (solver_gpu.js:338-345): stop_at = iter + nIter, clear the convergence flag
template <typename Real>
__global__ void k_begin(Ctl* ctls, const int32_t* sel, const int32_t* counts, int32_t nsel, int32_t n_iter, int32_t force) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsel) return;
  Ctl* c = ctls + sel[t];
  c->stop_at = c->iter + (counts ? counts[t] : n_iter);
  if (force) c->done = 0;
}

// field-wise copies of a side record (keeps the record in registers: no aggregate copy through private memory)
template <typename P, typename Real>
__device__ __forceinline__ void side_load(P p, SideRec<Real>& r) {
  r.seg_pre = p->seg_pre; r.seg_post = p->seg_post; r.t = p->t; r.kth = p->kth;
  r.F[0] = p->F[0]; r.F[1] = p->F[1]; r.F[2] = p->F[2]; r.T[0] = p->T[0]; r.T[1] = p->T[1]; r.T[2] = p->T[2];
}
template <typename P, typename Real>
__device__ __forceinline__ void side_store(const SideRec<Real>& r, P p) {
  p->seg_pre = r.seg_pre; p->seg_post = r.seg_post; p->t = r.t; p->kth = r.kth;
  p->F[0] = r.F[0]; p->F[1] = r.F[1]; p->F[2] = r.F[2]; p->T[0] = r.T[0]; p->T[1] = r.T[1]; p->T[2] = r.T[2];
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
  const int ns = gl(I.k_nseg)[2 * c + side];
  const int noff = gl(I.k_noff)[2 * c + side];
  const int nnr = gl(I.k_nn)[2 * c + side];
  const bool closed = (nnr >> 30) & 1;
  const int nn = nnr & 0x3fffffff;
  const int soff = gl(I.k_seg_off)[2 * c + side];
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
  // positions and frames are interleaved (x y z per node, w x y z per segment): a side's four nodes and three
  // segments are two short contiguous runs, a few cache lines
  const auto P = gl(buf(I.X, cur));
  const double xa = P[3 * E.a], ya = P[3 * E.a + 1], za = P[3 * E.a + 2];
  const double xb = P[3 * E.b], yb = P[3 * E.b + 1], zb = P[3 * E.b + 2];
  double xm = 0, ym = 0, zm = 0, xp = 0, yp = 0, zp = 0;
  if (hasPrevSeg) { const int k = 3 * (noff + iam); xm = P[k]; ym = P[k + 1]; zm = P[k + 2]; }
  if (hasNextSeg) { const int k = 3 * (noff + ibp); xp = P[k]; yp = P[k + 1]; zp = P[k + 2]; }
  const auto Qv = gl(buf(I.Q, cur));
  G q0[4] = {G(Qv[4 * E.seg]), G(Qv[4 * E.seg + 1]), G(Qv[4 * E.seg + 2]), G(Qv[4 * E.seg + 3])};
  G qm[4] = {0, 0, 0, 0}, qp[4] = {0, 0, 0, 0};
  if (hasPrevSeg) { const int j = 4 * (soff + sm); qm[0] = G(Qv[j]); qm[1] = G(Qv[j + 1]); qm[2] = G(Qv[j + 2]); qm[3] = G(Qv[j + 3]); }
  if (hasNextSeg) { const int j = 4 * (soff + sp); qp[0] = G(Qv[j]); qp[1] = G(Qv[j + 1]); qp[2] = G(Qv[j + 2]); qp[3] = G(Qv[j + 3]); }
  auto unitAdd = [](G* T, double dx, double dy, double dz) {
    G ex = G(dx), ey = G(dy), ez = G(dz);
    G il = rcp_(dmax(fsqrt_(ex * ex + ey * ey + ez * ez), G(1e-12)));
    T[0] += ex * il; T[1] += ey * il; T[2] += ez * il;
  };
  auto normT = [](G* T) {
    G il = rcp_(dmax(fsqrt_(T[0] * T[0] + T[1] * T[1] + T[2] * T[2]), G(1e-12)));
    T[0] *= il; T[1] *= il; T[2] *= il;
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
    const G inq = rcp_(nq);
    for (int k = 0; k < 4; k++) out[k] *= inq;
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

// half-extent h(n) of a side's section facing sign·n (f64: it enters the penetration) and its smoothed support
// point r (precision H: it only sets the lever of the contact force) (solver_cpu.js:274-282)
template <typename G, typename H>
__device__ __forceinline__ G support(const G* R, const G* n, G sign, G w, G h_, H* r) {
  G nx = (n[0] * R[0] + n[1] * R[1] + n[2] * R[2]) * sign, ny = (n[0] * R[3] + n[1] * R[4] + n[2] * R[5]) * sign;
  G h = G(0.5) * w * fabs(nx) + G(0.5) * h_ * fabs(ny);
  H ax = H(0.5) * H(w) * ftanh_(H(nx) * H(INV_SMOOTH)), ay = H(0.5) * H(h_) * ftanh_(H(ny) * H(INV_SMOOTH));
  for (int k = 0; k < 3; k++) r[k] = ax * H(R[k]) + ay * H(R[3 + k]);
  return h;
}

// W3 wall slope tan ψ (solver_cpu.js:285-293)
template <typename H>
__device__ __forceinline__ H walls(int sideCut, const H* R, const H* n, const H* mm) {
  H sg = sideCut == 0 ? H(1) : H(-1);
  H nx = (n[0] * R[0] + n[1] * R[1] + n[2] * R[2]) * sg, ny = (n[0] * R[3] + n[1] * R[4] + n[2] * R[5]) * sg;
  H fx[3], fy[3];
  for (int k = 0; k < 3; k++) { fx[k] = jsSign(nx) * R[k]; fy[k] = jsSign(ny) * R[3 + k]; }
  const bool useX = (fx[0] * mm[0] + fx[1] * mm[1] + fx[2] * mm[2]) > H(0);
  H f0 = useX ? fx[0] : fy[0], f1 = useX ? fx[1] : fy[1], f2 = useX ? fx[2] : fy[2];
  H cp = clampv(H(fabs(f0 * n[0] + f1 * n[1] + f2 * n[2])), H(1e-6), H(1));
  return ftan_(facos_(cp));
}

// share of a side's force carried by free nodes, (1 − t)² + t² over the nodes not fixed in all three directions, at
// the side's current parameter u (freeWeight, solver_cpu.js:483-487)
template <typename Real>
__device__ __forceinline__ double side_free_weight(const Inst<Real>& I, int q) {
  const int ns = gl(I.k_nseg)[q], nnr = gl(I.k_nn)[q], nn = nnr & 0x3fffffff;
  const bool closed = (nnr >> 30) & 1;
  const double uu = gl(I.u)[q];
  const int i = int(fmin(fmax(floor(uu), 0.0), double(max(ns - 1, 0))));
  const double t = fmin(fmax(uu - i, 0.0), 1.0);
  const int a = gl(I.k_noff)[q] + i, b = gl(I.k_noff)[q] + (closed ? (i + 1) % nn : i + 1);
  const bool fa = (gl(I.fix)[a] & 16) != 0, fb = (gl(I.fix)[b] & 16) != 0;
  return (fa ? 0.0 : (1 - t) * (1 - t)) + (fb ? 0.0 : t * t);
}

// the multiplier after an AL update (alValue + updateMultipliers, solver_cpu.js:492-506): λ + k_n · mean penetration
// since the last update, kept ≥ 0 unless locked, clipped to the crushing cap, zero when disabled or pinned.  The
// ramp is complete whenever the multipliers update (updates start after the ramp).
template <typename Real>
__device__ __forceinline__ double al_updated(const Inst<Real>& I, int c, double lam, double penacc, int pen_n) {
  const bool enabled = gl(I.k_en)[c] != 0, locked = gl(I.k_lock)[c] != 0;
  const bool pinned = side_free_weight(I, 2 * c) < AL_FREE_MIN || side_free_weight(I, 2 * c + 1) < AL_FREE_MIN;
  const double cap = fmin(double(gl(I.fc90)[2 * c]), double(gl(I.fc90)[2 * c + 1])) * double(gl(I.wk)[2 * c]) * double(gl(I.wk)[2 * c + 1]) /
                     fmax(sin(double(gl(I.theta)[c])), 1e-6);
  double v = lam + I.strength * (penacc / double(max(pen_n, 1)));
  if (!locked) v = fmax(v, 0.0);
  v = fmin(fmax(v, -cap), cap);
  return (!enabled || pinned) ? 0.0 : v;
}

__device__ __forceinline__ double xchg(double v) { return __shfl_xor(v, 1, 64); }
__device__ __forceinline__ float xchg(float v) { return __shfl_xor(v, 1, 64); }
__device__ __forceinline__ int xchg(int v) { return __shfl_xor(v, 1, 64); }

// XCD-aware block mapping: consecutive logical blocks run on the same XCD and its L2 (the hardware deals workgroups
// round-robin over the 8 XCDs).  Neighbouring strand tiles of one lath share halo nodes; the work lists are ordered by
// instance, so an instance's crossing blocks and control workgroups land about where its strand tiles ran
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int nx = 8;
  const int per = (nblk + nx - 1) / nx;
  const int x = bid % nx, r = bid / nx;
  const int full = nblk % nx == 0 ? nx : nblk % nx;   // XCDs that receive `per` blocks
  int base = x < full ? x * per : full * per + (x - full) * (per - 1);
  return base + r;
}

// LANES = 2: lane 2j evaluates side A of crossing c0 + j, lane 2j + 1 side B.  The side evaluations (the neighbour
// loads, tangents, node frames and the Hermite point) run in parallel; the lanes then swap the partner's point,
// tangent and frame and both compute the crossing's normal, penetration and friction law (the same instructions on
// the same data, so the two copies agree bit for bit).  Each lane writes its own side's u and record; lane 2j writes
// the crossing's state.  Short dependency chains for batches too small to fill the chip.
// LANES = 1: lane j evaluates both sides of crossing c0 + j one after the other and writes everything: half the
// instructions per crossing, for batches with enough crossings to fill the chip.  Both layouts run the same
// arithmetic on each side, so they give the same results.  Contact kinematics run in f64 in both forms (the penetration
// D − g cancels mm-scale lengths); the classification, the lever arms and the W2 / W3 trigonometry run in Real.
// ONE: the batch holds a single instance (slot 0) and block b takes its crossings [b·n, b·n + n), n the crossings per
// block, K = oneK: the work item follows from the block index and the descriptor and control record from slot 0, so
// they load together instead of one after the other
#ifndef WS_CROSS_WPE
#define WS_CROSS_WPE 2   // resident wavefronts per SIMD the crossing kernel is compiled for
#endif
template <typename Real, int LANES, bool ONE>
__global__ __launch_bounds__(CROSS_THREADS) __attribute__((amdgpu_waves_per_eu(WS_CROSS_WPE))) void k_cross(const Inst<Real>* __restrict__ insts, const CrossItem* __restrict__ items,
                                                        int32_t nitems, const Ctl* __restrict__ ctls, int32_t oneK) {
  static_assert(LANES == 1 || LANES == 2, "one or two lanes per crossing");
  using G = double;
  using H = Real;
  const int bid = xcd_tile(blockIdx.x, gridDim.x);
  if (bid >= nitems) return;
  // the work list holds blocks of occupied slots only (it.inst >= 0); the liveness test waits until the first
  // crossing loads are out, so the item, the control record with the descriptor, and the crossing state cost three
  // memory latencies, not one per test
  constexpr int PER_BLK = CROSS_THREADS / LANES;
  const CrossItem it = ONE ? CrossItem{0, bid * PER_BLK, min(PER_BLK, oneK - bid * PER_BLK), bid} : items[bid];
  const Inst<Real>& I = insts[ONE ? 0 : it.inst];
  const Ctl* ctl = ctls + (ONE ? 0 : it.inst);
  const CtlHead ch = ctl_head(ctl);
  const int cur = ch.cur, iter = ch.iter;
  const bool live = (ch.done == 0) & (iter < ch.stop_at);
  const bool al_update = ch.al_update != 0;
  const int pen_n = ctl->pen_n;
  const int tid = threadIdx.x;
  const int side = LANES == 2 ? (tid & 1) : 0, jj = LANES == 2 ? (tid >> 1) : tid;
  const bool valid = jj < it.count;
  const int c = it.c0 + (valid ? jj : 0);     // idle lanes shadow the block's first crossing and write nothing
  const bool lead = valid && side == 0;
  double eC = 0.0, mxFn = -INFINITY;

  const int rampN = max(I.ramp_iter, 0);
  const double rampD = rampN == 0 ? 1.0 : fmin(1.0, 0.01 + 0.99 * iter / rampN);
  const G kn = G(rampD) * G(I.strength), kt = G(rampD) * G(I.stick);
  const bool enabled = gl(I.k_en)[c] != 0;
  const bool locked = gl(I.k_lock)[c] != 0;
  const bool seat = I.seat != 0, aug = I.aug != 0;
  const G mu0 = G(I.mu);
  // ---- own side, relative to the first node of A's segment
  const double uA = gl(I.u)[2 * c], uB = gl(I.u)[2 * c + 1];
  const int noffA = gl(I.k_noff)[2 * c], nsegA = gl(I.k_nseg)[2 * c];
  if (!live) return;                               // uniform over the workgroup (one instance per block)
  // every other per-crossing value the law reads (sections, classification constants, the previous normal, the slip
  // total, the side records' slots), issued here with the first loads: read where it is used, each would cost a
  // memory latency of its own in the middle of the arithmetic
  const Real wk0 = gl(I.wk)[2 * c], wk1 = gl(I.wk)[2 * c + 1], hk0 = gl(I.hk)[2 * c], hk1 = gl(I.hk)[2 * c + 1];
  const Real Hs0 = gl(I.Hs)[2 * c], Hs1 = gl(I.Hs)[2 * c + 1], fc0 = gl(I.fc90)[2 * c], fc1 = gl(I.fc90)[2 * c + 1];
  const Real E900 = gl(I.E90)[2 * c], E901 = gl(I.E90)[2 * c + 1], kwc = gl(I.kw)[c], khc = gl(I.kh)[c];
  const Real np0 = gl(I.nprev)[3 * c], np1 = gl(I.nprev)[3 * c + 1], np2 = gl(I.nprev)[3 * c + 2];
  const double slip0 = gl(I.slip)[c];
  const int nsegB = gl(I.k_nseg)[2 * c + 1], spos0 = gl(I.side_pos)[2 * c], spos1 = gl(I.side_pos)[2 * c + 1];
  const double uOwn = side == 0 ? uA : uB;
  const int o = noffA + int(fmin(fmax(floor(uA), 0.0), double(max(nsegA - 1, 0))));
  const auto Xc = gl(buf(I.X, cur));
  const double ox = Xc[3 * o], oy = Xc[3 * o + 1], oz = Xc[3 * o + 2];
  SideEv<G> E, E2;                                // own side; side B when one lane takes both
  eval_side(I, cur, c, side, uOwn, ox, oy, oz, E);
  if (LANES == 1) eval_side(I, cur, c, 1, uB, ox, oy, oz, E2);
  // ---- partner's side
  G oP[3], oT[3], oR[6];
  if (LANES == 2) {
    for (int k = 0; k < 3; k++) { oP[k] = xchg(E.P[k]); oT[k] = xchg(E.tg[k]); }
    for (int k = 0; k < 6; k++) oR[k] = xchg(E.R[k]);
  } else {
    for (int k = 0; k < 3; k++) { oP[k] = E2.P[k]; oT[k] = E2.tg[k]; }
    for (int k = 0; k < 6; k++) oR[k] = E2.R[k];
  }
  G AP[3], BP[3], tA[3], tB[3], AR[6], BR[6];
  for (int k = 0; k < 3; k++) {
    AP[k] = side == 0 ? E.P[k] : oP[k]; BP[k] = side == 0 ? oP[k] : E.P[k];
    tA[k] = side == 0 ? E.tg[k] : oT[k]; tB[k] = side == 0 ? oT[k] : E.tg[k];
  }
  for (int k = 0; k < 6; k++) { AR[k] = side == 0 ? E.R[k] : oR[k]; BR[k] = side == 0 ? oR[k] : E.R[k]; }
  const G wA = wk0, wB = wk1, hA = hk0, hB = hk1;
  // ---- AL multiplier update deferred from the end of the previous iteration (updateMultipliers, :499-506)
  G lam = aug ? G(gl(I.lam)[c]) : G(0);
  G penacc = aug ? G(gl(I.penacc)[c]) : G(0);
  if (aug && al_update) {
    lam = al_updated(I, c, lam, penacc, pen_n);
    penacc = G(0);
  }
  G np[3] = {np0, np1, np2};
  G n[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
  G nn = fsqrt_(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nn < G(1e-6)) { n[0] = np[0]; n[1] = np[1]; n[2] = np[2]; }
  else { G id = rcp_(dmax(nn, G(1e-12))); n[0] *= id; n[1] *= id; n[2] *= id; }
  if (n[0] * np[0] + n[1] * np[1] + n[2] * np[2] < G(0)) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
  G S[3] = {BP[0] - AP[0], BP[1] - AP[1], BP[2] - AP[2]};
  G g = S[0] * n[0] + S[1] * n[1] + S[2] * n[2];
  H rA[3], rB[3];
  G supA = support(AR, n, G(1), wA, hA, rA);
  G supB = support(BR, n, G(-1), wB, hB, rB);
  G D = supA + supB;
  if (seat && I.seat_gap) D -= G(gl(I.dent)[2 * c]) + G(gl(I.dent)[2 * c + 1]) + G(gl(I.dente)[c]);
  G pen = D - g;
  if (aug && iter >= rampN) penacc += pen;
  G Fn = 0; bool act;
  if (aug) {
    G trial = lam + kn * pen;
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
  const G idet = ok ? rcp_(det) : G(0);
  G ca = ok ? (sa * tbb - sb * tab) * idet : G(0), cb = ok ? (sb * taa - sa * tab) * idet : G(0);
  G Ftv = fsqrt_((kt * St[0]) * (kt * St[0]) + (kt * St[1]) * (kt * St[1]) + (kt * St[2]) * (kt * St[2]));
  // Real copies of the frames and the normal for the classification and the W2 / W3 branches
  H nR[3] = {H(n[0]), H(n[1]), H(n[2])}, tAR[3] = {H(tA[0]), H(tA[1]), H(tA[2])}, tBR[3] = {H(tB[0]), H(tB[1]), H(tB[2])};
  H ARr[6], BRr[6];
  for (int k = 0; k < 6; k++) { ARr[k] = H(AR[k]); BRr[k] = H(BR[k]); }
  // ---- W3 directional seat friction (_seatFriction, :295-315)
  G mu = mu0;
  H tanpsi[2] = {H(0), H(0)};
  if (seat) {
    H Str[3] = {H(St[0]), H(St[1]), H(St[2])};
    H norm = fsqrt_(Str[0] * Str[0] + Str[1] * Str[1] + Str[2] * Str[2]);
    H dn = dmax(norm, H(1e-12));
    const H idn = rcp_(dn);
    H dh[3] = {Str[0] * idn, Str[1] * idn, Str[2] * idn};
    bool has = norm > H(1e-12);
    for (int j = 0; j < 2; j++) {
      tanpsi[j] = gl(I.tanpsi)[2 * c + j];
      if (!(gl(I.dent)[2 * c + j] > Real(0) && has)) continue;
      const H* tCut = j == 0 ? tBR : tAR;
      const H* Rcut = j == 0 ? BRr : ARr;
      H mm[3] = {nR[1] * tCut[2] - nR[2] * tCut[1], nR[2] * tCut[0] - nR[0] * tCut[2], nR[0] * tCut[1] - nR[1] * tCut[0]};
      H ml = dmax(fsqrt_(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), H(1e-12));
      const H iml = rcp_(ml);
      mm[0] *= iml; mm[1] *= iml; mm[2] *= iml;
      H sgn = j == 0 ? H(1) : H(-1);
      H dj = (dh[0] * mm[0] + dh[1] * mm[1] + dh[2] * mm[2]) * sgn;
      H wd[3] = {jsSign(dj) * mm[0] * sgn, jsSign(dj) * mm[1] * sgn, jsSign(dj) * mm[2] * sgn};
      H tp = walls(1 - j, Rcut, nR, wd);
      tanpsi[j] = tp;
      H m0 = H(mu0);
      H add = fabs(dj) * (m0 * tp < H(1) ? (m0 + tp) / dmax(H(1) - m0 * tp, H(1e-6)) - m0 : H(9) * m0);
      mu += G(dmin(add, H(9) * m0));
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
  G Se[3] = {0, 0, 0};
  if (act) { Se[0] = St[0] - da * tA[0] - db * tB[0]; Se[1] = St[1] - da * tA[1] - db * tB[1]; Se[2] = St[2] - da * tA[2] - db * tB[2]; }
  G fT[3] = {kt * Se[0], kt * Se[1], kt * Se[2]};
  G FA[3], FB[3];
  for (int k = 0; k < 3; k++) { FA[k] = -Fn * n[k] + fT[k]; FB[k] = Fn * n[k] - fT[k]; }
  // ---- W1 classification (_classify, :231-271)
  H cface[2], z0[2], phiv[2]; int edge[2];
  for (int sd = 0; sd < 2; sd++) {
    const H* R = sd == 0 ? ARr : BRr;
    H sg = sd == 0 ? H(1) : H(-1);
    H wv = H(sd == 0 ? wA : wB), hv = H(sd == 0 ? hA : hB);
    H nx = (nR[0] * R[0] + nR[1] * R[1] + nR[2] * R[2]) * sg, ny = (nR[0] * R[3] + nR[1] * R[4] + nR[2] * R[5]) * sg;
    H sx = fabs(ftanh_(nx * H(INV_SMOOTH))), sy = fabs(ftanh_(ny * H(INV_SMOOTH)));
    edge[sd] = sx * sy > H(I.edge_thr) ? 1 : 0;
    bool yf = fabs(ny) >= fabs(nx);
    cface[sd] = yf ? wv : hv; z0[sd] = yf ? hv : wv;
    phiv[sd] = fatan_(dmin(H(fabs(nx)), H(fabs(ny))) / dmax(dmax(H(fabs(nx)), H(fabs(ny))), H(1e-12)));
  }
  H cr[3] = {tAR[1] * tBR[2] - tAR[2] * tBR[1], tAR[2] * tBR[0] - tAR[0] * tBR[2], tAR[0] * tBR[1] - tAR[1] * tBR[0]};
  H sinT = clampv(fsqrt_(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]), H(1e-6), H(1));
  H theta = fasin_(sinT);
  int nEdge = edge[0] + edge[1];
  int faceSide = nEdge == 1 ? (edge[1] ? 0 : 1) : -1;
  H Fa = H(fabs(Fn));
  H Aa = cface[0] * cface[1] / sinT, pres = Fa / dmax(Aa, H(1e-12));
  H ut = pres / dmax(H(dmin(Hs0, Hs1)), H(1e-12));
  if (nEdge == 1) {
    int fs = faceSide, es = 1 - fs;
    H ell = cface[fs] / sinT, kap = H(fs ? E901 : E900) / dmax(z0[fs], H(1e-12));
    H phiE = dmax(phiv[es], H(1e-4));
    H bw = fsqrt_(H(2) * Fa / dmax(ell * kap * phiE, H(1e-12)));
    bool full = bw >= cface[es];
    H Afe = full ? cface[0] * cface[1] / sinT : bw * ell;
    H pfe = full ? Fa / dmax(Afe, H(1e-12)) : kap * phiE * bw;
    Aa = Afe; pres = pfe; ut = pfe / dmax(H(es ? fc1 : fc0), H(1e-12));
  } else if (nEdge == 2) {
    H fc = H(dmin(fc0, fc1));
    Aa = Fa / dmax(fc, H(1e-12)); pres = Fa > H(0) ? fc : H(0); ut = Fa > H(0) ? H(1) : H(0);
  }
  // ---- W3 seat update (_seatUpdate, :317-344)
  H dent[2] = {H(0), H(0)}, climb[2] = {H(0), H(0)}, dente = H(0);
  if (seat) {
    dent[0] = gl(I.dent)[2 * c]; dent[1] = gl(I.dent)[2 * c + 1]; climb[0] = gl(I.climb)[2 * c]; climb[1] = gl(I.climb)[2 * c + 1]; dente = gl(I.dente)[c];
    H dar = H(da), dbr = H(db);
    H inc[3] = {dar * tAR[0] - dbr * tBR[0], dar * tAR[1] - dbr * tBR[1], dar * tAR[2] - dbr * tBR[2]};
    for (int j = 0; j < 2; j++) {
      const H* tCut = j == 0 ? tBR : tAR;
      H mm[3] = {nR[1] * tCut[2] - nR[2] * tCut[1], nR[2] * tCut[0] - nR[0] * tCut[2], nR[0] * tCut[1] - nR[1] * tCut[0]};
      H ml = dmax(fsqrt_(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]), H(1e-12));
      const H iml = rcp_(ml);
      mm[0] *= iml; mm[1] *= iml; mm[2] *= iml;
      H stp = (inc[0] * mm[0] + inc[1] * mm[1] + inc[2] * mm[2]) * (j == 0 ? H(1) : H(-1));
      bool has = dent[j] > H(0);
      if (has && slip) climb[j] += stp;
      if (has && fabs(climb[j]) >= dent[j] / dmax(tanpsi[j], H(1e-6))) { dent[j] = H(0); climb[j] = H(0); }
    }
    if (iter >= I.dent_after) {
      H sp0 = fsin_(phiv[0]) * fcos_(phiv[0]), sp1 = fsin_(phiv[1]) * fcos_(phiv[1]);
      H sT = fsin_(theta);
      if (nEdge == 2 && act && !locked) {
        H fc = H(dmin(fc0, fc1));
        H dEst = fsqrt_(Fa / dmax(fc, H(1e-12)) * sT * sp0 * sp1);
        for (int j = 0; j < 2; j++) { H cu = dent[j]; dent[j] = dmax(cu, dmin(cu + H(I.dent_rate) * dEst, dEst)); }
      }
      if (nEdge == 1 && act && !locked) {
        int fs = faceSide, es = 1 - fs;
        H ell = cface[fs] / dmax(sT, H(1e-6));
        H dE = Fa / dmax(H(es ? fc1 : fc0) * ell, H(1e-12)) * (es == 0 ? sp0 : sp1);
        dente = dmax(dente, dE);
      }
    }
  }
  // ---- lever torques, W2 rotational friction (_rotationalFriction, :347-379)
  H tf = H(I.torsion_factor);
  H FAr[3] = {H(FA[0]), H(FA[1]), H(FA[2])}, FBr[3] = {H(FB[0]), H(FB[1]), H(FB[2])};
  H TA[3] = {tf * (rA[1] * FAr[2] - rA[2] * FAr[1]), tf * (rA[2] * FAr[0] - rA[0] * FAr[2]), tf * (rA[0] * FAr[1] - rA[1] * FAr[0])};
  H TB[3] = {tf * (rB[1] * FBr[2] - rB[2] * FBr[1]), tf * (rB[2] * FBr[0] - rB[0] * FBr[2]), tf * (rB[0] * FBr[1] - rB[1] * FBr[0])};
  // the relative rotation θ − θ0 cancels two angles near each other, and k_r is large: the W2 angle law runs in f64
  H eRot = 0, Mr = H(0);
  double th0 = 0, rotslip = 0;
  if (I.rot_fric) {
    G crd[3] = {tA[1] * tB[2] - tA[2] * tB[1], tA[2] * tB[0] - tA[0] * tB[2], tA[0] * tB[1] - tA[1] * tB[0]};
    G thc = atan2(crd[0] * n[0] + crd[1] * n[1] + crd[2] * n[2], tab);
    th0 = gl(I.theta0)[c];
    rotslip = gl(I.rotslip)[c];
    if (isnan(th0) && act) th0 = thc;
    bool okr = act && !isnan(th0);
    G sT2 = fmax(sin(G(theta)), G(1e-6));
    G req = G(0.38) * sqrt(wA * wB / sT2);
    if (nEdge == 1) req = G(0.25) * (faceSide == 0 ? wA : wB) / sT2;
    else if (nEdge == 2) req = G(0.38) * sqrt(fmax(G(Aa), G(1e-12)));
    G kr = kt * req * req;
    G dth = okr ? thc - th0 : G(0);
    const G PI = 3.141592653589793;
    dth = fmod(fmod(dth + PI, 2 * PI) + 2 * PI, 2 * PI) - PI;
    G Mc = kr * dth;
    G Mlim = locked ? G(INFINITY) : G(I.rot_factor) * mu0 * G(Fa) * req;
    if (okr && fabs(Mc) > Mlim) {
      G adv = clampv(clampv(G(I.slip_rate), G(1e-3), G(1)) * (G(1) - Mlim / fmax(fabs(Mc), G(1e-12))) * dth, G(-0.05), G(0.05));
      th0 += adv; rotslip += fabs(adv); dth -= adv;
      Mc = kr * dth;
    }
    Mr = H(okr ? Mc : G(0));
    for (int k = 0; k < 3; k++) { TA[k] += Mr * nR[k]; TB[k] -= Mr * nR[k]; }
    eRot = H(0.5 * kr * dth * dth);
  }
  // ---- righting stiffness (addTwistStiffness, :170-176) on the own side's segment after the slip
  G tfk = dmax(G(I.torsion_factor), G(1e-3));
  G kth = tfk * ((Fn + kn * G(0.25) * G(khc)) * (G(0.5) * G(kwc)) / G(SUPPORT_SMOOTH) + kt * (G(0.5) * G(khc)) * (G(0.5) * G(khc)));
  // a side's parameter after the slip and its record
  auto finish_side = [&](int sd, const SideEv<G>& Es, double uOld) {
    const int nsS = sd ? nsegB : nsegA;
    double uNew = uOld;
    if (moved) {
      const double iL = rcp_(double(dmax(Es.L, G(1e-12))));
      const double du = sd == 0 ? double(da) * iL : -double(db) * iL;
      uNew = fmin(fmax(uOld + du, 0.0), double(nsS));
      gl(I.u)[2 * c + sd] = uNew;
    }
    SideRec<Real> rec;
    rec.seg_pre = Es.loc;
    rec.seg_post = int(fmin(fmax(floor(uNew), 0.0), double(max(nsS - 1, 0))));
    rec.t = Real(Es.t); rec.kth = Real(kth);
    for (int k = 0; k < 3; k++) {
      rec.F[k] = Real(sd == 0 ? FA[k] : FB[k]);
      rec.T[k] = Real(sd == 0 ? TA[k] : TB[k]);
    }
    side_store(rec, gl(I.side) + (sd ? spos1 : spos0));
  };
  if (valid) {
    finish_side(side, E, uOwn);
    if (LANES == 1) finish_side(1, E2, uB);
  }
  if (lead) {
    gl(I.nprev)[3 * c] = Real(n[0]); gl(I.nprev)[3 * c + 1] = Real(n[1]); gl(I.nprev)[3 * c + 2] = Real(n[2]);
    if (slip) {
      G q0 = da * tA[0] + db * tB[0], q1 = da * tA[1] + db * tB[1], q2 = da * tA[2] + db * tB[2];
      gl(I.slip)[c] = slip0 + double(fsqrt_(q0 * q0 + q1 * q1 + q2 * q2));
    }
    if (aug) { gl(I.lam)[c] = lam; gl(I.penacc)[c] = penacc; }
    gl(I.fn)[c] = Real(Fn);
    gl(I.ft)[c] = Real(fsqrt_(fT[0] * fT[0] + fT[1] * fT[1] + fT[2] * fT[2]));
    gl(I.gap)[c] = Real(g);
    gl(I.kflags)[c] = uint8_t((slip ? 1 : 0) | (moved ? 0 : 2));
    gl(I.mid)[3 * c] = Real(G(0.5) * (AP[0] + BP[0]) + ox);
    gl(I.mid)[3 * c + 1] = Real(G(0.5) * (AP[1] + BP[1]) + oy);
    gl(I.mid)[3 * c + 2] = Real(G(0.5) * (AP[2] + BP[2]) + oz);
    gl(I.theta)[c] = theta; gl(I.anom)[c] = Aa; gl(I.pres)[c] = pres; gl(I.util)[c] = ut;
    gl(I.phi)[2 * c] = phiv[0]; gl(I.phi)[2 * c + 1] = phiv[1];
    gl(I.ctype)[c] = int8_t(nEdge); gl(I.face)[c] = int8_t(faceSide);
    if (seat) {
      gl(I.tanpsi)[2 * c] = tanpsi[0]; gl(I.tanpsi)[2 * c + 1] = tanpsi[1];
      gl(I.dent)[2 * c] = dent[0]; gl(I.dent)[2 * c + 1] = dent[1]; gl(I.climb)[2 * c] = climb[0]; gl(I.climb)[2 * c + 1] = climb[1];
      gl(I.dente)[c] = dente;
    }
    if (I.rot_fric) { gl(I.theta0)[c] = th0; gl(I.rotslip)[c] = rotslip; gl(I.mrot)[c] = Mr; }
    G ePen = act ? (aug ? Fn * Fn / dmax(kn, G(1e-30)) : kn * pen * pen) : G(0);
    eC = 0.5 * double(ePen) + 0.5 * double(kt) * double(Se[0] * Se[0] + Se[1] * Se[1] + Se[2] * Se[2]) + double(eRot);
    mxFn = double(Fn);
  }
  // ---- block partial (one wavefront; fixed DPP order)
  static_assert(CROSS_THREADS == 64, "the crossing partial is a single-wavefront reduction");
  eC = wave_reduce63<true>(eC);
  mxFn = wave_reduce63<false>(fmax(mxFn, 0.0));
  if (tid == 63) { gl(I.cpart)[it.blk] = eC; gl(I.cpart)[I.ncblk + it.blk] = mxFn; }   // [value][block]
}

// ================================================================================================= k_strand
template <typename Real, int TILE>
struct StrandLDS {
  // structure-of-arrays: consecutive threads touch consecutive 8-byte words (no bank conflicts)
  union {
    struct {
      double Xx[TILE + 2], Xy[TILE + 2], Xz[TILE + 2];                   // nodes k0-1 .. k1
      double Qw[TILE + 2], Qx[TILE + 2], Qy[TILE + 2], Qz[TILE + 2];     // segments k0-1 .. k1
    };
    double part[SPART][TILE];                                            // the lanes' partials, once the state is dead
  };
  int32_t rpre[TILE], rpost[TILE];                                   // gathered crossing sides: segments
  SideRec<Real> rec[TILE];                                           // ... and their records
  int32_t glo[TILE], ghi[TILE];                                      // per slot: first / last gathered record it needs
  int wcnt[TILE / 64];
  Real bnd[TILE / 64][9];            // last lane of each wavefront: segment force, second-segment torque, kDyn
  double rbw[RBP][TILE / 64];        // rigid-body sums per wavefront (rigid instances)
  double gmx[TILE / 64];             // largest ground Fn per wavefront (ground on, tiles of several wavefronts)
};

// the value of lane − 1 (DPP wave_shr:1; lane 0 reads 0)
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double from_left(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// f64 frame [X | Y | Z] of a quaternion with one reciprocal: R = I + 2/|q|² (…) (quat.js toFrame up to rounding)
__device__ __forceinline__ void qframe_d(double w, double x, double y, double z, double* R) {
  double n2 = w * w + x * x + y * y + z * z;
  double s = n2 > 0 ? 2.0 * rcp_(n2) : 2.0;
  R[0] = 1 - s * (y * y + z * z); R[1] = s * (x * y + z * w); R[2] = s * (x * z - y * w);
  R[3] = s * (x * y - z * w); R[4] = 1 - s * (x * x + z * z); R[5] = s * (y * z + x * w);
  R[6] = s * (x * z + y * w); R[7] = s * (y * z - x * w); R[8] = 1 - s * (x * x + y * y);
}

// Ψ and θ of the relative rotation d = ±(q_a* ⊗ q_b) (relRotation, rod_cosserat.js:33-45).  The product is formed in
// f64 (its vector part carries the full precision of a small relative rotation); θ = 2 atan2(|v|, w) in Real.
template <typename Real>
__device__ __forceinline__ Real relrot_R(double aw, double ax, double ay, double az, double bw, double bx, double by, double bz,
                                        Real* psi, double* d) {
  ax = -ax; ay = -ay; az = -az;
  double w = aw * bw - ax * bx - ay * by - az * bz;
  double x = aw * bx + ax * bw + ay * bz - az * by;
  double y = aw * by - ax * bz + ay * bw + az * bx;
  double z = aw * bz + ax * by - ay * bx + az * bw;
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  d[0] = w; d[1] = x; d[2] = y; d[3] = z;
  if constexpr (std::is_same<Real, double>::value) {
    double s = sqrt(x * x + y * y + z * z), th = 2 * atan2(s, w);
    double f = s > 1e-12 ? th / s : 2 / fmax(w, 1e-12);
    psi[0] = f * x; psi[1] = f * y; psi[2] = f * z;
    return th;
  } else {
    float xf = float(x), yf = float(y), zf = float(z), wf = float(w);
    const float s2 = xf * xf + yf * yf + zf * zf, s = __builtin_amdgcn_sqrtf(s2);   // hardware square root (1 ulp)
    const float iw = rcp_(fmaxf(wf, 1e-12f));
    const float r2 = s2 * iw * iw;                 // (|v| / w)²: w ≥ 0 here
    float f, th;
    if (r2 < 0.01f) {
      // small relative rotation (θ < 0.2): Ψ = (2/w)·atan(r)/r · v with the series of atan(r)/r (error < r⁸/9)
      f = 2.0f * iw * (1.0f + r2 * (-1.0f / 3 + r2 * (1.0f / 5 + r2 * (-1.0f / 7))));
      th = f * s;
    } else {
      th = 2.0f * atan2f(s, wf);
      f = th * rcp_(s);
    }
    psi[0] = f * xf; psi[1] = f * yf; psi[2] = f * zf;
    return th;
  }
}

// R(q)ᵀ v: the components of v along the axes X | Y | Z of the frame of q/|q| (qframe's columns).  With s = 2/|q|²
// the frame is I + s(w [q]× + [q]×²), so Rᵀ v = v + s (q × (q × v) − w q × v): two cross products instead of the nine
// matrix entries
template <typename Real>
__device__ __forceinline__ void rotT_q(Real w, Real x, Real y, Real z, const Real* v, Real* out) {
  const Real n2 = w * w + x * x + y * y + z * z;
  const Real s = n2 > Real(0) ? Real(2) * rcp_(n2) : Real(2);
  const Real c0 = y * v[2] - z * v[1], c1 = z * v[0] - x * v[2], c2 = x * v[1] - y * v[0];
  const Real e0 = y * c2 - z * c1, e1 = z * c0 - x * c2, e2 = x * c1 - y * c0;
  out[0] = v[0] + s * (e0 - w * c0);
  out[1] = v[1] + s * (e1 - w * c1);
  out[2] = v[2] + s * (e2 - w * c2);
}
// R(d)ᵀ v of a relative rotation d (unit up to rounding)
template <typename Real>
__device__ __forceinline__ void rotT2(const double* d, const Real* v, Real* out) {
  rotT_q(Real(d[0]), Real(d[1]), Real(d[2]), Real(d[3]), v, out);
}

// one segment's stretch / shear (CosseratRod.apply, rod_cosserat.js:118-133): world force sf; body torque
// (Rᵀe) × (Rᵀn) and the section forces returned for an owned segment.  σ is formed in f64 (it cancels in f32).
template <typename Real, int TILE>
__device__ __forceinline__ void seg_force(StrandLDS<Real, TILE>& L, int i, double l0, Real EA, Real GA, Real* sf, Real* st,
                                          Real* nb, double* eS, double* eSh) {
  double R[9];
  qframe_d(L.Qw[i], L.Qx[i], L.Qy[i], L.Qz[i], R);
  double ex = L.Xx[i + 1] - L.Xx[i], ey = L.Xy[i + 1] - L.Xy[i], ez = L.Xz[i + 1] - L.Xz[i];
  double b0 = R[0] * ex + R[1] * ey + R[2] * ez, b1 = R[3] * ex + R[4] * ey + R[5] * ez, b2 = R[6] * ex + R[7] * ey + R[8] * ez;
  double il = rcp_(l0);
  double s1 = b0 * il, s2 = b1 * il, s3 = b2 * il - 1.0;
  Real n1 = GA * Real(s1), n2 = GA * Real(s2), n3 = EA * Real(s3);
  sf[0] = n1 * Real(R[0]) + n2 * Real(R[3]) + n3 * Real(R[6]);
  sf[1] = n1 * Real(R[1]) + n2 * Real(R[4]) + n3 * Real(R[7]);
  sf[2] = n1 * Real(R[2]) + n2 * Real(R[5]) + n3 * Real(R[8]);
  Real e0 = Real(b0), e1 = Real(b1), e2 = Real(b2);
  st[0] = e1 * n3 - e2 * n2; st[1] = e2 * n1 - e0 * n3; st[2] = e0 * n2 - e1 * n1;
  nb[0] = n1; nb[1] = n2; nb[2] = n3;
  *eS = 0.5 * l0 * double(n3) * s3; *eSh = 0.5 * l0 * (double(n1) * s1 + double(n2) * s2);
}

// one joint (rod_cosserat.js:136-159) between LDS segments i and i+1: J on the first (ja), −R(Δq)ᵀ J on the second
// (jb), the bending-coupling inertia kDyn (kd)
template <typename Real, int TILE>
__device__ __forceinline__ void joint_torque(StrandLDS<Real, TILE>& L, int i, Real lb, const Real* om, Real Bx, Real By, Real Bz,
                                             Real* mm, Real* d, Real* ja, Real* jb, Real* kd) {
  Real ps[3]; double dq[4];
  Real th = relrot_R<Real>(L.Qw[i], L.Qx[i], L.Qy[i], L.Qz[i], L.Qw[i + 1], L.Qx[i + 1], L.Qy[i + 1], L.Qz[i + 1], ps, dq);
  Real il = rcp_(lb);
  d[0] = ps[0] * il - om[0]; d[1] = ps[1] * il - om[1]; d[2] = ps[2] * il - om[2];
  mm[0] = Bx * d[0]; mm[1] = By * d[1]; mm[2] = Bz * d[2];
  Real q0 = ps[0] * ps[0], q1 = ps[1] * ps[1], q2 = ps[2] * ps[2], f = Real(0.25) * il;
  kd[0] = f * (By * q2 + Bz * q1); kd[1] = f * (Bx * q2 + Bz * q0); kd[2] = f * (Bx * q1 + By * q0);
  jlt(ps, th, mm, ja);
  Real Jb[3];
  rotT2(dq, ja, Jb);
  jb[0] = -Jb[0]; jb[1] = -Jb[1]; jb[2] = -Jb[2];
}

// One strand tile.  A tile owns cnt ≤ TILE − 1 consecutive nodes k0 .. k1−1 of one lath and the segments starting
// at them; lane t works on "slot" t: node k0−1+t, segment k0−1+t and joint k0−1+t.  Slot 0 is the halo (the segment
// and joint just before the tile, whose forces the first owned node and frame need): every lane computes one segment
// and one joint, and lanes 1 .. cnt integrate their node and frame.
template <typename Real, int TILE, bool RB>
__device__ void strand_control_tail(const Inst<Real>& I, Ctl* ctl, const TileItem& tl, StrandLDS<Real, TILE>& L);

// a value the control step reduces (a tile's partial, a rigid-body sum).  When the strand launch runs the control
// step (FUSED) it is an agent-scope store, written through the L2, so the tile that reduces it reads it from any XCD
// with agent-scope loads and no tile has to write back its L2; otherwise a plain store.
template <bool FUSED>
__device__ __forceinline__ void part_store(double* p, double x) {
  if (FUSED) __hip_atomic_store(gl(p), x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *gl(p) = x;
}

// FUSED: the control step runs in this launch, in the last tile of the instance to finish (strand_control_tail)
template <typename Real, int TILE, bool OUT, bool RB, bool FUSED = false>
__device__ void strand_tile(const Inst<Real>& I, const Ctl* ctl, const TileItem& tl, StrandLDS<Real, TILE>& L,
                            const OutPtrs& out) {
  const int tid = threadIdx.x;
  const CtlHead ch = ctl_head(ctl);
  int cur = ch.cur;
  const bool rbk = ch.rollback != 0;
  // the state the last iteration evaluated (downloads), or the current one (evaluations: out.F set)
  if (OUT) cur = (rbk || out.F) ? cur : (cur ^ 1);
  const int iter = ch.iter;
  // an iteration runs only on a live instance (not stopped, below its stop); tested after the first loads are issued
  const bool live = OUT || ((ch.done == 0) & (iter < ch.stop_at));
  const int noff = tl.noff, nn = tl.nn, soff = tl.soff, nseg = tl.nseg, joff = tl.joff, njs = tl.njs;
  const bool closed = tl.closed != 0;
  const int k0 = tl.k0, k1 = tl.k1, cnt = k1 - k0;
  auto wrap = [](int v, int n) { int r = v % n; return r < 0 ? r + n : r; };
  auto node_ok = [&](int k, int& kk) { if (closed) { kk = wrap(k, nn); return true; } kk = k; return k >= 0 && k < nn; };
  auto seg_ok = [&](int j, int& jj) { if (closed) { if (nseg <= 0) { jj = 0; return false; } jj = wrap(j, nseg); return true; } jj = j; return j >= 0 && j < nseg; };
  auto joint_ok = [&](int j, int& jj) { if (closed) { if (njs <= 0) { jj = 0; return false; } jj = wrap(j, njs); return true; } jj = j; return j >= 0 && j < njs; };

  // ---------------------------------------------------------------- issue every independent load first
  // the tile's cnt + 2 ≤ TILE + 1 slots of positions and frames: slot tid, and slot TILE for lane 0 of a full tile,
  // loaded into registers (written to LDS once the parameter and side-record loads below are out too)
  const auto Xc = gl(buf(I.X, cur)), Qc = gl(buf(I.Q, cur));
  double sx[2], sy[2], sz[2], sw[2], qx[2], qy[2], qz[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int h = tid + r * TILE;
    int kk, jj;
    const bool inh = h < cnt + 2;
    if (inh && node_ok(k0 - 1 + h, kk)) { const int i = 3 * (noff + kk); sx[r] = Xc[i]; sy[r] = Xc[i + 1]; sz[r] = Xc[i + 2]; }
    else { sx[r] = sy[r] = sz[r] = 0.0; }
    if (inh && seg_ok(k0 - 1 + h, jj)) { const int j = 4 * (soff + jj); sw[r] = Qc[j]; qx[r] = Qc[j + 1]; qy[r] = Qc[j + 2]; qz[r] = Qc[j + 3]; }
    else { sw[r] = 1.0; qx[r] = qy[r] = qz[r] = 0.0; }
  }
  const int t = tid;
  const int k = k0 - 1 + t;                        // node / segment / joint of this slot (strand-local)
  const bool inSlot = t <= cnt;
  const bool ownNode = t >= 1 && t <= cnt;
  int kkN = 0, jself = 0, jjo = 0;
  if (ownNode) node_ok(k, kkN);
  const bool hasSeg = inSlot && seg_ok(k, jself);
  const bool hasJoint = inSlot && joint_ok(k, jjo);
  const bool ownSeg = ownNode && hasSeg;
  const int i = noff + kkN, j = soff + jself;
  // node parameters and velocity
  uint8_t fx = 0; Real ms = 1, ldx = 0, ldy = 0, ldz = 0, llx = 0, lly = 0, llz = 0, vx = 0, vy = 0, vz = 0;
  if (ownNode) {
    fx = gl(I.fix)[i]; ms = gl(I.mass)[i];
    const int ax = I.load_axes;                    // the load components that are +0 everywhere are not read
    if (ax & 1) ldx = gl(I.lx)[i];
    if (ax & 2) ldy = gl(I.ly)[i];
    if (ax & 4) ldz = gl(I.lz)[i];
    if (I.npl) { llx = gl(I.llx)[i]; lly = gl(I.lly)[i]; llz = gl(I.llz)[i]; }
    if (!OUT && !rbk) { vx = gl(I.vx)[i]; vy = gl(I.vy)[i]; vz = gl(I.vz)[i]; }
  }
  // segment and joint parameters of the slot; a strand whose segments share their section reads it once per tile
  // (a uniform load) instead of per segment
  const bool secS = I.sec_strand != 0;
  Real sEA = 0, sGA = 0, sBx = 0, sBy = 0, sBz = 0;
  if (secS) {
    const auto sp = gl(I.sec) + 5 * tl.strand;
    sEA = sp[0]; sGA = sp[1]; sBx = sp[2]; sBy = sp[3]; sBz = sp[4];
  }
  double l0 = 1.0; Real EA = 0, GA = 0, Bx = 0, By = 0, Bz = 0, krx = 0, kry = 0, krz = 0, wx = 0, wy = 0, wz = 0;
  if (hasSeg) {
    l0 = gl(I.l0)[j];
    if (secS) { EA = sEA; GA = sGA; } else { EA = gl(I.EA)[j]; GA = gl(I.GA)[j]; }
    if (ownSeg) {
      krx = gl(I.krx)[j]; kry = gl(I.kry)[j]; krz = gl(I.krz)[j];
      if (!OUT && !rbk) { wx = gl(I.wx)[j]; wy = gl(I.wy)[j]; wz = gl(I.wz)[j]; }
    }
  }
  Real lb = 1, om[3] = {0, 0, 0};
  if (hasJoint) {
    const int jg = joff + jjo;
    lb = gl(I.jlen)[jg]; om[0] = gl(I.om0)[3 * jg]; om[1] = gl(I.om0)[3 * jg + 1]; om[2] = gl(I.om0)[3 * jg + 2];
    if (secS) { Bx = sBx; By = sBy; Bz = sBz; } else { Bx = gl(I.EIx)[j]; By = gl(I.EIy)[j]; Bz = gl(I.GJ)[j]; }
  }
  const int so = tl.so, ns = tl.ns;
  SideRec<Real> r0;
  if (tid < ns) side_load(gl(I.side) + so + tid, r0);
  if (!live) {                                     // uniform over the workgroup (one instance per tile)
    // no iteration, so no rigid step after it (the control kernel clears the flag where it runs)
    if (FUSED && RB && I.rigid && tl.tile_local == 0 && tid == 0) const_cast<Ctl*>(ctl)->rb_move = 0;
    return;
  }
  // the staged positions and frames go to LDS only now, after every load above is out (one memory latency for all)
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int h = tid + r * TILE;
    if (h < cnt + 2) {
      L.Xx[h] = sx[r]; L.Xy[h] = sy[r]; L.Xz[h] = sz[r];
      L.Qw[h] = sw[r]; L.Qx[h] = qx[r]; L.Qy[h] = qy[r]; L.Qz[h] = qz[r];
    }
  }
  __syncthreads();

  // ---------------------------------------------------------------- segments and joints
  double eS = 0, eSh = 0, eB = 0, eT = 0, mxN = 0, mxM = 0;
  Real ST[3] = {0, 0, 0}, SF[3] = {0, 0, 0}, JA[3] = {0, 0, 0}, JB[3] = {0, 0, 0}, KD[3] = {0, 0, 0};
  if (hasSeg) {
    Real nb[3]; double a1, a2;
    seg_force(L, t, l0, EA, GA, SF, ST, nb, &a1, &a2);
    if (ownSeg) {
      eS = a1; eSh = a2;
      mxN = fabs(double(nb[2]));
      if (OUT && out.N) { out.N[j] = double(nb[2]); out.V[2 * j] = double(nb[0]); out.V[2 * j + 1] = double(nb[1]); }
    }
  }
  if (hasJoint) {
    Real mm[3], d[3];
    joint_torque(L, t, lb, om, Bx, By, Bz, mm, d, JA, JB, KD);
    if (ownNode) {
      eB = 0.5 * double(lb) * (double(mm[0]) * double(d[0]) + double(mm[1]) * double(d[1]));
      eT = 0.5 * double(lb) * double(mm[2]) * double(d[2]);
      mxM = fmax(fabs(double(mm[0])), fmax(fabs(double(mm[1])), fabs(double(mm[2]))));
      if (OUT && out.M) { const int jg = joff + jjo; out.M[3 * jg] = double(mm[0]); out.M[3 * jg + 1] = double(mm[1]); out.M[3 * jg + 2] = double(mm[2]); }
    }
  }
  // what lane t needs from lane t − 1: the force of segment t − 1, joint t − 1's torque on segment t and its kDyn —
  // a one-lane shift inside the wavefront, through LDS across a wavefront boundary
  Real pSF[3], pJB[3], pKD[3];
#pragma unroll
  for (int c = 0; c < 3; c++) { pSF[c] = from_left(SF[c]); pJB[c] = from_left(JB[c]); pKD[c] = from_left(KD[c]); }
  if (TILE > 64) {
    if ((tid & 63) == 63) {
#pragma unroll
      for (int c = 0; c < 3; c++) { L.bnd[tid >> 6][c] = SF[c]; L.bnd[tid >> 6][3 + c] = JB[c]; L.bnd[tid >> 6][6 + c] = KD[c]; }
    }
    __syncthreads();
    if ((tid & 63) == 0 && tid > 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) { pSF[c] = L.bnd[(tid >> 6) - 1][c]; pJB[c] = L.bnd[(tid >> 6) - 1][3 + c]; pKD[c] = L.bnd[(tid >> 6) - 1][6 + c]; }
    }
  }

  // ---------------------------------------------------------------- crossing sides of this strand touching the tile,
  // summed in the strand's fixed side order (independent of the tiling).  Each gathered record registers the range
  // of gathered indices per owned slot (integer min / max, order-free), so a lane visits only the window that can
  // hold its records, in increasing order: the same sums as a scan over all of them.
  Real Fc[3] = {0, 0, 0}, Tc[3] = {0, 0, 0}, kTw = 0;
  const int lane = tid & 63, wv = tid >> 6;
  const int nz = nseg > 0 ? nseg : 1;
  const int km1 = closed ? wrap(k - 1, nz) : k - 1;
  auto slot_of = [&](int v) { return closed ? wrap(v - (k0 - 1), nz) : v - (k0 - 1); };
  for (int base = 0; base < ns; base += TILE) {
    const int q = base + tid;
    bool rel = false;
    SideRec<Real> r = r0;
    if (q < ns) {
      if (base != 0) side_load(gl(I.side) + so + q, r);
      if (closed) {
        int dp = wrap(r.seg_pre - (k0 - 1), nz), dq = wrap(r.seg_post - k0, nz);
        rel = dp < cnt + 1 || dq < cnt;
      } else {
        rel = (r.seg_pre >= k0 - 1 && r.seg_pre < k1) || (r.seg_post >= k0 && r.seg_post < k1);
      }
    }
    unsigned long long mask = __ballot(rel);
    int pos = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) L.wcnt[wv] = __popcll(mask);
    L.glo[t] = 0x7fffffff; L.ghi[t] = -1;
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < TILE / 64; w++) { if (w < wv) off += L.wcnt[w]; tot += L.wcnt[w]; }
    if (rel) {
      const int g = off + pos;
      side_store(r, L.rec + g);
      L.rpre[g] = r.seg_pre; L.rpost[g] = r.seg_post;
      const int sa = slot_of(r.seg_pre), sb = slot_of(closed ? wrap(r.seg_pre + 1, nz) : r.seg_pre + 1), sc = slot_of(r.seg_post);
      if (sa >= 1 && sa <= cnt) { atomicMin(&L.glo[sa], g); atomicMax(&L.ghi[sa], g); }
      if (sb >= 1 && sb <= cnt) { atomicMin(&L.glo[sb], g); atomicMax(&L.ghi[sb], g); }
      if (sc >= 1 && sc <= cnt) { atomicMin(&L.glo[sc], g); atomicMax(&L.ghi[sc], g); }
    }
    __syncthreads();
    if (ownNode) {
      const bool hasStart = closed || k < nseg, hasEnd = closed || k >= 1;
      const int glo = L.glo[t], ghi = L.ghi[t];
      for (int g = glo; g <= ghi; g++) {
        const int sp = L.rpre[g], sq = L.rpost[g];
        const bool asA = hasStart && sp == k, asB = hasEnd && sp == km1;
        if (asA | asB) {
          const Real tt = L.rec[g].t;
          const Real w = asA ? Real(1) - tt : tt;
          Fc[0] += w * L.rec[g].F[0]; Fc[1] += w * L.rec[g].F[1]; Fc[2] += w * L.rec[g].F[2];
          if (asA && asB) { Fc[0] += tt * L.rec[g].F[0]; Fc[1] += tt * L.rec[g].F[1]; Fc[2] += tt * L.rec[g].F[2]; }
        }
        if (ownSeg) {
          if (sp == k) { Tc[0] += L.rec[g].T[0]; Tc[1] += L.rec[g].T[1]; Tc[2] += L.rec[g].T[2]; }
          if (sq == k) kTw += L.rec[g].kth;
        }
      }
    }
    __syncthreads();
  }

  // ---------------------------------------------------------------- nodes: loads, reactions, anchors, v += F/m, x += v
  double keN = 0, keR = 0, mxF = 0, mxT = 0, mxL = 0, eG = 0, mxG = 0;
  double rb[RBP];                                  // rigid-body sums of the lane's node (rigid instances)
#pragma unroll
  for (int q = 0; q < RBP; q++) rb[q] = 0.0;
  const bool warm = iter < I.warmup && I.contacts;
  const Real damping = Real(I.damping);
  if (ownNode) {
    const bool hasEnd = closed || k >= 1, hasStart = closed || k < nseg;
    Real F[3];
    if (hasEnd && hasStart) { F[0] = SF[0] - pSF[0]; F[1] = SF[1] - pSF[1]; F[2] = SF[2] - pSF[2]; }
    else if (hasStart) { F[0] = SF[0]; F[1] = SF[1]; F[2] = SF[2]; }
    else if (hasEnd) { F[0] = -pSF[0]; F[1] = -pSF[1]; F[2] = -pSF[2]; }
    else { F[0] = F[1] = F[2] = Real(0); }
    Real ld[3] = {ldx, ldy, ldz};
    const bool localLoad = I.npl && (llx != Real(0) || lly != Real(0) || llz != Real(0));
    double Rn[9];
    if (localLoad || I.floor_on) {
      // mean node frame from the adjacent segments (nodeQuats: the starting segment is the reference)
      const int is = hasStart ? t : t - 1;
      double acc[4] = {0, 0, 0, 0};
      if (hasStart) { acc[0] += L.Qw[t]; acc[1] += L.Qx[t]; acc[2] += L.Qy[t]; acc[3] += L.Qz[t]; }
      if (hasEnd) {
        double dd = L.Qw[t - 1] * L.Qw[is] + L.Qx[t - 1] * L.Qx[is] + L.Qy[t - 1] * L.Qy[is] + L.Qz[t - 1] * L.Qz[is];
        double sg = dd < 0 ? -1.0 : 1.0;
        acc[0] += sg * L.Qw[t - 1]; acc[1] += sg * L.Qx[t - 1]; acc[2] += sg * L.Qy[t - 1]; acc[3] += sg * L.Qz[t - 1];
      }
      qframe_d(acc[0], acc[1], acc[2], acc[3], Rn);
    }
    if (localLoad) for (int c = 0; c < 3; c++) ld[c] += Real(Rn[c]) * llx + Real(Rn[3 + c]) * lly + Real(Rn[6 + c]) * llz;
    if (I.proto == 1 && i == I.probe) for (int c = 0; c < 3; c++) ld[c] += Real(ctl->pf[c]);   // the replay's probe force
    mxL = fmax(fabs(double(ld[0])), fmax(fabs(double(ld[1])), fabs(double(ld[2]))));
    for (int c = 0; c < 3; c++) F[c] += ld[c] + Fc[c];
    if (I.floor_on) {
      // the ground (Floor.apply, solver_cpu.js:583-627): penalty against the plane under the section's lowest
      // point, Coulomb friction against a stick point that follows a sliding node.  In f64 (positions, stick
      // points); the output pass recomputes the force from the moved stick point, which gives the same force.
      // a download evaluates the state the last iteration evaluated, at that iteration's ramp (Floor.rampAt(iter − 1),
      // solver_gpu.js:462)
      const int rampN = max(I.ramp_iter, 0);
      const int itg = (OUT && !out.F) ? max(iter - 1, 0) : iter;
      const double rmp = rampN <= 0 ? 1.0 : fmin(1.0, 0.01 + 0.99 * itg / rampN);
      const double kn = rmp * I.strength, kt = rmp * I.stick;
      const double pen = (I.floor_z - L.Xz[t]) + (0.5 * double(gl(I.nodew)[i]) * fabs(Rn[2]) + 0.5 * double(gl(I.nodeh)[i]) * fabs(Rn[5]));
      double fg[3] = {0, 0, 0};
      if (RB && I.rigid && !OUT) {
        // the resultant of loads + ground and its moment about the centroid; Σ d² and count of the nodes on the
        // ground (the rigid masses)
        const double dx = L.Xx[t] - ctl->rc[0], dy = L.Xy[t] - ctl->rc[1], dz = L.Xz[t] - ctl->rc[2];
        if (pen > 0) { rb[6] = dx * dx; rb[7] = dy * dy; rb[8] = dz * dz; rb[9] = 1.0; }
        rb[3] = dx; rb[4] = dy; rb[5] = dz;          // lever, until the force is known
      }
      if (pen > 0) {
        const double fn = kn * pen;
        double ax = gl(I.fax)[i], ay = gl(I.fay)[i];
        const double sx = L.Xx[t] - ax, sy = L.Xy[t] - ay;
        const double Ftv = kt * sqrt(sx * sx + sy * sy), limit = I.floor_mu * fn;
        const double scale = Ftv > limit ? 1.0 - limit / fmax(Ftv, 1e-12) : 0.0;
        const double dax = sx * scale, day = sy * scale;
        if (!OUT && scale > 0) { gl(I.fax)[i] = ax + dax; gl(I.fay)[i] = ay + day; }
        if (!OUT) gl(I.gstate)[i] = scale > 0 ? 2 : 1;
        const double ex = sx - dax, ey = sy - day;
        fg[0] = -kt * ex; fg[1] = -kt * ey; fg[2] = fn;
        eG = 0.5 * kn * pen * pen + 0.5 * kt * (ex * ex + ey * ey);
        mxG = fn;
      } else if (!OUT) {
        gl(I.fax)[i] = L.Xx[t]; gl(I.fay)[i] = L.Xy[t];
        gl(I.gstate)[i] = 0;
      }
      if (OUT && out.G) {
        // the stick point has moved with a slide already, so the state is the one the iteration recorded
        out.G[3 * i] = fg[0]; out.G[3 * i + 1] = fg[1]; out.G[3 * i + 2] = fg[2];
        out.GS[i] = pen > 0 ? gl(I.gstate)[i] : 0;
      }
      for (int c = 0; c < 3; c++) F[c] += Real(fg[c]);
      if (RB && I.rigid && !OUT) {
        const double gx = double(ld[0]) + fg[0], gy = double(ld[1]) + fg[1], gz = double(ld[2]) + fg[2];
        const double dx = rb[3], dy = rb[4], dz = rb[5];
        rb[0] = gx; rb[1] = gy; rb[2] = gz;
        rb[3] = dy * gz - dz * gy; rb[4] = dz * gx - dx * gz; rb[5] = dx * gy - dy * gx;
      }
    }
    if (OUT) {
      if (out.R) {
        if (fx & 8) { out.R[3 * i] = -double(F[0]); out.R[3 * i + 1] = -double(F[1]); out.R[3 * i + 2] = -double(F[2]); }
        else { out.R[3 * i] = out.R[3 * i + 1] = out.R[3 * i + 2] = 0.0; }
      }
      if (out.F) { out.F[3 * i] = double(F[0]); out.F[3 * i + 1] = double(F[1]); out.F[3 * i + 2] = double(F[2]); }
    } else if (I.proto == 2 && i == I.probe) {
      gl(I.preact)[0] = -double(F[0]); gl(I.preact)[1] = -double(F[1]); gl(I.preact)[2] = -double(F[2]);
    }
    if (fx & 1) F[0] = Real(0);
    if (fx & 2) F[1] = Real(0);
    if (fx & 4) F[2] = Real(0);
    mxF = double(F[0]) * double(F[0]) + double(F[1]) * double(F[1]) + double(F[2]) * double(F[2]);   // squared: √ after the max
    if (!OUT) {
      Real V[3] = {vx, vy, vz};
      if (warm) { V[0] = V[1] = V[2] = Real(0); F[0] = F[1] = F[2] = Real(0); }
      const Real im = rcp_(ms);
      for (int c = 0; c < 3; c++) V[c] += F[c] * im;
      if (damping != Real(1)) for (int c = 0; c < 3; c++) V[c] *= damping;
      if (fx & 1) V[0] = Real(0);
      if (fx & 2) V[1] = Real(0);
      if (fx & 4) V[2] = Real(0);
      const int nx = cur ^ 1;
      const double xn = L.Xx[t] + double(V[0]), yn = L.Xy[t] + double(V[1]), zn = L.Xz[t] + double(V[2]);
      const auto Xn = gl(buf(I.X, nx));
      // the new state is not read again in this launch: streaming stores (measured: the 10 m dome's strand launch
      // 13.1 → 12.1 µs, the 256 variants' 167.2 → 166.2 µs)
      __builtin_nontemporal_store(xn, &Xn[3 * i]);
      __builtin_nontemporal_store(yn, &Xn[3 * i + 1]);
      __builtin_nontemporal_store(zn, &Xn[3 * i + 2]);
      if (RB) { rb[10] = xn; rb[11] = yn; rb[12] = zn; }
      __builtin_nontemporal_store(V[0], &gl(I.vx)[i]); __builtin_nontemporal_store(V[1], &gl(I.vy)[i]);
      __builtin_nontemporal_store(V[2], &gl(I.vz)[i]);
      keN = double(ms) * (double(V[0]) * double(V[0]) + double(V[1]) * double(V[1]) + double(V[2]) * double(V[2]));
    }
  }
  // ---------------------------------------------------------------- frames: body torque, inertia, ω, q ← q ⊗ exp(ω)
  if (ownSeg) {
    Real tm[3] = {ST[0] + JA[0] + pJB[0], ST[1] + JA[1] + pJB[1], ST[2] + JA[2] + pJB[2]};
    if (I.C > 0) {
      for (int e = gl(I.clamp_off)[j]; e < gl(I.clamp_off)[j + 1]; e++) {
        const int cc = gl(I.clamp_idx)[e];
        Real ps[3]; double dq[4];
        Real th = relrot_R<Real>(gl(I.c_Q)[4 * cc], gl(I.c_Q)[4 * cc + 1], gl(I.c_Q)[4 * cc + 2], gl(I.c_Q)[4 * cc + 3],
                                 L.Qw[t], L.Qx[t], L.Qy[t], L.Qz[t], ps, dq);
        Real lc = gl(I.c_len)[cc];
        Real d[3] = {ps[0] / lc, ps[1] / lc, ps[2] / lc};
        Real mm[3] = {gl(I.c_B)[3 * cc] * d[0], gl(I.c_B)[3 * cc + 1] * d[1], gl(I.c_B)[3 * cc + 2] * d[2]};
        Real J[3], Jb[3];
        jlt(ps, th, mm, J);
        rotT2(dq, J, Jb);
        for (int c = 0; c < 3; c++) tm[c] -= Jb[c];
        eB += 0.5 * double(lc) * (double(mm[0]) * double(d[0]) + double(mm[1]) * double(d[1]));
        eT += 0.5 * double(lc) * double(mm[2]) * double(d[2]);
        if (OUT && out.CM) { out.CM[2 * cc] = double(mm[0]); out.CM[2 * cc + 1] = double(mm[1]); }
      }
    }
    if (Tc[0] != Real(0) || Tc[1] != Real(0) || Tc[2] != Real(0)) {
      // the crossings' world torques on the segment's body axes (in Real: the torque is Real)
      Real tb[3];
      rotT_q(Real(L.Qw[t]), Real(L.Qx[t]), Real(L.Qy[t]), Real(L.Qz[t]), Tc, tb);
      tm[0] += tb[0]; tm[1] += tb[1]; tm[2] += tb[2];
    }
    mxT = double(tm[0]) * double(tm[0]) + double(tm[1]) * double(tm[1]) + double(tm[2]) * double(tm[2]);   // squared
    if (OUT && out.T) {
      // the generalised forces of an evaluation: body-axis torques, and the frame inertia of this state
      out.T[3 * j] = double(tm[0]); out.T[3 * j + 1] = double(tm[1]); out.T[3 * j + 2] = double(tm[2]);
      const double msc = I.mass_scale;
      out.I3[3 * j] = msc * fmax(double(krx + kTw + (pKD[0] + KD[0])), 1e-12);
      out.I3[3 * j + 1] = msc * fmax(double(kry + kTw + (pKD[1] + KD[1])), 1e-12);
      out.I3[3 * j + 2] = msc * fmax(double(krz + kTw + (pKD[2] + KD[2])), 1e-12);
    }
    if (!OUT) {
      const Real msc = Real(I.mass_scale);
      Real I3[3] = {msc * dmax(krx + kTw + (pKD[0] + KD[0]), Real(1e-12)),
                    msc * dmax(kry + kTw + (pKD[1] + KD[1]), Real(1e-12)),
                    msc * dmax(krz + kTw + (pKD[2] + KD[2]), Real(1e-12))};
      Real W[3] = {wx + tm[0] * rcp_(I3[0]), wy + tm[1] * rcp_(I3[1]), wz + tm[2] * rcp_(I3[2])};
      W[2] = clampv(W[2], Real(-TWIST_RATE_MAX), Real(TWIST_RATE_MAX));
      if (kTw > Real(0)) W[2] *= Real(ROLL_DAMP);
      if (damping != Real(1)) for (int c = 0; c < 3; c++) W[c] *= damping;
      __builtin_nontemporal_store(W[0], &gl(I.wx)[j]); __builtin_nontemporal_store(W[1], &gl(I.wy)[j]);
      __builtin_nontemporal_store(W[2], &gl(I.wz)[j]);
      // q ⊗ exp(ω): the increment exp(ω) − 1 = (−2 sin²(a/4), ω sin(a/2)/a) carries the precision of the small
      // rotation itself (the delta form of wgsl.js:962-980); the product and the renormalisation run in f64
      double e0, e1, e2, e3;
      if constexpr (std::is_same<Real, double>::value) {
        double a = sqrt(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
        double sh = sin(0.25 * a), sc = a > 1e-12 ? sin(0.5 * a) / a : 0.5;
        e0 = -2.0 * sh * sh; e1 = W[0] * sc; e2 = W[1] * sc; e3 = W[2] * sc;
      } else {
        const float a2 = W[0] * W[0] + W[1] * W[1] + W[2] * W[2];
        float c0, sc;
        if (a2 < 0.0625f) {
          // a < 0.25 rad per iteration: cos(a/2) − 1 and sin(a/2)/a by their series (truncation below FP32 rounding)
          c0 = a2 * (-1.0f / 8 + a2 * (1.0f / 384 + a2 * (-1.0f / 46080)));
          sc = 0.5f + a2 * (-1.0f / 48 + a2 * (1.0f / 3840));
        } else {
          const float a = sqrtf(a2), sh = sinf(0.25f * a);
          c0 = -2.0f * sh * sh;
          sc = sinf(0.5f * a) * rcp_(a);
        }
        e0 = double(c0); e1 = double(W[0] * sc); e2 = double(W[1] * sc); e3 = double(W[2] * sc);
      }
      const double qw = L.Qw[t], qx = L.Qx[t], qy = L.Qy[t], qz = L.Qz[t];
      double rw = qw + (qw * e0 - qx * e1 - qy * e2 - qz * e3);
      double rx = qx + (qw * e1 + qx * e0 + qy * e3 - qz * e2);
      double ry = qy + (qw * e2 - qx * e3 + qy * e0 + qz * e1);
      double rz = qz + (qw * e3 + qx * e2 - qy * e1 + qz * e0);
      double n2 = rw * rw + rx * rx + ry * ry + rz * rz;
      double inv = n2 > 0 ? rsqrt_d(n2) : 1.0;
      const int nx = cur ^ 1;
      const auto Qn = gl(buf(I.Q, nx));
      __builtin_nontemporal_store(rw * inv, &Qn[4 * j]); __builtin_nontemporal_store(rx * inv, &Qn[4 * j + 1]);
      __builtin_nontemporal_store(ry * inv, &Qn[4 * j + 2]); __builtin_nontemporal_store(rz * inv, &Qn[4 * j + 3]);
      keR = double(I3[0]) * double(W[0]) * double(W[0]) + double(I3[1]) * double(W[1]) * double(W[1]) + double(I3[2]) * double(W[2]) * double(W[2]);
    }
  }
  if (OUT) return;

  // ---------------------------------------------------------------- partial of the tile (fixed order; reduced by k_ctrl):
  // sums Σm|v|² + ΣI3ω², stretch, shear, bending, torsion energies | maxima |F_free|², |t_m|², and the node /
  // segment / joint / ground terms of F_ref: max(|load|, |N|, |M| / l̄_min, ground Fn) (forceRef,
  // solver_cpu.js:843-853) | the ground's energy (written only when the ground is on)
  // Transposed through LDS: every lane leaves its values; then lane l of the first wavefront takes value l / 8
  // over lanes l % 8, l % 8 + 8, … in order, and the eight lanes of each value combine by a butterfly within the
  // 8-lane group (xor 1, xor 2 within quads, then the half-row mirror).  The ninth value (the ground's energy) is
  // reduced the same way by lanes 0–7 when the ground is on.  Fixed order, no atomics.
  const double fr = fmax(fmax(mxL, mxG), fmax(mxN, mxM * I.inv_jlen_min));
  const double v[SPART] = {keN + keR, eS, eSh, eB, eT, mxF, mxT, fr, eG};
  __syncthreads();                                 // the staged state is dead: its space takes the partials
#pragma unroll
  for (int q = 0; q < SPART; q++) L.part[q][t] = v[q];
  __syncthreads();
  if (tid < 64) {
    static_assert(SPART == 9, "eight values in 8-lane groups, the ninth in a second pass");
    const int k = lane & 7;
    auto group = [&](int q, bool sum) {
      double x = L.part[q][k];
      for (int i = k + 8; i < TILE; i += 8) { const double y = L.part[q][i]; x = sum ? x + y : fmax(x, y); }
      double y = dpp_<0xB1, 0xF>(x); x = sum ? x + y : fmax(x, y);   // quad_perm [1,0,3,2]
      y = dpp_<0x4E, 0xF>(x); x = sum ? x + y : fmax(x, y);          // quad_perm [2,3,0,1]
      y = dpp_<0x141, 0xF>(x); x = sum ? x + y : fmax(x, y);         // row_half_mirror
      return x;
    };
    const int q = lane >> 3;
    const double x = group(q, q < 5);
    if (k == 0) part_store<FUSED>(I.spart + size_t(tl.tile_local) * SPART_STRIDE + q, x);   // [tile][value]: vector loads in k_ctrl
    if (I.floor_on) {
      const double g = group(8, true);
      if (lane == 0) part_store<FUSED>(I.spart + size_t(tl.tile_local) * SPART_STRIDE + 8, g);
    }
  }
  if (I.floor_on) {
    // the tile's largest ground Fn (an exact maximum, so any order): each wavefront's, then the wavefronts'
    const double g = wave_reduce63<false>(mxG);
    if (TILE == 64) {
      if (lane == 63) part_store<FUSED>(I.spart + size_t(tl.tile_local) * SPART_STRIDE + 9, g);
    } else {
      if (lane == 63) L.gmx[wv] = g;
      __syncthreads();
      if (tid == 0) {
        double x = L.gmx[0];
        for (int w = 1; w < TILE / 64; w++) x = fmax(x, L.gmx[w]);
        part_store<FUSED>(I.spart + size_t(tl.tile_local) * SPART_STRIDE + 9, x);
      }
    }
  }
  if (RB && I.rigid) {
    // the tile's rigid-body sums: each wavefront's in a fixed DPP order, then the wavefronts in order
#pragma unroll
    for (int q = 0; q < RBP; q++) {
      const double x = wave_reduce63<true>(rb[q]);
      if (lane == 63) L.rbw[q][wv] = x;
    }
    __syncthreads();
    if (tid < RBP) {
      double x = L.rbw[tid][0];
      for (int w = 1; w < TILE / 64; w++) x += L.rbw[tid][w];
      part_store<FUSED>(I.rbpart + size_t(tid) * I.ntiles + tl.tile_local, x);
    }
  }
  if (FUSED) strand_control_tail<Real, TILE, RB>(I, const_cast<Ctl*>(ctl), tl, L);
}

// fixed-order block reduction of NV doubles (the first nsum summed, the rest max-reduced over non-negative values);
// the result in every thread
template <int TILE, int NV>
__device__ __forceinline__ void block_reduce(double (*red)[TILE / 64], double* v, int nsum) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double x = k < nsum ? wave_reduce63<true>(v[k]) : wave_reduce63<false>(v[k]);
    if (lane == 63) red[k][wv] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double x = red[k][0];
    for (int w = 1; w < TILE / 64; w++) x = k < nsum ? x + red[k][w] : fmax(x, red[k][w]);
    v[k] = x;
  }
  __syncthreads();
}

// The control step of one instance after its reductions (lane 0; solver_cpu.js:809-837; wgsl.js:990-1047): acc
// holds [0] Σm|v|² + ΣI3ω², [1..4] stretch / shear / bending / torsion energy, [5] ground energy, [6] contact energy
// (sums), [7] max |F_free|², [8] max |t_m|², [9] max(|load|, |N|, |M|/l̄_min, ground Fn), [10] max Fn, [11] max
// ground Fn (maxima); rbs the
// instance's rigid-body sums.  c is the control record as the iteration found it; the new record is written to ctl.
template <typename Real>
__device__ void control_step(const Inst<Real>& I, Ctl* ctl, Ctl& c, const double* acc, const double* rbs) {
  const double eCont = acc[6], eGround = acc[5];
  double ke = 0.5 * acc[0];
  const int iterPre = c.iter;
  // rigid-body step (CpuSolver.step, solver_cpu.js:784-790; rigidMasses :565-572): velocities from this iteration's
  // resultant, v ← clip((v + Q/M)·damping), zero while warming up; their kinetic energy counts for the roll-back
  double rv[3] = {0, 0, 0}, rw[3] = {0, 0, 0};
  if (I.rigid) {
    const int rampN = max(I.ramp_iter, 0);
    const double rmp = rampN <= 0 ? 1.0 : fmin(1.0, 0.01 + 0.99 * iterPre / rampN);
    const double kn = rmp * I.strength, kt = rmp * I.stick, f = I.mass_scale * RB_MASS;
    const double nc = fmax(rbs[9], 1.0), rmin = (RB_LEVER * I.rb_R) * (RB_LEVER * I.rb_R) * kn;
    double Mr[6];
    Mr[0] = fmax(f * nc * kt, 1e-30); Mr[1] = Mr[0]; Mr[2] = fmax(f * nc * kn, 1e-30);
    Mr[3] = fmax(f * fmax(kn * rbs[7] + kt * rbs[8], rmin), 1e-30);
    Mr[4] = fmax(f * fmax(kn * rbs[6] + kt * rbs[8], rmin), 1e-30);
    Mr[5] = fmax(f * fmax(kt * (rbs[6] + rbs[7]), rmin), 1e-30);
    if (!(iterPre < I.warmup && I.contacts)) {
      for (int k = 0; k < 3; k++) {
        rv[k] = fmin(fmax((c.rv[k] + rbs[k] / Mr[k]) * I.damping, -I.rb_vmax), I.rb_vmax);
        rw[k] = fmin(fmax((c.rw[k] + rbs[3 + k] / Mr[3 + k]) * I.damping, -I.rb_wmax), I.rb_wmax);
      }
    }
    for (int k = 0; k < 3; k++) ke += 0.5 * (Mr[k] * rv[k] * rv[k] + Mr[3 + k] * rw[k] * rw[k]);
  }
  int rollback = 0;
  if (I.kin_damp && ke < c.ke_prev) {
    rollback = 1; ke = 0.0;
    for (int k = 0; k < 3; k++) { rv[k] = 0.0; rw[k] = 0.0; }
  }
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
  double fref = acc[9];
  if (I.contacts) fref = fmax(fref, acc[10]);
  fref = fmax(fref, 1e-12);
  // the strand partials carry max |F_free|² and max |t_m|²: the square root of the maximum is the maximum norm
  const double rf = I.n ? sqrt(acc[7]) / fref : 0.0, rt = I.m ? sqrt(acc[8]) / (fref * I.l_ref) : 0.0;
  double resid = fmax(rf, rt);
  if (I.rigid) {
    // the rigid part of the residual (rigidResidual, :575), then the rigid record: the next reference point is the
    // centroid of the new positions, moved by the rigid translation (the rotation turns them about the old one)
    const double rrb = fmax(sqrt(rbs[0] * rbs[0] + rbs[1] * rbs[1] + rbs[2] * rbs[2]) / fref,
                            sqrt(rbs[3] * rbs[3] + rbs[4] * rbs[4] + rbs[5] * rbs[5]) / (fref * I.rb_R));
    resid = fmax(resid, rrb);
    const bool move = rv[0] != 0 || rv[1] != 0 || rv[2] != 0 || rw[0] != 0 || rw[1] != 0 || rw[2] != 0;
    c.rb_move = move ? 1 : 0;
    for (int k = 0; k < 3; k++) {
      c.rc_prev[k] = c.rc[k];
      if (!rollback) c.rc[k] = rbs[10 + k] / I.n + (move ? rv[k] : 0.0);
      c.rv[k] = rv[k]; c.rw[k] = rw[k];
    }
    c.res_rb = rrb;
  }
  if (!(c.iter > I.warmup)) resid = INFINITY;
  c.resid = resid; c.fref = fref; c.res_f = rf; c.res_t = rt;
  c.e_stretch = acc[1]; c.e_shear = acc[2]; c.e_bend = acc[3]; c.e_tors = acc[4]; c.e_cont = eCont;
  c.e_floor = eGround; c.fn_floor = acc[11];
  const bool conv = resid <= I.tol;
  if (I.proto == 0 && conv) c.done = 1;
  if (!rollback) c.cur = c.cur ^ 1;
  if (I.proto != 0) {
    // UTM replay (the protocol branch of the reference's control kernel, wgsl.js:1048-1093; calibrate.js:154-215):
    // settle at most `settle` iterations (ending at the residual), then average the probe's travel from P0 and its
    // reaction along the axis over `avg` iterations — the state after each iteration, i.e. before it when rolled
    // back — record them, and go on with the next schedule value
    c.it_phase += 1;
    if (c.phase == 0) {
      if (conv || c.it_phase >= I.settle) {
        c.phase = 1; c.it_phase = 0; c.acc_n = 0;
        c.ad[0] = c.ad[1] = c.ad[2] = 0.0; c.ar = 0.0;
      }
    } else if (c.phase == 1) {
      const int pr = I.probe;
      c.ad[0] += gl(I.X[c.cur])[3 * pr] - I.p0x; c.ad[1] += gl(I.X[c.cur])[3 * pr + 1] - I.p0y; c.ad[2] += gl(I.X[c.cur])[3 * pr + 2] - I.p0z;
      if (I.proto == 2) c.ar += gl(I.preact)[0] * I.pax + gl(I.preact)[1] * I.pay + gl(I.preact)[2] * I.paz;
      c.acc_n += 1;
      if (c.it_phase >= I.avg) {
        const double nn = double(c.acc_n);
        const auto o = gl(I.pout) + size_t(c.inc) * 4;
        o[0] = (c.ad[0] * I.pax + c.ad[1] * I.pay + c.ad[2] * I.paz) / nn;
        o[1] = c.ar / nn; o[2] = c.cur_f; o[3] = double(c.iter);
        c.inc += 1; c.it_phase = 0; c.phase = 0;
        if (c.inc >= I.n_inc) { c.phase = 2; c.done = 1; }
        else {
          const double v = gl(I.sched)[c.inc];
          c.cur_f = v;
          if (I.proto == 1) { c.pf[0] = I.pax * v; c.pf[1] = I.pay * v; c.pf[2] = I.paz * v; }
          else {
            // displacement control: the probe node at P0 + v·axis, at rest, in the state the next iteration reads
            gl(I.X[c.cur])[3 * pr] = I.p0x + I.pax * v; gl(I.X[c.cur])[3 * pr + 1] = I.p0y + I.pay * v; gl(I.X[c.cur])[3 * pr + 2] = I.p0z + I.paz * v;
            gl(I.vx)[pr] = Real(0); gl(I.vy)[pr] = Real(0); gl(I.vz)[pr] = Real(0);
          }
        }
      }
    }
  }
  c.live_flag = 1;   // an iteration has been evaluated
  if (I.hist_len > 0) {
    const auto h = gl(I.hist) + size_t((c.iter - 1) % I.hist_len) * HIST_STRIDE;
    h[0] = ke; h[1] = acc[1] + acc[2] + acc[3] + acc[4] + eCont + eGround; h[2] = acc[3]; h[3] = acc[4];
    h[4] = acc[1]; h[5] = acc[2]; h[6] = eCont + eGround; h[7] = resid; h[8] = double(c.iter);
  }
  c.hist_n += 1;
  *ctl = c;
}

// The release before a tile (or a chunk) counts in and the acquire after its count.  The values the control step
// reduces were written through the L2 (part_store), so the release only waits for this wavefront's stores to
// complete (s_waitcnt vmcnt(0), the part of a release that orders agent-scope stores) and the reads — agent-scope
// loads — need no acquire.  An instance whose control step reads positions this launch wrote (a UTM replay: `fence`)
// takes device-scope fences, which on gfx950 write back the XCD's L2 (as every tile doing so was measured to cost
// ~15x the strand launch on the 256 design variants, only those instances pay it).
constexpr int WAIT_VMEM = 0x0F70;   // s_waitcnt vmcnt(0), expcnt and lgkmcnt left alone (gfx9 encoding)
__device__ __forceinline__ void part_release(bool fence) {
  if (fence) __threadfence();
  else __builtin_amdgcn_s_waitcnt(WAIT_VMEM);
}
__device__ __forceinline__ void part_acquire(bool fence) {
  if (fence) __threadfence();
}

// The control step fused into the strand launch.  Every tile counts itself into its chunk (ctrl_chunk consecutive
// tiles of the instance); the last tile of a chunk to finish reduces the chunk's tile partials in tile order, and the
// chunk's share of the crossing-block partials, into a chunk partial; with several chunks the last chunk to finish
// combines the chunk partials in chunk order.  That workgroup then runs the control step.  Every order is fixed: the
// result does not depend on which workgroup arrives last.  Counters are reset by the workgroup that read them last,
// so the next iteration starts from zero.  Partials of this launch are written through the L2 and read with
// device-scope loads; the writers release theirs before counting in (part_release).
template <typename Real, int TILE, bool RB>
__device__ void strand_control_tail(const Inst<Real>& I, Ctl* ctl, const TileItem& tl, StrandLDS<Real, TILE>& L) {
  constexpr int NACC = RPART, NSUM = 7, FPER = 4;
  constexpr int FSTAGE = (SPART * TILE) / RPART;   // chunk partials staged at a time in the tile's dead partial space
  __shared__ double red[NACC][TILE / 64];
  __shared__ double red_rb[RBP][TILE / 64];
  __shared__ int last;
  const int tid = threadIdx.x;
  const int np = I.ntiles, per = I.ctrl_chunk, nch = I.ctrl_chunks;
  const int chunk = tl.tile_local / per;
  const int p0 = chunk * per, p1 = min(np, p0 + per);
  const bool fence = I.proto != 0;
  part_release(fence);                      // release this tile's partial
  __syncthreads();
  if (tid == 0) last = atomicAdd(I.ccount + chunk, 1) == p1 - p0 - 1;
  __syncthreads();
  if (!last) return;
  part_acquire(fence);                      // acquire the chunk's partials
  auto ld = [](const double* q) { return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
  const bool fon = I.floor_on != 0;
  double acc[NACC];
#pragma unroll
  for (int q = 0; q < NACC; q++) acc[q] = 0.0;
  for (int w0 = p0; w0 < p1; w0 += TILE * FPER) {
    double x[FPER][SPART_STRIDE];
#pragma unroll
    for (int k = 0; k < FPER; k++) {
      const int w = w0 + k * TILE + tid;
      const bool in = w < p1;
      const double* sp = I.spart + size_t(in ? w : p0) * SPART_STRIDE;
#pragma unroll
      for (int q = 0; q < SPART - 1; q++) x[k][q] = in ? ld(sp + q) : 0.0;
      x[k][SPART - 1] = (in && fon) ? ld(sp + SPART - 1) : 0.0;
      x[k][SPART] = (in && fon) ? ld(sp + SPART) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < FPER; k++) {
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] += x[k][q];
#pragma unroll
      for (int q = 5; q < 8; q++) acc[q + 2] = fmax(acc[q + 2], x[k][q]);
      acc[5] += x[k][8];
      acc[11] = fmax(acc[11], x[k][9]);
    }
  }
  // the chunk's share of the crossing-block partials (written by the crossing launch before this one)
  const int cper = (I.ncblk + nch - 1) / nch;
  const int c0 = chunk * cper, c1 = min(I.ncblk, c0 + cper);
  for (int w = c0 + tid; w < c1; w += TILE) {
    acc[6] += gl(I.cpart)[w];
    acc[10] = fmax(acc[10], gl(I.cpart)[I.ncblk + w]);
  }
  block_reduce<TILE, NACC>(red, acc, NSUM);
  if (tid == 0) gl(I.ccount)[chunk] = 0;
  if (nch > 1) {
    if (tid == 0) {
      for (int q = 0; q < NACC; q++) part_store<true>(I.rpart + size_t(chunk) * RPART + q, acc[q]);
      part_release(fence);                  // release the chunk partial before counting in
      last = atomicAdd(&ctl->counter, 1) == nch - 1;
    }
    __syncthreads();
    if (!last) return;
    part_acquire(fence);                    // acquire every chunk partial
    double* stage = &L.part[0][0];
    for (int q = 0; q < NACC; q++) acc[q] = 0.0;
    for (int k0 = 0; k0 < nch; k0 += FSTAGE) {
      const int nk = min(FSTAGE, nch - k0);
      __syncthreads();
      for (int v = tid; v < nk * RPART; v += TILE) stage[v] = ld(I.rpart + size_t(k0) * RPART + v);
      __syncthreads();
      if (tid == 0) {
        for (int k = 0; k < nk; k++) {
          for (int q = 0; q < NSUM; q++) acc[q] += stage[k * RPART + q];
          for (int q = NSUM; q < NACC; q++) acc[q] = fmax(acc[q], stage[k * RPART + q]);
        }
      }
    }
  }
  // rigid-body sums of the instance's tiles (a structure only the ground holds), in tile order per lane, then the
  // fixed block order
  double rbs[RBP];
#pragma unroll
  for (int q = 0; q < RBP; q++) rbs[q] = 0.0;
  if (RB && I.rigid) {
    for (int w = tid; w < np; w += TILE)
#pragma unroll
      for (int q = 0; q < RBP; q++) rbs[q] += ld(I.rbpart + size_t(q) * np + w);
    block_reduce<TILE, RBP>(red_rb, rbs, RBP);
  }
  if (tid != 0) return;
  Ctl c = *ctl;
  c.counter = 0;
  control_step(I, ctl, c, acc, rbs);
}

// ================================================================================================= k_ctrl
// one workgroup per instance: fixed-order reduction of the tile and crossing-block partials, then kinetic damping,
// the residual, the stop test and the history ring (solver_cpu.js:809-837; wgsl.js:990-1047)
// ONE: the batch holds a single instance (slot 0), workgroup b reduces its chunk b of gridDim.x chunks
template <typename Real, bool ONE>
__global__ __launch_bounds__(CTRL_THREADS) void k_ctrl(const Inst<Real>* __restrict__ insts, const CtrlItem* __restrict__ items,
                                                      int32_t nitems, Ctl* __restrict__ ctls) {
  // accumulator: [0] Σm|v|² + ΣI3ω², [1..4] stretch / shear / bending / torsion energy, [5] ground energy,
  // [6] contact energy (sums); [7] max |F_free|², [8] max |t_m|², [9] max(|load|, |N|, |M|/l̄_min, ground Fn),
  // [10] max Fn, [11] max ground Fn (maxima)
  constexpr int NACC = RPART, NSUM = 7;
  __shared__ double red[NACC][CTRL_THREADS / 64];
  __shared__ double stage[CTRL_STAGE * RPART];
  __shared__ double red_rb[RBP][CTRL_THREADS / 64];
  __shared__ int last;
  const int cb = xcd_tile(blockIdx.x, gridDim.x);
  if (cb >= nitems) return;
  const CtrlItem it = ONE ? CtrlItem{0, cb, int(gridDim.x), 0} : items[cb];
  const Inst<Real>& I = insts[ONE ? 0 : it.inst];
  Ctl* ctl = ctls + (ONE ? 0 : it.inst);           // read beside the descriptor, not through it
  const int tid = threadIdx.x;
  Ctl c = *ctl;
  // this workgroup's share of the strand-tile and crossing-block partials (value-major: coalesced loads).  A round
  // issues every load of CTRL_PER partials per lane before using any, so a round costs one memory latency; the
  // loads are issued before the liveness test (a stopped instance's partials are read and discarded).  All maxima
  // are of non-negative quantities (Fn enters clamped at 0, below F_ref's other terms), so 0 is their neutral value.
  double acc[NACC];
#pragma unroll
  for (int q = 0; q < NACC; q++) acc[q] = 0.0;
  const int np = I.ntiles;
  const int p0 = it.chunk * I.ctrl_chunk, p1 = min(np, p0 + I.ctrl_chunk);
  // the first crossing-block partials of the lane go out with the first round of strand partials (independent loads;
  // added below in the same order as the rest)
  const int cper = (I.ncblk + it.nchunks - 1) / it.nchunks;
  const int c0 = it.chunk * cper, c1 = min(I.ncblk, c0 + cper);
  const bool cin = c0 + tid < c1, cin1 = c0 + tid + CTRL_THREADS < c1;
  // unconditional loads at an index kept in range (the array holds at least one block's partials), so the
  // descriptor's pointer is not read behind a branch; two per lane, so up to 512 crossing blocks cost no second
  // memory latency
  const int ci = cin ? c0 + tid : 0, ci1 = cin1 ? c0 + tid + CTRL_THREADS : 0;
  const double ce0v = gl(I.cpart)[ci], cm0v = gl(I.cpart)[I.ncblk + ci];
  const double ce1v = gl(I.cpart)[ci1], cm1v = gl(I.cpart)[I.ncblk + ci1];
  const double ce0 = cin ? ce0v : 0.0, cm0 = cin ? cm0v : 0.0;
  const bool fon = I.floor_on != 0;
  for (int w0 = p0; w0 < p1; w0 += CTRL_THREADS * CTRL_PER) {
    double x[CTRL_PER][SPART_STRIDE];
#pragma unroll
    for (int k = 0; k < CTRL_PER; k++) {
      const int w = w0 + k * CTRL_THREADS + tid;
      const bool in = w < p1;
      // one tile's record: five 16-byte loads (the ninth value is the ground's energy, the tenth its largest Fn)
      typedef double dvec2 __attribute__((ext_vector_type(2)));
      const auto sp = gl(reinterpret_cast<const dvec2*>(I.spart + size_t(in ? w : p0) * SPART_STRIDE));
      dvec2 v2[SPART_STRIDE / 2];
#pragma unroll
      for (int h = 0; h < SPART_STRIDE / 2; h++) v2[h] = sp[h];
#pragma unroll
      for (int q = 0; q < SPART - 1; q++) x[k][q] = in ? (q & 1 ? v2[q >> 1].y : v2[q >> 1].x) : 0.0;
      x[k][SPART - 1] = (in && fon) ? v2[(SPART - 1) >> 1].x : 0.0;
      x[k][SPART] = (in && fon) ? v2[SPART >> 1].y : 0.0;
    }
#pragma unroll
    for (int k = 0; k < CTRL_PER; k++) {
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] += x[k][q];
#pragma unroll
      for (int q = 5; q < 8; q++) acc[q + 2] = fmax(acc[q + 2], x[k][q]);
      acc[5] += x[k][8];
      acc[11] = fmax(acc[11], x[k][9]);
    }
  }
  if (cin) { acc[6] += ce0; acc[10] = fmax(acc[10], cm0); }
  if (cin1) { acc[6] += ce1v; acc[10] = fmax(acc[10], cm1v); }
  for (int w = c0 + tid + 2 * CTRL_THREADS; w < c1; w += CTRL_THREADS) {
    acc[6] += gl(I.cpart)[w];
    acc[10] = fmax(acc[10], gl(I.cpart)[I.ncblk + w]);
  }
  if (!(c.done == 0 && c.iter < c.stop_at)) {        // uniform over the workgroup
    if (I.rigid && it.chunk == 0 && tid == 0 && c.rb_move) ctl->rb_move = 0;   // no rigid step without an iteration
    return;
  }
  block_reduce<CTRL_THREADS, NACC>(red, acc, NSUM);
  if (it.nchunks > 1) {
    // several workgroups: each leaves its partial; the last to arrive reduces them in chunk order and goes on
    if (tid == 0) {
      for (int q = 0; q < NACC; q++) gl(I.rpart)[size_t(it.chunk) * RPART + q] = acc[q];
      __threadfence();                             // release the partial before counting in
      last = atomicAdd(&ctl->counter, 1) == it.nchunks - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();                               // acquire: every lane reads the other workgroups' partials
    // the chunk partials come in through LDS, every lane loading one value (one memory latency for a batch of
    // CTRL_STAGE chunks, where one lane reading them in turn would pay one per value); lane 0 combines them in
    // chunk order
    for (int q = 0; q < NACC; q++) acc[q] = 0.0;
    for (int k0 = 0; k0 < it.nchunks; k0 += CTRL_STAGE) {
      const int nk = min(CTRL_STAGE, it.nchunks - k0);
      for (int v = tid; v < nk * RPART; v += CTRL_THREADS)
        stage[v] = __hip_atomic_load(I.rpart + size_t(k0) * RPART + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      if (tid == 0) {
        for (int k = 0; k < nk; k++) {
          for (int q = 0; q < NSUM; q++) acc[q] += stage[k * RPART + q];
          for (int q = NSUM; q < NACC; q++) acc[q] = fmax(acc[q], stage[k * RPART + q]);
        }
      }
      __syncthreads();
    }
    c.counter = 0;                                 // the other workgroups change nothing else in the record
  }
  // rigid-body sums of the instance's tiles (a structure only the ground holds), in tile order per lane, then the
  // fixed block order
  double rbs[RBP];
#pragma unroll
  for (int q = 0; q < RBP; q++) rbs[q] = 0.0;
  if (I.rigid) {
    for (int w = tid; w < np; w += CTRL_THREADS)
#pragma unroll
      for (int q = 0; q < RBP; q++) rbs[q] += gl(I.rbpart)[size_t(q) * np + w];
    block_reduce<CTRL_THREADS, RBP>(red_rb, rbs, RBP);
  }
  if (tid != 0) return;
  control_step(I, ctl, c, acc, rbs);
}


// RB: the batch holds an instance that only the ground holds (its tiles also leave the rigid-body sums); the other
// batches run the form without them
// ONE: the batch holds a single instance (slot 0): its descriptor and control record load beside the tile item.
// WPE > 1 asks for that many resident wavefronts per SIMD (the FP64 form of the 64-lane tile: 168 registers and a
// 36-byte spill for three wavefronts instead of two; faster on batches of thousands of tiles, slower on small ones)
// FUSED: the last tiles of an instance to finish run its control step (strand_control_tail) in this launch, and no
// control launch follows
template <typename Real, int TILE, bool RB, bool ONE, int WPE, bool FUSED>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(WPE))) void k_strand(const Inst<Real>* __restrict__ insts,
                                                 const TileItem* __restrict__ tiles, int32_t ntiles, Ctl* __restrict__ ctls) {
  __shared__ StrandLDS<Real, TILE> L;
  int lt = xcd_tile(blockIdx.x, gridDim.x);
  if (lt >= ntiles) return;
  // the work list holds tiles of occupied slots only (tl.inst >= 0); the liveness test waits until the tile's first
  // loads are out (strand_tile), so the item, the control record, the descriptor and the state cost three memory
  // latencies, not one per test
  const TileItem tl = tiles[lt];
  const int b = ONE ? 0 : tl.inst;
  strand_tile<Real, TILE, false, RB, FUSED>(insts[b], ctls + b, tl, L, OutPtrs{});
}

// The rigid step of a structure only the ground holds (CpuSolver.step, solver_cpu.js:791-805), after the control
// kernel has decided the rigid velocities and kept the iteration: every owned node moves by v + w × (x − c) from
// its position before the node step, every frame turns by exp(w) in world axes, q ← normalize(exp(w) ⊗ q), in f64.
// Same tiles and lanes as the strand kernel; nothing to do when the instance is not rigid or did not move.
template <typename Real, int TILE>
__global__ __launch_bounds__(TILE) void k_rigid(const Inst<Real>* __restrict__ insts, const TileItem* __restrict__ tiles, int32_t ntiles,
                                                const Ctl* __restrict__ ctls) {
  const int lt = blockIdx.x;
  if (lt >= ntiles) return;
  const TileItem tl = tiles[lt];
  if (tl.inst < 0) return;
  const Inst<Real>& I = insts[tl.inst];
  const Ctl* ctl = ctls + tl.inst;
  if (!I.rigid || !ctl->rb_move) return;
  const int t = threadIdx.x, cnt = tl.k1 - tl.k0;
  if (t < 1 || t > cnt) return;                    // lane t owns node and segment k0 − 1 + t
  const int k = tl.k0 - 1 + t;
  const int cur = ctl->cur, old = cur ^ 1;          // the kept iteration flipped the buffers: cur holds the new state
  const double v0 = ctl->rv[0], v1 = ctl->rv[1], v2 = ctl->rv[2], w0 = ctl->rw[0], w1 = ctl->rw[1], w2 = ctl->rw[2];
  const int i = tl.noff + k;
  const double dx = gl(I.X[old])[3 * i] - ctl->rc_prev[0], dy = gl(I.X[old])[3 * i + 1] - ctl->rc_prev[1], dz = gl(I.X[old])[3 * i + 2] - ctl->rc_prev[2];
  gl(I.X[cur])[3 * i] += v0 + w1 * dz - w2 * dy;
  gl(I.X[cur])[3 * i + 1] += v1 + w2 * dx - w0 * dz;
  gl(I.X[cur])[3 * i + 2] += v2 + w0 * dy - w1 * dx;
  if (tl.nseg > 0 && (tl.closed || k < tl.nseg)) {
    const int j = tl.soff + k;
    const double a = sqrt(w0 * w0 + w1 * w1 + w2 * w2), half = 0.5 * a;
    const double sc = a > 1e-12 ? sin(half) / a : 0.5;
    const double ew = cos(half), ex = w0 * sc, ey = w1 * sc, ez = w2 * sc;      // exp(w) (quat.js qexp)
    const double qw = gl(I.Q[cur])[4 * j], qx = gl(I.Q[cur])[4 * j + 1], qy = gl(I.Q[cur])[4 * j + 2], qz = gl(I.Q[cur])[4 * j + 3];
    const double rw = ew * qw - ex * qx - ey * qy - ez * qz;
    const double rx = ew * qx + ex * qw + ey * qz - ez * qy;
    const double ry = ew * qy - ex * qz + ey * qw + ez * qx;
    const double rz = ew * qz + ex * qy - ey * qx + ez * qw;
    const double n2 = rw * rw + rx * rx + ry * ry + rz * rz;
    const double inv = n2 > 0 ? 1.0 / sqrt(n2) : 1.0;
    gl(I.Q[cur])[4 * j] = rw * inv; gl(I.Q[cur])[4 * j + 1] = rx * inv; gl(I.Q[cur])[4 * j + 2] = ry * inv; gl(I.Q[cur])[4 * j + 3] = rz * inv;
  }
}

// generalised forces of every instance of the batch at its current state in one launch over all strand tiles (the
// stability check's evaluations): node forces, body-axis torques and frame inertia into one buffer each, instance b
// at element offsets fofs[b] (3n per instance) and tofs[b] (3m)
template <typename Real, int TILE>
__global__ __launch_bounds__(TILE) void k_evaluate(const Inst<Real>* __restrict__ insts, const TileItem* __restrict__ tiles,
                                                   int32_t ntiles, double* F, double* T, double* I3,
                                                   const int64_t* __restrict__ fofs, const int64_t* __restrict__ tofs) {
  __shared__ StrandLDS<Real, TILE> L;
  if (int(blockIdx.x) >= ntiles) return;
  const TileItem tl = tiles[blockIdx.x];
  if (tl.inst < 0) return;
  const Inst<Real>& I = insts[tl.inst];
  OutPtrs o{};
  o.F = F + fofs[tl.inst]; o.T = T + tofs[tl.inst]; o.I3 = I3 + tofs[tl.inst];
  strand_tile<Real, TILE, true, false>(I, I.ctl, tl, L, o);
}

// positions X [3n] and frames Q [4m] of instance b (interleaved, as the host holds them) into both of its buffers
template <typename Real>
__global__ __launch_bounds__(256) void k_set_xq(const Inst<Real>* __restrict__ insts, int32_t b, const double* __restrict__ X,
                                               const double* __restrict__ Q) {
  const Inst<Real>& I = insts[b];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < I.n) {
    const double x = X[3 * t], y = X[3 * t + 1], z = X[3 * t + 2];
    for (int k = 0; k < 2; k++) { gl(I.X[k])[3 * t] = x; gl(I.X[k])[3 * t + 1] = y; gl(I.X[k])[3 * t + 2] = z; }
  }
  if (t < I.m) {
    const double w = Q[4 * t], qx = Q[4 * t + 1], qy = Q[4 * t + 2], qz = Q[4 * t + 3];
    for (int k = 0; k < 2; k++) { gl(I.Q[k])[4 * t] = w; gl(I.Q[k])[4 * t + 1] = qx; gl(I.Q[k])[4 * t + 2] = qy; gl(I.Q[k])[4 * t + 3] = qz; }
  }
}

// outputs of the last evaluated state of one instance: N, V, M, clamp moments, reactions (download)
template <typename Real, int TILE>
__global__ __launch_bounds__(TILE) void k_outputs(const Inst<Real>* __restrict__ insts, const TileItem* __restrict__ tiles, int32_t inst,
                                                  OutPtrs out) {
  __shared__ StrandLDS<Real, TILE> L;
  const Inst<Real>& I = insts[inst];
  if (int(blockIdx.x) >= I.ntiles) return;
  const TileItem tl = tiles[I.tile0 + blockIdx.x];
  strand_tile<Real, TILE, true, false>(I, I.ctl, tl, L, out);
}

// the multipliers as the host solver holds them after an iteration: an AL update decided by the control kernel is
// applied at the start of the next iteration on the device, so a download between the two applies it here
template <typename Real>
__global__ __launch_bounds__(256) void k_al_preview(const Inst<Real>* __restrict__ insts, int32_t inst, double* out) {
  const Inst<Real>& I = insts[inst];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= I.K) return;
  const Ctl* ctl = I.ctl;
  const double lam = gl(I.lam)[c];
  out[c] = (I.aug && I.contacts && ctl->al_update) ? al_updated(I, c, lam, gl(I.penacc)[c], ctl->pen_n) : lam;
}

// the plain copy the benchmark measures beside the strand kernel (the practical HBM ceiling) and calibrates the
// traffic counters with: each workgroup moves COPY_PER 16-byte elements, every load of a lane issued before its
// stores, streaming loads and stores (measured on MI355X at 1 GiB: 6.2 TB/s, where a grid-stride loop of 16,384
// workgroups reached 4.6)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  constexpr int U = COPY_PER / 256;
  const u32x4* s = reinterpret_cast<const u32x4*>(src);
  u32x4* d = reinterpret_cast<u32x4*>(dst);
  const int64_t b0 = int64_t(blockIdx.x) * COPY_PER, base = b0 + threadIdx.x;
  if (b0 + COPY_PER <= n) {                        // a whole block: every load in flight before the stores
    u32x4 v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = __builtin_nontemporal_load(&s[base + k * 256]);
#pragma unroll
    for (int k = 0; k < U; k++) __builtin_nontemporal_store(v[k], &d[base + k * 256]);
  } else {                                         // the last, partial block
    for (int64_t i = base; i < n; i += 256) d[i] = s[i];
  }
}

// explicit instantiations
#define WSD_STRAND_F(R, T, W, F) \
  template __global__ void k_strand<R, T, false, false, W, F>(const Inst<R>*, const TileItem*, int32_t, Ctl*); \
  template __global__ void k_strand<R, T, true, false, W, F>(const Inst<R>*, const TileItem*, int32_t, Ctl*); \
  template __global__ void k_strand<R, T, false, true, W, F>(const Inst<R>*, const TileItem*, int32_t, Ctl*); \
  template __global__ void k_strand<R, T, true, true, W, F>(const Inst<R>*, const TileItem*, int32_t, Ctl*);
#define WSD_STRAND(R, T, W) WSD_STRAND_F(R, T, W, false) WSD_STRAND_F(R, T, W, true)
WSD_STRAND(double, 64, 3)
#define WSD_INST(R, T) \
  WSD_STRAND(R, T, 1) \
  template __global__ void k_rigid<R, T>(const Inst<R>*, const TileItem*, int32_t, const Ctl*); \
  template __global__ void k_outputs<R, T>(const Inst<R>*, const TileItem*, int32_t, OutPtrs); \
  template __global__ void k_evaluate<R, T>(const Inst<R>*, const TileItem*, int32_t, double*, double*, double*, const int64_t*, const int64_t*);
WSD_INST(float, 64) WSD_INST(float, 128) WSD_INST(float, 256)
WSD_INST(double, 64) WSD_INST(double, 128) WSD_INST(double, 256)
#define WSD_CROSS(R, N) \
  template __global__ void k_cross<R, N, false>(const Inst<R>*, const CrossItem*, int32_t, const Ctl*, int32_t); \
  template __global__ void k_cross<R, N, true>(const Inst<R>*, const CrossItem*, int32_t, const Ctl*, int32_t);
WSD_CROSS(float, 1) WSD_CROSS(float, 2) WSD_CROSS(double, 1) WSD_CROSS(double, 2)
template __global__ void k_ctrl<float, false>(const Inst<float>*, const CtrlItem*, int32_t, Ctl*);
template __global__ void k_ctrl<float, true>(const Inst<float>*, const CtrlItem*, int32_t, Ctl*);
template __global__ void k_ctrl<double, false>(const Inst<double>*, const CtrlItem*, int32_t, Ctl*);
template __global__ void k_ctrl<double, true>(const Inst<double>*, const CtrlItem*, int32_t, Ctl*);
template __global__ void k_set_xq<float>(const Inst<float>*, int32_t, const double*, const double*);
template __global__ void k_set_xq<double>(const Inst<double>*, int32_t, const double*, const double*);
template __global__ void k_al_preview<float>(const Inst<float>*, int32_t, double*);
template __global__ void k_al_preview<double>(const Inst<double>*, int32_t, double*);
template __global__ void k_begin<float>(Ctl*, const int32_t*, const int32_t*, int32_t, int32_t, int32_t);
template __global__ void k_begin<double>(Ctl*, const int32_t*, const int32_t*, int32_t, int32_t, int32_t);

}  // namespace wsd
