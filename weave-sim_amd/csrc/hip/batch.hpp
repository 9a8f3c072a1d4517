// Host side of the device solver: the per-instance device arena, the upload of a host model (the reference's
// GpuSolver upload path, web/js/solver_gpu.js:162-310), and Batch — a set of instance slots relaxed by the same
// launches (GpuSolver, solver_gpu.js:62-523).  Shared by the batch C ABI (solver.hip) and the request server
// (server.hip); the kernels are in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/weave.h"
#include "device.hpp"

namespace wsd {
template <typename Real> __global__ void k_begin(Ctl*, const int32_t*, const int32_t*, int32_t, int32_t, int32_t);
template <typename Real, int LANES, bool ONE> __global__ void k_cross(const Inst<Real>*, const CrossItem*, int32_t, const Ctl*, int32_t);
template <typename Real, int TILE, bool RB, bool ONE, int WPE, bool FUSED> __global__ void k_strand(const Inst<Real>*, const TileItem*, int32_t, Ctl*);
template <typename Real, int TILE> __global__ void k_rigid(const Inst<Real>*, const TileItem*, int32_t, const Ctl*);
template <typename Real, bool ONE> __global__ void k_ctrl(const Inst<Real>*, const CtrlItem*, int32_t, Ctl*);
__global__ void k_copy(const uint4*, uint4*, int64_t);
template <typename Real> __global__ void k_al_preview(const Inst<Real>*, int32_t, double*);
template <typename Real> __global__ void k_set_xq(const Inst<Real>*, int32_t, const double*, const double*);
template <typename Real, int TILE> __global__ void k_outputs(const Inst<Real>*, const TileItem*, int32_t, OutPtrs);
template <typename Real, int TILE> __global__ void k_evaluate(const Inst<Real>*, const TileItem*, int32_t, double*, double*, double*, const int64_t*, const int64_t*);
}  // namespace wsd

using namespace wsd;

struct ws_ctx {
  int device = 0;
  std::string err;
  hipDeviceProp_t prop{};
};

namespace {

struct Fail { int code; std::string msg; };
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) throw Fail{WS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)}; } while (0)

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// ------------------------------------------------------------------------------------------------ device arena
struct Arena {
  size_t size = 0;
  char* base = nullptr;
  std::vector<std::pair<void**, size_t>> pending;   // (pointer slot, offset)
  template <typename T> void add(T*& slot, size_t count) {
    size_t off = size;
    size += align256(std::max<size_t>(count, 1) * sizeof(T));
    pending.push_back({reinterpret_cast<void**>(const_cast<typename std::remove_const<T>::type**>(&slot)), off});
  }
  template <typename T> void addc(const T*& slot, size_t count) {
    size_t off = size;
    size += align256(std::max<size_t>(count, 1) * sizeof(T));
    pending.push_back({reinterpret_cast<void**>(const_cast<T**>(&slot)), off});
  }
  void alloc() {
    HIPCHK(hipMalloc(&base, size));
    HIPCHK(hipMemset(base, 0, size));
    for (auto& p : pending) *p.first = base + p.second;
    pending.clear();
  }
  void release() { if (base) (void)hipFree(base); base = nullptr; }
};

template <typename T> void h2d(T* dst, const std::vector<T>& src) {
  if (!src.empty()) HIPCHK(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
}
template <typename T> void h2d(const T* dst, const std::vector<T>& src) { h2d(const_cast<T*>(dst), src); }
template <typename T> std::vector<T> d2h(const T* src, size_t n) {
  std::vector<T> v(n);
  if (n) HIPCHK(hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  return v;
}

// ------------------------------------------------------------------------------------------------ one instance
template <typename Real>
struct InstHost {
  Inst<Real> d{};                     // descriptor with device pointers (copied to the device table)
  Arena arena;
  int n = 0, m = 0, p = 0, K = 0, C = 0, S = 0;
  std::vector<int32_t> cseg;          // clamp segments (fixed for the batch)
  std::vector<double> X0, Q0;         // displacement origin (host; internalize moves it)
  std::vector<TileItem> tiles;
  int ncblk = 0;
  double alg_bytes = 0;               // algorithmic HBM bytes per iteration
  int hist_read = 0;
  double tol = 0;
  double outN_dummy = 0;
};

template <typename Real>
std::vector<Real> toReal(const double* p, size_t n) {
  std::vector<Real> v(n);
  for (size_t i = 0; i < n; i++) v[i] = p ? Real(p[i]) : Real(0);
  return v;
}

// a node's fixed freedoms: bit0..2 x, y, z fixed, bit3 any (a reaction is reported), bit4 all three (the contact's
// setFixed); a replay's pinned probe (displacement control) has bits 0..3 only: pinned after the contact goal was set
// up, as calibrate.js:190-193 pins it (solver_gpu.js:215)
inline uint8_t fix_bits(const ws_model& md, int i) {
  const uint8_t* f = md.node_free + 3 * i;
  uint8_t bits = uint8_t((f[0] ? 0 : 1) | (f[1] ? 0 : 2) | (f[2] ? 0 : 4));
  if (bits) bits |= 8;
  if ((bits & 7) == 7) bits |= 16;
  if (md.proto_mode == 2 && i == md.proto_probe) bits = uint8_t(bits | 15);
  return bits;
}

// the section stiffness per strand, when every segment of each strand carries the same (EA, GA, EIx, EIy, GJ) bit for
// bit in the device's precision (a lath's section is its group's, model.js:236-252): the strand kernel then reads 20
// bytes per strand instead of per segment.  And the axes along which some constant load is not +0: the others are
// not read.  Uploaded at build and by every parameter update (a scene edit may change either).
template <typename Real>
void upload_sections(Inst<Real>& d, const ws_model& md) {
  const int S = md.S;
  const double* src[5] = {md.EA, md.GA, md.EIx, md.EIy, md.GJ};
  std::vector<Real> tab(5 * size_t(S), Real(0));
  bool uni = true;
  for (int s = 0; s < S && uni; s++) {
    const int j0 = md.seg_offset[s], j1 = md.seg_offset[s + 1];
    for (int q = 0; q < 5 && uni; q++) {
      const Real v = j1 > j0 ? Real(src[q][j0]) : Real(0);
      tab[5 * size_t(s) + q] = v;
      for (int j = j0 + 1; j < j1 && uni; j++) { const Real w = Real(src[q][j]); uni = std::memcmp(&w, &v, sizeof(Real)) == 0; }
    }
  }
  d.sec_strand = uni ? 1 : 0;
  if (uni) h2d(d.sec, tab);
  int axes = 0;
  const Real zero = Real(0);
  for (int i = 0; i < md.n && axes != 7; i++)
    for (int c = 0; c < 3; c++) { const Real v = Real(md.load_const[3 * i + c]); if (std::memcmp(&v, &zero, sizeof(Real)) != 0) axes |= 1 << c; }
  d.load_axes = axes;
}

// build the device image of one host model (GpuSolver.upload, solver_gpu.js:191-310)
// fused: the strand launch runs the control step (chunks of FUSED_CHUNK tiles by default); else the control kernel
// (measured: fusing is slower on gfx950 — each tile's device-scope release fence writes back its XCD's L2)
template <typename Real>
void build_instance(InstHost<Real>& H, const ws_model& md, const ws_state& st, int tile, int hist_len, Ctl* ctl,
                    int ctrl_chunk = 0, bool layout_only = false, bool fused = false) {
  const int n = md.n, m = md.m, p = md.p, C = md.C, S = md.S;
  const int K = md.par.contacts ? md.K : 0;   // no contact goal: no crossings on the device
  if (n <= 0 || m < 0 || S <= 0) throw Fail{WS_ERR_ARG, "model has no nodes or strands"};
  // the rigid-body freedoms: on when the ground alone holds the structure (no node with a fixed freedom, no clamp;
  // CpuSolver.step, solver_cpu.js:745)
  bool anchored = md.C > 0;
  for (int i = 0; i < md.n && !anchored; i++) anchored = !(md.node_free[3 * i] && md.node_free[3 * i + 1] && md.node_free[3 * i + 2]);
  if (md.proto_mode == 2) anchored = true;          // the pinned probe (solver_gpu.js:224)
  const bool rigid = md.par.floor_on && !anchored;
  if (md.proto_mode < 0 || md.proto_mode > 2) throw Fail{WS_ERR_ARG, "protocol mode must be 0, 1 or 2"};
  if (md.proto_mode) {
    if (md.proto_probe < 0 || md.proto_probe >= md.n) throw Fail{WS_ERR_ARG, "protocol probe node out of range"};
    if (md.proto_n < 1 || !md.proto_sched) throw Fail{WS_ERR_ARG, "GPU replay: empty load schedule"};
    if (md.proto_settle < 1 || md.proto_avg < 1) throw Fail{WS_ERR_ARG, "protocol settle and avg must be at least 1"};
  }
  for (int q = 0; q < 2 * (md.par.contacts ? md.K : 0); q++)
    if (md.k_nseg[q] < 1 || md.k_strand[q] < 0 || md.k_strand[q] >= S) throw Fail{WS_ERR_ARG, "a crossing lies on a strand without segments"};
  for (int s2 = 0; s2 < S; s2++)
    if (md.node_offset[s2 + 1] - md.node_offset[s2] < 1 || (md.strand_closed[s2] && md.seg_offset[s2 + 1] == md.seg_offset[s2]))
      throw Fail{WS_ERR_ARG, "a strand has no nodes (or is closed without segments)"};
  H.n = n; H.m = m; H.p = p; H.K = K; H.C = C; H.S = S;
  H.tol = md.par.tol;
  Inst<Real>& d = H.d;
  d.n = n; d.m = m; d.p = p; d.K = K; d.C = C; d.S = S; d.npl = md.n_point_local > 0 ? 1 : 0;
  const ws_params& P = md.par;
  d.mass_scale = P.mass_scale; d.damping = P.damping; d.tol = P.tol; d.mu = P.mu; d.strength = P.strength; d.stick = P.stick_strength;
  d.torsion_factor = P.torsion_factor; d.epsilon = P.epsilon; d.slip_rate = P.slip_rate; d.rot_factor = P.rot_factor;
  d.dent_rate = P.dent_rate; d.edge_thr = P.edge_threshold; d.l_ref = md.l_ref;
  double jmin = INFINITY; for (int j = 0; j < p; j++) jmin = std::min(jmin, md.j_len[j]);
  d.jlen_min = p ? jmin : 1.0;
  d.inv_jlen_min = p ? 1.0 / std::max(jmin, 1e-12) : 0.0;
  d.kin_damp = P.kinetic_damping; d.warmup = P.warmup_iter; d.contacts = K > 0 ? 1 : 0; d.ramp_iter = P.ramp_iter;
  d.rot_fric = P.rot_friction; d.seat = P.seat_law; d.seat_gap = P.seat_gap; d.dent_after = P.dent_after;
  d.aug = P.augmented; d.al_every = P.al_every; d.hist_len = hist_len;
  d.floor_on = P.floor_on ? 1 : 0; d.floor_z = P.floor_z; d.floor_mu = std::max(P.floor_mu, 0.0);
  d.rigid = rigid ? 1 : 0;
  d.proto = md.proto_mode; d.probe = md.proto_probe; d.n_inc = md.proto_mode ? md.proto_n : 0;
  d.settle = md.proto_settle; d.avg = md.proto_avg;
  d.pax = md.proto_axis[0]; d.pay = md.proto_axis[1]; d.paz = md.proto_axis[2];
  if (md.proto_mode) { d.p0x = st.X[3 * md.proto_probe]; d.p0y = st.X[3 * md.proto_probe + 1]; d.p0z = st.X[3 * md.proto_probe + 2]; }
  if (rigid) {
    // model radius about the centroid and the rigid speed limits (the Floor constructor, solver_cpu.js:546-562): the
    // radius from the uploaded positions unless the state carries one (a downloaded state's, as the reference's
    // md.floor keeps Rb from its construction)
    double c[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) c[k] += st.X[3 * i + k] / std::max(n, 1);
    double R2 = 0, hmin = INFINITY;
    for (int i = 0; i < n; i++) {
      const double dx = st.X[3 * i] - c[0], dy = st.X[3 * i + 1] - c[1], dz = st.X[3 * i + 2] - c[2];
      R2 = std::max(R2, dx * dx + dy * dy + dz * dz);
      hmin = std::min(hmin, std::min(md.node_h[i], md.node_w[i]));
    }
    d.rb_R = st.rb_radius > 0 ? st.rb_radius : std::max(std::sqrt(R2), 1e-9);
    d.rb_vmax = RB_VMAX * (std::isfinite(hmin) && hmin > 0 ? hmin : 1.0);
    d.rb_wmax = d.rb_vmax / d.rb_R;
  }
  // ---- strand tiles and crossing blocks
  H.tiles.clear();
  for (int s = 0; s < S; s++) {
    int nn = md.node_offset[s + 1] - md.node_offset[s];
    // a tile of `tile` lanes owns tile − 1 nodes: lane 0 computes the halo segment and joint before them
    for (int k0 = 0; k0 < nn; k0 += tile - 1) {
      TileItem t{};
      t.inst = 0; t.strand = s; t.k0 = k0; t.k1 = std::min(nn, k0 + tile - 1);
      t.noff = md.node_offset[s]; t.nn = nn;
      t.soff = md.seg_offset[s]; t.nseg = md.seg_offset[s + 1] - md.seg_offset[s];
      t.joff = md.joint_offset[s]; t.njs = md.joint_offset[s + 1] - md.joint_offset[s];
      t.closed = md.strand_closed[s] ? 1 : 0;
      t.tile_local = int(H.tiles.size());
      H.tiles.push_back(t);
    }
  }
  d.ntiles = int(H.tiles.size());
  H.ncblk = (K + CROSS_BLOCK - 1) / CROSS_BLOCK;   // at two lanes per crossing: the most blocks a launch uses
  d.ncblk = H.ncblk;
  // ---- arena layout
  Arena& A = H.arena;
  A.addc(d.node_off, S + 1); A.addc(d.seg_off, S + 1); A.addc(d.joint_off, S + 1); A.addc(d.closed, S);
  A.addc(d.seg_i0, m); A.addc(d.seg_i1, m); A.addc(d.nsa, n); A.addc(d.nsb, n);
  A.addc(d.side_off, S + 1); A.addc(d.side_pos, 2 * K);
  if (C) { A.addc(d.clamp_off, m + 1); A.addc(d.clamp_idx, C); }
  for (int b = 0; b < 2; b++) A.add(d.X[b], 3 * size_t(n));
  A.add(d.vx, n); A.add(d.vy, n); A.add(d.vz, n);
  A.addc(d.mass, n); A.addc(d.lx, n); A.addc(d.ly, n); A.addc(d.lz, n);
  if (d.npl) { A.addc(d.llx, n); A.addc(d.lly, n); A.addc(d.llz, n); }
  A.addc(d.fix, n);
  for (int b = 0; b < 2; b++) A.add(d.Q[b], 4 * size_t(m));
  A.add(d.wx, m); A.add(d.wy, m); A.add(d.wz, m);
  A.addc(d.l0, m); A.addc(d.EA, m); A.addc(d.GA, m); A.addc(d.EIx, m); A.addc(d.EIy, m); A.addc(d.GJ, m);
  A.addc(d.sec, 5 * size_t(S));
  A.addc(d.krx, m); A.addc(d.kry, m); A.addc(d.krz, m);
  if (d.floor_on) { A.addc(d.nodew, n); A.addc(d.nodeh, n); A.add(d.fax, n); A.add(d.fay, n); A.add(d.gstate, n); }
  A.addc(d.jlen, p); A.addc(d.om0, 3 * p);
  if (C) { A.addc(d.c_len, C); A.addc(d.c_B, 3 * C); A.addc(d.c_Q, 4 * C); }
  A.addc(d.k_seg_off, 2 * K); A.addc(d.k_nseg, 2 * K); A.addc(d.k_noff, 2 * K); A.addc(d.k_nn, 2 * K); A.addc(d.wk, 2 * K); A.addc(d.hk, 2 * K);
  A.addc(d.fc90, 2 * K); A.addc(d.E90, 2 * K); A.addc(d.Hs, 2 * K); A.addc(d.kw, K); A.addc(d.kh, K); A.addc(d.k_en, K); A.add(d.k_lock, K);
  A.add(d.u, 2 * K); A.add(d.slip, K); A.add(d.nprev, 3 * K); A.add(d.kflags, K); A.add(d.fn, K); A.add(d.ft, K); A.add(d.gap, K);
  A.add(d.lam, K); A.add(d.penacc, K); A.add(d.theta0, K); A.add(d.rotslip, K); A.add(d.mrot, K);
  A.add(d.dent, 2 * K); A.add(d.climb, 2 * K); A.add(d.tanpsi, 2 * K); A.add(d.dente, K);
  A.add(d.theta, K); A.add(d.anom, K); A.add(d.pres, K); A.add(d.util, K); A.add(d.phi, 2 * K); A.add(d.mid, 3 * K);
  A.add(d.ctype, K); A.add(d.face, K); A.add(d.side, 2 * K);
  d.ctl = ctl; A.add(d.hist, size_t(std::max(hist_len, 1)) * HIST_STRIDE);
  d.tile_waves = tile / 64;
  A.add(d.spart, size_t(d.ntiles) * SPART_STRIDE); A.add(d.cpart, size_t(std::max(H.ncblk, 1)) * CPART);
  {
    const int np = std::max(1, d.ntiles);
    d.ctrl_chunk = ctrl_chunk > 0 ? ctrl_chunk : (fused ? FUSED_CHUNK : (np <= CTRL_ONE ? np : CTRL_CHUNK));
    d.ctrl_chunks = (np + d.ctrl_chunk - 1) / d.ctrl_chunk;
  }
  A.add(d.rpart, size_t(d.ctrl_chunks) * RPART);
  A.add(d.ccount, size_t(d.ctrl_chunks));
  if (d.rigid) A.add(d.rbpart, size_t(d.ntiles) * RBP);
  if (d.proto) { A.addc(d.sched, d.n_inc); A.add(d.pout, 4 * size_t(d.n_inc)); A.add(d.preact, 3); }
  if (layout_only) return;   // A.size: the device bytes the instance would take
  A.alloc();
  if (d.proto) h2d(d.sched, std::vector<double>(md.proto_sched, md.proto_sched + d.n_inc));
  // ---- topology
  h2d(d.node_off, std::vector<int32_t>(md.node_offset, md.node_offset + S + 1));
  h2d(d.seg_off, std::vector<int32_t>(md.seg_offset, md.seg_offset + S + 1));
  h2d(d.joint_off, std::vector<int32_t>(md.joint_offset, md.joint_offset + S + 1));
  h2d(d.closed, std::vector<uint8_t>(md.strand_closed, md.strand_closed + S));
  h2d(d.seg_i0, std::vector<int32_t>(md.seg_i0, md.seg_i0 + m));
  h2d(d.seg_i1, std::vector<int32_t>(md.seg_i1, md.seg_i1 + m));
  std::vector<int32_t> nsa(n, -1), nsb(n, -1);
  for (int j = 0; j < m; j++) { nsa[md.seg_i0[j]] = j; nsb[md.seg_i1[j]] = j; }
  h2d(d.nsa, nsa); h2d(d.nsb, nsb);
  // crossing sides per strand, ordered by their position along the strand (then by side id)
  std::vector<std::vector<std::pair<double, int>>> per(S);
  for (int c = 0; c < K; c++) for (int sd = 0; sd < 2; sd++) per[md.k_strand[2 * c + sd]].push_back({st.u[2 * c + sd], 2 * c + sd});
  std::vector<int32_t> soff(S + 1, 0), spos(2 * K, 0);
  int slot = 0;
  for (int s = 0; s < S; s++) { std::sort(per[s].begin(), per[s].end()); for (auto& e : per[s]) spos[e.second] = slot++; soff[s + 1] = slot; }
  h2d(d.side_off, soff); h2d(d.side_pos, spos);
  for (auto& t : H.tiles) { t.so = soff[t.strand]; t.ns = soff[t.strand + 1] - soff[t.strand]; }
  H.cseg.assign(md.c_seg, md.c_seg + C);
  if (C) {
    std::vector<int32_t> coff(m + 1, 0), cidx(C);
    for (int c = 0; c < C; c++) coff[md.c_seg[c] + 1]++;
    for (int j = 0; j < m; j++) coff[j + 1] += coff[j];
    std::vector<int32_t> fill(coff.begin(), coff.end() - 1);
    for (int c = 0; c < C; c++) cidx[fill[md.c_seg[c]]++] = c;
    h2d(d.clamp_off, coff); h2d(d.clamp_idx, cidx);
  }
  // ---- goals, parameters, crossing constants
  h2d(d.mass, toReal<Real>(md.mass, n));
  if (d.floor_on) { h2d(d.nodew, toReal<Real>(md.node_w, n)); h2d(d.nodeh, toReal<Real>(md.node_h, n)); }
  std::vector<Real> lx(n), ly(n), lz(n);
  for (int i = 0; i < n; i++) { lx[i] = Real(md.load_const[3 * i]); ly[i] = Real(md.load_const[3 * i + 1]); lz[i] = Real(md.load_const[3 * i + 2]); }
  h2d(d.lx, lx); h2d(d.ly, ly); h2d(d.lz, lz);
  if (d.npl) {
    std::vector<double> acc(3 * n, 0.0);
    for (int q = 0; q < md.n_point_local; q++) for (int c = 0; c < 3; c++) acc[3 * md.point_local_node[q] + c] += md.point_local_F[3 * q + c];
    std::vector<Real> a(n), b(n), cc(n);
    for (int i = 0; i < n; i++) { a[i] = Real(acc[3 * i]); b[i] = Real(acc[3 * i + 1]); cc[i] = Real(acc[3 * i + 2]); }
    h2d(d.llx, a); h2d(d.lly, b); h2d(d.llz, cc);
  }
  std::vector<uint8_t> fx(n);
  for (int i = 0; i < n; i++) {
    fx[i] = fix_bits(md, i);
  }
  h2d(d.fix, fx);
  h2d(d.l0, std::vector<double>(md.l0, md.l0 + m));
  h2d(d.EA, toReal<Real>(md.EA, m)); h2d(d.GA, toReal<Real>(md.GA, m)); h2d(d.EIx, toReal<Real>(md.EIx, m));
  h2d(d.EIy, toReal<Real>(md.EIy, m)); h2d(d.GJ, toReal<Real>(md.GJ, m));
  upload_sections(d, md);
  std::vector<Real> kx(m), ky(m), kz(m);
  for (int j = 0; j < m; j++) { kx[j] = Real(md.krot_static[3 * j]); ky[j] = Real(md.krot_static[3 * j + 1]); kz[j] = Real(md.krot_static[3 * j + 2]); }
  h2d(d.krx, kx); h2d(d.kry, ky); h2d(d.krz, kz);
  h2d(d.jlen, toReal<Real>(md.j_len, p)); h2d(d.om0, toReal<Real>(md.omega0, 3 * p));
  if (C) { h2d(d.c_len, toReal<Real>(md.c_len, C)); h2d(d.c_B, toReal<Real>(md.c_B, 3 * C)); h2d(d.c_Q, std::vector<double>(md.c_Q, md.c_Q + 4 * C)); }
  if (K) {
    h2d(d.k_seg_off, std::vector<int32_t>(md.k_seg_off, md.k_seg_off + 2 * K));
    h2d(d.k_nseg, std::vector<int32_t>(md.k_nseg, md.k_nseg + 2 * K));
    std::vector<int32_t> knoff(2 * K), knn(2 * K);
    for (int q = 0; q < 2 * K; q++) {
      int s2 = md.k_strand[q];
      knoff[q] = md.node_offset[s2];
      knn[q] = (md.node_offset[s2 + 1] - md.node_offset[s2]) | (md.strand_closed[s2] ? (1 << 30) : 0);
    }
    h2d(d.k_noff, knoff); h2d(d.k_nn, knn);
    h2d(d.wk, toReal<Real>(md.k_wk, 2 * K)); h2d(d.hk, toReal<Real>(md.k_hk, 2 * K));
    h2d(d.fc90, toReal<Real>(md.k_fc90, 2 * K)); h2d(d.E90, toReal<Real>(md.k_E90, 2 * K)); h2d(d.Hs, toReal<Real>(md.k_Hs, 2 * K));
    h2d(d.kw, toReal<Real>(md.k_w, K)); h2d(d.kh, toReal<Real>(md.k_h, K));
    h2d(d.k_en, std::vector<uint8_t>(md.k_enabled, md.k_enabled + K));
    h2d(d.k_lock, std::vector<uint8_t>(md.k_locked, md.k_locked + K));
  }
  // ---- algorithmic bytes per iteration: every per-iteration array read once, written once if it changes
  const double R = double(sizeof(Real));
  H.alg_bytes = n * (48.0 + 6 * R + R + 3 * R + 1) + m * (64.0 + 6 * R + 8 + 5 * R + 3 * R) + p * (4 * R) +
                K * (32.0 + 6 * R + 16 + R + 1 + 16 + 4 * R + 2 * R + 1);
}

// dynamic state (GpuSolver._uploadState, solver_gpu.js:271-310)
// device bytes of one instance (its arena), without allocating: the server's admission test
template <typename Real>
size_t instance_bytes(const ws_model& md, const ws_state& st, int tile, int hist_len) {
  InstHost<Real> H;
  build_instance(H, md, st, tile, hist_len, nullptr, 0, true);
  return H.arena.size;
}

template <typename Real>
void upload_state(InstHost<Real>& H, const ws_state& st) {
  Inst<Real>& d = H.d;
  const int n = H.n, m = H.m, K = H.K;
  std::vector<double> xs(st.X, st.X + 3 * size_t(n));
  double v0 = 0;
  if (d.proto) HIPCHK(hipMemcpy(&v0, d.sched, sizeof(double), hipMemcpyDeviceToHost));
  if (d.proto == 2) {   // displacement control starts with the probe at P0 + sched[0]·axis (solver_gpu.js:309, 314-318)
    xs[3 * d.probe] = d.p0x + d.pax * v0; xs[3 * d.probe + 1] = d.p0y + d.pay * v0; xs[3 * d.probe + 2] = d.p0z + d.paz * v0;
  }
  for (int b = 0; b < 2; b++) h2d(d.X[b], xs);
  std::vector<Real> vx(n), vy(n), vz(n);
  for (int i = 0; i < n; i++) { vx[i] = st.V ? Real(st.V[3 * i]) : Real(0); vy[i] = st.V ? Real(st.V[3 * i + 1]) : Real(0); vz[i] = st.V ? Real(st.V[3 * i + 2]) : Real(0); }
  if (d.proto == 2) { vx[d.probe] = vy[d.probe] = vz[d.probe] = Real(0); }
  h2d(d.vx, vx); h2d(d.vy, vy); h2d(d.vz, vz);
  if (d.floor_on) {   // stick points: the state's, else under the nodes (Floor constructor, solver_cpu.js:540-545)
    std::vector<double> ax(n), ay(n);
    for (int i = 0; i < n; i++) { ax[i] = st.floor_a ? st.floor_a[2 * i] : xs[3 * i]; ay[i] = st.floor_a ? st.floor_a[2 * i + 1] : xs[3 * i + 1]; }
    h2d(d.fax, ax); h2d(d.fay, ay);
  }
  const std::vector<double> qs(st.Q, st.Q + 4 * size_t(m));
  for (int b = 0; b < 2; b++) h2d(d.Q[b], qs);
  std::vector<Real> wx(m), wy(m), wz(m);
  for (int j = 0; j < m; j++) { wx[j] = st.W ? Real(st.W[3 * j]) : Real(0); wy[j] = st.W ? Real(st.W[3 * j + 1]) : Real(0); wz[j] = st.W ? Real(st.W[3 * j + 2]) : Real(0); }
  h2d(d.wx, wx); h2d(d.wy, wy); h2d(d.wz, wz);
  H.X0.assign(st.X0 ? st.X0 : st.X, (st.X0 ? st.X0 : st.X) + 3 * n);
  H.Q0.assign(st.Q0 ? st.Q0 : st.Q, (st.Q0 ? st.Q0 : st.Q) + 4 * m);
  if (K) {
    h2d(d.u, std::vector<double>(st.u, st.u + 2 * K));
    h2d(d.slip, st.slip ? std::vector<double>(st.slip, st.slip + K) : std::vector<double>(K, 0.0));
    h2d(d.nprev, toReal<Real>(st.n_prev, 3 * K));
    std::vector<uint8_t> fl(K, 2);
    if (st.stuck) for (int c = 0; c < K; c++) fl[c] = st.stuck[c] ? 2 : 0;
    h2d(d.kflags, fl);
    h2d(d.fn, toReal<Real>(st.Fn, K));
    h2d(d.lam, toReal<double>(st.lam, K)); h2d(d.penacc, toReal<double>(st.pen_acc, K));
    std::vector<double> th0(K, NAN);
    if (st.theta0) for (int c = 0; c < K; c++) th0[c] = st.theta0[c];
    h2d(d.theta0, th0); h2d(d.rotslip, toReal<double>(st.rot_slip, K));
    h2d(d.dent, toReal<Real>(st.dent, 2 * K)); h2d(d.climb, toReal<Real>(st.climb, 2 * K));
    std::vector<Real> tp(2 * K, Real(1));
    if (st.tanpsi) for (int c = 0; c < 2 * K; c++) tp[c] = Real(st.tanpsi[c]);
    h2d(d.tanpsi, tp); h2d(d.dente, toReal<Real>(st.dent_e, K));
    if (st.theta) h2d(d.theta, toReal<Real>(st.theta, K));
  }
  Ctl c{};
  c.iter = st.iteration; c.stop_at = st.iteration; c.done = 0; c.cur = 0; c.pen_n = st.pen_n;
  c.ke_prev = st.ke_prev; c.resid = st.residual;
  if (d.proto) {   // the replay starts settling under its first schedule value (solver_gpu.js:296-304)
    c.cur_f = v0;
    if (d.proto == 1) { c.pf[0] = d.pax * v0; c.pf[1] = d.pay * v0; c.pf[2] = d.paz * v0; }
  }
  if (d.rigid) {
    // the reference point of the first evaluation: the centroid of the uploaded nodes (Floor.apply,
    // solver_cpu.js:590); the rigid velocities as the state carries them (_uploadFloor, solver_gpu.js:236-237)
    double s[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) for (int k = 0; k < 3; k++) s[k] += st.X[3 * i + k];
    for (int k = 0; k < 3; k++) { c.rc[k] = s[k] / n; c.rc_prev[k] = c.rc[k]; c.rv[k] = st.rb_v[k]; c.rw[k] = st.rb_w[k]; }
  }
  HIPCHK(hipMemcpy(d.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice));
  H.hist_read = 0;
}

// ------------------------------------------------------------------------------------------------ batch
struct BatchBase {
  virtual ~BatchBase() {}
  ws_ctx* ctx = nullptr;
  int B = 0, precision = 0, tile = 128, per_submit = 64, hist_len = 4096;
  int ctrl_chunk = 0;                   // strand partials per control workgroup (0: by size)
  int ctrl_mode = 0;                    // 0: the control kernel; 1: the strand launch runs the control step
  bool has_rigid = false;               // an instance that only the ground holds: the rigid step runs
  int cross_lanes = 0;                  // lanes per crossing (0: by size at every relist)
  int lanes = 2;                        // ... as the current work lists use it
  // the instances split into two groups of about equal work, each replaying its own iteration graphs on its own
  // stream: one group's latency-bound launches (crossings, control) and kernel boundaries overlap the other's
  // bandwidth-bound strand launch.  Instances never read each other's memory, so the results are those of one stream.
  int streams_opt = 0;                  // 0: by size (two from SPLIT_TILES strand tiles and two instances), 1 or 2
  int nstreams = 1;                     // ... as the current work lists use it
  hipStream_t stream = nullptr;         // the batch's stream (group 0); every call outside a step uses it alone
  hipStream_t stream2 = nullptr;        // group 1's stream during a step (joined back into `stream` before it ends)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // captured iteration sequences by length, per group: per_submit iterations, and every power of two below it, so a
  // step of n iterations replays whole per_submit graphs and then the binary pieces of the remainder — only live
  // iterations are launched
  std::map<int, hipGraphExec_t> graphs, graphs2;
  bool eager_graphs = true;             // build every length at the first step (the request server builds lazily)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ms_last = 0, ms_cross = 0, ms_strand = 0, ms_ctrl = 0;
  bool profile = false;
  int32_t* d_sel = nullptr;
  int32_t* d_cnt = nullptr;   // per-selected iteration counts (step with counts)
  virtual int upload(int b, const ws_model& md, const ws_state& st) = 0;
  virtual int update_params(int b, const ws_model& md) = 0;
  // n_iter iterations on the selected instances (all when sel is null), or counts[q] on sel[q] when counts is given
  virtual int step(int n_iter, const int32_t* sel, int nsel, int force, ws_control* out, const int32_t* counts = nullptr) = 0;
  virtual int read_control(ws_control* out) = 0;
  virtual int read_history(int b, double* out, int cap, int* n) = 0;
  virtual int read_outputs(int b, double* out, int cap, int* n) = 0;
  virtual int evaluate(double* F, double* T, double* I3) = 0;
  virtual int set_xq(int count, const double* X, const double* Q) = 0;
  // stage: device memory for the outputs the download computes on the device (section forces, moments, reactions,
  // a pending AL update), stage_cap bytes; null: buffers are allocated for the call
  virtual int download(int b, ws_state& out, char* stage = nullptr, size_t stage_cap = 0) = 0;
  virtual size_t download_stage_bytes(int b) = 0;
  virtual int internalize(int b) = 0;
  virtual void kernel_bench(int reps, double* out) = 0;
  virtual double alg_bytes() = 0;
  virtual double dev_bytes() = 0;
  // slots (the request server): an empty batch of `capacity` slots, instances added to and removed from slots
  virtual void create_slots(int capacity) = 0;
  virtual void add_instance(int slot, const ws_model& md, const ws_state& st) = 0;   // relist() afterwards
  virtual void remove_instance(int slot) = 0;                                         // relist() afterwards
  virtual bool occupied(int slot) const = 0;
  virtual void relist() = 0;
};

template <typename Real>
struct Batch : BatchBase {
  std::vector<std::unique_ptr<InstHost<Real>>> inst;
  Inst<Real>* d_inst = nullptr;
  TileItem* d_tiles = nullptr; int ntiles = 0;
  CrossItem* d_cross = nullptr; int ncross = 0;
  CtrlItem* d_ctrl = nullptr; int nctrl = 0;
  Ctl* d_ctls = nullptr;            // [B] control records
  std::vector<Ctl> h_ctls;
  std::vector<TileItem> tiles;

  ~Batch() override {
    drop_graphs();
    for (auto& h : inst) if (h) h->arena.release();
    for (void* p : scratch_ptr) if (p) (void)hipFree(p);
    if (d_inst) (void)hipFree(d_inst);
    if (d_tiles) (void)hipFree(d_tiles);
    if (d_cross) (void)hipFree(d_cross);
    if (d_ctrl) (void)hipFree(d_ctrl);
    if (d_sel) (void)hipFree(d_sel);
    if (d_cnt) (void)hipFree(d_cnt);
    if (d_ctls) (void)hipFree(d_ctls);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
  }

  void make_streams() {
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0)); HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  }

  void create(const ws_model* const* models, const ws_state* const* states, int nB) {
    B = nB;
    make_streams();
    // the control records of all instances sit side by side: one copy reads every instance's state
    HIPCHK(hipMalloc(&d_ctls, sizeof(Ctl) * std::max(B, 1)));
    HIPCHK(hipMemset(d_ctls, 0, sizeof(Ctl) * std::max(B, 1)));
    for (int b = 0; b < B; b++) {
      inst.emplace_back(new InstHost<Real>());
      build_instance(*inst.back(), *models[b], *states[b], tile, hist_len, d_ctls + b, ctrl_chunk, false, ctrl_mode != 0);
      upload_state(*inst.back(), *states[b]);
    }
    relist();
    HIPCHK(hipMalloc(&d_sel, sizeof(int32_t) * std::max(B, 1)));
    HIPCHK(hipMalloc(&d_cnt, sizeof(int32_t) * std::max(B, 1)));
  }
  // an empty batch of `capacity` slots: an empty slot has a zero descriptor and a finished control record, so every
  // kernel skips it
  void create_slots(int capacity) override {
    B = capacity;
    make_streams();
    HIPCHK(hipMalloc(&d_ctls, sizeof(Ctl) * std::max(B, 1)));
    std::vector<Ctl> idle(std::max(B, 1));
    for (auto& c : idle) { c = Ctl{}; c.done = 1; }
    HIPCHK(hipMemcpy(d_ctls, idle.data(), sizeof(Ctl) * idle.size(), hipMemcpyHostToDevice));
    inst.resize(B);
    HIPCHK(hipMalloc(&d_sel, sizeof(int32_t) * std::max(B, 1)));
    HIPCHK(hipMalloc(&d_cnt, sizeof(int32_t) * std::max(B, 1)));
    relist();
  }

  void add_instance(int slot, const ws_model& md, const ws_state& st) override {
    if (slot < 0 || slot >= B || inst[slot]) throw Fail{WS_ERR_ARG, "add_instance: slot out of range or occupied"};
    HIPCHK(hipStreamSynchronize(stream));
    std::unique_ptr<InstHost<Real>> h(new InstHost<Real>());
    build_instance(*h, md, st, tile, hist_len, d_ctls + slot, ctrl_chunk, false, ctrl_mode != 0);
    upload_state(*h, st);
    inst[slot] = std::move(h);
  }

  void remove_instance(int slot) override {
    if (slot < 0 || slot >= B || !inst[slot]) throw Fail{WS_ERR_ARG, "remove_instance: empty slot"};
    HIPCHK(hipStreamSynchronize(stream));
    inst[slot]->arena.release();
    inst[slot].reset();
    Ctl c{}; c.done = 1;
    HIPCHK(hipMemcpy(d_ctls + slot, &c, sizeof(Ctl), hipMemcpyHostToDevice));
  }

  bool occupied(int slot) const override { return slot >= 0 && slot < B && inst[slot] != nullptr; }

  InstHost<Real>& at(int b) {
    if (b < 0 || b >= B || !inst[b]) throw Fail{WS_ERR_ARG, "instance index out of range or empty slot"};
    return *inst[b];
  }


  // work lists and the device descriptor table
  // device work lists and descriptor table (buffers grow, never shrink: the server relists every round)
  template <typename T> void put_list(T*& dptr, int& cap, const std::vector<T>& v) {
    const int need = std::max(int(v.size()), 1);
    if (need > cap) {
      if (dptr) HIPCHK(hipFree(dptr));
      HIPCHK(hipMalloc(&dptr, sizeof(T) * need));
      cap = need;
    }
    if (!v.empty()) HIPCHK(hipMemcpy(dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
  int cap_tiles = 0, cap_cross = 0, cap_ctrl = 0, cap_inst = 0;

  // a contiguous range of the work lists: every item of the instances of one group (or of all of them)
  struct Range { int t0 = 0, nt = 0, c0 = 0, nc = 0, k0 = 0, nk = 0; bool rigid = false; };
  Range group[2];
  Range all_range() const {
    Range r; r.nt = ntiles; r.nc = ncross; r.nk = nctrl; r.rigid = has_rigid; return r;
  }
  // two streams from this many strand tiles: each group then still fills the chip (65,370 tiles of 256 variants:
  // 6.5 % faster per iteration with two than with one; four streams were no faster than one)
  static constexpr long SPLIT_TILES = 8192;

  void relist() override {
    tiles.clear();
    std::vector<CrossItem> cross;
    std::vector<CtrlItem> ctrl;
    long total_k = 0, total_t = 0;
    int ninst = 0;
    for (int b = 0; b < B; b++) if (inst[b]) { total_k += inst[b]->K; total_t += long(inst[b]->tiles.size()); ninst++; }
    lanes = cross_lanes > 0 ? cross_lanes : (total_k >= CROSS_ONE_LANE ? 1 : 2);
    nstreams = streams_opt > 0 ? streams_opt : (total_t >= SPLIT_TILES ? 2 : 1);
    if (ninst < 2) nstreams = 1;
    // groups of about equal strand tiles: largest instance first into the lighter group
    std::vector<int> grp(std::max(B, 1), 0);
    if (nstreams == 2) {
      std::vector<int> order;
      for (int b = 0; b < B; b++) if (inst[b]) order.push_back(b);
      std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return inst[a]->tiles.size() > inst[c]->tiles.size(); });
      long load[2] = {0, 0};
      for (int b : order) { const int g = load[1] < load[0] ? 1 : 0; grp[b] = g; load[g] += long(inst[b]->tiles.size()); }
    }
    const int per_blk = CROSS_THREADS / lanes;
    for (int g = 0; g < nstreams; g++) {
      Range& R = group[g];
      R = Range{};
      R.t0 = int(tiles.size()); R.c0 = int(cross.size()); R.k0 = int(ctrl.size());
      for (int b = 0; b < B; b++) {
        if (!inst[b] || grp[b] != g) continue;
        InstHost<Real>& H = at(b);
        H.d.tile0 = int(tiles.size());
        for (auto t : H.tiles) { t.inst = b; tiles.push_back(t); }
        H.d.cblk0 = int(cross.size());
        H.d.ncblk = (H.K + per_blk - 1) / per_blk;   // cpart holds H.ncblk ≥ this many blocks
        for (int q = 0; q < H.d.ncblk; q++) cross.push_back(CrossItem{b, q * per_blk, std::min(per_blk, H.K - q * per_blk), q});
        for (int q = 0; q < H.d.ctrl_chunks; q++) ctrl.push_back(CtrlItem{b, q, H.d.ctrl_chunks, 0});
        if (H.d.rigid) R.rigid = true;
      }
      R.nt = int(tiles.size()) - R.t0; R.nc = int(cross.size()) - R.c0; R.nk = int(ctrl.size()) - R.k0;
    }
    ntiles = int(tiles.size()); ncross = int(cross.size()); nctrl = int(ctrl.size());
    has_rigid = false;
    for (int b = 0; b < B; b++) if (inst[b] && inst[b]->d.rigid) has_rigid = true;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipStreamSynchronize(stream2));
    put_list(d_tiles, cap_tiles, tiles);
    put_list(d_cross, cap_cross, cross);
    put_list(d_ctrl, cap_ctrl, ctrl);
    std::vector<Inst<Real>> table(B);
    for (int b = 0; b < B; b++) table[b] = inst[b] ? inst[b]->d : Inst<Real>{};
    put_list(d_inst, cap_inst, table);
    drop_graphs();
  }


  // a batch of one instance launches the kernels' single-instance form (work item from the block index, descriptor
  // and control record of slot 0 loaded beside it)
  bool one_inst() const { return B == 1 && inst[0] != nullptr; }
  // the kernels over a range of the work lists (the single-instance forms take the whole batch of one)
  void launch_cross(hipStream_t s, const Range& R) {
    if (!R.nc) return;
    const bool one = one_inst();
    const int32_t K = one ? inst[0]->K : 0;
    const CrossItem* items = d_cross + R.c0;
    if (lanes == 1) {
      if (one) hipLaunchKernelGGL((k_cross<Real, 1, true>), dim3(R.nc), dim3(CROSS_THREADS), 0, s, d_inst, items, R.nc, d_ctls, K);
      else hipLaunchKernelGGL((k_cross<Real, 1, false>), dim3(R.nc), dim3(CROSS_THREADS), 0, s, d_inst, items, R.nc, d_ctls, K);
    } else {
      if (one) hipLaunchKernelGGL((k_cross<Real, 2, true>), dim3(R.nc), dim3(CROSS_THREADS), 0, s, d_inst, items, R.nc, d_ctls, K);
      else hipLaunchKernelGGL((k_cross<Real, 2, false>), dim3(R.nc), dim3(CROSS_THREADS), 0, s, d_inst, items, R.nc, d_ctls, K);
    }
  }
  template <int T, bool RB, int W, bool F> void strand_wf(hipStream_t s, const Range& R) {
    const TileItem* items = d_tiles + R.t0;
    if (one_inst()) hipLaunchKernelGGL((k_strand<Real, T, RB, true, W, F>), dim3(R.nt), dim3(T), 0, s, d_inst, items, R.nt, d_ctls);
    else hipLaunchKernelGGL((k_strand<Real, T, RB, false, W, F>), dim3(R.nt), dim3(T), 0, s, d_inst, items, R.nt, d_ctls);
  }
  template <int T, bool RB, int W> void strand_w(hipStream_t s, const Range& R) {
    if (ctrl_mode == 1) strand_wf<T, RB, W, true>(s, R); else strand_wf<T, RB, W, false>(s, R);
  }
  // the FP64 form of the 64-lane tile runs three resident wavefronts per SIMD once the batch has enough tiles to keep
  // them busy (measured: 10 % faster on the 30 m dome's 14k tiles, 4 % slower on the 10 m dome's 1,350)
  static constexpr int FP64_WIDE_TILES = 4096;
  template <int T, bool RB> void strand_as(hipStream_t s, const Range& R) {
    if (std::is_same<Real, double>::value && T == 64 && ntiles >= FP64_WIDE_TILES) strand_w<T, RB, (T == 64 && std::is_same<Real, double>::value) ? 3 : 1>(s, R);
    else strand_w<T, RB, 1>(s, R);
  }
  void launch_strand(hipStream_t s, const Range& R) {
    if (!R.nt) return;
    switch (tile) {
      case 64: if (R.rigid) strand_as<64, true>(s, R); else strand_as<64, false>(s, R); break;
      case 128: if (R.rigid) strand_as<128, true>(s, R); else strand_as<128, false>(s, R); break;
      default: if (R.rigid) strand_as<256, true>(s, R); else strand_as<256, false>(s, R); break;
    }
  }
  void launch_ctrl(hipStream_t s, const Range& R) {
    if (R.nk && ctrl_mode == 0) {
      const CtrlItem* items = d_ctrl + R.k0;
      if (one_inst()) hipLaunchKernelGGL((k_ctrl<Real, true>), dim3(R.nk), dim3(CTRL_THREADS), 0, s, d_inst, items, R.nk, d_ctls);
      else hipLaunchKernelGGL((k_ctrl<Real, false>), dim3(R.nk), dim3(CTRL_THREADS), 0, s, d_inst, items, R.nk, d_ctls);
    }
    if (R.rigid && R.nt) {
      const TileItem* items = d_tiles + R.t0;
      switch (tile) {
        case 64: hipLaunchKernelGGL((k_rigid<Real, 64>), dim3(R.nt), dim3(64), 0, s, d_inst, items, R.nt, d_ctls); break;
        case 128: hipLaunchKernelGGL((k_rigid<Real, 128>), dim3(R.nt), dim3(128), 0, s, d_inst, items, R.nt, d_ctls); break;
        default: hipLaunchKernelGGL((k_rigid<Real, 256>), dim3(R.nt), dim3(256), 0, s, d_inst, items, R.nt, d_ctls); break;
      }
    }
  }
  void launch_cross(hipStream_t s) { launch_cross(s, all_range()); }
  void launch_strand(hipStream_t s) { launch_strand(s, all_range()); }
  void launch_ctrl(hipStream_t s) { launch_ctrl(s, all_range()); }
  void launch_iteration(hipStream_t s, const Range& R) { launch_cross(s, R); launch_strand(s, R); launch_ctrl(s, R); }

  void drop_graphs() {
    for (auto* m : {&graphs, &graphs2}) {
      for (auto& g : *m) (void)hipGraphExecDestroy(g.second);
      m->clear();
    }
  }
  // the graph of `len` iterations of group g (the whole batch with one stream), captured on that group's stream,
  // instantiated and uploaded on first use
  hipGraphExec_t graph_of(int len, int g = 0) {
    auto& m = g ? graphs2 : graphs;
    auto f = m.find(len);
    if (f != m.end()) return f->second;
    hipStream_t s = g ? stream2 : stream;
    const Range R = nstreams == 1 ? all_range() : group[g];
    hipGraph_t gr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < len; it++) launch_iteration(s, R);
    HIPCHK(hipStreamEndCapture(s, &gr));
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCHK(r);
    m[len] = e;
    HIPCHK(hipGraphUpload(e, s));
    return e;
  }
  void ensure_graphs() {
    if (!eager_graphs || graphs.count(per_submit)) return;
    for (int g = 0; g < nstreams; g++) {
      graph_of(per_submit, g);
      for (int len = 1; len < per_submit; len *= 2) graph_of(len, g);
    }
    // one replay of every graph with no live iteration (stop_at = iter on every instance, `done` left alone: each
    // kernel exits after its first loads).  A graph's first launch costs more than its later ones even after
    // hipGraphUpload (measured on the 256 design variants: ~160 µs once for the 16-iteration graph); it is paid here
    // instead of inside the first step that replays that graph
    std::vector<int32_t> all;
    for (int b = 0; b < B; b++) if (inst[b]) all.push_back(b);
    if (all.empty()) return;
    HIPCHK(hipMemcpyAsync(d_sel, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_begin<Real>, dim3((all.size() + 63) / 64), dim3(64), 0, stream, d_ctls, d_sel, nullptr,
                       int(all.size()), 0, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    hipStream_t ss[2] = {stream, stream2};
    for (int g = 0; g < nstreams; g++)
      for (auto& e : (g ? graphs2 : graphs)) HIPCHK(hipGraphLaunch(e.second, ss[g]));
    for (int g = 0; g < nstreams; g++) HIPCHK(hipStreamSynchronize(ss[g]));
  }

  void fetch_ctls() {
    h_ctls.resize(std::max(B, 1));
    HIPCHK(hipMemcpy(h_ctls.data(), d_ctls, sizeof(Ctl) * B, hipMemcpyDeviceToHost));
  }

  bool all_done(const int32_t* sel, int nsel) {
    fetch_ctls();
    for (int q = 0; q < nsel; q++) {
      const Ctl& c = h_ctls[sel[q]];
      if (c.done == 0 && c.iter < c.stop_at) return false;
    }
    return true;
  }

  int step(int n_iter, const int32_t* sel_in, int nsel, int force, ws_control* out, const int32_t* counts = nullptr) override {
    std::vector<int32_t> sel;
    if (sel_in && nsel > 0) {
      sel.assign(sel_in, sel_in + nsel);
      for (int b : sel) if (b < 0 || b >= B || !inst[b]) throw Fail{WS_ERR_ARG, "instance index out of range or empty slot"};
    } else {
      if (counts) throw Fail{WS_ERR_ARG, "per-instance counts need an instance list"};
      for (int b = 0; b < B; b++) if (inst[b]) sel.push_back(b);
    }
    if (counts) { n_iter = 0; for (int q = 0; q < nsel; q++) n_iter = std::max(n_iter, counts[q]); }
    if (n_iter > 0 && !sel.empty()) {
      if (!profile) ensure_graphs();   // instantiated and uploaded before the timed region
      HIPCHK(hipMemcpyAsync(d_sel, sel.data(), sizeof(int32_t) * sel.size(), hipMemcpyHostToDevice, stream));
      if (counts) HIPCHK(hipMemcpyAsync(d_cnt, counts, sizeof(int32_t) * sel.size(), hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(k_begin<Real>, dim3((sel.size() + 63) / 64), dim3(64), 0, stream, d_ctls, d_sel, counts ? d_cnt : nullptr,
                         int(sel.size()), n_iter, force);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(ev0, stream));
      if (profile) {
        // direct launches with events around each kernel: per-kernel durations on this stream
        std::vector<hipEvent_t> ev(3 * n_iter + 1);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(ev[0], stream));
        for (int it = 0; it < n_iter; it++) {
          launch_cross(stream);
          HIPCHK(hipEventRecord(ev[3 * it + 1], stream));
          launch_strand(stream);
          HIPCHK(hipEventRecord(ev[3 * it + 2], stream));
          launch_ctrl(stream);
          HIPCHK(hipEventRecord(ev[3 * it + 3], stream));
        }
        HIPCHK(hipStreamSynchronize(stream));
        double sc = 0, ss = 0, sk = 0;
        for (int it = 0; it < n_iter; it++) {
          float a = 0, b2 = 0, c2 = 0;
          HIPCHK(hipEventElapsedTime(&a, ev[3 * it], ev[3 * it + 1]));
          HIPCHK(hipEventElapsedTime(&b2, ev[3 * it + 1], ev[3 * it + 2]));
          HIPCHK(hipEventElapsedTime(&c2, ev[3 * it + 2], ev[3 * it + 3]));
          sc += a; ss += b2; sk += c2;
        }
        ms_cross = sc / n_iter; ms_strand = ss / n_iter; ms_ctrl = sk / n_iter;
        for (auto& e : ev) (void)hipEventDestroy(e);
      } else {
        // with two groups, group 1's stream starts after k_begin and `stream` waits for it before the step ends
        const int ns = nstreams;
        hipStream_t ss[2] = {stream, stream2};
        if (ns == 2) { HIPCHK(hipEventRecord(ev_fork, stream)); HIPCHK(hipStreamWaitEvent(stream2, ev_fork, 0)); }
        const int full = n_iter / per_submit, rem = n_iter % per_submit;
        for (int r = 0; r < full; r++) {
          for (int g = 0; g < ns; g++) HIPCHK(hipGraphLaunch(graph_of(per_submit, g), ss[g]));
          if (full > 8 && (r % 8) == 7 && r + 1 < full) {
            for (int g = 0; g < ns; g++) HIPCHK(hipStreamSynchronize(ss[g]));
            if (all_done(sel.data(), int(sel.size()))) break;
          }
        }
        // the remainder in binary pieces, largest first
        int hb = 1;
        while (hb * 2 <= rem) hb *= 2;
        for (int len = hb; rem > 0 && len >= 1; len /= 2)
          if (rem & len) for (int g = 0; g < ns; g++) HIPCHK(hipGraphLaunch(graph_of(len, g), ss[g]));
        if (ns == 2) { HIPCHK(hipEventRecord(ev_join, stream2)); HIPCHK(hipStreamWaitEvent(stream, ev_join, 0)); }
      }
      HIPCHK(hipEventRecord(ev1, stream));
      HIPCHK(hipStreamSynchronize(stream));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
      ms_last = ms;
    }
    if (out) return read_control(out);
    return WS_OK;
  }

  static ws_control to_control(const Ctl& c) {
    ws_control o{};
    o.iter = c.iter; o.done = c.done; o.rollback = c.rollback; o.al_update = c.al_update; o.stop_at = c.stop_at; o.hist_n = c.hist_n;
    o.ke_prev = c.ke_prev; o.ke = c.ke; o.resid = c.resid; o.fref = c.fref; o.e_stretch = c.e_stretch; o.e_bend = c.e_bend;
    o.e_tors = c.e_tors; o.e_cont = c.e_cont; o.e_shear = c.e_shear; o.res_t = c.res_t; o.res_f = c.res_f;
    o.res_rb = c.res_rb;
    for (int k = 0; k < 3; k++) { o.rv[k] = c.rv[k]; o.rw[k] = c.rw[k]; o.rc[k] = c.rc[k]; }
    o.phase = c.phase; o.inc = c.inc;
    o.e_floor = c.e_floor; o.fn_floor = c.fn_floor;
    return o;
  }

  int read_control(ws_control* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    fetch_ctls();
    for (int b = 0; b < B; b++) out[b] = to_control(h_ctls[b]);
    return WS_OK;
  }

  // generalised forces of every instance at its current state (the evaluator of the reference's stability check,
  // stability.js:27-52): the crossing pass and the strand pass in output mode, node forces F [3n] (rod, loads,
  // contacts, ground; before the anchors), body-axis torques T [3m] and frame inertia I3 [3m], concatenated in
  // instance order.  The crossing pass advances the contact state as one iteration would; every arena and control
  // record is restored afterwards, so the batch is as before the call.
  int evaluate(double* F, double* T, double* I3) override {
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<int32_t> all;
    std::vector<int64_t> fofs(std::max(B, 1), 0), tofs(std::max(B, 1), 0);
    size_t total = 0, nF = 0, nT = 0;
    for (int b = 0; b < B; b++) if (inst[b]) {
      all.push_back(b); total += inst[b]->arena.size;
      fofs[b] = int64_t(nF); tofs[b] = int64_t(nT);
      nF += 3 * size_t(inst[b]->n); nT += 3 * size_t(inst[b]->m);
    }
    if (all.empty()) return WS_OK;
    // scratch kept by the batch between calls (grown when needed): the arena snapshot, the outputs, the offsets
    char* save = static_cast<char*>(eval_scratch(0, total));
    double* dF = static_cast<double*>(eval_scratch(1, sizeof(double) * nF));
    double* dT = static_cast<double*>(eval_scratch(2, sizeof(double) * nT));
    double* dI = static_cast<double*>(eval_scratch(3, sizeof(double) * nT));
    int64_t* dofs = static_cast<int64_t*>(eval_scratch(4, sizeof(int64_t) * 2 * fofs.size()));
    std::vector<int64_t> ofs(fofs); ofs.insert(ofs.end(), tofs.begin(), tofs.end());
    HIPCHK(hipMemcpyAsync(dofs, ofs.data(), sizeof(int64_t) * ofs.size(), hipMemcpyHostToDevice, stream));
    std::vector<Ctl> ctl_save(B);
    HIPCHK(hipMemcpy(ctl_save.data(), d_ctls, sizeof(Ctl) * B, hipMemcpyDeviceToHost));
    size_t off = 0;
    for (int b : all) { HIPCHK(hipMemcpyAsync(save + off, inst[b]->arena.base, inst[b]->arena.size, hipMemcpyDeviceToDevice, stream)); off += inst[b]->arena.size; }
    HIPCHK(hipMemcpyAsync(d_sel, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_begin<Real>, dim3((all.size() + 63) / 64), dim3(64), 0, stream, d_ctls, d_sel, nullptr, int(all.size()), 1, 1);
    launch_cross(stream);
    const int64_t* dfo = dofs; const int64_t* dto = dofs + fofs.size();
    switch (tile) {
      case 64: hipLaunchKernelGGL((k_evaluate<Real, 64>), dim3(ntiles), dim3(64), 0, stream, d_inst, d_tiles, ntiles, dF, dT, dI, dfo, dto); break;
      case 128: hipLaunchKernelGGL((k_evaluate<Real, 128>), dim3(ntiles), dim3(128), 0, stream, d_inst, d_tiles, ntiles, dF, dT, dI, dfo, dto); break;
      default: hipLaunchKernelGGL((k_evaluate<Real, 256>), dim3(ntiles), dim3(256), 0, stream, d_inst, d_tiles, ntiles, dF, dT, dI, dfo, dto); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(F, dF, sizeof(double) * nF, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(T, dT, sizeof(double) * nT, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(I3, dI, sizeof(double) * nT, hipMemcpyDeviceToHost, stream));
    off = 0;
    for (int b : all) { HIPCHK(hipMemcpyAsync(inst[b]->arena.base, save + off, inst[b]->arena.size, hipMemcpyDeviceToDevice, stream)); off += inst[b]->arena.size; }
    HIPCHK(hipMemcpyAsync(d_ctls, ctl_save.data(), sizeof(Ctl) * B, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return WS_OK;
  }

  void* scratch_ptr[7] = {};
  size_t scratch_cap[7] = {};
  void* eval_scratch(int k, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > scratch_cap[k]) {
      if (scratch_ptr[k]) HIPCHK(hipFree(scratch_ptr[k]));
      scratch_ptr[k] = nullptr; scratch_cap[k] = 0;
      HIPCHK(hipMalloc(&scratch_ptr[k], bytes));
      scratch_cap[k] = bytes;
    }
    return scratch_ptr[k];
  }

  // positions and frames of instances 0 .. count−1 (both buffers), the rest of their state as it is (the stability
  // check's perturbed copies): X and Q hold the instances' arrays one after the other; one copy to the device, one
  // scatter launch per instance
  int set_xq(int count, const double* X, const double* Q) override {
    if (count < 0 || count > B) throw Fail{WS_ERR_ARG, "set_positions: count out of range"};
    HIPCHK(hipStreamSynchronize(stream));
    size_t nx = 0, nq = 0;
    for (int b = 0; b < count; b++) { InstHost<Real>& H = at(b); nx += 3 * size_t(H.n); nq += 4 * size_t(H.m); }
    double* dX = static_cast<double*>(eval_scratch(5, sizeof(double) * nx));
    double* dQ = static_cast<double*>(eval_scratch(6, sizeof(double) * nq));
    HIPCHK(hipMemcpyAsync(dX, X, sizeof(double) * nx, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dQ, Q, sizeof(double) * nq, hipMemcpyHostToDevice, stream));
    size_t ox = 0, oq = 0;
    for (int b = 0; b < count; b++) {
      const InstHost<Real>& H = *inst[b];
      const int nt = std::max(H.n, H.m);
      hipLaunchKernelGGL(k_set_xq<Real>, dim3((nt + 255) / 256), dim3(256), 0, stream, d_inst, b, dX + ox, dQ + oq);
      ox += 3 * size_t(H.n); oq += 4 * size_t(H.m);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    return WS_OK;
  }

  // replay outputs recorded so far (readOutputs, solver_gpu.js:408-411): rows of travel, reaction, value, iteration
  int read_outputs(int b, double* out, int cap, int* nrows) override {
    InstHost<Real>& H = at(b);
    Ctl c;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&c, H.d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const int nr = H.d.proto ? std::max(0, std::min(c.inc, cap)) : 0;
    if (nr > 0 && out) HIPCHK(hipMemcpy(out, H.d.pout, sizeof(double) * 4 * nr, hipMemcpyDeviceToHost));
    if (nrows) *nrows = nr;
    return WS_OK;
  }

  int read_history(int b, double* out, int cap, int* nrows) override {
    InstHost<Real>& H = at(b);
    Ctl c;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(&c, H.d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    int avail = std::min(c.hist_n - H.hist_read, hist_len);
    int nr = std::max(0, std::min(avail, cap));
    if (nr > 0) {
      auto ring = d2h(H.d.hist, size_t(hist_len) * HIST_STRIDE);
      // the k-th returned row is iteration c.iter - (avail - 1) + k (readHistory, solver_gpu.js:389-405)
      int first = c.iter - avail + 1;
      for (int k = 0; k < nr; k++) {
        int it = first + k;
        std::memcpy(out + size_t(k) * HIST_STRIDE, ring.data() + size_t((it - 1) % hist_len) * HIST_STRIDE, sizeof(double) * HIST_STRIDE);
      }
    }
    H.hist_read = c.hist_n;
    *nrows = nr;
    return WS_OK;
  }

  // the device staging one download of instance b needs: the pending AL multipliers [K], N [m], V [2m], M [3p],
  // clamp moments [2C] and reactions [3n], each 256-byte aligned
  size_t download_stage_bytes(int b) override {
    const InstHost<Real>& H = at(b);
    size_t s = align256(sizeof(double) * std::max(H.K, 1)) + align256(sizeof(double) * std::max(H.m, 1)) +
               align256(sizeof(double) * std::max(2 * H.m, 1)) + align256(sizeof(double) * std::max(3 * H.p, 1)) +
               align256(sizeof(double) * std::max(2 * H.C, 1)) + align256(sizeof(double) * std::max(3 * H.n, 1));
    if (H.d.floor_on) s += align256(sizeof(double) * std::max(3 * H.n, 1)) + align256(sizeof(double) * std::max((H.n + 7) / 8, 1));
    return s;
  }

  int download(int b, ws_state& o, char* stage = nullptr, size_t stage_cap = 0) override {
    InstHost<Real>& H = at(b);
    Inst<Real>& d = H.d;
    const int n = H.n, m = H.m, p = H.p, K = H.K, C = H.C;
    if (stage && download_stage_bytes(b) > stage_cap)
      throw Fail{WS_ERR_CAPACITY, "download without allocation: the instance needs " + std::to_string(download_stage_bytes(b)) +
                                  " bytes of device staging, more than the " + std::to_string(stage_cap) + " the library holds for it"};
    // carves the staging region in the order download_stage_bytes counts it
    size_t stage_off = 0;
    auto carve = [&](size_t count) {
      double* q = reinterpret_cast<double*>(stage + stage_off);
      stage_off += align256(sizeof(double) * std::max<size_t>(count, 1));
      return q;
    };
    double* st_lam = stage ? carve(K) : nullptr;
    HIPCHK(hipStreamSynchronize(stream));
    Ctl c;
    HIPCHK(hipMemcpy(&c, d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    const int cur = c.cur;
    const bool rb = c.rollback != 0;
    if (o.X) HIPCHK(hipMemcpy(o.X, d.X[cur], sizeof(double) * 3 * size_t(n), hipMemcpyDeviceToHost));
    if (o.floor_a && d.floor_on) {
      auto ax = d2h(d.fax, n), ay = d2h(d.fay, n);
      for (int i = 0; i < n; i++) { o.floor_a[2 * i] = ax[i]; o.floor_a[2 * i + 1] = ay[i]; }
    }
    if (o.X0) std::copy(H.X0.begin(), H.X0.end(), o.X0);
    if (o.Q0) std::copy(H.Q0.begin(), H.Q0.end(), o.Q0);
    if (o.Q) HIPCHK(hipMemcpy(o.Q, d.Q[cur], sizeof(double) * 4 * size_t(m), hipMemcpyDeviceToHost));
    if (o.V) {
      if (rb) std::fill(o.V, o.V + 3 * n, 0.0);
      else { auto a = d2h(d.vx, n), bb = d2h(d.vy, n), cc = d2h(d.vz, n); for (int i = 0; i < n; i++) { o.V[3 * i] = a[i]; o.V[3 * i + 1] = bb[i]; o.V[3 * i + 2] = cc[i]; } }
    }
    if (o.W) {
      if (rb) std::fill(o.W, o.W + 3 * m, 0.0);
      else { auto a = d2h(d.wx, m), bb = d2h(d.wy, m), cc = d2h(d.wz, m); for (int j = 0; j < m; j++) { o.W[3 * j] = a[j]; o.W[3 * j + 1] = bb[j]; o.W[3 * j + 2] = cc[j]; } }
    }
    auto putR = [](double* dst, const Real* src, size_t nn) { if (dst && nn) { auto v = d2h(src, nn); for (size_t i = 0; i < nn; i++) dst[i] = double(v[i]); } };
    if (K) {
      if (o.u) { auto v = d2h(d.u, 2 * K); std::copy(v.begin(), v.end(), o.u); }
      if (o.slip) { auto v = d2h(d.slip, K); std::copy(v.begin(), v.end(), o.slip); }
      putR(o.n_prev, d.nprev, 3 * K);
      auto fl = d2h(d.kflags, K);
      if (o.stuck) for (int q = 0; q < K; q++) o.stuck[q] = (fl[q] >> 1) & 1;
      if (o.kstate) for (int q = 0; q < K; q++) o.kstate[q] = fl[q] & 1;
      putR(o.Fn, d.fn, K); putR(o.Ft, d.ft, K); putR(o.gap, d.gap, K);
      if (c.al_update && d.aug) {
        // the multiplier update decided at the end of the last iteration runs at the start of the next one here;
        // the host solver has already applied it (updateMultipliers, solver_cpu.js:499-506)
        double* dl = st_lam;
        if (!stage) HIPCHK(hipMalloc(&dl, sizeof(double) * K));
        hipLaunchKernelGGL(k_al_preview<Real>, dim3((K + 255) / 256), dim3(256), 0, stream, d_inst, b, dl);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        if (o.lam) HIPCHK(hipMemcpy(o.lam, dl, sizeof(double) * K, hipMemcpyDeviceToHost));
        if (!stage) (void)hipFree(dl);
        if (o.pen_acc) std::fill(o.pen_acc, o.pen_acc + K, 0.0);
      } else {
        if (o.lam) { auto v = d2h(d.lam, K); std::copy(v.begin(), v.end(), o.lam); }
        if (o.pen_acc) { auto v = d2h(d.penacc, K); std::copy(v.begin(), v.end(), o.pen_acc); }
      }
      if (o.theta0) { auto v = d2h(d.theta0, K); std::copy(v.begin(), v.end(), o.theta0); }
      if (o.rot_slip) { auto v = d2h(d.rotslip, K); std::copy(v.begin(), v.end(), o.rot_slip); }
      putR(o.dent, d.dent, 2 * K); putR(o.climb, d.climb, 2 * K); putR(o.tanpsi, d.tanpsi, 2 * K); putR(o.dent_e, d.dente, K);
      putR(o.theta, d.theta, K); putR(o.Anom, d.anom, K); putR(o.pressure, d.pres, K); putR(o.util, d.util, K);
      putR(o.Mrot, d.mrot, K); putR(o.phi, d.phi, 2 * K); putR(o.mid, d.mid, 3 * K);
      if (o.ctype || o.face_side) {
        auto ct = d2h(d.ctype, K), fs = d2h(d.face, K);
        if (o.ctype) for (int q = 0; q < K; q++) o.ctype[q] = ct[q];
        if (o.face_side) for (int q = 0; q < K; q++) o.face_side[q] = fs[q];
      }
    }
    // section forces, moments, reactions and the ground's forces of the last evaluated state
    const bool ground = d.floor_on && (o.floor_F || o.floor_state);
    if (o.N || o.Vsh || o.M || o.cM || o.reactions || ground) {
      double *dN = nullptr, *dV = nullptr, *dM = nullptr, *dC = nullptr, *dR = nullptr, *dG = nullptr;
      uint8_t* dGS = nullptr;
      const size_t bN = sizeof(double) * std::max(m, 1), bV = sizeof(double) * std::max(2 * m, 1), bM = sizeof(double) * std::max(3 * p, 1),
                   bC = sizeof(double) * std::max(2 * C, 1), bR = sizeof(double) * std::max(3 * n, 1), bGS = size_t(std::max(n, 1));
      if (stage) {
        // the staging region keeps the last call's values: clear the outputs (k_outputs leaves entries untouched)
        dN = carve(m); dV = carve(2 * m); dM = carve(3 * p); dC = carve(2 * C); dR = carve(3 * n);
        if (ground) { dG = carve(3 * n); dGS = reinterpret_cast<uint8_t*>(carve((n + 7) / 8)); }
      } else {
        HIPCHK(hipMalloc(&dN, bN)); HIPCHK(hipMalloc(&dV, bV)); HIPCHK(hipMalloc(&dM, bM)); HIPCHK(hipMalloc(&dC, bC));
        HIPCHK(hipMalloc(&dR, bR));
        if (ground) { HIPCHK(hipMalloc(&dG, bR)); HIPCHK(hipMalloc(&dGS, bGS)); }
      }
      HIPCHK(hipMemsetAsync(dN, 0, bN, stream)); HIPCHK(hipMemsetAsync(dV, 0, bV, stream)); HIPCHK(hipMemsetAsync(dM, 0, bM, stream));
      HIPCHK(hipMemsetAsync(dC, 0, bC, stream)); HIPCHK(hipMemsetAsync(dR, 0, bR, stream));
      if (ground) { HIPCHK(hipMemsetAsync(dG, 0, bR, stream)); HIPCHK(hipMemsetAsync(dGS, 0, bGS, stream)); }
      HIPCHK(hipStreamSynchronize(stream));
      if (c.live_flag) {
        const OutPtrs op{dN, dV, dM, dC, dR, nullptr, nullptr, nullptr, dG, dGS};
        switch (tile) {
          case 64: hipLaunchKernelGGL((k_outputs<Real, 64>), dim3(d.ntiles), dim3(64), 0, stream, d_inst, d_tiles, b, op); break;
          case 128: hipLaunchKernelGGL((k_outputs<Real, 128>), dim3(d.ntiles), dim3(128), 0, stream, d_inst, d_tiles, b, op); break;
          default: hipLaunchKernelGGL((k_outputs<Real, 256>), dim3(d.ntiles), dim3(256), 0, stream, d_inst, d_tiles, b, op); break;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
      }
      auto cp = [](double* dst, double* src, size_t nn) { if (dst && nn) HIPCHK(hipMemcpy(dst, src, nn * sizeof(double), hipMemcpyDeviceToHost)); };
      cp(o.N, dN, m); cp(o.Vsh, dV, 2 * m); cp(o.M, dM, 3 * p); cp(o.cM, dC, 2 * C); cp(o.reactions, dR, 3 * n);
      if (ground) {
        cp(o.floor_F, dG, 3 * n);
        if (o.floor_state && n) HIPCHK(hipMemcpy(o.floor_state, dGS, n, hipMemcpyDeviceToHost));
      }
      if (!stage) {
        (void)hipFree(dN); (void)hipFree(dV); (void)hipFree(dM); (void)hipFree(dC); (void)hipFree(dR);
        if (ground) { (void)hipFree(dG); (void)hipFree(dGS); }
      }
    }
    o.iteration = c.iter; o.ke_prev = c.ke_prev; o.residual = c.resid; o.pen_n = (c.al_update && d.aug && K) ? 0 : c.pen_n;
    o.converged = (c.done == 1 && c.resid <= H.tol) ? 1 : 0;
    o.res_force = c.res_f; o.res_torque = c.res_t; o.fref = c.fref; o.ke = c.ke; o.res_rigid = c.res_rb;
    o.e_stretch = c.e_stretch; o.e_shear = c.e_shear; o.e_bend = c.e_bend; o.e_twist = c.e_tors; o.e_contact = c.e_cont;
    o.e_floor = d.floor_on ? c.e_floor : 0.0; o.fn_floor = d.floor_on ? c.fn_floor : 0.0;
    // the rigid freedoms (Floor.rv / rw / c, solver_gpu.js:461) and the model radius the instance runs with
    for (int k = 0; k < 3; k++) {
      o.rb_v[k] = d.rigid ? c.rv[k] : 0.0; o.rb_w[k] = d.rigid ? c.rw[k] : 0.0; o.rb_c[k] = d.rigid ? c.rc[k] : 0.0;
    }
    o.rb_radius = d.rigid ? d.rb_R : 0.0;
    return WS_OK;
  }

  int upload(int b, const ws_model& md, const ws_state& st) override {
    InstHost<Real>& H = at(b);
    if (md.n != H.n || md.m != H.m || md.p != H.p || (md.par.contacts ? md.K : 0) != H.K || md.C != H.C || md.S != H.S)
      throw Fail{WS_ERR_TOPOLOGY, "upload: the model's sizes differ from the instance's"};
    HIPCHK(hipStreamSynchronize(stream));
    H.arena.release();
    H.arena = Arena();
    build_instance(H, md, st, tile, hist_len, d_ctls + b, ctrl_chunk, false, ctrl_mode != 0);
    upload_state(H, st);
    relist();
    return WS_OK;
  }

  int update_params(int b, const ws_model& md) override {
    InstHost<Real>& H = at(b);
    if (md.C != H.C || !std::equal(H.cseg.begin(), H.cseg.end(), md.c_seg))
      throw Fail{WS_ERR_TOPOLOGY, "the anchors changed the rotation clamps: rebuild the GPU solver"};
    if (md.n != H.n || md.m != H.m || md.p != H.p) throw Fail{WS_ERR_TOPOLOGY, "update_params: sizes changed"};
    HIPCHK(hipStreamSynchronize(stream));
    Inst<Real>& d = H.d;
    const int n = H.n, m = H.m, p = H.p, C = H.C, K = H.K;
    const ws_params& P = md.par;
    d.mass_scale = P.mass_scale; d.damping = P.damping; d.tol = P.tol; d.mu = P.mu; d.strength = P.strength; d.stick = P.stick_strength;
    d.torsion_factor = P.torsion_factor; d.epsilon = P.epsilon; d.slip_rate = P.slip_rate; d.rot_factor = P.rot_factor;
    d.dent_rate = P.dent_rate; d.edge_thr = P.edge_threshold; d.l_ref = md.l_ref; d.kin_damp = P.kinetic_damping;
    d.warmup = P.warmup_iter; d.ramp_iter = P.ramp_iter; d.rot_fric = P.rot_friction; d.seat = P.seat_law; d.seat_gap = P.seat_gap;
    d.dent_after = P.dent_after; d.aug = P.augmented; d.al_every = P.al_every;
    if ((P.floor_on != 0) != (d.floor_on != 0)) throw Fail{WS_ERR_TOPOLOGY, "the ground was switched on or off: rebuild the GPU solver"};
    d.floor_z = P.floor_z; d.floor_mu = std::max(P.floor_mu, 0.0);
    H.tol = P.tol;
    h2d(d.mass, toReal<Real>(md.mass, n));
    std::vector<Real> lx(n), ly(n), lz(n);
    for (int i = 0; i < n; i++) { lx[i] = Real(md.load_const[3 * i]); ly[i] = Real(md.load_const[3 * i + 1]); lz[i] = Real(md.load_const[3 * i + 2]); }
    h2d(d.lx, lx); h2d(d.ly, ly); h2d(d.lz, lz);
    std::vector<uint8_t> fx(n);
    for (int i = 0; i < n; i++) {
      fx[i] = fix_bits(md, i);
    }
    h2d(d.fix, fx);
    h2d(d.EA, toReal<Real>(md.EA, m)); h2d(d.GA, toReal<Real>(md.GA, m)); h2d(d.EIx, toReal<Real>(md.EIx, m));
    h2d(d.EIy, toReal<Real>(md.EIy, m)); h2d(d.GJ, toReal<Real>(md.GJ, m));
    upload_sections(d, md);
    std::vector<Real> kx(m), ky(m), kz(m);
    for (int j = 0; j < m; j++) { kx[j] = Real(md.krot_static[3 * j]); ky[j] = Real(md.krot_static[3 * j + 1]); kz[j] = Real(md.krot_static[3 * j + 2]); }
    h2d(d.krx, kx); h2d(d.kry, ky); h2d(d.krz, kz);
    h2d(d.om0, toReal<Real>(md.omega0, 3 * p));
    if (C) { h2d(d.c_B, toReal<Real>(md.c_B, 3 * C)); h2d(d.c_Q, std::vector<double>(md.c_Q, md.c_Q + 4 * C)); }
    if (K) h2d(d.k_lock, std::vector<uint8_t>(md.k_locked, md.k_locked + K));
    // a rebuild resets the kinetic-energy peak and the convergence flag (solver_gpu.js:331-332)
    Ctl c;
    HIPCHK(hipMemcpy(&c, d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    c.ke_prev = 0; c.done = 0;
    HIPCHK(hipMemcpy(d.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice));
    relist();
    return WS_OK;
  }

  int internalize(int b) override {
    InstHost<Real>& H = at(b);
    ws_state tmp{};
    std::vector<double> X(3 * H.n), Q(4 * H.m);
    tmp.X = X.data(); tmp.Q = Q.data();
    download(b, tmp);
    H.X0 = X; H.Q0 = Q;
    if (H.K) HIPCHK(hipMemset(H.d.slip, 0, sizeof(double) * H.K));
    return WS_OK;
  }

  void kernel_bench(int reps, double* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    // make every instance live for the timing launches (stop_at beyond all of them, convergence flag cleared)
    std::vector<int32_t> all;
    for (int b = 0; b < B; b++) if (inst[b]) all.push_back(b);
    if (all.empty()) { out[0] = out[1] = out[2] = 0; return; }
    HIPCHK(hipMemcpyAsync(d_sel, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_begin<Real>, dim3((all.size() + 63) / 64), dim3(64), 0, stream, d_ctls, d_sel, nullptr,
                       int(all.size()), 3 * reps + 8, 1);
    HIPCHK(hipGetLastError());
    for (int k = 0; k < 3; k++) {
      HIPCHK(hipEventRecord(ev0, stream));
      for (int r = 0; r < reps; r++) { if (k == 0) launch_cross(stream); else if (k == 1) launch_strand(stream); else launch_ctrl(stream); }
      HIPCHK(hipEventRecord(ev1, stream));
      HIPCHK(hipStreamSynchronize(stream));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
      out[k] = double(ms) / reps;
    }
  }
  double alg_bytes() override { double s = 0; for (auto& h : inst) if (h) s += h->alg_bytes; return s; }
  double dev_bytes() override { double s = 0; for (auto& h : inst) if (h) s += double(h->arena.size); return s; }
};

int fail(ws_ctx* ctx, const Fail& f) { if (ctx) ctx->err = f.msg; return f.code; }


// the batch of the given options (tile already resolved to 64 / 128 / 256)
std::unique_ptr<BatchBase> make_batch(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states, int B,
                                      int prec, int tile, int per_submit, int history, int ctrl_chunk = 0,
                                      int cross_lanes = 0, int ctrl_mode = 0, int streams = 0) {
  std::unique_ptr<BatchBase> b;
  if (prec == WS_FP64) b.reset(new Batch<double>()); else b.reset(new Batch<float>());
  b->ctx = ctx; b->precision = prec;
  if (per_submit > 0) b->per_submit = per_submit;
  if (history > 0) b->hist_len = history;
  b->ctrl_chunk = std::max(ctrl_chunk, 0);
  b->cross_lanes = cross_lanes;
  b->ctrl_mode = ctrl_mode;
  b->streams_opt = streams;
  b->tile = tile;
  if (prec == WS_FP64) static_cast<Batch<double>*>(b.get())->create(models, states, B);
  else static_cast<Batch<float>*>(b.get())->create(models, states, B);
  return b;
}

// the default strand tile: one wavefront (the neighbour exchange is DPP alone, no barrier between wavefronts, and the
// last tile of a lath wastes fewest lanes); fastest or tied with 128 / 256 on the 10 m dome, the 30 m dome and a
// batch of 64 design variants
constexpr int DEFAULT_TILE = 64;

// Autotuner: every candidate (strand tile × iterations per captured graph) relaxes the same inputs from the same
// state for `iters` iterations after a warm-up of two graph replays; the score is executed node-iterations per
// second of the timed region (HIP events), so an instance that stops early does not flatter a candidate.
void autotune(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states, int B, int prec, int iters,
              int history, std::vector<ws_tune_entry>& table, ws_batch_opts& best) {
  static const int tiles[3] = {64, 128, 256};
  static const int subs[3] = {32, 64, 128};
  double best_rate = -1;
  std::vector<ws_control> c0(B), c1(B);
  for (int t : tiles) for (int ps : subs) {
    auto b = make_batch(ctx, models, states, B, prec, t, ps, std::max(history, 1));
    b->eager_graphs = false;   // the candidate steps whole graphs only
    b->step(2 * ps, nullptr, 0, 1, c0.data());
    const int n = std::max(ps, (iters + ps - 1) / ps * ps);
    b->step(n, nullptr, 0, 1, c1.data());
    double node_iters = 0;
    for (int q = 0; q < B; q++) node_iters += double(models[q]->n) * double(c1[q].iter - c0[q].iter);
    const double ms = b->ms_last;
    const double rate = ms > 0 ? node_iters / (ms * 1e-3) : 0;
    table.push_back(ws_tune_entry{t, ps, ms / n, rate, b->lanes});
    if (rate > best_rate) { best_rate = rate; best.tile = t; best.per_submit = ps; best.cross_lanes = b->lanes; }
  }
  // the other crossing layout under the best tile and graph length
  {
    const int other = best.cross_lanes == 1 ? 2 : 1, t = best.tile, ps = best.per_submit;
    auto b = make_batch(ctx, models, states, B, prec, t, ps, std::max(history, 1), 0, other);
    b->eager_graphs = false;
    b->step(2 * ps, nullptr, 0, 1, c0.data());
    const int n = std::max(ps, (iters + ps - 1) / ps * ps);
    b->step(n, nullptr, 0, 1, c1.data());
    double node_iters = 0;
    for (int q = 0; q < B; q++) node_iters += double(models[q]->n) * double(c1[q].iter - c0[q].iter);
    const double ms = b->ms_last;
    const double rate = ms > 0 ? node_iters / (ms * 1e-3) : 0;
    table.push_back(ws_tune_entry{t, ps, ms / n, rate, other});
    if (rate > best_rate) { best_rate = rate; best.cross_lanes = other; }
  }
  best.precision = prec;
  best.history = history;
}

}  // namespace

