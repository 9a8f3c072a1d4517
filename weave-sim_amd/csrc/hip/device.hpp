// Device-side layout of the MI355X dynamic-relaxation solver (shared by the kernels and the host ABI code).
//
// One instance = one gridshell.  Its arrays live in one device arena, structure-of-arrays:
//   nodes     x, y, z (f64, double-buffered: [cur] is the state, [cur^1] receives the next iterate or keeps the
//             previous one for a kinetic-damping roll-back), vx, vy, vz (Real), mass, load (Real), fix bits (u8)
//   segments  qw, qx, qy, qz (f64, double-buffered), wx, wy, wz (Real), l0 (f64), EA, GA, EIx, EIy, GJ, krot (Real)
//   joints    jlen, Ω0 (Real)
//   crossings u (f64 ×2), slip (f64), nprev (Real ×3), flags (u8), contact outputs, side records
// Real is float in the FP32+FP64 form and double in the plain FP64 form.
#pragma once
#include <stdint.h>

namespace wsd {

constexpr int HIST_STRIDE = 9;   // history ring entry: ke, strain, bending, torsion, stretch, shear, contact, residual, iteration
constexpr int SPART = 9;         // doubles per strand-tile partial (a tenth, the largest ground Fn, when the ground is on)
constexpr int SPART_STRIDE = 10; // stored tile-major, 80 bytes: five 16-byte loads per tile
constexpr int CPART = 2;         // doubles per crossing-block partial (stored value-major)
constexpr int CTRL_THREADS = 256;  // lanes of the control kernel's workgroup
constexpr int CTRL_PER = 8;         // strand partials per lane per round (all loads of a round in flight together)
// up to CTRL_ONE strand partials one control workgroup reduces them all; beyond, workgroups of CTRL_CHUNK each (the
// last to finish combines them): one CU streams partials at ~80 GB/s, so a large instance's partials are spread
// over several CUs (measured on the 10 m and 30 m domes)
constexpr int CTRL_ONE = 2048;
constexpr int CTRL_CHUNK = 1024;
constexpr int CTRL_STAGE = 64;    // chunk partials the combining workgroup stages in LDS at a time
constexpr int RPART = 12;         // doubles per control-chunk partial
constexpr int FUSED_CHUNK = 256;  // strand tiles per chunk when the strand launch runs the control step (four partials
                                  // per lane of a 64-lane tile: one memory latency for the chunk's reduction)
// rigid-body sums per strand tile: resultant of loads + ground (3), its moment about the centroid (3), Σ d² of the
// ground-contact nodes per axis (3), their count (1), Σ x of the new positions (3)
constexpr int RBP = 13;
constexpr double RB_MASS = 2.0, RB_VMAX = 0.05, RB_LEVER = 0.05;   // solver_cpu.js:37-39
constexpr int CROSS_THREADS = 64;                 // lanes per crossing-kernel block (two per crossing)
constexpr int CROSS_BLOCK = CROSS_THREADS / 2;    // crossings per crossing-kernel block at two lanes per crossing
// crossings in a batch from which one lane takes both sides: from here the one-lane layout alone fills every SIMD
// at the kernel's two-wavefront occupancy (1,024 SIMDs × 2 × 64 lanes); below it the two-lane layout's shorter chains
// measured faster (10 m dome, 30 m dome, 64 variants)
constexpr long CROSS_ONE_LANE = 131072;
constexpr int COPY_PER = 2048;   // 16-byte elements per workgroup of the copy kernel (eight per lane)
constexpr int MAX_REL = 1024;    // crossing sides a strand tile may gather (checked on the host per strand)

// instance control record, device-written (the reference's CT struct, wgsl.js:154-172)
struct Ctl {
  int32_t iter, done, rollback, al_update;
  int32_t stop_at, hist_n, cur, counter;
  int32_t pen_n, live_flag, rb_move, pad2;   // rb_move: the rigid step of the iteration just evaluated is pending
  double ke_prev, ke, resid, fref;
  double e_stretch, e_bend, e_tors, e_cont, e_shear, res_t, res_f;
  // rigid-body freedoms of a structure only the ground holds (Floor, solver_cpu.js:546-551): the reference point
  // (centroid) the next iteration uses, the one the last iteration used, translation / rotation velocities, and the
  // rigid part of the residual
  double rc[3], rc_prev[3], rv[3], rw[3], res_rb;
  // UTM replay: phase (0 settle, 1 average, 2 finished), iterations in the phase, schedule values recorded,
  // samples averaged; the probe force (force control), the current schedule value, Σ probe travel per axis and
  // Σ reaction along the axis over the averaging window
  int32_t phase, it_phase, inc, acc_n;
  double pf[3], cur_f, ad[3], ar;
  double e_floor, fn_floor;   // the ground's energy and largest normal force of the last iteration (Floor.energy, FnMax)
};

// one crossing side's contribution of this iteration, written by the crossing kernel, gathered per strand tile
template <typename Real>
struct SideRec {
  int32_t seg_pre;      // segment (strand-local) the force acts on: the side's segment before this iteration's slip
  int32_t seg_post;     // segment after the slip: receives the righting stiffness
  Real t;               // position on seg_pre (weights 1 − t, t)
  Real kth;             // contact righting stiffness
  Real F[3];            // force on the side's centreline point (world)
  Real T[3];            // lever + rotational-friction torque (world)
};

template <typename Real>
struct Inst {
  // sizes
  int32_t n, m, p, K, C, S, npl, nclamp_entries;
  // work-list ranges of this instance
  int32_t tile0, ntiles, cblk0, ncblk;
  int32_t tile_waves, ctrl_chunks;  // wavefronts per strand tile; control workgroups of the instance
  int32_t ctrl_chunk, pad_c;        // strand partials per control workgroup
  // scalar parameters (ws_params)
  double mass_scale, damping, tol, mu, strength, stick, torsion_factor, epsilon, slip_rate, rot_factor, dent_rate;
  double edge_thr, l_ref, jlen_min, inv_jlen_min;   // inv_jlen_min: 1/l̄_min, 0 without joints
  int32_t kin_damp, warmup, contacts, ramp_iter, rot_fric, seat, seat_gap, dent_after, aug, al_every, hist_len, floor_on;
  int32_t rigid, pad_r;              // rigid: the ground holds the structure alone (no anchor, no clamp)
  int32_t sec_strand;                // every strand's segments share (EA, GA, EIx, EIy, GJ): read from sec, not per segment
  int32_t load_axes;                 // bit c: some node's constant load has a component along axis c (else not read)
  double floor_z, floor_mu;          // the ground plane
  double rb_R, rb_vmax, rb_wmax;     // rigid body: model radius about the centroid, speed limits per iteration
  int32_t proto, probe, n_inc, settle, avg, pad_p;   // UTM replay: mode (0 off, 1 force, 2 displacement), probe node, …
  double pax, pay, paz, p0x, p0y, p0z;               // … load axis, probe position at upload
  // topology
  const int32_t* node_off;   // [S+1]
  const int32_t* seg_off;    // [S+1]
  const int32_t* joint_off;  // [S+1]
  const uint8_t* closed;     // [S]
  const int32_t* seg_i0;     // [m]
  const int32_t* seg_i1;     // [m]
  const int32_t* nsa;        // [n] segment starting at the node, -1 if none
  const int32_t* nsb;        // [n] segment ending at the node, -1 if none
  const int32_t* side_off;   // [S+1] crossing sides per strand (CSR over the strand-ordered side records)
  const int32_t* side_pos;   // [2K] slot of side 2c + s in the strand-ordered side records
  const int32_t* clamp_off;  // [m+1] clamp joints per segment (CSR), null when C == 0
  const int32_t* clamp_idx;  // [C]
  // nodes
  double* X[2];              // [3n] positions, x y z interleaved (double buffer)
  Real* vx; Real* vy; Real* vz;
  const Real* mass;
  const Real* lx; const Real* ly; const Real* lz;            // constant loads
  const Real* llx; const Real* lly; const Real* llz;         // frame-local point loads (null when npl == 0)
  const uint8_t* fix;        // bit0 x, bit1 y, bit2 z fixed; bit3 any fixed (reaction); bit4 all three fixed
  // segments
  double* Q[2];              // [4m] frames, w x y z interleaved (double buffer)
  Real* wx; Real* wy; Real* wz;
  const double* l0;
  const Real* EA; const Real* GA; const Real* EIx; const Real* EIy; const Real* GJ;
  const Real* sec;           // [5S] EA, GA, EIx, EIy, GJ per strand (sec_strand)
  const Real* krx; const Real* kry; const Real* krz;
  // ground (floor_on): node section width / height, stick points (f64)
  const Real* nodew; const Real* nodeh;
  double* fax; double* fay;
  uint8_t* gstate;           // [n] the ground's state of the last iteration per node: 0 clear, 1 stick, 2 slip
  // joints
  const Real* jlen; const Real* om0;                          // om0 [3p]
  // clamps
  const Real* c_len; const Real* c_B; const double* c_Q;      // [C], [3C], [4C]
  // crossings (static)
  const int32_t* k_seg_off; const int32_t* k_nseg;            // [2K]
  const int32_t* k_noff; const int32_t* k_nn;                 // [2K] first node and node count (bit 30: closed) of the side's strand
  const Real* wk; const Real* hk;                             // [2K]
  const Real* fc90; const Real* E90; const Real* Hs;          // [2K]
  const Real* kw; const Real* kh;                             // [K]
  const uint8_t* k_en; uint8_t* k_lock;                       // [K]
  // crossings (state)
  double* u;                 // [2K]
  double* slip;              // [K]
  Real* nprev;               // [3K]
  uint8_t* kflags;           // [K] bit0 slip, bit1 stuck
  Real* fn; Real* ft; Real* gap;                              // [K]
  double* lam; double* penacc;                                // [K] AL multipliers and accumulated penetration (f64 in both forms)
  double* theta0; double* rotslip;                            // [K] W2 reference angle and rotational slip (f64)
  Real* mrot;                                                 // [K]
  Real* dent; Real* climb; Real* tanpsi; Real* dente;         // [2K], [2K], [2K], [K]
  // crossing classification outputs (W1)
  Real* theta; Real* anom; Real* pres; Real* util; Real* phi; Real* mid;   // [K], [K], [K], [K], [2K], [3K]
  int8_t* ctype; int8_t* face;                                // [K]
  SideRec<Real>* side;       // [2K] in strand order (side_off / side_pos)
  // control
  Ctl* ctl;
  double* hist;              // [hist_len * HIST_STRIDE]
  double* spart;             // [ntiles][SPART_STRIDE] one partial per strand tile
  double* cpart;             // [CPART][ncblk]
  double* rpart;             // [ctrl_chunks][RPART] control-chunk partials (ctrl_chunks > 1)
  int32_t* ccount;           // [ctrl_chunks] tiles of each chunk finished in this launch (the control step fused into
                             // the strand launch; zero between launches)
  double* rbpart;            // [RBP][ntiles] rigid-body sums per strand tile (rigid only)
  const double* sched;       // [n_inc] replay schedule (proto != 0)
  double* pout;              // [4 n_inc] replay outputs: travel, reaction, value, iteration
  double* preact;            // [3] the pinned probe's reaction of the last iteration (displacement control)
};

// output arrays of the strand tile in output mode (downloads, evaluations): N [m], V [2m], M [3p], clamp moments
// [2C], reactions [3n], the ground's force on every node G [3n] and its state GS [n] (ground on); for an evaluation
// also the node forces F [3n], body-axis torques T [3m] and frame inertia I3 [3m] (null otherwise)
struct OutPtrs { double *N, *V, *M, *CM, *R, *F, *T, *I3, *G; uint8_t* GS; };

// one strand tile: strand-local node range [k0, k1) plus the strand's layout, so that a tile reads its geometry in
// one 64-byte load
struct TileItem {
  int32_t inst, strand, k0, k1;
  int32_t noff, nn, soff, nseg;      // first node / node count, first segment / segment count of the strand
  int32_t joff, njs, so, ns;         // first joint / joint count, first side record / side count
  int32_t closed, tile_local, pad0, pad1;
};
struct CrossItem { int32_t inst, c0, count, blk; };  // crossings [c0, c0 + count) of one instance, block index in the instance
struct CtrlItem { int32_t inst, chunk, nchunks, pad; };   // one control workgroup: a chunk of an instance's partials

}  // namespace wsd
