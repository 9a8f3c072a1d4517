/* weave.h — C ABI of the MI355X dynamic-relaxation solver for woven elastic-rod gridshells.
 *
 * This is the drop-in boundary for the reference's batched GPU solver class `GpuSolver`
 * (web/js/solver_gpu.js:62-523).  Every entry point below replaces one method of that class; the
 * citation after each declaration names it.  The JavaScript host binds these symbols through the
 * N-API addon in weave-sim_amd/csrc/napi (a `GpuSolver`-shaped class, weave-sim_amd/js/solver_mi355x.js);
 * Python binds them with ctypes (weave-sim_amd/__init__.py).
 *
 * Conventions
 *  - plain C: pointers and sizes only; every function returns a status (WS_OK = 0) and leaves a message
 *    for ws_last_error() on failure;
 *  - units and array layouts are those of the reference's host model (model.js:10-75): positions X[3n]
 *    (x, y, z per node), quaternions Q[4m] (w, x, y, z per segment), body angular velocities W[3m];
 *  - the host model (ws_model) is read on upload; device memory is owned by the library; downloads write
 *    into caller-owned buffers (ws_state), as GpuSolver.download overwrites the CpuSolver's typed arrays
 *    (solver_gpu.js:431-503);
 *  - calls on one batch or server handle are serialised by the library (a call made while another runs on a
 *    second thread waits for it); a context may hold many batches (they run on one HIP stream each).
 */
#ifndef WEAVE_H
#define WEAVE_H
#include <stdint.h>

#ifdef __cplusplus
#define WS_EXTERN extern "C"
#else
#define WS_EXTERN
#endif
#define WS_API WS_EXTERN __attribute__((visibility("default")))

enum ws_status {
  WS_OK = 0,
  WS_ERR_ARG = 1,       /* bad argument (the reference throws a JS Error) */
  WS_ERR_DEVICE = 2,    /* HIP runtime error */
  WS_ERR_NOMEM = 3,     /* device allocation failed */
  WS_ERR_TOPOLOGY = 4,  /* instances/clamps incompatible with the batch (solver_gpu.js:70,325) */
  WS_ERR_CAPACITY = 5,  /* admission refused: over the HBM budget or out of slots */
  WS_ERR_NOTFOUND = 6,  /* unknown request / instance */
  WS_ERR_STATE = 7      /* call out of order */
};

/* arithmetic of the device path */
enum ws_precision {
  WS_FP32_FP64 = 0,     /* FP32 forces/velocities; FP64 positions, frames, kinetic energy, residual */
  WS_FP64 = 1           /* plain FP64 everywhere */
};

/* Scalar parameters of one instance: the scene's SOLVER and CONTACT groups (params.js:8-23) with the
 * derived values the reference passes in its IP struct (solver_gpu.js:244-269). */
typedef struct ws_params {
  double mass_scale;      /* SOLVER.mass_scale   1.5 */
  double damping;         /* SOLVER.damping      1.0 */
  double tol;             /* SOLVER.tol          1e-4 */
  double mu;              /* CONTACT.mu          0.35 */
  double strength;        /* CONTACT.strength    1e4   (penalty k_n) */
  double stick_strength;  /* CONTACT.stick_strength 1e4 (tangential k_t) */
  double torsion_factor;  /* CONTACT.torsion_factor 1.0 */
  double epsilon;         /* CONTACT.epsilon     15.0  (slip step clip) */
  double slip_rate;       /* CONTACT.slip_rate   1.0 */
  double rot_factor;      /* CONTACT.rot_factor  1.0 */
  double dent_rate;       /* CONTACT.dent_rate   0.01 */
  double edge_threshold;  /* CONTACT.edge_threshold 0.5 */
  double floor_z;         /* scene.floor.z (floor_on only) */
  double floor_mu;        /* scene.floor.mu */
  int32_t kinetic_damping;/* SOLVER.kinetic_damping 1 */
  int32_t warmup_iter;    /* SOLVER.warmup_iter 300 */
  int32_t max_iter;       /* SOLVER.max_iter 5000 (host loop only) */
  int32_t contacts;       /* 1 when the instance has a crossing-contact goal (CpuSolver.contact != null) */
  int32_t ramp_iter;      /* CONTACT.ramp_iter 1500 */
  int32_t rot_friction;   /* CONTACT.rot_friction (W2) */
  int32_t seat_law;       /* CONTACT.seat_law != 'off' (W3) */
  int32_t seat_gap;       /* CONTACT.seat_gap */
  int32_t dent_after;     /* resolved CONTACT.dent_after_iter (PolylineContact.dentAfterDefault when < 0) */
  int32_t augmented;      /* CONTACT.augmented (AL) */
  int32_t al_every;       /* CONTACT.al_every 600 */
  int32_t floor_on;       /* scene.floor enabled */
} ws_params;

/* One instance as the host model holds it after goal setup (CpuSolver._buildGoals, solver_cpu.js:670-710).
 * Topology arrays follow buildState (model.js:10-75); strands are contiguous node/segment/joint ranges. */
typedef struct ws_model {
  int32_t n, m, p, K, C, S;           /* nodes, segments, joints, crossings, clamp joints, strands */
  int32_t n_point_local;              /* frame-local point loads (Loads.pointLocal) */
  double l_ref;                       /* min(l̄ ∪ l0): torque residual scale (CpuSolver.lRef) */
  ws_params par;
  /* strands */
  const int32_t* node_offset;         /* [S+1] */
  const int32_t* seg_offset;          /* [S+1] */
  const int32_t* joint_offset;        /* [S+1] */
  const uint8_t* strand_closed;       /* [S] */
  /* connectivity */
  const int32_t* seg_i0;              /* [m] first node of each segment */
  const int32_t* seg_i1;              /* [m] second node */
  const int32_t* j_seg0;              /* [p] */
  const int32_t* j_seg1;              /* [p] */
  const int32_t* j_node;              /* [p] */
  const double* l0;                   /* [m] rest lengths */
  const double* j_len;                /* [p] Voronoi lengths l̄ */
  /* rod (CosseratRod) */
  const double* EA; const double* GA; const double* EIx; const double* EIy; const double* GJ;  /* [m] */
  const double* omega0;               /* [3p] rest curvature Ω0 */
  const int32_t* c_seg;               /* [C] clamp joints (setClamps) */
  const int32_t* c_node;              /* [C] */
  const double* c_len;                /* [C] */
  const double* c_B;                  /* [3C] per-axis stiffness, 0 on a free axis */
  const double* c_Q;                  /* [4C] frozen frames */
  /* goals */
  const double* mass;                 /* [n] node masses ms·Σk */
  const double* krot_static;          /* [3m] static rotational stiffness (addRotStiffness) */
  const double* load_const;           /* [3n] constant loads (Loads.const) */
  const int32_t* point_local_node;    /* [n_point_local] */
  const double* point_local_F;        /* [3 n_point_local] force in the node frame */
  const uint8_t* node_free;           /* [3n] 1 = free DOF (Anchors.nodeFree) */
  /* crossings (PolylineContact) */
  const int32_t* k_strand;            /* [2K] strand of side A / B */
  const int32_t* k_seg_off;           /* [2K] first segment of that strand */
  const int32_t* k_nseg;              /* [2K] segments of that strand */
  const double* k_wk; const double* k_hk;                    /* [2K] section width / height per side */
  const double* k_fc90; const double* k_E90; const double* k_Hs;  /* [2K] material per side */
  const double* k_w; const double* k_h;                      /* [K] mean width / height */
  const uint8_t* k_enabled;           /* [K] */
  const uint8_t* k_locked;            /* [K] */
  /* floor (Floor goal): node sections */
  const double* node_w; const double* node_h;                /* [n] (floor_on only) */
  /* UTM load-test replay (md.protocol, solver_gpu.js:256-318; calibrate.js replayForce / replayDisplacement):
   * 0 off (relax to the residual), 1 force control (the probe node carries axis·value), 2 displacement control (the
   * probe node pinned at P0 + value·axis, P0 its uploaded position).  Per schedule value: settle (at most `settle`
   * iterations, ending early at the residual), then average the probe's travel and reaction along the axis over `avg`
   * iterations and record them (ws_read_outputs); the instance is done after the last value. */
  int32_t proto_mode, proto_probe, proto_n, proto_settle, proto_avg, proto_pad;
  double proto_axis[3];
  const double* proto_sched;          /* [proto_n] forces (mode 1) or probe displacements (mode 2) */
} ws_model;

/* Dynamic state and outputs of one instance: uploaded from / downloaded into caller buffers.
 * Arrays marked (out) may be NULL on download; state arrays are read on upload and written on download. */
typedef struct ws_state {
  int32_t iteration;        /* st.iteration */
  int32_t converged;        /* CpuSolver.converged */
  int32_t pen_n;            /* PolylineContact.penN (AL window length) */
  int32_t pad0;
  double ke_prev;           /* CpuSolver.kePrev */
  double residual;          /* st.residual (inf while warming up) */
  double res_force, res_torque, fref;               /* residualParts and F_ref of the last iteration (out) */
  double ke, e_stretch, e_shear, e_bend, e_twist, e_contact;  /* energies of the last iteration (out) */
  double res_rigid;         /* residualParts.rigid (out) */
  double *X, *X0, *V;       /* [3n] */
  double *Q, *Q0, *W;       /* [4m], [4m], [3m] */
  double *u;                /* [2K] contact parameters (segment index + fraction) */
  double *n_prev;           /* [3K] persistent contact normals */
  double *slip;             /* [K] accumulated slip length */
  uint8_t *stuck;           /* [K] */
  double *lam;              /* [K] AL multipliers */
  double *pen_acc;          /* [K] AL penetration window sums */
  double *theta0, *rot_slip;/* [K] W2 state (theta0 NaN = unset) */
  double *dent, *climb, *tanpsi; /* [2K] W3 state */
  double *dent_e;           /* [K] W3 state */
  double *floor_a;          /* [2n] floor stick points (x, y) */
  /* outputs */
  double *N, *Vsh;          /* [m], [2m] axial / shear forces (out) */
  double *M;                /* [3p] joint moments (out) */
  double *cM;               /* [2C] clamp moments (out) */
  double *reactions;        /* [3n] (out) */
  double *Fn, *Ft, *gap, *kstate;  /* [K] contact normal/friction force, gap, state 0 stick / 1 slip (out) */
  double *theta, *Anom, *pressure, *util, *Mrot;  /* [K] W1 classification, W2 moment (out) */
  double *phi;              /* [2K] (out) */
  double *mid;              /* [3K] crossing mid points (out) */
  int32_t *ctype, *face_side; /* [K] (out) */
  /* the ground (floor_on; Floor.F / state / energy / FnMax, filled by download, solver_gpu.js:454-461) */
  double *floor_F;          /* [3n] the ground's force on every node, world axes (out) */
  uint8_t *floor_state;     /* [n] 0 clear of the ground, 1 stick, 2 slip (out) */
  double e_floor, fn_floor; /* the ground's energy and largest normal force of the last iteration (out) */
  /* the rigid-body freedoms of a structure only the ground holds (Floor.rv / rw / c / Rb; the reference's upload
     writes rv and rw, solver_gpu.js:228-240, and its download reads all three back, :454-461).  An upload takes the
     velocities, and the model radius when positive (0: from the uploaded positions, as the Floor constructor
     computes it, solver_cpu.js:546-550); the reference point starts at the centroid of the uploaded positions.  A
     download returns all four; they are 0 for an instance without these freedoms. */
  double rb_v[3], rb_w[3];  /* rigid translation / rotation per iteration (in/out) */
  double rb_c[3];           /* reference point of the rigid rotation: the centroid of the positions (out) */
  double rb_radius;         /* model radius about the centroid (in/out; 0 on input: from X) */
} ws_state;

/* Per-instance control record, the reference's CT struct as readControl returns it (solver_gpu.js:374-386). */
typedef struct ws_control {
  int32_t iter, done, rollback, al_update, stop_at, hist_n, phase, inc;   /* phase / inc: replay settle 0, average 1,
                                                                             finished 2; schedule values recorded */
  double ke_prev, ke, resid, fref, e_stretch, e_bend, e_tors, e_cont, e_shear, res_t, res_f;
  double res_rb, rv[3], rw[3];   /* rigid-body freedoms (ground only): residual part, translation / rotation per iteration */
  double e_floor, fn_floor;      /* the ground's energy and largest normal force (eFloor, fnFloor) */
  double rc[3];                  /* reference point of the rigid rotation (rc; ground only, else 0) */
} ws_control;

typedef struct ws_batch_opts {
  int32_t precision;       /* ws_precision */
  int32_t per_submit;      /* iterations per captured graph (GpuSolver perSubmit, default 64) */
  int32_t tile;            /* lanes per strand tile: 64 / 128 / 256; 0 = default (64); -1 = autotune */
  int32_t history;         /* history ring length per instance (0 = 4096, the reference's HN) */
  int32_t ctrl_chunk;      /* strand partials per control reduction group (0 = by size: with a control launch one
                              workgroup up to 2048, else 1024 each, the last to finish combining them; 256 when the
                              strand launch runs the control step) */
  int32_t cross_lanes;     /* lanes per crossing in the contact kernel: 2 (the sides split, shorter chains), 1 (one lane
                              takes both sides, fewer instructions per crossing), 0 = by size (1 from 131,072
                              crossings in the batch) or the autotuner's pick */
  int32_t ctrl_mode;       /* 0: a control launch after the strand launch (the default); 1: the strand launch's last
                              tiles per instance run the control step, no control launch (the partials written
                              through the L2 with agent-scope stores and read with agent-scope loads; a UTM replay
                              instance releases them with device-scope fences).  Slower than 0 on gfx950 */
  int32_t streams;         /* 1: every launch on the batch's stream; 2: the instances in two groups of about equal
                              work, each replaying its own iteration graphs on its own stream (one group's
                              latency-bound launches overlap the other's strand launch; results as with one);
                              0 = by size (2 from two instances and 8,192 strand tiles) */
} ws_batch_opts;

#define WS_TILE_AUTOTUNE (-1)

/* one autotuner candidate: its options, mean ms per iteration and executed node-iterations per second */
typedef struct ws_tune_entry {
  int32_t tile;
  int32_t per_submit;
  double ms_per_iter;
  double node_iters_per_s;
  int32_t cross_lanes;
} ws_tune_entry;

typedef struct ws_ctx ws_ctx;
typedef struct ws_batch ws_batch;
typedef struct ws_server ws_server;

/* ---- device and context ------------------------------------------------------------------------ */
WS_API int ws_ctx_create(int32_t device, ws_ctx** out);      /* getDevice() (solver_gpu.js:17-29) */
WS_API void ws_ctx_destroy(ws_ctx* ctx);
WS_API const char* ws_last_error(const ws_ctx* ctx);         /* message of the last failed call */
WS_API int ws_device_info(ws_ctx* ctx, char* name, int32_t cap, int64_t* hbm_bytes, int32_t* n_cu);
/* a plain device-to-device copy kernel over `bytes` (read once, written once), `reps` launches: the calibration
 * of the HBM traffic counters and the achievable copy bandwidth; returns the mean ms per launch */
WS_API int ws_copy_probe(ws_ctx* ctx, int64_t bytes, int32_t reps, double* ms);

/* ---- batch (GpuSolver) ---------------------------------------------------------------------------
 * B instances, heterogeneous topologies allowed (the reference requires one topology per batch,
 * solver_gpu.js:70; here each instance keeps its own sizes).  Clamp segments are fixed per instance
 * for the life of the batch (updateParams throws otherwise, solver_gpu.js:325). */
WS_API int ws_batch_create(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states,
                           int32_t B, const ws_batch_opts* opts, ws_batch** out);  /* constructor :65-123 */
WS_API int ws_batch_destroy(ws_batch* batch);                                       /* dispose() :522 */
WS_API int ws_batch_size(ws_batch* batch, int32_t* B);
WS_API int ws_upload(ws_batch* batch, int32_t b, const ws_model* model, const ws_state* state); /* upload(b) :191-198 */
WS_API int ws_update_params(ws_batch* batch, int32_t b, const ws_model* model);    /* updateParams(b) :323-333 */
WS_API int ws_step(ws_batch* batch, int32_t n_iter, const int32_t* instances, int32_t n_inst,
                   int32_t force, ws_control* ctl_out);                             /* step(nIter, opts) :338-361 */
WS_API int ws_read_control(ws_batch* batch, ws_control* out);                       /* readControl() :374-386 */
WS_API int ws_read_history(ws_batch* batch, int32_t b, double* out, int32_t cap_rows, int32_t* n_rows); /* readHistory(b) :389-405 */
WS_API int ws_download(ws_batch* batch, int32_t b, ws_state* out);                  /* download(b) :431-503 */
WS_API int ws_read_outputs(ws_batch* batch, int32_t b, double* out, int32_t cap_rows, int32_t* n_rows);
                          /* readOutputs(b) :408-411: per recorded schedule value (travel, reaction, value, iteration) */
WS_API int ws_internalize(ws_batch* batch, int32_t b);                              /* internalize(b) :507-520 */
/* ws_download without device allocation — no counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same outputs as
 * ws_download, staged in one fixed region of WS_NOALLOC_STAGE_BYTES per GPU that the library holds from
 * ws_ctx_create until the process ends; an instance whose staging needs more (about 72 bytes per node, 97 with the
 * ground: instances up to ~930k nodes, ~690k with the ground) is refused with WS_ERR_CAPACITY.  Calls that overlap on one GPU share that region and race: one
 * host thread per GPU at a time. */
#define WS_NOALLOC_STAGE_BYTES (64LL << 20)
WS_API int ws_download_noalloc(ws_batch* batch, int32_t b, ws_state* out);
/* generalised forces of every instance at its current state — the evaluator of the reference's stability check
 * (stability.js:27-52, on the CPU there; no GpuSolver counterpart): node forces F [3n] (rod, constant and frame-local
 * loads, contacts, ground; before the anchors), body-axis torques T [3m] and the frame inertia I3 [3m] of that state,
 * concatenated in instance order.  The batch's state is unchanged by the call. */
WS_API int ws_evaluate(ws_batch* batch, double* F, double* T, double* I3);
/* positions and frames of instances 0 .. count−1, the rest of their state kept (the stability check's perturbed
 * copies; no counterpart in the reference): X holds each instance's [3n] one after the other, Q each [4m] */
WS_API int ws_set_positions(ws_batch* batch, int32_t count, const double* X, const double* Q);
WS_API int ws_batch_timing(ws_batch* batch, double* ms_last_step, double* ms_kernel_cross, double* ms_kernel_strand,
                           double* ms_kernel_ctrl);   /* hipEvent timing of the last step(); per-kernel means of the last profiled step */
WS_API int ws_batch_bytes(ws_batch* batch, double* alg_bytes_per_iter, double* device_bytes);  /* algorithmic HBM bytes per iteration, device footprint */
/* on = 1: step() launches the kernels one by one (no graph) and times each with hipEvents (ws_batch_timing) */
WS_API int ws_batch_profile(ws_batch* batch, int32_t on);
/* mean duration of one launch of each kernel (crossing, strand, control) from `reps` back-to-back launches of that
 * kernel alone, bracketed by two hipEvents on the batch's stream; for timing only: it advances the state */
WS_API int ws_batch_kernel_bench(ws_batch* batch, int32_t reps, double* ms_cross, double* ms_strand, double* ms_ctrl);
/* the options in effect (the tile chosen when the batch was created with tile 0 or WS_TILE_AUTOTUNE) */
WS_API int ws_batch_get_opts(ws_batch* batch, ws_batch_opts* out);
/* autotuner (no counterpart in the reference): relaxes the inputs `iters` iterations under every strand tile
 * (64 / 128 / 256) and graph length (32 / 64 / 128 iterations), then the other crossing layout (one or two lanes
 * per crossing) under the best of those, on temporary batches, and returns the fastest in node-iterations per
 * second; `table` (optional, `cap` rows) receives every candidate */
WS_API int ws_autotune(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states, int32_t B,
                       int32_t precision, int32_t iters, ws_batch_opts* best, ws_tune_entry* table, int32_t cap,
                       int32_t* n_entries);

/* ---- serving (continuous batching; no counterpart in the reference, whose batched caller is the calibration
 * loop, web/js/calibrate.js:438-520) ------------------------------------------------------------------------
 * Requests (one instance each) are queued by ws_submit and admitted in arrival order by ws_pump while a slot is
 * free and their device bytes fit the HBM budget; every pump relaxes all resident instances one chunk in shared
 * launches and retires the ones that converged, reached their iteration limit or went non-finite, writing the final
 * state into the request's ws_state.  Model and state stay caller-owned and valid until the result is polled.
 * Calls on one server are serialised like those on a batch. */
typedef struct ws_server_opts {
  int32_t precision;       /* ws_precision */
  int32_t chunk;           /* iterations per pump (0 = 64) */
  int32_t tile;            /* lanes per strand tile: 64 / 128 / 256 (0 = 64) */
  int32_t max_slots;       /* instances resident at once (0 = 256) */
  int64_t hbm_budget;      /* bytes of instance state the server may hold (0 = 80 % of the free memory) */
} ws_server_opts;

typedef struct ws_result {
  int64_t id;              /* request id returned by ws_submit */
  int32_t status;          /* 0 converged, 1 iteration limit reached, 2 cancelled, 3 non-finite state */
  int32_t iterations;      /* the instance's iteration count at the end */
  double residual;
  double t_submit_ms, t_admit_ms, t_done_ms;   /* server clock (ws_server_clock_ms) */
  int32_t n; int32_t K;
} ws_result;

WS_API int ws_server_create(ws_ctx* ctx, const ws_server_opts* opts, ws_server** out);
WS_API int ws_server_destroy(ws_server* srv);
/* max_iter: the request's iteration limit (0 = the model's par.max_iter); the request is priced (device bytes) and
 * validated here, WS_ERR_ARG when it can never fit the budget */
WS_API int ws_submit(ws_server* srv, const ws_model* model, ws_state* state, int32_t max_iter, int64_t* id_out);
WS_API int ws_cancel(ws_server* srv, int64_t id);          /* queued or resident; the result has status 2 */
WS_API int ws_pump(ws_server* srv, int32_t* n_active, int32_t* n_queued);   /* admit, relax one chunk, retire */
WS_API int ws_poll(ws_server* srv, ws_result* out, int32_t cap, int32_t* n_out);
WS_API int ws_fetch(ws_server* srv, int64_t id, ws_state* out);             /* snapshot of a resident request */
WS_API double ws_server_clock_ms(ws_server* srv);
WS_API int ws_server_stats(ws_server* srv, int32_t* n_active, int32_t* n_queued, int64_t* bytes_in_use, int64_t* hbm_budget);

#endif
