/* weave_model.h — host-side model builder (C ABI), the caller side of the solver boundary.
 *
 * The reference builds every solver instance on the host before the GPU sees it:
 *   Mesh.fromArrays / weld / edgeTopology        web/js/mesh.js:8-163
 *   buildTopology (trace, parity, polylines,      web/js/topology.js:78-345
 *                  frames)
 *   buildState, Anchors, Loads, strandStiffness   web/js/model.js:10-252
 *   CosseratRod setup, PolylineContact setup      web/js/rod_cosserat.js:49-111, web/js/solver_cpu.js:62-139
 *   CpuSolver._buildGoals (masses, inertias)      web/js/solver_cpu.js:670-710
 * This library restates that pipeline in C++ (f64) so the same instance description (ws_model + ws_state,
 * weave.h) feeds the device solver, the CPU oracle and the N-API addon.  It also holds the synthetic
 * scene generator of the benchmark workloads (hexagonal kagome domes, seeded).
 */
#ifndef WEAVE_MODEL_H
#define WEAVE_MODEL_H
#include "weave.h"

/* lath group: MATERIAL, SECTION and BEAM of params.js:6-19 (the scene's own in wsm_scene.grp; the further groups of a
 * hybrid weave go to wsm_from_mesh_groups) */
typedef struct wsm_group {
  double E, G, density, E90, fc90, H;      /* N/mm², kg/m³ (density is SI, as units.js) */
  double width, height;                    /* section (mm) */
  double Ix, Iy, J;                        /* <= 0: derived from the rectangle (params.js:39-41) */
  int32_t beam_advanced;                   /* BEAM.mode == 'advanced' */
  double activation;                       /* simple mode */
  double stiffness;                        /* simple mode */
  double act_flat, act_edge, act_twist;    /* advanced mode */
  double k_axial, k_flat, k_edge, k_twist, k_shear;  /* advanced mode */
} wsm_group;

typedef struct wsm_scene {
  wsm_group grp;
  ws_params par;                           /* SOLVER / CONTACT / floor */
  int32_t contacts;                        /* CpuSolver option {contacts} */
  int32_t locked;                          /* CONTACT.locked: every crossing starts locked */
  int32_t prestress_asbuilt;               /* SOLVER.prestress == 'asbuilt' */
  /* WEAVE (params.js:24) */
  int32_t subdivisions;
  double overhang;
  int32_t offset_half_height;              /* offset_mode == 'half_height' */
  int32_t seed_strand_phase;
  /* units (units.js): scene length → m, scene force → N */
  double length_to_m, force_to_n;
  /* loads (scene.js:40) */
  int32_t gravity;
  double gravity_dir[3];
  int32_t udl;                             /* UDL: 0 off, 1 per length (model.js:215), 2 per area of the mesh (:216-219) */
  double udl_q[3];
} wsm_scene;

enum wsm_selector { WSM_FRAME = 0, WSM_ALL_START = 1, WSM_ALL_END = 2, WSM_ALL_ENDS = 3, WSM_STRAND = 4, WSM_BOUNDARY = 5,
                    WSM_REGION = 6 };
/* model.js:115-189.  WSM_FRAME: frame (s, k); WSM_STRAND: strand s; WSM_REGION: the frames of every knot whose mesh
 * edge lies within `radius` of the knot of frame (s, k), by graph distance along the mesh edges (regionFrames,
 * model.js:127-145; no frames for a model without a mesh) */
typedef struct wsm_anchor { int32_t type; int32_t s, k; int32_t fix[6]; int32_t pad; double radius; } wsm_anchor;
typedef struct wsm_point_load { int32_t s, k; int32_t local; int32_t pad; double F[3]; } wsm_point_load;

/* synthetic preparation applied after buildState and before the goals: seeded out-of-plane noise, a smooth field
 * z += noise · Σ w_k sin(2π (x cos θ_k + y sin θ_k) / λ_k + φ_k) / Σ w_k of 8 plane waves with λ_k in
 * [wavelength, 4·wavelength] (bounded by ±noise; drawn from xoshiro128**, rng.js) */
typedef struct wsm_prep {
  double noise;                            /* ± amplitude on node z (scene length) */
  double wavelength;                       /* shortest wavelength of the field (<= 0: white noise per node) */
  uint32_t seed;                           /* xoshiro128** seed (rng.js) */
  int32_t pad;
} wsm_prep;

/* hexagonal kagome dome of the benchmark workloads */
typedef struct wsm_hex_spec {
  double span;          /* developed width of the flat weave (scene length) */
  double spacing;       /* strand spacing of the flat weave */
  double segment;       /* target segment length */
  double contraction;   /* plan contraction of the developed width: sets the rise of the spherical cap */
  double noise;         /* ± out-of-plane noise on the initial nodes */
  uint32_t seed;
  int32_t pad;
} wsm_hex_spec;

typedef struct wsm_info {
  int32_t n, m, p, K, C, S;
  int32_t nV, nF, nE, subdivisions, rings, parity_conflicts;
  double span_plan;     /* plan diameter of the built model (for the 1e-4·span tolerance) */
  double rise;
} wsm_info;

typedef struct wsm_model wsm_model;

WS_API void wsm_scene_defaults(wsm_scene* sc);            /* params.js defaults, mm·N */
WS_API int wsm_from_mesh(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t weld,
                         const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                         const wsm_point_load* loads, int32_t n_loads, const int32_t* flip_knots, int32_t n_flips,
                         const wsm_prep* prep, wsm_model** out);
/* a hybrid weave (Scene.hybrid with strand assignments, scene.js:50-62): groups[0 .. n_groups) are lath groups
 * 2, 3, …; strand_group[2a], strand_group[2a+1] = strand id, group (n_assign pairs); an unassigned strand, or one
 * assigned outside 2 … 1 + n_groups, is in group 1 = sc->grp.  Each segment's stiffness, activation and self-weight,
 * each crossing side's section and material, each node's floor section and each knot's half-height lift follow
 * the group of the strand (model.js:203-252, solver_cpu.js:72-78 and 558-560, topology.js:225).  Per-piece groups
 * (Scene.hybrid.pieces): wsm_from_mesh_pieces. */
WS_API int wsm_from_mesh_groups(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t weld,
                                const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                                const wsm_point_load* loads, int32_t n_loads, const int32_t* flip_knots,
                                int32_t n_flips, const wsm_prep* prep, const wsm_group* groups, int32_t n_groups,
                                const int32_t* strand_group, int32_t n_assign, wsm_model** out);
/* a hybrid weave whose groups are also assigned to pieces of a discretisation (Scene.joints and Scene.hybrid.pieces,
 * scene.js:57-90; discretize.js:334-358): joint q sits at arc length joint_s[q] along strand joint_strand[q];
 * piece_group[3q..3q+2] = strand, piece (0-based, counted from the strand's start between its sorted joints),
 * group.  A segment takes the group of the piece that holds its midpoint, a node (floor section, crossing side
 * section) the piece that holds it, by arc length on the built positions; a piece not assigned, or assigned
 * outside 2 … 1 + n_groups, takes its strand's group.  With n_pieces = 0 the model is wsm_from_mesh_groups'. */
WS_API int wsm_from_mesh_pieces(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t weld,
                                const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                                const wsm_point_load* loads, int32_t n_loads, const int32_t* flip_knots,
                                int32_t n_flips, const wsm_prep* prep, const wsm_group* groups, int32_t n_groups,
                                const int32_t* strand_group, int32_t n_assign, const int32_t* joint_strand,
                                const double* joint_s, int32_t n_joints, const int32_t* piece_group, int32_t n_pieces,
                                wsm_model** out);
/* stand-alone laths (topologyFromPolylines, topology.js:348-363) with optional explicit crossings
 * (knots[4c..4c+3] = sA, kA, sB, kB frame ids) as the test fixtures of web/tests/cases.mjs build them */
WS_API int wsm_from_polylines(const double* P, const int32_t* counts, int32_t n_strands,
                              const int32_t* knot_rows, const int32_t* knot_counts,
                              const double* normals,          /* NULL = +z, else [3 * total nodes] */
                              const int32_t* knots, int32_t n_knots,
                              const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                              const wsm_point_load* loads, int32_t n_loads, wsm_model** out);
WS_API int wsm_hex_dome(const wsm_hex_spec* spec, const wsm_scene* sc, wsm_model** out);
WS_API int wsm_get(wsm_model* mdl, ws_model* model, ws_state* state);   /* pointers into the model's arrays */
WS_API int wsm_info_get(wsm_model* mdl, wsm_info* info);
WS_API int wsm_add_load(wsm_model* mdl, int32_t node, const double* F);   /* Loads.const[node] += F (tests) */
WS_API int wsm_set_strength(wsm_model* mdl, double strength, double stick_strength); /* re-derive masses (rebuildGoals) */
WS_API int wsm_clone(wsm_model* mdl, wsm_model** out);
/* segment moments of the results (buildResults, solver_cpu.js:878-891): node moments from the joint moments M [3p] and
 * the clamp moments cM [2C] of a downloaded state, averaged over each segment's two nodes, an end without a joint
 * (without a joint or clamp for Mx, My) taking its neighbour's value; Mx, My, T [m] */
WS_API int wsm_segment_moments(const ws_model* md, const double* M, const double* cM, double* Mx, double* My, double* T);
WS_API int wsm_mesh_counts(wsm_model* mdl, int32_t* nV, int32_t* nF);
/* the stability check's layout (dofList, stencil, colour; stability.js:17-23, 57-108): the DOFs in order (the free
 * components of every node, then the three rotations of every segment: kind 0 / 1, node or segment, component), each
 * DOF's stencil row (the nodes' and segments' DOFs within W along its lath; a crossing joins the windows of its two
 * sides within W + 2 of floor(u)) as CSR row_ptr [n_dofs + 1] / cols [nnz], and a greedy distance-2 colouring in DOF
 * order.  Call with dof_kind null for the sizes, then with arrays of at least those sizes (*n_dofs, *nnz in). */
WS_API int wsm_stability_layout(const ws_model* md, const double* u, int32_t W, int32_t* n_dofs, int64_t* nnz,
                                int32_t* n_colours, int32_t* dof_kind, int32_t* dof_index, int32_t* dof_comp,
                                int64_t* row_ptr, int32_t* cols, int32_t* colour);
/* the crossings' identities (st.knotEdges and topo.knots, model.js:49-63, topology.js:50, 145): per crossing its mesh
 * edge id (the key of the reference's knot results; explicit crossings of stand-alone laths are keyed by their
 * index), the frames [strand, k] of sides A and B, and the strand of the frame that passes over — the columns
 * knot_edge, strand_a, k_a, strand_b, k_b, over of the reference's knots.csv.  out [6 · min(cap_rows, K)] */
WS_API int wsm_knots(wsm_model* mdl, int32_t* out, int32_t cap_rows, int32_t* K);
/* the frames of the model (st.nodeFrameIds, model.js:24-25): rows (strand, frame k, node) in frame-key order, the
 * keys of the reference's reactions ("s,k", app.js:366); out [3 · min(cap_rows, n_frames)] */
WS_API int wsm_frames(wsm_model* mdl, int32_t* out, int32_t cap_rows, int32_t* n_frames);
/* CpuSolver.lockContacts (lock = 1) / unlockContacts (lock = 0) (solver_cpu.js:869-871): the crossings on the listed
 * mesh edges, all of them when edges is NULL; edges without a crossing are ignored.  ws_update_params takes the new
 * locks to the device (app.js:192 does the same with updateParams). */
WS_API int wsm_lock_contacts(wsm_model* mdl, const int32_t* edges, int32_t n_edges, int32_t lock);
WS_API const char* wsm_last_error(void);
/* sizeof of the ABI structs, for binding checks: ws_params, ws_model, ws_state, ws_control, ws_batch_opts,
 * ws_server_opts, ws_result, wsm_scene, wsm_anchor, wsm_point_load, wsm_prep, wsm_hex_spec, wsm_info */
WS_API int32_t wsm_abi_sizes(int64_t* out, int32_t cap);
WS_API void wsm_free(wsm_model* mdl);

#endif
