"""ctypes mirror of the C ABI in include/weave.h and include/weave_model.h.

The structures here must match the C layouts field for field; tests/test_abi.py checks their sizes against the
sizes the libraries report.
"""
import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

d, i32, u32, i64, vp = C.c_double, C.c_int32, C.c_uint32, C.c_int64, C.c_void_p
P_d, P_i32, P_u8 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)


class ws_params(C.Structure):
    _fields_ = [(n, d) for n in ("mass_scale", "damping", "tol", "mu", "strength", "stick_strength", "torsion_factor",
                                 "epsilon", "slip_rate", "rot_factor", "dent_rate", "edge_threshold", "floor_z",
                                 "floor_mu")] + \
               [(n, i32) for n in ("kinetic_damping", "warmup_iter", "max_iter", "contacts", "ramp_iter",
                                   "rot_friction", "seat_law", "seat_gap", "dent_after", "augmented", "al_every",
                                   "floor_on")]


class ws_model(C.Structure):
    _fields_ = [(n, i32) for n in ("n", "m", "p", "K", "C", "S", "n_point_local")] + [
        ("l_ref", d), ("par", ws_params),
        ("node_offset", P_i32), ("seg_offset", P_i32), ("joint_offset", P_i32), ("strand_closed", P_u8),
        ("seg_i0", P_i32), ("seg_i1", P_i32), ("j_seg0", P_i32), ("j_seg1", P_i32), ("j_node", P_i32),
        ("l0", P_d), ("j_len", P_d),
        ("EA", P_d), ("GA", P_d), ("EIx", P_d), ("EIy", P_d), ("GJ", P_d), ("omega0", P_d),
        ("c_seg", P_i32), ("c_node", P_i32), ("c_len", P_d), ("c_B", P_d), ("c_Q", P_d),
        ("mass", P_d), ("krot_static", P_d), ("load_const", P_d), ("point_local_node", P_i32), ("point_local_F", P_d),
        ("node_free", P_u8),
        ("k_strand", P_i32), ("k_seg_off", P_i32), ("k_nseg", P_i32),
        ("k_wk", P_d), ("k_hk", P_d), ("k_fc90", P_d), ("k_E90", P_d), ("k_Hs", P_d), ("k_w", P_d), ("k_h", P_d),
        ("k_enabled", P_u8), ("k_locked", P_u8),
        ("node_w", P_d), ("node_h", P_d),
        ("proto_mode", i32), ("proto_probe", i32), ("proto_n", i32), ("proto_settle", i32), ("proto_avg", i32),
        ("proto_pad", i32), ("proto_axis", d * 3), ("proto_sched", P_d),
    ]


# (name, ctype, length key) of every array in ws_state; the length key is resolved against the model sizes
STATE_ARRAYS = [
    ("X", P_d, "3n"), ("X0", P_d, "3n"), ("V", P_d, "3n"), ("Q", P_d, "4m"), ("Q0", P_d, "4m"), ("W", P_d, "3m"),
    ("u", P_d, "2K"), ("n_prev", P_d, "3K"), ("slip", P_d, "K"), ("stuck", P_u8, "K"), ("lam", P_d, "K"),
    ("pen_acc", P_d, "K"), ("theta0", P_d, "K"), ("rot_slip", P_d, "K"), ("dent", P_d, "2K"), ("climb", P_d, "2K"),
    ("tanpsi", P_d, "2K"), ("dent_e", P_d, "K"), ("floor_a", P_d, "2n"),
    ("N", P_d, "m"), ("Vsh", P_d, "2m"), ("M", P_d, "3p"), ("cM", P_d, "2C"), ("reactions", P_d, "3n"),
    ("Fn", P_d, "K"), ("Ft", P_d, "K"), ("gap", P_d, "K"), ("kstate", P_d, "K"), ("theta", P_d, "K"),
    ("Anom", P_d, "K"), ("pressure", P_d, "K"), ("util", P_d, "K"), ("Mrot", P_d, "K"), ("phi", P_d, "2K"),
    ("mid", P_d, "3K"), ("ctype", P_i32, "K"), ("face_side", P_i32, "K"), ("floor_F", P_d, "3n"),
    ("floor_state", P_u8, "n"),
]


class ws_state(C.Structure):
    _fields_ = [(n, i32) for n in ("iteration", "converged", "pen_n", "pad0")] + \
               [(n, d) for n in ("ke_prev", "residual", "res_force", "res_torque", "fref", "ke", "e_stretch",
                                 "e_shear", "e_bend", "e_twist", "e_contact", "res_rigid")] + \
               [(n, t) for (n, t, _) in STATE_ARRAYS] + [("e_floor", d), ("fn_floor", d)] + \
               [("rb_v", d * 3), ("rb_w", d * 3), ("rb_c", d * 3), ("rb_radius", d)]


class ws_control(C.Structure):
    _fields_ = [(n, i32) for n in ("iter", "done", "rollback", "al_update", "stop_at", "hist_n", "phase", "inc")] + \
               [(n, d) for n in ("ke_prev", "ke", "resid", "fref", "e_stretch", "e_bend", "e_tors", "e_cont",
                                 "e_shear", "res_t", "res_f", "res_rb")] + \
               [("rv", d * 3), ("rw", d * 3), ("e_floor", d), ("fn_floor", d), ("rc", d * 3)]


class ws_batch_opts(C.Structure):
    _fields_ = [("precision", i32), ("per_submit", i32), ("tile", i32), ("history", i32), ("ctrl_chunk", i32),
                ("cross_lanes", i32), ("ctrl_mode", i32), ("streams", i32)]


WS_TILE_AUTOTUNE = -1


class ws_tune_entry(C.Structure):
    _fields_ = [("tile", i32), ("per_submit", i32), ("ms_per_iter", d), ("node_iters_per_s", d),
                ("cross_lanes", i32)]


class ws_server_opts(C.Structure):
    _fields_ = [("precision", i32), ("chunk", i32), ("tile", i32), ("max_slots", i32), ("hbm_budget", i64)]


class ws_result(C.Structure):
    _fields_ = [("id", i64), ("status", i32), ("iterations", i32), ("residual", d), ("t_submit_ms", d),
                ("t_admit_ms", d), ("t_done_ms", d), ("n", i32), ("K", i32)]


class wsm_group(C.Structure):
    _fields_ = [(n, d) for n in ("E", "G", "density", "E90", "fc90", "H", "width", "height", "Ix", "Iy", "J")] + \
               [("beam_advanced", i32)] + \
               [(n, d) for n in ("activation", "stiffness", "act_flat", "act_edge", "act_twist", "k_axial", "k_flat",
                                 "k_edge", "k_twist", "k_shear")]


class wsm_scene(C.Structure):
    _fields_ = [("grp", wsm_group), ("par", ws_params), ("contacts", i32), ("locked", i32),
                ("prestress_asbuilt", i32), ("subdivisions", i32), ("overhang", d), ("offset_half_height", i32),
                ("seed_strand_phase", i32), ("length_to_m", d), ("force_to_n", d), ("gravity", i32),
                ("gravity_dir", d * 3), ("udl", i32), ("udl_q", d * 3)]


class wsm_anchor(C.Structure):
    _fields_ = [("type", i32), ("s", i32), ("k", i32), ("fix", i32 * 6), ("pad", i32), ("radius", d)]


class wsm_point_load(C.Structure):
    _fields_ = [("s", i32), ("k", i32), ("local", i32), ("pad", i32), ("F", d * 3)]


class wsm_prep(C.Structure):
    _fields_ = [("noise", d), ("wavelength", d), ("seed", u32), ("pad", i32)]


class wsm_hex_spec(C.Structure):
    _fields_ = [("span", d), ("spacing", d), ("segment", d), ("contraction", d), ("noise", d), ("seed", u32),
                ("pad", i32)]


class wsm_info(C.Structure):
    _fields_ = [(n, i32) for n in ("n", "m", "p", "K", "C", "S", "nV", "nF", "nE", "subdivisions", "rings",
                                   "parity_conflicts")] + [("span_plan", d), ("rise", d)]


WSM_FRAME, WSM_ALL_START, WSM_ALL_END, WSM_ALL_ENDS, WSM_STRAND, WSM_BOUNDARY, WSM_REGION = range(7)
WS_FP32_FP64, WS_FP64 = 0, 1


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise OSError(f"{path} is not built: run __graft_entry__.build()")
    return C.CDLL(path)


_model_lib = None
_hip_lib = None


_load_lock = threading.Lock()     # the loaders below may be first called from several threads at once


def model_lib():
    """The host model builder (libweave_model.so)."""
    if _model_lib is None:
        with _load_lock:
            return _open_model_lib()
    return _model_lib


def hip_lib():
    """The device solver (libweave_hip.so).  Raises when it is not built: there is no CPU fallback."""
    if _hip_lib is None:
        with _load_lock:
            return _open_hip_lib()
    return _hip_lib


def _open_model_lib():
    global _model_lib
    if _model_lib is None:
        L = _load("libweave_model.so")
        L.wsm_scene_defaults.argtypes = [C.POINTER(wsm_scene)]
        L.wsm_scene_defaults.restype = None
        L.wsm_from_mesh.argtypes = [P_d, i32, P_i32, i32, i32, C.POINTER(wsm_scene), C.POINTER(wsm_anchor), i32,
                                    C.POINTER(wsm_point_load), i32, P_i32, i32, C.POINTER(wsm_prep), C.POINTER(vp)]
        L.wsm_from_mesh_groups.argtypes = [P_d, i32, P_i32, i32, i32, C.POINTER(wsm_scene), C.POINTER(wsm_anchor), i32,
                                           C.POINTER(wsm_point_load), i32, P_i32, i32, C.POINTER(wsm_prep),
                                           C.POINTER(wsm_group), i32, P_i32, i32, C.POINTER(vp)]
        L.wsm_from_mesh_pieces.argtypes = [P_d, i32, P_i32, i32, i32, C.POINTER(wsm_scene), C.POINTER(wsm_anchor), i32,
                                           C.POINTER(wsm_point_load), i32, P_i32, i32, C.POINTER(wsm_prep),
                                           C.POINTER(wsm_group), i32, P_i32, i32, P_i32, P_d, i32, P_i32, i32,
                                           C.POINTER(vp)]
        L.wsm_from_polylines.argtypes = [P_d, P_i32, i32, P_i32, P_i32, P_d, P_i32, i32, C.POINTER(wsm_scene),
                                         C.POINTER(wsm_anchor), i32, C.POINTER(wsm_point_load), i32, C.POINTER(vp)]
        L.wsm_hex_dome.argtypes = [C.POINTER(wsm_hex_spec), C.POINTER(wsm_scene), C.POINTER(vp)]
        L.wsm_get.argtypes = [vp, C.POINTER(ws_model), C.POINTER(ws_state)]
        L.wsm_info_get.argtypes = [vp, C.POINTER(wsm_info)]
        L.wsm_add_load.argtypes = [vp, i32, P_d]
        L.wsm_set_strength.argtypes = [vp, d, d]
        L.wsm_clone.argtypes = [vp, C.POINTER(vp)]
        L.wsm_segment_moments.argtypes = [C.POINTER(ws_model), P_d, P_d, P_d, P_d, P_d]
        L.wsm_stability_layout.argtypes = [C.POINTER(ws_model), P_d, i32, C.POINTER(i32), C.POINTER(i64), C.POINTER(i32),
                                           P_i32, P_i32, P_i32, C.POINTER(i64), P_i32, P_i32]
        L.wsm_knots.argtypes = [vp, P_i32, i32, C.POINTER(i32)]
        L.wsm_frames.argtypes = [vp, P_i32, i32, C.POINTER(i32)]
        L.wsm_lock_contacts.argtypes = [vp, P_i32, i32, i32]
        L.wsm_last_error.restype = C.c_char_p
        L.wsm_free.argtypes = [vp]
        L.wsm_free.restype = None
        _model_lib = L
    return _model_lib


def _open_hip_lib():
    global _hip_lib
    if _hip_lib is None:
        L = _load("libweave_hip.so")
        L.ws_ctx_create.argtypes = [i32, C.POINTER(vp)]
        L.ws_ctx_destroy.argtypes = [vp]
        L.ws_ctx_destroy.restype = None
        L.ws_last_error.argtypes = [vp]
        L.ws_last_error.restype = C.c_char_p
        L.ws_device_info.argtypes = [vp, C.c_char_p, i32, C.POINTER(i64), C.POINTER(i32)]
        L.ws_copy_probe.argtypes = [vp, i64, i32, P_d]
        L.ws_batch_create.argtypes = [vp, C.POINTER(C.POINTER(ws_model)), C.POINTER(C.POINTER(ws_state)), i32,
                                      C.POINTER(ws_batch_opts), C.POINTER(vp)]
        L.ws_batch_destroy.argtypes = [vp]
        L.ws_batch_size.argtypes = [vp, C.POINTER(i32)]
        L.ws_upload.argtypes = [vp, i32, C.POINTER(ws_model), C.POINTER(ws_state)]
        L.ws_update_params.argtypes = [vp, i32, C.POINTER(ws_model)]
        L.ws_step.argtypes = [vp, i32, P_i32, i32, i32, C.POINTER(ws_control)]
        L.ws_read_control.argtypes = [vp, C.POINTER(ws_control)]
        L.ws_read_history.argtypes = [vp, i32, P_d, i32, C.POINTER(i32)]
        L.ws_read_outputs.argtypes = [vp, i32, P_d, i32, C.POINTER(i32)]
        L.ws_evaluate.argtypes = [vp, P_d, P_d, P_d]
        L.ws_set_positions.argtypes = [vp, i32, P_d, P_d]
        L.ws_download.argtypes = [vp, i32, C.POINTER(ws_state)]
        L.ws_internalize.argtypes = [vp, i32]
        L.ws_download_noalloc.argtypes = [vp, i32, C.POINTER(ws_state)]
        L.ws_batch_timing.argtypes = [vp, P_d, P_d, P_d, P_d]
        L.ws_batch_bytes.argtypes = [vp, P_d, P_d]
        L.ws_batch_profile.argtypes = [vp, i32]
        L.ws_batch_kernel_bench.argtypes = [vp, i32, P_d, P_d, P_d]
        L.ws_batch_get_opts.argtypes = [vp, C.POINTER(ws_batch_opts)]
        L.ws_autotune.argtypes = [vp, C.POINTER(C.POINTER(ws_model)), C.POINTER(C.POINTER(ws_state)), i32, i32, i32,
                                  C.POINTER(ws_batch_opts), C.POINTER(ws_tune_entry), i32, C.POINTER(i32)]
        L.ws_server_create.argtypes = [vp, C.POINTER(ws_server_opts), C.POINTER(vp)]
        L.ws_server_destroy.argtypes = [vp]
        L.ws_submit.argtypes = [vp, C.POINTER(ws_model), C.POINTER(ws_state), i32, C.POINTER(i64)]
        L.ws_cancel.argtypes = [vp, i64]
        L.ws_pump.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
        L.ws_poll.argtypes = [vp, C.POINTER(ws_result), i32, C.POINTER(i32)]
        L.ws_fetch.argtypes = [vp, i64, C.POINTER(ws_state)]
        L.ws_server_clock_ms.argtypes = [vp]
        L.ws_server_clock_ms.restype = d
        L.ws_server_stats.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64), C.POINTER(i64)]
        _hip_lib = L
    return _hip_lib


# every symbol include/weave.h declares (checked by tests/test_abi.py)
WEAVE_H_SYMBOLS = [
    "ws_ctx_create", "ws_ctx_destroy", "ws_last_error", "ws_device_info", "ws_copy_probe", "ws_batch_create", "ws_batch_destroy",
    "ws_batch_size", "ws_upload", "ws_update_params", "ws_step", "ws_read_control", "ws_read_history", "ws_read_outputs", "ws_evaluate", "ws_set_positions", "ws_download",
    "ws_internalize", "ws_download_noalloc", "ws_batch_timing", "ws_batch_bytes", "ws_batch_profile", "ws_batch_kernel_bench",
    "ws_batch_get_opts", "ws_autotune", "ws_server_create", "ws_server_destroy", "ws_submit",
    "ws_cancel", "ws_pump", "ws_poll", "ws_fetch", "ws_server_clock_ms", "ws_server_stats",
]
WEAVE_MODEL_H_SYMBOLS = [
    "wsm_scene_defaults", "wsm_from_mesh", "wsm_from_mesh_groups", "wsm_from_mesh_pieces", "wsm_from_polylines", "wsm_hex_dome", "wsm_get", "wsm_info_get",
    "wsm_add_load", "wsm_set_strength", "wsm_clone", "wsm_mesh_counts", "wsm_last_error", "wsm_free",
    "wsm_abi_sizes", "wsm_segment_moments", "wsm_stability_layout", "wsm_knots", "wsm_frames", "wsm_lock_contacts",
]
ABI_STRUCTS = ["ws_params", "ws_model", "ws_state", "ws_control", "ws_batch_opts", "ws_server_opts", "ws_result",
               "wsm_scene", "wsm_anchor", "wsm_point_load", "wsm_prep", "wsm_hex_spec", "wsm_info", "ws_tune_entry"]
