// Node N-API addon: the thin binding between the JavaScript host (weave-sim_amd/js/solver_mi355x.js, the GpuSolver
// interface of web/js/solver_gpu.js) and the C ABI (include/weave.h, include/weave_model.h).  Device calls that
// wait for the GPU (step, readControl, readHistory, download, server pump) run on libuv worker threads and resolve
// Promises, as the reference's mapAsync-based methods do; the rest are synchronous.  Model state arrays are exposed
// to JavaScript as typed arrays over the model's own memory, so a download overwrites them in place.
#include <node_api.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/weave.h"
#include "../../../include/weave_model.h"

namespace {

#define NAPI_CALL(env, call)                                                         \
  do {                                                                               \
    if ((call) != napi_ok) {                                                         \
      const napi_extended_error_info* ei = nullptr;                                  \
      napi_get_last_error_info((env), &ei);                                          \
      napi_throw_error((env), nullptr, ei && ei->error_message ? ei->error_message : "napi call failed"); \
      return nullptr;                                                                \
    }                                                                                \
  } while (0)

struct Model {
  wsm_model* h = nullptr;      // the C++ builder's model (null: a model over the caller's typed arrays)
  ws_model md{};
  ws_state st{};
  wsm_info info{};
  std::vector<double> sched;   // the replay schedule md.proto_sched points at
  std::vector<napi_ref> refs;  // the caller's typed arrays md / st point into (modelFromFields), kept alive
};

void model_finalize(napi_env env, void* p, void*) {
  Model* mm = static_cast<Model*>(p);
  for (napi_ref r : mm->refs) napi_delete_reference(env, r);
  if (mm->h) wsm_free(mm->h);
  delete mm;
}
struct Batch { ws_ctx* ctx = nullptr; ws_batch* b = nullptr; int32_t B = 0; };
struct Server { ws_ctx* ctx = nullptr; ws_server* s = nullptr; };

// ---------------------------------------------------------------------------------------------- helpers
napi_value undef(napi_env env) { napi_value v; napi_get_undefined(env, &v); return v; }
napi_value num(napi_env env, double x) { napi_value v; napi_create_double(env, x, &v); return v; }
napi_value arr3(napi_env env, const double* v) {
  napi_value a; napi_create_array_with_length(env, 3, &a);
  for (uint32_t k = 0; k < 3; k++) napi_set_element(env, a, k, num(env, v[k]));
  return a;
}
napi_value i32(napi_env env, int32_t x) { napi_value v; napi_create_int32(env, x, &v); return v; }
napi_value str(napi_env env, const char* s) { napi_value v; napi_create_string_utf8(env, s, NAPI_AUTO_LENGTH, &v); return v; }
void set(napi_env env, napi_value o, const char* k, napi_value v) { napi_set_named_property(env, o, k, v); }

bool get_args(napi_env env, napi_callback_info info, size_t n, napi_value* argv) {
  size_t argc = n;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  for (size_t i = argc; i < n; i++) argv[i] = undef(env);
  return true;
}
template <typename T> T* ext(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return static_cast<T*>(p);
}
double getd(napi_env env, napi_value o, const char* k, double dflt) {
  bool has = false; napi_value v; double x = dflt;
  if (napi_has_named_property(env, o, k, &has) == napi_ok && has && napi_get_named_property(env, o, k, &v) == napi_ok) {
    napi_valuetype t; napi_typeof(env, v, &t);
    if (t == napi_number) napi_get_value_double(env, v, &x);
    else if (t == napi_boolean) { bool b; napi_get_value_bool(env, v, &b); x = b ? 1 : 0; }
  }
  return x;
}
bool is_object(napi_env env, napi_value v) { napi_valuetype t; napi_typeof(env, v, &t); return t == napi_object; }

napi_value throw_status(napi_env env, int rc, const char* what) {
  std::string m = std::string(what) + " failed with status " + std::to_string(rc);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

napi_value ctl_object(napi_env env, const ws_control& c) {
  napi_value o; napi_create_object(env, &o);
  set(env, o, "iter", i32(env, c.iter)); set(env, o, "done", i32(env, c.done)); set(env, o, "rollback", i32(env, c.rollback));
  set(env, o, "alUpdate", i32(env, c.al_update)); set(env, o, "stopAt", i32(env, c.stop_at)); set(env, o, "histN", i32(env, c.hist_n));
  set(env, o, "phase", i32(env, c.phase)); set(env, o, "inc", i32(env, c.inc));
  set(env, o, "kePrev", num(env, c.ke_prev)); set(env, o, "ke", num(env, c.ke)); set(env, o, "resid", num(env, c.resid));
  set(env, o, "fref", num(env, c.fref)); set(env, o, "eStretch", num(env, c.e_stretch)); set(env, o, "eBend", num(env, c.e_bend));
  set(env, o, "eTors", num(env, c.e_tors)); set(env, o, "eCont", num(env, c.e_cont)); set(env, o, "eShear", num(env, c.e_shear));
  set(env, o, "resT", num(env, c.res_t)); set(env, o, "resF", num(env, c.res_f)); set(env, o, "resRb", num(env, c.res_rb));
  napi_value rv, rw; napi_create_array_with_length(env, 3, &rv); napi_create_array_with_length(env, 3, &rw);
  for (uint32_t k = 0; k < 3; k++) { napi_set_element(env, rv, k, num(env, c.rv[k])); napi_set_element(env, rw, k, num(env, c.rw[k])); }
  set(env, o, "rv", rv); set(env, o, "rw", rw);
  set(env, o, "eFloor", num(env, c.e_floor)); set(env, o, "fnFloor", num(env, c.fn_floor));
  set(env, o, "rc", arr3(env, c.rc));
  return o;
}

// ---------------------------------------------------------------------------------------------- async work
struct Job {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  std::function<int()> run;                               // worker thread: C ABI calls only
  std::function<napi_value(napi_env)> done;               // main thread: builds the resolved value
  std::function<std::string()> error;                     // main thread: message of a failed run
  std::vector<napi_ref> keep;                             // the handles (batch, server, models) the run uses
  int rc = 0;
};

void job_execute(napi_env, void* data) { Job* j = static_cast<Job*>(data); j->rc = j->run(); }
void job_complete(napi_env env, napi_status, void* data) {
  Job* j = static_cast<Job*>(data);
  if (j->rc == 0) {
    napi_resolve_deferred(env, j->deferred, j->done(env));
  } else {
    napi_value msg, err;
    std::string m = j->error ? j->error() : ("status " + std::to_string(j->rc));
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_async_work(env, j->work);
  for (napi_ref r : j->keep) napi_delete_reference(env, r);
  delete j;
}
// queue the job on a libuv worker; `keep` are the JS handles whose native objects the run uses: referenced until
// the job completes, so a handle dropped meanwhile is not finalised under a running call
napi_value queue_job(napi_env env, Job* j, const char* name, std::initializer_list<napi_value> keep = {}) {
  for (napi_value v : keep) {
    napi_ref r = nullptr;
    if (napi_create_reference(env, v, 1, &r) == napi_ok) j->keep.push_back(r);
  }
  napi_value promise, res;
  NAPI_CALL(env, napi_create_promise(env, &j->deferred, &promise));
  NAPI_CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &res));
  NAPI_CALL(env, napi_create_async_work(env, nullptr, res, job_execute, job_complete, j, &j->work));
  NAPI_CALL(env, napi_queue_async_work(env, j->work));
  return promise;
}
std::function<std::string()> ctx_error(ws_ctx* ctx) { return [ctx]() { return std::string(ws_last_error(ctx)); }; }

// ---------------------------------------------------------------------------------------------- device
napi_value DeviceCreate(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  int32_t dev = 0; napi_get_value_int32(env, a[0], &dev);
  ws_ctx* ctx = nullptr;
  if (ws_ctx_create(dev, &ctx) != WS_OK) { napi_value n; napi_get_null(env, &n); return n; }   // getDevice() → null
  napi_value v; NAPI_CALL(env, napi_create_external(env, ctx, [](napi_env, void* p, void*) { ws_ctx_destroy(static_cast<ws_ctx*>(p)); }, nullptr, &v));
  return v;
}
napi_value DeviceInfo(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  ws_ctx* ctx = ext<ws_ctx>(env, a[0]);
  char name[256] = {0}; int64_t hbm = 0; int32_t ncu = 0;
  int rc = ws_device_info(ctx, name, 256, &hbm, &ncu);
  if (rc) return throw_status(env, rc, "ws_device_info");
  napi_value o; napi_create_object(env, &o);
  set(env, o, "name", str(env, name)); set(env, o, "hbmBytes", num(env, double(hbm))); set(env, o, "cu", i32(env, ncu));
  return o;
}

// ---------------------------------------------------------------------------------------------- host models
// three numbers under key k: an array or a typed array of length 3 (absent or another shape: out unchanged)
void get3(napi_env env, napi_value o, const char* k, double* out) {
  bool has = false; napi_value v; bool is = false, ta = false; uint32_t n = 0;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has || napi_get_named_property(env, o, k, &v) != napi_ok) return;
  if (napi_is_array(env, v, &is) == napi_ok && is) {
    if (napi_get_array_length(env, v, &n) != napi_ok) return;
  } else if (napi_is_typedarray(env, v, &ta) == napi_ok && ta) {
    size_t len = 0;
    if (napi_get_typedarray_info(env, v, nullptr, &len, nullptr, nullptr, nullptr) != napi_ok) return;
    n = uint32_t(len);
  }
  if (n != 3) return;
  for (uint32_t i = 0; i < 3; i++) {
    napi_value e; double x = 0;
    if (napi_get_element(env, v, i, &e) == napi_ok && napi_get_value_double(env, e, &x) == napi_ok) out[i] = x;
  }
}

// a lath group's material, section and beam factors, over what g holds
void group_from(napi_env env, napi_value o, wsm_group& g) {
  if (!is_object(env, o)) return;
#define D(field, key) g.field = getd(env, o, key, g.field)
  D(E, "E"); D(G, "G"); D(density, "density"); D(width, "width"); D(height, "height");
  D(E90, "E90"); D(fc90, "fc90"); D(H, "H");
  D(Ix, "Ix"); D(Iy, "Iy"); D(J, "J"); g.beam_advanced = int32_t(getd(env, o, "beam_advanced", g.beam_advanced));
  D(activation, "activation"); D(stiffness, "stiffness"); D(act_flat, "act_flat"); D(act_edge, "act_edge");
  D(act_twist, "act_twist"); D(k_axial, "k_axial"); D(k_flat, "k_flat"); D(k_edge, "k_edge");
  D(k_twist, "k_twist"); D(k_shear, "k_shear");
#undef D
}

void scene_from(napi_env env, napi_value o, wsm_scene& sc) {
  wsm_scene_defaults(&sc);
  if (!is_object(env, o)) return;
  group_from(env, o, sc.grp);
#define D(field, key) sc.field = getd(env, o, key, sc.field)
#define I(field, key) sc.field = int32_t(getd(env, o, key, sc.field))
  D(par.mass_scale, "mass_scale"); D(par.damping, "damping"); D(par.tol, "tol"); D(par.mu, "mu");
  D(par.strength, "strength"); D(par.stick_strength, "stick_strength"); D(par.torsion_factor, "torsion_factor");
  D(par.epsilon, "epsilon"); D(par.slip_rate, "slip_rate"); D(par.rot_factor, "rot_factor"); D(par.dent_rate, "dent_rate");
  D(par.edge_threshold, "edge_threshold");
  I(par.kinetic_damping, "kinetic_damping"); I(par.warmup_iter, "warmup_iter"); I(par.max_iter, "max_iter");
  I(par.ramp_iter, "ramp_iter"); I(par.rot_friction, "rot_friction"); I(par.seat_law, "seat_law"); I(par.seat_gap, "seat_gap");
  I(par.dent_after, "dent_after"); I(par.augmented, "augmented"); I(par.al_every, "al_every");
  I(contacts, "contacts"); I(locked, "locked"); I(gravity, "gravity");
  // the rest of a scene (scene.json through the JS reader): ground, weave, units, UDL; vectors as 3-element arrays
  I(par.floor_on, "floor_on"); D(par.floor_z, "floor_z"); D(par.floor_mu, "floor_mu");
  I(prestress_asbuilt, "prestress_asbuilt"); I(subdivisions, "subdivisions"); D(overhang, "overhang");
  I(offset_half_height, "offset_half_height"); I(seed_strand_phase, "seed_strand_phase");
  D(length_to_m, "length_to_m"); D(force_to_n, "force_to_n"); I(udl, "udl");
  get3(env, o, "gravity_dir", sc.gravity_dir); get3(env, o, "udl_q", sc.udl_q);
#undef D
#undef I
}

// a typed array's data and element count (null / 0 when v is not one of the expected type)
template <typename T> T* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* n) {
  bool is = false; *n = 0;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return nullptr;
  napi_typedarray_type t; void* data = nullptr; napi_value ab; size_t off = 0;
  if (napi_get_typedarray_info(env, v, &t, n, &data, &ab, &off) != napi_ok || t != want) { *n = 0; return nullptr; }
  return static_cast<T*>(data);
}
uint32_t array_len(napi_env env, napi_value v) {
  bool is = false; uint32_t n = 0;
  if (napi_is_array(env, v, &is) == napi_ok && is) napi_get_array_length(env, v, &n);
  return n;
}
napi_value elem(napi_env env, napi_value v, uint32_t i) { napi_value e; napi_get_element(env, v, i, &e); return e; }
double num_at(napi_env env, napi_value v, uint32_t i) { double x = 0; napi_get_value_double(env, elem(env, v, i), &x); return x; }

napi_value wrap_model(napi_env env, wsm_model* h) {
  Model* m = new Model();
  m->h = h;
  wsm_get(h, &m->md, &m->st);
  wsm_info_get(h, &m->info);
  napi_value v;
  NAPI_CALL(env, napi_create_external(env, m, model_finalize, nullptr, &v));
  return v;
}

// modelFromFields(f): a model over the caller's typed arrays, no copy — the binding of the reference's CpuSolver
// objects (solver_mi355x.js maps a CpuSolver's fields onto f).  f holds the sizes, `par` (the ws_params fields), the
// ws_model arrays and the ws_state arrays under their include/weave.h names, each a typed array of the C type and
// length; state arrays are written in place by a download, as GpuSolver.download overwrites the CpuSolver's arrays
// (web/js/solver_gpu.js:431-503).  The arrays are referenced for the life of the model.
napi_value ModelFromFields(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  napi_value f = a[0];
  if (!is_object(env, f)) { napi_throw_type_error(env, nullptr, "modelFromFields(fields)"); return nullptr; }
  std::unique_ptr<Model> m(new Model());
  ws_model& md = m->md;
  ws_state& st = m->st;
  auto I = [&](const char* k) { return int32_t(getd(env, f, k, 0)); };
  md.n = I("n"); md.m = I("m"); md.p = I("p"); md.K = I("K"); md.C = I("C"); md.S = I("S");
  md.n_point_local = I("n_point_local"); md.l_ref = getd(env, f, "l_ref", 1.0);
  const size_t n = size_t(std::max(md.n, 0)), mm = size_t(std::max(md.m, 0)), p = size_t(std::max(md.p, 0));
  const size_t K = size_t(std::max(md.K, 0)), C = size_t(std::max(md.C, 0)), S = size_t(std::max(md.S, 0));
  const size_t npl = size_t(std::max(md.n_point_local, 0));
  std::string bad;
  // binds key to *slot: a typed array of type `want` and `count` elements (absent: null, unless required)
  auto bind = [&](const char* key, void** slot, napi_typedarray_type want, size_t count, bool required) {
    if (!bad.empty()) return;
    bool has = false; napi_value v;
    if (napi_has_named_property(env, f, key, &has) != napi_ok || !has || napi_get_named_property(env, f, key, &v) != napi_ok) {
      if (required && count) bad = std::string(key) + " is missing";
      return;
    }
    napi_valuetype t; napi_typeof(env, v, &t);
    if (t == napi_undefined || t == napi_null) { if (required && count) bad = std::string(key) + " is missing"; return; }
    size_t len = 0;
    void* data = typed<void>(env, v, want, &len);
    if (!data && count) { bad = std::string(key) + " has the wrong typed-array type"; return; }
    if (len != count) { bad = std::string(key) + " has " + std::to_string(len) + " elements, expected " + std::to_string(count); return; }
    napi_ref r = nullptr;
    if (napi_create_reference(env, v, 1, &r) == napi_ok) m->refs.push_back(r);
    *slot = data;
  };
#define BM(field, type, count, req) bind(#field, reinterpret_cast<void**>(const_cast<void**>(reinterpret_cast<const void* const*>(&md.field))), type, count, req)
#define BS(field, type, count) bind(#field, reinterpret_cast<void**>(&st.field), type, count, false)
  const auto F64 = napi_float64_array, I32 = napi_int32_array, U8 = napi_uint8_array;
  BM(node_offset, I32, S + 1, true); BM(seg_offset, I32, S + 1, true); BM(joint_offset, I32, S + 1, true);
  BM(strand_closed, U8, S, true); BM(seg_i0, I32, mm, true); BM(seg_i1, I32, mm, true);
  BM(j_seg0, I32, p, true); BM(j_seg1, I32, p, true); BM(j_node, I32, p, true); BM(l0, F64, mm, true); BM(j_len, F64, p, true);
  BM(EA, F64, mm, true); BM(GA, F64, mm, true); BM(EIx, F64, mm, true); BM(EIy, F64, mm, true); BM(GJ, F64, mm, true);
  BM(omega0, F64, 3 * p, true);
  BM(c_seg, I32, C, true); BM(c_node, I32, C, true); BM(c_len, F64, C, true); BM(c_B, F64, 3 * C, true); BM(c_Q, F64, 4 * C, true);
  BM(mass, F64, n, true); BM(krot_static, F64, 3 * mm, true); BM(load_const, F64, 3 * n, true);
  BM(point_local_node, I32, npl, true); BM(point_local_F, F64, 3 * npl, true); BM(node_free, U8, 3 * n, true);
  BM(k_strand, I32, 2 * K, true); BM(k_seg_off, I32, 2 * K, true); BM(k_nseg, I32, 2 * K, true);
  BM(k_wk, F64, 2 * K, true); BM(k_hk, F64, 2 * K, true); BM(k_fc90, F64, 2 * K, true); BM(k_E90, F64, 2 * K, true);
  BM(k_Hs, F64, 2 * K, true); BM(k_w, F64, K, true); BM(k_h, F64, K, true); BM(k_enabled, U8, K, true); BM(k_locked, U8, K, true);
  BM(node_w, F64, n, false); BM(node_h, F64, n, false);
  // parameters (params.js SOLVER / CONTACT as the reference's _uploadParams reads them, solver_gpu.js:246-284)
  napi_value par; bool hp = false;
  if (napi_has_named_property(env, f, "par", &hp) == napi_ok && hp && napi_get_named_property(env, f, "par", &par) == napi_ok) {
    ws_params& P = md.par;
#define PD(k) P.k = getd(env, par, #k, P.k)
#define PI(k) P.k = int32_t(getd(env, par, #k, P.k))
    PD(mass_scale); PD(damping); PD(tol); PD(mu); PD(strength); PD(stick_strength); PD(torsion_factor); PD(epsilon);
    PD(slip_rate); PD(rot_factor); PD(dent_rate); PD(edge_threshold); PD(floor_z); PD(floor_mu);
    PI(kinetic_damping); PI(warmup_iter); PI(max_iter); PI(contacts); PI(ramp_iter); PI(rot_friction); PI(seat_law);
    PI(seat_gap); PI(dent_after); PI(augmented); PI(al_every); PI(floor_on);
#undef PD
#undef PI
  } else bad = "par is missing";
  if (md.par.floor_on && (!md.node_w || !md.node_h)) bad = "node_w and node_h are needed with the ground";
  // the replay protocol (md.protocol)
  md.proto_mode = I("proto_mode");
  if (md.proto_mode) {
    md.proto_probe = I("proto_probe"); md.proto_n = I("proto_n"); md.proto_settle = I("proto_settle"); md.proto_avg = I("proto_avg");
    get3(env, f, "proto_axis", md.proto_axis);
    BM(proto_sched, F64, size_t(std::max(md.proto_n, 0)), true);
  }
  // state: read on upload, written by a download
  BS(X, F64, 3 * n); BS(X0, F64, 3 * n); BS(V, F64, 3 * n); BS(Q, F64, 4 * mm); BS(Q0, F64, 4 * mm); BS(W, F64, 3 * mm);
  BS(u, F64, 2 * K); BS(n_prev, F64, 3 * K); BS(slip, F64, K); BS(stuck, U8, K); BS(lam, F64, K); BS(pen_acc, F64, K);
  BS(theta0, F64, K); BS(rot_slip, F64, K); BS(dent, F64, 2 * K); BS(climb, F64, 2 * K); BS(tanpsi, F64, 2 * K);
  BS(dent_e, F64, K); BS(floor_a, F64, 2 * n);
  BS(N, F64, mm); BS(Vsh, F64, 2 * mm); BS(M, F64, 3 * p); BS(cM, F64, 2 * C); BS(reactions, F64, 3 * n);
  BS(Fn, F64, K); BS(Ft, F64, K); BS(gap, F64, K); BS(kstate, F64, K); BS(theta, F64, K); BS(Anom, F64, K);
  BS(pressure, F64, K); BS(util, F64, K); BS(Mrot, F64, K); BS(phi, F64, 2 * K); BS(mid, F64, 3 * K);
  BS(ctype, I32, K); BS(face_side, I32, K); BS(floor_F, F64, 3 * n); BS(floor_state, U8, n);
#undef BM
#undef BS
  if (bad.empty() && (!st.X || !st.Q || (K && !st.u))) bad = "X, Q (and u with crossings) are needed";
  st.iteration = I("iteration"); st.pen_n = I("pen_n"); st.ke_prev = getd(env, f, "ke_prev", 0.0);
  st.residual = getd(env, f, "residual", INFINITY);
  // the rigid freedoms of a structure only the ground holds (Floor.rv / rw / Rb): [3] arrays and a number
  get3(env, f, "rb_v", st.rb_v); get3(env, f, "rb_w", st.rb_w); st.rb_radius = getd(env, f, "rb_radius", 0.0);
  if (!bad.empty()) {
    for (napi_ref r : m->refs) napi_delete_reference(env, r);
    napi_throw_type_error(env, nullptr, ("modelFromFields: " + bad).c_str());
    return nullptr;
  }
  m->info.n = md.n; m->info.m = md.m; m->info.p = md.p; m->info.K = md.K; m->info.C = md.C; m->info.S = md.S;
  napi_value v;
  Model* raw = m.release();
  if (napi_create_external(env, raw, model_finalize, nullptr, &v) != napi_ok) { model_finalize(env, raw, nullptr); return nullptr; }
  return v;
}

// modelFields(model): copies of a built model's ws_model / ws_state arrays and scalars under their include/weave.h
// names (the shape modelFromFields takes), e.g. to assemble an object with the fields of the reference's CpuSolver
template <typename T> napi_value copy_typed(napi_env env, const T* p, size_t n, napi_typedarray_type type) {
  napi_value ab, ta; void* data = nullptr;
  napi_create_arraybuffer(env, n * sizeof(T), &data, &ab);
  if (n && p) std::memcpy(data, p, n * sizeof(T));
  else if (n) std::memset(data, 0, n * sizeof(T));
  napi_create_typedarray(env, type, n, ab, 0, &ta);
  return ta;
}
napi_value ModelFields(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "modelFields(model)"); return nullptr; }
  const ws_model& md = m->md; const ws_state& st = m->st;
  const size_t n = md.n, mm = md.m, p = md.p, K = md.par.contacts ? md.K : 0, C = md.C, S = md.S, npl = md.n_point_local;
  napi_value o; napi_create_object(env, &o);
  set(env, o, "n", i32(env, md.n)); set(env, o, "m", i32(env, md.m)); set(env, o, "p", i32(env, md.p));
  set(env, o, "K", i32(env, int32_t(K))); set(env, o, "C", i32(env, md.C)); set(env, o, "S", i32(env, md.S));
  set(env, o, "n_point_local", i32(env, md.n_point_local)); set(env, o, "l_ref", num(env, md.l_ref));
  napi_value par; napi_create_object(env, &par);
  const ws_params& P = md.par;
#define PD(k) set(env, par, #k, num(env, P.k))
#define PI(k) set(env, par, #k, i32(env, P.k))
  PD(mass_scale); PD(damping); PD(tol); PD(mu); PD(strength); PD(stick_strength); PD(torsion_factor); PD(epsilon);
  PD(slip_rate); PD(rot_factor); PD(dent_rate); PD(edge_threshold); PD(floor_z); PD(floor_mu);
  PI(kinetic_damping); PI(warmup_iter); PI(max_iter); PI(contacts); PI(ramp_iter); PI(rot_friction); PI(seat_law);
  PI(seat_gap); PI(dent_after); PI(augmented); PI(al_every); PI(floor_on);
#undef PD
#undef PI
  set(env, o, "par", par);
  const auto F64 = napi_float64_array, I32 = napi_int32_array, U8 = napi_uint8_array;
#define CI(key, ptr, cnt) set(env, o, key, copy_typed<int32_t>(env, ptr, cnt, I32))
#define CD(key, ptr, cnt) set(env, o, key, copy_typed<double>(env, ptr, cnt, F64))
#define CU(key, ptr, cnt) set(env, o, key, copy_typed<uint8_t>(env, ptr, cnt, U8))
  CI("node_offset", md.node_offset, S + 1); CI("seg_offset", md.seg_offset, S + 1); CI("joint_offset", md.joint_offset, S + 1);
  CU("strand_closed", md.strand_closed, S); CI("seg_i0", md.seg_i0, mm); CI("seg_i1", md.seg_i1, mm);
  CI("j_seg0", md.j_seg0, p); CI("j_seg1", md.j_seg1, p); CI("j_node", md.j_node, p); CD("l0", md.l0, mm); CD("j_len", md.j_len, p);
  CD("EA", md.EA, mm); CD("GA", md.GA, mm); CD("EIx", md.EIx, mm); CD("EIy", md.EIy, mm); CD("GJ", md.GJ, mm);
  CD("omega0", md.omega0, 3 * p); CI("c_seg", md.c_seg, C); CI("c_node", md.c_node, C); CD("c_len", md.c_len, C);
  CD("c_B", md.c_B, 3 * C); CD("c_Q", md.c_Q, 4 * C); CD("mass", md.mass, n); CD("krot_static", md.krot_static, 3 * mm);
  CD("load_const", md.load_const, 3 * n); CI("point_local_node", md.point_local_node, npl);
  CD("point_local_F", md.point_local_F, 3 * npl); CU("node_free", md.node_free, 3 * n);
  CI("k_strand", md.k_strand, 2 * K); CI("k_seg_off", md.k_seg_off, 2 * K); CI("k_nseg", md.k_nseg, 2 * K);
  CD("k_wk", md.k_wk, 2 * K); CD("k_hk", md.k_hk, 2 * K); CD("k_fc90", md.k_fc90, 2 * K); CD("k_E90", md.k_E90, 2 * K);
  CD("k_Hs", md.k_Hs, 2 * K); CD("k_w", md.k_w, K); CD("k_h", md.k_h, K); CU("k_enabled", md.k_enabled, K); CU("k_locked", md.k_locked, K);
  CD("X", st.X, 3 * n); CD("X0", st.X0, 3 * n); CD("V", st.V, 3 * n); CD("Q", st.Q, 4 * mm); CD("Q0", st.Q0, 4 * mm);
  CD("W", st.W, 3 * mm); CD("u", st.u, 2 * K); CD("n_prev", st.n_prev, 3 * K); CD("slip", st.slip, K); CU("stuck", st.stuck, K);
  CD("lam", st.lam, K); CD("pen_acc", st.pen_acc, K); CD("theta0", st.theta0, K); CD("rot_slip", st.rot_slip, K);
  CD("dent", st.dent, 2 * K); CD("climb", st.climb, 2 * K); CD("tanpsi", st.tanpsi, 2 * K); CD("dent_e", st.dent_e, K);
  CD("Fn", st.Fn, K); CD("Ft", st.Ft, K); CD("gap", st.gap, K); CD("kstate", st.kstate, K); CD("theta", st.theta, K);
  CD("Anom", st.Anom, K); CD("pressure", st.pressure, K); CD("util", st.util, K); CD("Mrot", st.Mrot, K);
  CD("phi", st.phi, 2 * K); CD("mid", st.mid, 3 * K); CI("ctype", st.ctype, K); CI("face_side", st.face_side, K);
  CD("rb_v", st.rb_v, 3); CD("rb_w", st.rb_w, 3);
#undef CI
#undef CD
#undef CU
  set(env, o, "iteration", i32(env, st.iteration)); set(env, o, "residual", num(env, st.residual));
  set(env, o, "ke_prev", num(env, st.ke_prev)); set(env, o, "pen_n", i32(env, st.pen_n));
  set(env, o, "rb_radius", num(env, st.rb_radius));
  return o;
}

// modelKnots(model) → Int32Array [6K] (wsm_knots); modelFrames(model) → Int32Array [3·frames] (wsm_frames);
// modelLockContacts(model, edges | null, lock) (wsm_lock_contacts)
napi_value ModelKnots(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m || !m->h) { napi_throw_type_error(env, nullptr, "modelKnots(model of the builder)"); return nullptr; }
  int32_t K = 0; wsm_knots(m->h, nullptr, 0, &K);
  std::vector<int32_t> v(6 * size_t(K)); wsm_knots(m->h, v.data(), K, &K);
  return copy_typed<int32_t>(env, v.data(), v.size(), napi_int32_array);
}
napi_value ModelFrames(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m || !m->h) { napi_throw_type_error(env, nullptr, "modelFrames(model of the builder)"); return nullptr; }
  int32_t nf = 0; wsm_frames(m->h, nullptr, 0, &nf);
  std::vector<int32_t> v(3 * size_t(nf)); wsm_frames(m->h, v.data(), nf, &nf);
  return copy_typed<int32_t>(env, v.data(), v.size(), napi_int32_array);
}
napi_value ModelLockContacts(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m || !m->h) { napi_throw_type_error(env, nullptr, "modelLockContacts(model of the builder, edges, lock)"); return nullptr; }
  std::vector<int32_t> e; bool all = true;
  uint32_t ne = array_len(env, a[1]);
  bool isarr = false; napi_is_array(env, a[1], &isarr);
  if (isarr) { all = false; for (uint32_t i = 0; i < ne; i++) e.push_back(int32_t(num_at(env, a[1], i))); }
  bool lock = true; napi_valuetype t; napi_typeof(env, a[2], &t); if (t == napi_boolean) napi_get_value_bool(env, a[2], &lock);
  int rc = wsm_lock_contacts(m->h, all ? nullptr : e.data(), all ? 0 : int32_t(e.size()), lock ? 1 : 0);
  if (rc) return throw_status(env, rc, "wsm_lock_contacts");
  return undef(env);
}

napi_value ModelHexDome(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  wsm_hex_spec spec{};
  spec.span = getd(env, a[0], "span", 3000); spec.spacing = getd(env, a[0], "spacing", 300);
  spec.segment = getd(env, a[0], "segment", 50); spec.contraction = getd(env, a[0], "contraction", 0.15);
  spec.noise = getd(env, a[0], "noise", 1.0); spec.seed = uint32_t(getd(env, a[0], "seed", 1));
  wsm_scene sc; scene_from(env, a[1], sc);
  wsm_model* h = nullptr;
  int rc = wsm_hex_dome(&spec, &sc, &h);
  if (rc) { napi_throw_error(env, nullptr, wsm_last_error()); return nullptr; }
  return wrap_model(env, h);
}

// modelFromMesh(V: Float64Array [3nV], F: Int32Array [3nF], scene, anchors: [{type, s, k, fix[6]}],
//               loads: [{s, k, F[3], local}], flips: [[s, k], ...], weld) — buildTopology … _buildGoals on a mesh
napi_value ModelFromMesh(napi_env env, napi_callback_info info) {
  napi_value a[11]; get_args(env, info, 11, a);
  size_t nV = 0, nF = 0;
  const double* V = typed<double>(env, a[0], napi_float64_array, &nV);
  const int32_t* F = typed<int32_t>(env, a[1], napi_int32_array, &nF);
  if (!V || !F || nV % 3 || nF % 3 || !nV || !nF) {
    napi_throw_type_error(env, nullptr, "modelFromMesh: V must be a Float64Array of 3·nV, F an Int32Array of 3·nF");
    return nullptr;
  }
  wsm_scene sc; scene_from(env, a[2], sc);
  std::vector<wsm_anchor> an(array_len(env, a[3]));
  for (uint32_t i = 0; i < an.size(); i++) {
    napi_value e = elem(env, a[3], i);
    an[i].type = int32_t(getd(env, e, "type", 0)); an[i].s = int32_t(getd(env, e, "s", 0)); an[i].k = int32_t(getd(env, e, "k", 0));
    an[i].pad = 0; an[i].radius = getd(env, e, "radius", 0);
    napi_value fx; bool has = false;
    napi_has_named_property(env, e, "fix", &has);
    if (has && napi_get_named_property(env, e, "fix", &fx) == napi_ok)
      for (uint32_t c = 0; c < 6; c++) { bool b = false; napi_value x = elem(env, fx, c); napi_valuetype t; napi_typeof(env, x, &t);
        if (t == napi_boolean) napi_get_value_bool(env, x, &b); else if (t == napi_number) { double d = 0; napi_get_value_double(env, x, &d); b = d != 0; }
        an[i].fix[c] = b ? 1 : 0; }
  }
  std::vector<wsm_point_load> ld(array_len(env, a[4]));
  for (uint32_t i = 0; i < ld.size(); i++) {
    napi_value e = elem(env, a[4], i);
    ld[i].s = int32_t(getd(env, e, "s", 0)); ld[i].k = int32_t(getd(env, e, "k", 0)); ld[i].local = getd(env, e, "local", 0) != 0;
    napi_value f;
    if (napi_get_named_property(env, e, "F", &f) == napi_ok && array_len(env, f) == 3)
      for (uint32_t c = 0; c < 3; c++) ld[i].F[c] = num_at(env, f, c);
  }
  std::vector<int32_t> fl;
  for (uint32_t i = 0, n = array_len(env, a[5]); i < n; i++) {
    napi_value e = elem(env, a[5], i);
    fl.push_back(int32_t(num_at(env, e, 0))); fl.push_back(int32_t(num_at(env, e, 1)));
  }
  bool weld = true; napi_valuetype wt; napi_typeof(env, a[6], &wt);
  if (wt == napi_boolean) napi_get_value_bool(env, a[6], &weld);
  wsm_model* h = nullptr;
  int rc;
  bool hybrid = false;
  napi_is_array(env, a[7], &hybrid);
  if (hybrid) {
    // a hybrid weave: lath groups 2, 3, … (objects with the scene's group keys) and [strand, group] pairs
    std::vector<wsm_group> gr(array_len(env, a[7]));
    for (uint32_t i = 0; i < gr.size(); i++) { wsm_scene d; wsm_scene_defaults(&d); gr[i] = d.grp; group_from(env, elem(env, a[7], i), gr[i]); }
    std::vector<int32_t> asg;
    for (uint32_t i = 0, n = array_len(env, a[8]); i < n; i++) {
      napi_value e = elem(env, a[8], i);
      asg.push_back(int32_t(num_at(env, e, 0))); asg.push_back(int32_t(num_at(env, e, 1)));
    }
    // groups of pieces of a discretisation: joints [[strand, arc length], ...] and [strand, piece, group] triples
    std::vector<int32_t> jst, pcs; std::vector<double> jsv;
    for (uint32_t i = 0, n = array_len(env, a[9]); i < n; i++) {
      napi_value e = elem(env, a[9], i);
      jst.push_back(int32_t(num_at(env, e, 0))); jsv.push_back(num_at(env, e, 1));
    }
    for (uint32_t i = 0, n = array_len(env, a[10]); i < n; i++) {
      napi_value e = elem(env, a[10], i);
      for (uint32_t c = 0; c < 3; c++) pcs.push_back(int32_t(num_at(env, e, c)));
    }
    if (pcs.empty())
      rc = wsm_from_mesh_groups(V, int32_t(nV / 3), F, int32_t(nF / 3), weld ? 1 : 0, &sc, an.data(), int32_t(an.size()),
                                ld.data(), int32_t(ld.size()), fl.data(), int32_t(fl.size() / 2), nullptr, gr.data(),
                                int32_t(gr.size()), asg.data(), int32_t(asg.size() / 2), &h);
    else
      rc = wsm_from_mesh_pieces(V, int32_t(nV / 3), F, int32_t(nF / 3), weld ? 1 : 0, &sc, an.data(), int32_t(an.size()),
                                ld.data(), int32_t(ld.size()), fl.data(), int32_t(fl.size() / 2), nullptr, gr.data(),
                                int32_t(gr.size()), asg.data(), int32_t(asg.size() / 2), jst.data(), jsv.data(),
                                int32_t(jst.size()), pcs.data(), int32_t(pcs.size() / 3), &h);
  } else {
    rc = wsm_from_mesh(V, int32_t(nV / 3), F, int32_t(nF / 3), weld ? 1 : 0, &sc, an.data(), int32_t(an.size()),
                       ld.data(), int32_t(ld.size()), fl.data(), int32_t(fl.size() / 2), nullptr, &h);
  }
  if (rc) { napi_throw_error(env, nullptr, wsm_last_error()); return nullptr; }
  return wrap_model(env, h);
}

napi_value ModelClone(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "not a model"); return nullptr; }
  wsm_model* h = nullptr;
  if (wsm_clone(m->h, &h)) { napi_throw_error(env, nullptr, wsm_last_error()); return nullptr; }
  return wrap_model(env, h);
}

// modelSetProtocol(model, { mode: 'force' | 'displacement' | null, probe, axis: [x, y, z], sched: [...], settle, avg })
// (md.protocol of the reference's GPU solver, solver_gpu.js:256-318)
napi_value ModelSetProtocol(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "not a model"); return nullptr; }
  napi_value v; bool has = false;
  std::string mode;
  if (napi_has_named_property(env, a[1], "mode", &has) == napi_ok && has && napi_get_named_property(env, a[1], "mode", &v) == napi_ok) {
    napi_valuetype t; napi_typeof(env, v, &t);
    if (t == napi_string) { char buf[32]; size_t len = 0; napi_get_value_string_utf8(env, v, buf, sizeof buf, &len); mode.assign(buf, len); }
  }
  if (mode.empty()) { m->md.proto_mode = 0; m->sched.clear(); m->md.proto_sched = nullptr; return undef(env); }
  if (mode != "force" && mode != "displacement") { napi_throw_type_error(env, nullptr, "mode: 'force' or 'displacement'"); return nullptr; }
  m->sched.clear();
  if (napi_get_named_property(env, a[1], "sched", &v) == napi_ok) {
    uint32_t len = 0; napi_get_array_length(env, v, &len);
    for (uint32_t k = 0; k < len; k++) { napi_value e; double x = 0; napi_get_element(env, v, k, &e); napi_get_value_double(env, e, &x); m->sched.push_back(x); }
  }
  double ax[3] = {0, 0, 0};
  if (napi_get_named_property(env, a[1], "axis", &v) == napi_ok) {
    for (uint32_t k = 0; k < 3; k++) { napi_value e; napi_get_element(env, v, k, &e); napi_get_value_double(env, e, &ax[k]); }
  }
  m->md.proto_mode = mode == "force" ? 1 : 2;
  m->md.proto_probe = int32_t(getd(env, a[1], "probe", 0));
  m->md.proto_settle = int32_t(getd(env, a[1], "settle", 1));
  m->md.proto_avg = int32_t(getd(env, a[1], "avg", 1));
  m->md.proto_n = int32_t(m->sched.size());
  m->md.proto_sched = m->sched.data();
  for (int k = 0; k < 3; k++) m->md.proto_axis[k] = ax[k];
  return undef(env);
}

// modelSegmentMoments(model) → { Mx, My, T } (Float64Array[m]) from the model's downloaded joint and clamp moments
napi_value ModelSegmentMoments(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "not a model"); return nullptr; }
  const size_t nm = size_t(m->md.m);
  napi_value o; napi_create_object(env, &o);
  double* out[3] = {nullptr, nullptr, nullptr};
  napi_value arrs[3];
  for (int k = 0; k < 3; k++) {
    napi_value ab; void* data = nullptr;
    napi_create_arraybuffer(env, std::max<size_t>(nm, 1) * sizeof(double), &data, &ab);
    out[k] = static_cast<double*>(data);
    napi_create_typedarray(env, napi_float64_array, nm, ab, 0, &arrs[k]);
  }
  if (wsm_segment_moments(&m->md, m->st.M, m->st.cM, out[0], out[1], out[2])) { napi_throw_error(env, nullptr, wsm_last_error()); return nullptr; }
  set(env, o, "Mx", arrs[0]); set(env, o, "My", arrs[1]); set(env, o, "T", arrs[2]);
  return o;
}

napi_value ModelInfo(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "not a model"); return nullptr; }
  napi_value o; napi_create_object(env, &o);
  const wsm_info& f = m->info;
  set(env, o, "n", i32(env, f.n)); set(env, o, "m", i32(env, f.m)); set(env, o, "p", i32(env, f.p));
  set(env, o, "K", i32(env, f.K)); set(env, o, "C", i32(env, f.C)); set(env, o, "S", i32(env, f.S));
  set(env, o, "rings", i32(env, f.rings)); set(env, o, "subdivisions", i32(env, f.subdivisions));
  set(env, o, "span", num(env, f.span_plan)); set(env, o, "rise", num(env, f.rise));
  const ws_params& P = m->md.par;
  set(env, o, "contacts", i32(env, P.contacts)); set(env, o, "seatLaw", i32(env, P.seat_law));
  set(env, o, "rotFriction", i32(env, P.rot_friction)); set(env, o, "augmented", i32(env, P.augmented));
  set(env, o, "floorOn", i32(env, P.floor_on)); set(env, o, "floorZ", num(env, P.floor_z)); set(env, o, "floorMu", num(env, P.floor_mu));
  // strand layout and free freedoms (copies): what the results need to split arrays per strand and find anchors
  auto copy = [&](const char* name, const void* src, size_t count, napi_typedarray_type t, size_t esz) {
    napi_value ab, ta; void* data = nullptr;
    napi_create_arraybuffer(env, std::max<size_t>(count, 1) * esz, &data, &ab);
    if (count && src) std::memcpy(data, src, count * esz);
    napi_create_typedarray(env, t, count, ab, 0, &ta);
    set(env, o, name, ta);
  };
  copy("nodeOffset", m->md.node_offset, size_t(f.S) + 1, napi_int32_array, 4);
  copy("segOffset", m->md.seg_offset, size_t(f.S) + 1, napi_int32_array, 4);
  copy("nodeFree", m->md.node_free, 3 * size_t(f.n), napi_uint8_array, 1);
  copy("loadConst", m->md.load_const, 3 * size_t(f.n), napi_float64_array, 8);
  napi_value pr; napi_create_object(env, &pr);
  set(env, pr, "mass_scale", num(env, P.mass_scale)); set(env, pr, "damping", num(env, P.damping)); set(env, pr, "tol", num(env, P.tol));
  set(env, pr, "mu", num(env, P.mu)); set(env, pr, "strength", num(env, P.strength)); set(env, pr, "stick_strength", num(env, P.stick_strength));
  set(env, pr, "torsion_factor", num(env, P.torsion_factor)); set(env, pr, "epsilon", num(env, P.epsilon));
  set(env, pr, "slip_rate", num(env, P.slip_rate)); set(env, pr, "rot_factor", num(env, P.rot_factor)); set(env, pr, "dent_rate", num(env, P.dent_rate));
  set(env, pr, "edge_threshold", num(env, P.edge_threshold)); set(env, pr, "warmup_iter", i32(env, P.warmup_iter));
  set(env, pr, "ramp_iter", i32(env, P.ramp_iter)); set(env, pr, "dent_after", i32(env, P.dent_after)); set(env, pr, "al_every", i32(env, P.al_every));
  set(env, pr, "seat_gap", i32(env, P.seat_gap)); set(env, pr, "floor_on", i32(env, P.floor_on)); set(env, pr, "kinetic_damping", i32(env, P.kinetic_damping));
  set(env, o, "params", pr);
  return o;
}

// typed arrays over the model's state (the arrays a download overwrites), and its scalars
napi_value ModelState(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Model* m = ext<Model>(env, a[0]);
  if (!m) { napi_throw_type_error(env, nullptr, "not a model"); return nullptr; }
  const ws_model& md = m->md;
  const size_t n = md.n, mm = md.m, p = md.p, K = md.K, C = md.C;
  napi_value o; napi_create_object(env, &o);
  auto arr = [&](const char* name, void* ptr, size_t count, napi_typedarray_type t, size_t esz) {
    if (!ptr || count == 0) return;
    napi_value ab, ta;
    if (napi_create_external_arraybuffer(env, ptr, count * esz, nullptr, nullptr, &ab) != napi_ok) return;
    if (napi_create_typedarray(env, t, count, ab, 0, &ta) != napi_ok) return;
    set(env, o, name, ta);
  };
  const ws_state& s = m->st;
  arr("X", s.X, 3 * n, napi_float64_array, 8); arr("X0", s.X0, 3 * n, napi_float64_array, 8);
  arr("V", s.V, 3 * n, napi_float64_array, 8); arr("Q", s.Q, 4 * mm, napi_float64_array, 8);
  arr("Q0", s.Q0, 4 * mm, napi_float64_array, 8); arr("Wr", s.W, 3 * mm, napi_float64_array, 8);
  arr("u", s.u, 2 * K, napi_float64_array, 8); arr("nPrev", s.n_prev, 3 * K, napi_float64_array, 8);
  arr("slip", s.slip, K, napi_float64_array, 8); arr("stuck", s.stuck, K, napi_uint8_array, 1);
  arr("lam", s.lam, K, napi_float64_array, 8); arr("penAcc", s.pen_acc, K, napi_float64_array, 8);
  arr("N", s.N, mm, napi_float64_array, 8); arr("V2", s.Vsh, 2 * mm, napi_float64_array, 8);
  arr("M", s.M, 3 * p, napi_float64_array, 8); arr("cM", s.cM, 2 * C, napi_float64_array, 8);
  arr("reactions", s.reactions, 3 * n, napi_float64_array, 8);
  arr("Fn", s.Fn, K, napi_float64_array, 8); arr("Ft", s.Ft, K, napi_float64_array, 8); arr("gap", s.gap, K, napi_float64_array, 8);
  arr("state", s.kstate, K, napi_float64_array, 8); arr("theta", s.theta, K, napi_float64_array, 8);
  arr("Anom", s.Anom, K, napi_float64_array, 8); arr("pressure", s.pressure, K, napi_float64_array, 8);
  arr("util", s.util, K, napi_float64_array, 8); arr("phi", s.phi, 2 * K, napi_float64_array, 8);
  arr("mid", s.mid, 3 * K, napi_float64_array, 8); arr("ctype", s.ctype, K, napi_int32_array, 4);
  arr("faceSide", s.face_side, K, napi_int32_array, 4);
  if (md.par.floor_on) { arr("floorF", s.floor_F, 3 * n, napi_float64_array, 8); arr("floorState", s.floor_state, n, napi_uint8_array, 1); }
  set(env, o, "iteration", i32(env, s.iteration)); set(env, o, "residual", num(env, s.residual));
  set(env, o, "converged", i32(env, s.converged));
  return o;
}

napi_value scalars(napi_env env, const ws_state& s) {
  napi_value o; napi_create_object(env, &o);
  set(env, o, "iteration", i32(env, s.iteration)); set(env, o, "residual", num(env, s.residual));
  set(env, o, "converged", i32(env, s.converged)); set(env, o, "kePrev", num(env, s.ke_prev)); set(env, o, "ke", num(env, s.ke));
  set(env, o, "resForce", num(env, s.res_force)); set(env, o, "resTorque", num(env, s.res_torque)); set(env, o, "fref", num(env, s.fref));
  set(env, o, "eStretch", num(env, s.e_stretch)); set(env, o, "eShear", num(env, s.e_shear)); set(env, o, "eBend", num(env, s.e_bend));
  set(env, o, "eTwist", num(env, s.e_twist)); set(env, o, "eContact", num(env, s.e_contact));
  set(env, o, "resRigid", num(env, s.res_rigid)); set(env, o, "penN", i32(env, s.pen_n));
  set(env, o, "eFloor", num(env, s.e_floor)); set(env, o, "fnFloor", num(env, s.fn_floor));
  set(env, o, "rbV", arr3(env, s.rb_v)); set(env, o, "rbW", arr3(env, s.rb_w)); set(env, o, "rbC", arr3(env, s.rb_c));
  set(env, o, "rbRadius", num(env, s.rb_radius));
  return o;
}

// ---------------------------------------------------------------------------------------------- batches
bool models_from(napi_env env, napi_value arr, std::vector<Model*>& out) {
  bool isarr = false; napi_is_array(env, arr, &isarr);
  if (!isarr) return false;
  uint32_t n = 0; napi_get_array_length(env, arr, &n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e; napi_get_element(env, arr, i, &e);
    Model* m = ext<Model>(env, e);
    if (!m) return false;
    out.push_back(m);
  }
  return !out.empty();
}

napi_value BatchCreate(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  ws_ctx* ctx = ext<ws_ctx>(env, a[0]);
  std::vector<Model*> ms;
  if (!ctx || !models_from(env, a[1], ms)) { napi_throw_type_error(env, nullptr, "batchCreate(ctx, models[], opts)"); return nullptr; }
  ws_batch_opts o{};
  o.precision = int32_t(getd(env, a[2], "precision", WS_FP32_FP64)); o.per_submit = int32_t(getd(env, a[2], "perSubmit", 0));
  o.tile = int32_t(getd(env, a[2], "tile", 0)); o.history = int32_t(getd(env, a[2], "history", 0));
  o.ctrl_chunk = int32_t(getd(env, a[2], "ctrlChunk", 0));
  o.cross_lanes = int32_t(getd(env, a[2], "crossLanes", 0));
  o.ctrl_mode = int32_t(getd(env, a[2], "ctrlMode", 0));
  o.streams = int32_t(getd(env, a[2], "streams", 0));
  std::vector<const ws_model*> mp; std::vector<const ws_state*> sp;
  for (Model* m : ms) { mp.push_back(&m->md); sp.push_back(&m->st); }
  Batch* b = new Batch(); b->ctx = ctx; b->B = int32_t(ms.size());
  int rc = ws_batch_create(ctx, mp.data(), sp.data(), b->B, &o, &b->b);
  if (rc) { std::string e = ws_last_error(ctx); delete b; napi_throw_error(env, nullptr, e.c_str()); return nullptr; }
  napi_value v;
  NAPI_CALL(env, napi_create_external(env, b, [](napi_env, void* p, void*) {
    Batch* bb = static_cast<Batch*>(p); if (bb->b) ws_batch_destroy(bb->b); delete bb; }, nullptr, &v));
  return v;
}

napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Batch* b = ext<Batch>(env, a[0]);
  if (b && b->b) { ws_batch_destroy(b->b); b->b = nullptr; }
  return undef(env);
}

napi_value BatchOpts(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Batch* b = ext<Batch>(env, a[0]);
  ws_batch_opts o{};
  if (!b || !b->b || ws_batch_get_opts(b->b, &o)) { napi_throw_error(env, nullptr, "batchOpts: no batch"); return nullptr; }
  napi_value r; napi_create_object(env, &r);
  set(env, r, "precision", i32(env, o.precision)); set(env, r, "perSubmit", i32(env, o.per_submit));
  set(env, r, "tile", i32(env, o.tile)); set(env, r, "history", i32(env, o.history));
  set(env, r, "crossLanes", i32(env, o.cross_lanes)); set(env, r, "ctrlMode", i32(env, o.ctrl_mode));
  set(env, r, "streams", i32(env, o.streams));
  return r;
}

napi_value BatchStep(napi_env env, napi_callback_info info) {
  napi_value a[4]; get_args(env, info, 4, a);
  Batch* b = ext<Batch>(env, a[0]);
  if (!b || !b->b) { napi_throw_type_error(env, nullptr, "step: no batch"); return nullptr; }
  int32_t n_iter = 0; napi_get_value_int32(env, a[1], &n_iter);
  std::vector<int32_t> sel;
  bool isarr = false; napi_is_array(env, a[2], &isarr);
  if (isarr) {
    uint32_t n = 0; napi_get_array_length(env, a[2], &n);
    for (uint32_t i = 0; i < n; i++) { napi_value e; napi_get_element(env, a[2], i, &e); int32_t x; napi_get_value_int32(env, e, &x); sel.push_back(x); }
  }
  bool force = true; napi_valuetype t; napi_typeof(env, a[3], &t);
  if (t == napi_boolean) napi_get_value_bool(env, a[3], &force);
  auto ctl = std::make_shared<std::vector<ws_control>>(b->B);
  Job* j = new Job();
  ws_batch* bh = b->b;
  j->run = [=]() { return ws_step(bh, n_iter, sel.empty() ? nullptr : sel.data(), int32_t(sel.size()), force ? 1 : 0, ctl->data()); };
  j->done = [=](napi_env e) {
    napi_value arr; napi_create_array_with_length(e, ctl->size(), &arr);
    for (size_t i = 0; i < ctl->size(); i++) napi_set_element(e, arr, uint32_t(i), ctl_object(e, (*ctl)[i]));
    return arr;
  };
  j->error = ctx_error(b->ctx);
  return queue_job(env, j, "ws_step", {a[0]});
}

napi_value BatchReadControl(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Batch* b = ext<Batch>(env, a[0]);
  if (!b || !b->b) { napi_throw_type_error(env, nullptr, "readControl: no batch"); return nullptr; }
  auto ctl = std::make_shared<std::vector<ws_control>>(b->B);
  Job* j = new Job();
  ws_batch* bh = b->b;
  j->run = [=]() { return ws_read_control(bh, ctl->data()); };
  j->done = [=](napi_env e) {
    napi_value arr; napi_create_array_with_length(e, ctl->size(), &arr);
    for (size_t i = 0; i < ctl->size(); i++) napi_set_element(e, arr, uint32_t(i), ctl_object(e, (*ctl)[i]));
    return arr;
  };
  j->error = ctx_error(b->ctx);
  return queue_job(env, j, "ws_read_control", {a[0]});
}

napi_value BatchReadHistory(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0, cap = 4096; napi_get_value_int32(env, a[1], &idx);
  napi_valuetype t; napi_typeof(env, a[2], &t); if (t == napi_number) napi_get_value_int32(env, a[2], &cap);
  if (!b || !b->b) { napi_throw_type_error(env, nullptr, "readHistory: no batch"); return nullptr; }
  auto rows = std::make_shared<std::vector<double>>(size_t(cap) * 9);
  auto nrows = std::make_shared<int32_t>(0);
  Job* j = new Job();
  ws_batch* bh = b->b;
  j->run = [=]() { return ws_read_history(bh, idx, rows->data(), cap, nrows.get()); };
  j->done = [=](napi_env e) {
    napi_value ab, ta; void* data = nullptr;
    size_t bytes = size_t(*nrows) * 9 * sizeof(double);
    napi_create_arraybuffer(e, bytes, &data, &ab);
    if (bytes) std::memcpy(data, rows->data(), bytes);
    napi_create_typedarray(e, napi_float64_array, size_t(*nrows) * 9, ab, 0, &ta);
    return ta;
  };
  j->error = ctx_error(b->ctx);
  return queue_job(env, j, "ws_read_history", {a[0]});
}

napi_value BatchReadOutputs(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0; napi_get_value_int32(env, a[1], &idx);
  if (!b || !b->b) { napi_throw_type_error(env, nullptr, "readOutputs: no batch"); return nullptr; }
  const int32_t cap = 4096;
  auto rows = std::make_shared<std::vector<double>>(size_t(cap) * 4);
  auto nrows = std::make_shared<int32_t>(0);
  Job* j = new Job();
  ws_batch* bh = b->b;
  j->run = [=]() { return ws_read_outputs(bh, idx, rows->data(), cap, nrows.get()); };
  j->done = [=](napi_env e) {
    napi_value ab, ta; void* data = nullptr;
    size_t bytes = size_t(*nrows) * 4 * sizeof(double);
    napi_create_arraybuffer(e, bytes, &data, &ab);
    if (bytes) std::memcpy(data, rows->data(), bytes);
    napi_create_typedarray(e, napi_float64_array, size_t(*nrows) * 4, ab, 0, &ta);
    return ta;
  };
  j->error = ctx_error(b->ctx);
  return queue_job(env, j, "ws_read_outputs", {a[0]});
}

napi_value BatchDownload(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0; napi_get_value_int32(env, a[1], &idx);
  Model* m = ext<Model>(env, a[2]);
  if (!b || !b->b || !m) { napi_throw_type_error(env, nullptr, "download(batch, b, model)"); return nullptr; }
  Job* j = new Job();
  ws_batch* bh = b->b;
  j->run = [=]() { return ws_download(bh, idx, &m->st); };
  j->done = [=](napi_env e) { return scalars(e, m->st); };
  j->error = ctx_error(b->ctx);
  return queue_job(env, j, "ws_download", {a[0], a[2]});
}

napi_value BatchUpload(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0; napi_get_value_int32(env, a[1], &idx);
  Model* m = ext<Model>(env, a[2]);
  if (!b || !b->b || !m) { napi_throw_type_error(env, nullptr, "upload(batch, b, model)"); return nullptr; }
  int rc = ws_upload(b->b, idx, &m->md, &m->st);
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(b->ctx)); return nullptr; }
  return undef(env);
}

napi_value BatchUpdateParams(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0; napi_get_value_int32(env, a[1], &idx);
  Model* m = ext<Model>(env, a[2]);
  if (!b || !b->b || !m) { napi_throw_type_error(env, nullptr, "updateParams(batch, b, model)"); return nullptr; }
  int rc = ws_update_params(b->b, idx, &m->md);
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(b->ctx)); return nullptr; }
  return undef(env);
}

napi_value BatchInternalize(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  Batch* b = ext<Batch>(env, a[0]);
  int32_t idx = 0; napi_get_value_int32(env, a[1], &idx);
  if (!b || !b->b) { napi_throw_type_error(env, nullptr, "internalize: no batch"); return nullptr; }
  int rc = ws_internalize(b->b, idx);
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(b->ctx)); return nullptr; }
  return undef(env);
}

// ---------------------------------------------------------------------------------------------- server
napi_value ServerCreate(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  ws_ctx* ctx = ext<ws_ctx>(env, a[0]);
  if (!ctx) { napi_throw_type_error(env, nullptr, "serverCreate(ctx, opts)"); return nullptr; }
  ws_server_opts o{};
  o.precision = int32_t(getd(env, a[1], "precision", WS_FP32_FP64)); o.chunk = int32_t(getd(env, a[1], "chunk", 0));
  o.tile = int32_t(getd(env, a[1], "tile", 0)); o.max_slots = int32_t(getd(env, a[1], "maxSlots", 0));
  o.hbm_budget = int64_t(getd(env, a[1], "hbmBudget", 0));
  Server* s = new Server(); s->ctx = ctx;
  int rc = ws_server_create(ctx, &o, &s->s);
  if (rc) { std::string e = ws_last_error(ctx); delete s; napi_throw_error(env, nullptr, e.c_str()); return nullptr; }
  napi_value v;
  NAPI_CALL(env, napi_create_external(env, s, [](napi_env, void* p, void*) {
    Server* ss = static_cast<Server*>(p); if (ss->s) ws_server_destroy(ss->s); delete ss; }, nullptr, &v));
  return v;
}

napi_value ServerDestroy(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Server* s = ext<Server>(env, a[0]);
  if (s && s->s) { ws_server_destroy(s->s); s->s = nullptr; }
  return undef(env);
}

napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  napi_value a[3]; get_args(env, info, 3, a);
  Server* s = ext<Server>(env, a[0]);
  Model* m = ext<Model>(env, a[1]);
  int32_t max_iter = 0; napi_valuetype t; napi_typeof(env, a[2], &t); if (t == napi_number) napi_get_value_int32(env, a[2], &max_iter);
  if (!s || !s->s || !m) { napi_throw_type_error(env, nullptr, "submit(server, model, maxIter)"); return nullptr; }
  int64_t id = 0;
  int rc = ws_submit(s->s, &m->md, &m->st, max_iter, &id);
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(s->ctx)); return nullptr; }
  return num(env, double(id));
}

napi_value ServerPump(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Server* s = ext<Server>(env, a[0]);
  if (!s || !s->s) { napi_throw_type_error(env, nullptr, "pump: no server"); return nullptr; }
  auto out = std::make_shared<std::pair<int32_t, int32_t>>(0, 0);
  Job* j = new Job();
  ws_server* sh = s->s;
  j->run = [=]() { return ws_pump(sh, &out->first, &out->second); };
  j->done = [=](napi_env e) {
    napi_value o; napi_create_object(e, &o);
    set(e, o, "active", i32(e, out->first)); set(e, o, "queued", i32(e, out->second));
    return o;
  };
  j->error = ctx_error(s->ctx);
  return queue_job(env, j, "ws_pump", {a[0]});
}

napi_value ServerPoll(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Server* s = ext<Server>(env, a[0]);
  if (!s || !s->s) { napi_throw_type_error(env, nullptr, "poll: no server"); return nullptr; }
  std::vector<ws_result> r(1024); int32_t n = 0;
  int rc = ws_poll(s->s, r.data(), int32_t(r.size()), &n);
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(s->ctx)); return nullptr; }
  static const char* names[4] = {"converged", "max_iter", "cancelled", "non_finite"};
  napi_value arr; napi_create_array_with_length(env, n, &arr);
  for (int32_t i = 0; i < n; i++) {
    napi_value o; napi_create_object(env, &o);
    set(env, o, "id", num(env, double(r[i].id)));
    set(env, o, "status", str(env, r[i].status >= 0 && r[i].status < 4 ? names[r[i].status] : "unknown"));
    set(env, o, "iterations", i32(env, r[i].iterations)); set(env, o, "residual", num(env, r[i].residual));
    set(env, o, "tSubmitMs", num(env, r[i].t_submit_ms)); set(env, o, "tAdmitMs", num(env, r[i].t_admit_ms));
    set(env, o, "tDoneMs", num(env, r[i].t_done_ms)); set(env, o, "n", i32(env, r[i].n)); set(env, o, "K", i32(env, r[i].K));
    napi_set_element(env, arr, uint32_t(i), o);
  }
  return arr;
}

napi_value ServerCancel(napi_env env, napi_callback_info info) {
  napi_value a[2]; get_args(env, info, 2, a);
  Server* s = ext<Server>(env, a[0]);
  double id = 0; napi_get_value_double(env, a[1], &id);
  if (!s || !s->s) { napi_throw_type_error(env, nullptr, "cancel: no server"); return nullptr; }
  int rc = ws_cancel(s->s, int64_t(id));
  if (rc) { napi_throw_error(env, nullptr, ws_last_error(s->ctx)); return nullptr; }
  return undef(env);
}

napi_value ServerClock(napi_env env, napi_callback_info info) {
  napi_value a[1]; get_args(env, info, 1, a);
  Server* s = ext<Server>(env, a[0]);
  return num(env, s && s->s ? ws_server_clock_ms(s->s) : 0.0);
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn { const char* name; napi_callback cb; };
  static const Fn fns[] = {
      {"deviceCreate", DeviceCreate}, {"deviceInfo", DeviceInfo},
      {"modelHexDome", ModelHexDome}, {"modelFromMesh", ModelFromMesh}, {"modelFromFields", ModelFromFields}, {"modelFields", ModelFields}, {"modelKnots", ModelKnots}, {"modelFrames", ModelFrames}, {"modelLockContacts", ModelLockContacts}, {"modelClone", ModelClone}, {"modelInfo", ModelInfo}, {"modelState", ModelState},
      {"batchCreate", BatchCreate}, {"batchDestroy", BatchDestroy}, {"batchOpts", BatchOpts}, {"batchStep", BatchStep},
      {"batchReadControl", BatchReadControl}, {"batchReadHistory", BatchReadHistory}, {"batchDownload", BatchDownload},
      {"batchReadOutputs", BatchReadOutputs}, {"modelSetProtocol", ModelSetProtocol},
      {"modelSegmentMoments", ModelSegmentMoments},
      {"batchUpload", BatchUpload}, {"batchUpdateParams", BatchUpdateParams}, {"batchInternalize", BatchInternalize},
      {"serverCreate", ServerCreate}, {"serverDestroy", ServerDestroy}, {"serverSubmit", ServerSubmit},
      {"serverPump", ServerPump}, {"serverPoll", ServerPoll}, {"serverCancel", ServerCancel}, {"serverClockMs", ServerClock},
  };
  for (const Fn& f : fns) {
    napi_value v;
    NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &v));
    NAPI_CALL(env, napi_set_named_property(env, exports, f.name, v));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
