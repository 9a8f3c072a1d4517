// C ABI of the device solver (include/weave.h): contexts and batches of instances (GpuSolver,
// web/js/solver_gpu.js:62-523).  Host code only; the kernels are in kernels.hip.
#include "batch.hpp"

#include <mutex>

// a batch handle: calls on one handle are serialised by its mutex (a JS host may issue one from the main thread while
// another runs on a worker).  Defined in this file only: its member's type lives in batch.hpp's per-file namespace.
struct ws_batch { std::unique_ptr<BatchBase> impl; std::mutex mu; };

namespace {

// The staging region of ws_download_noalloc: section forces, moments, reactions and pending AL multipliers of one
// download.  A file-scope device array, so the entry point allocates nothing; every GPU the library runs on holds
// its own copy until the process ends.  Host code uses the per-device address from hipGetSymbolAddress, never the
// array's name (a host-side stand-in).
static __device__ __align__(256) char scratch[WS_NOALLOC_STAGE_BYTES];

std::mutex g_stage_mu;
std::map<int, char*> g_stage_ptr;   // device → the address of its copy of `scratch`

// once per device, with that device current (ws_ctx_create)
void stage_setup(int device) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  if (g_stage_ptr.count(device)) return;
  void* ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw Fail{WS_ERR_DEVICE, "hipGetSymbolAddress returned no address for the download staging region"};
  g_stage_ptr[device] = static_cast<char*>(ptr);
}

char* stage_of(int device) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  auto f = g_stage_ptr.find(device);
  if (f == g_stage_ptr.end()) throw Fail{WS_ERR_STATE, "download staging region not set up for this device"};
  return f->second;
}

}  // namespace


// ================================================================================================= C ABI
extern "C" {

WS_API int ws_ctx_create(int32_t device, ws_ctx** out) {
  if (!out) return WS_ERR_ARG;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return WS_ERR_DEVICE;
  if (device < 0 || device >= count) return WS_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return WS_ERR_DEVICE;
  auto* c = new ws_ctx();
  c->device = device;
  (void)hipGetDeviceProperties(&c->prop, device);
  try {
    stage_setup(device);
  } catch (const Fail& f) {
    delete c;
    return f.code;
  }
  *out = c;
  return WS_OK;
}

WS_API void ws_ctx_destroy(ws_ctx* ctx) { delete ctx; }
WS_API const char* ws_last_error(const ws_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

WS_API int ws_device_info(ws_ctx* ctx, char* name, int32_t cap, int64_t* hbm, int32_t* ncu) {
  if (!ctx) return WS_ERR_ARG;
  if (name && cap > 0) {
    char dn[256] = {0};
    if (hipDeviceGetName(dn, sizeof(dn) - 1, ctx->device) != hipSuccess || !dn[0]) std::snprintf(dn, sizeof(dn), "%s", ctx->prop.name);
    std::snprintf(name, cap, "%s (%s)", dn, ctx->prop.gcnArchName);
  }
  if (hbm) *hbm = int64_t(ctx->prop.totalGlobalMem);
  if (ncu) *ncu = ctx->prop.multiProcessorCount;
  return WS_OK;
}

WS_API int ws_copy_probe(ws_ctx* ctx, int64_t bytes, int32_t reps, double* ms) {
  if (!ctx || bytes < 16 || reps <= 0 || !ms) return WS_ERR_ARG;
  try {
    HIPCHK(hipSetDevice(ctx->device));
    int64_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n * 16)); HIPCHK(hipMalloc(&b, n * 16));
    HIPCHK(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
    const int64_t blocks = (n + COPY_PER - 1) / COPY_PER;
    if (blocks > INT32_MAX) throw Fail{WS_ERR_ARG, "ws_copy_probe: size too large"};
    const int grid = int(blocks);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    HIPCHK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    HIPCHK(hipEventRecord(e1, 0));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / reps;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    (void)hipFree(a); (void)hipFree(b);
    return WS_OK;
  } catch (const Fail& f) { return fail(ctx, f); }
}


WS_API int ws_batch_create(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states, int32_t B,
                           const ws_batch_opts* opts, ws_batch** out) {
  if (!ctx || !models || !states || B <= 0 || !out) return ctx ? (ctx->err = "ws_batch_create: bad arguments", WS_ERR_ARG) : WS_ERR_ARG;
  try {
    HIPCHK(hipSetDevice(ctx->device));
    int prec = opts ? opts->precision : WS_FP32_FP64;
    int ps = opts ? opts->per_submit : 0, hist = opts ? opts->history : 0;
    int t = opts ? opts->tile : 0;
    if (t != WS_TILE_AUTOTUNE && t != 0 && t != 64 && t != 128 && t != 256)
      throw Fail{WS_ERR_ARG, "tile must be 0 (default), -1 (autotune), 64, 128 or 256"};
    int lanes = opts ? opts->cross_lanes : 0;
    if (lanes < 0 || lanes > 2) throw Fail{WS_ERR_ARG, "cross_lanes must be 0 (by size), 1 or 2"};
    if (t == 0) t = DEFAULT_TILE;
    if (t == WS_TILE_AUTOTUNE) {
      std::vector<ws_tune_entry> table;
      ws_batch_opts best{};
      autotune(ctx, models, states, B, prec, 256, 16, table, best);
      t = best.tile; if (ps <= 0) ps = best.per_submit;
      if (lanes == 0) lanes = best.cross_lanes;
    }
    const int cm = opts ? opts->ctrl_mode : 0;
    if (cm < 0 || cm > 1) throw Fail{WS_ERR_ARG, "ctrl_mode must be 0 (a control launch) or 1 (in the strand launch)"};
    const int nsr = opts ? opts->streams : 0;
    if (nsr < 0 || nsr > 2) throw Fail{WS_ERR_ARG, "streams must be 0 (by size), 1 or 2"};
    *out = new ws_batch{make_batch(ctx, models, states, B, prec, t, ps, hist, opts ? opts->ctrl_chunk : 0, lanes, cm, nsr)};
    return WS_OK;
  } catch (const Fail& f) { return fail(ctx, f); }
}

WS_API int ws_autotune(ws_ctx* ctx, const ws_model* const* models, const ws_state* const* states, int32_t B,
                       int32_t precision, int32_t iters, ws_batch_opts* best, ws_tune_entry* table, int32_t cap,
                       int32_t* n_entries) {
  if (!ctx || !models || !states || B <= 0 || !best) return ctx ? (ctx->err = "ws_autotune: bad arguments", WS_ERR_ARG) : WS_ERR_ARG;
  try {
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<ws_tune_entry> tab;
    *best = ws_batch_opts{};
    autotune(ctx, models, states, B, precision, std::max(iters, 1), 16, tab, *best);
    best->history = 0;
    int n = int(tab.size());
    if (table) for (int q = 0; q < std::min(n, cap); q++) table[q] = tab[q];
    if (n_entries) *n_entries = n;
    return WS_OK;
  } catch (const Fail& f) { return fail(ctx, f); }
}

WS_API int ws_batch_get_opts(ws_batch* batch, ws_batch_opts* out) {
  if (!batch || !out) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(batch->mu);
  out->precision = batch->impl->precision; out->per_submit = batch->impl->per_submit;
  out->tile = batch->impl->tile; out->history = batch->impl->hist_len; out->ctrl_chunk = batch->impl->ctrl_chunk;
  out->cross_lanes = batch->impl->lanes;
  out->ctrl_mode = batch->impl->ctrl_mode;
  out->streams = batch->impl->nstreams;
  return WS_OK;
}

WS_API int ws_batch_destroy(ws_batch* batch) {
  if (!batch) return WS_OK;
  { std::lock_guard<std::mutex> lk(batch->mu); }   // a call running on another thread finishes first
  delete batch;
  return WS_OK;
}
WS_API int ws_batch_size(ws_batch* batch, int32_t* B) { if (!batch || !B) return WS_ERR_ARG; *B = batch->impl->B; return WS_OK; }

// every call makes the batch's device current on the calling thread (the N-API addon calls from worker threads)
#define BATCH_CALL(expr) \
  if (!batch) return WS_ERR_ARG; \
  std::lock_guard<std::mutex> lk_(batch->mu); \
  try { HIPCHK(hipSetDevice(batch->impl->ctx->device)); return (expr); } catch (const Fail& f) { return fail(batch->impl->ctx, f); }

WS_API int ws_upload(ws_batch* batch, int32_t b, const ws_model* model, const ws_state* state) {
  if (!batch || !model || !state || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->upload(b, *model, *state));
}
WS_API int ws_update_params(ws_batch* batch, int32_t b, const ws_model* model) {
  if (!batch || !model || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->update_params(b, *model));
}
WS_API int ws_step(ws_batch* batch, int32_t n_iter, const int32_t* instances, int32_t n_inst, int32_t force, ws_control* ctl_out) {
  if (!batch || n_iter < 0) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->step(n_iter, instances, n_inst, force, ctl_out));
}
WS_API int ws_read_control(ws_batch* batch, ws_control* out) {
  if (!batch || !out) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->read_control(out));
}
WS_API int ws_read_history(ws_batch* batch, int32_t b, double* out, int32_t cap, int32_t* n_rows) {
  if (!batch || !out || !n_rows || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->read_history(b, out, cap, n_rows));
}
WS_API int ws_evaluate(ws_batch* batch, double* F, double* T, double* I3) {
  if (!batch || !F || !T || !I3) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->evaluate(F, T, I3));
}
WS_API int ws_set_positions(ws_batch* batch, int32_t count, const double* X, const double* Q) {
  if (!batch || !X || !Q || count < 0 || count > batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->set_xq(count, X, Q));
}
WS_API int ws_read_outputs(ws_batch* batch, int32_t b, double* out, int32_t cap, int32_t* n_rows) {
  if (!batch || !n_rows || b < 0 || b >= batch->impl->B || cap < 0) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->read_outputs(b, out, cap, n_rows));
}
WS_API int ws_download(ws_batch* batch, int32_t b, ws_state* out) {
  if (!batch || !out || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->download(b, *out));
}
WS_API int ws_download_noalloc(ws_batch* batch, int32_t b, ws_state* out) {
  if (!batch || !out || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->download(b, *out, stage_of(batch->impl->ctx->device), WS_NOALLOC_STAGE_BYTES));
}
WS_API int ws_internalize(ws_batch* batch, int32_t b) {
  if (!batch || b < 0 || b >= batch->impl->B) return WS_ERR_ARG;
  BATCH_CALL(batch->impl->internalize(b));
}
WS_API int ws_batch_timing(ws_batch* batch, double* ms_last, double* ms_cross, double* ms_strand, double* ms_ctrl) {
  if (!batch) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(batch->mu);
  if (ms_ctrl) *ms_ctrl = batch->impl->ms_ctrl;
  if (ms_last) *ms_last = batch->impl->ms_last;
  if (ms_cross) *ms_cross = batch->impl->ms_cross;
  if (ms_strand) *ms_strand = batch->impl->ms_strand;
  return WS_OK;
}
WS_API int ws_batch_profile(ws_batch* batch, int32_t on) {
  if (!batch) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(batch->mu);
  batch->impl->profile = on != 0;
  return WS_OK;
}
WS_API int ws_batch_kernel_bench(ws_batch* batch, int32_t reps, double* ms_cross, double* ms_strand, double* ms_ctrl) {
  if (!batch || reps <= 0) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(batch->mu);
  try {
    double out[3] = {0, 0, 0};
    batch->impl->kernel_bench(reps, out);
    if (ms_cross) *ms_cross = out[0];
    if (ms_strand) *ms_strand = out[1];
    if (ms_ctrl) *ms_ctrl = out[2];
    return WS_OK;
  } catch (const Fail& f) { return fail(batch->impl->ctx, f); }
}
WS_API int ws_batch_bytes(ws_batch* batch, double* alg, double* dev) {
  if (!batch) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(batch->mu);
  if (alg) *alg = batch->impl->alg_bytes();
  if (dev) *dev = batch->impl->dev_bytes();
  return WS_OK;
}

// ---- serving: implemented in server.hip
}  // extern "C"
