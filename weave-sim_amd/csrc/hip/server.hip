// Continuous-batching request server for design-variant solves (include/weave.h, "serving").  The reference has
// no server: its batched caller is the calibration loop (web/js/calibrate.js:438-520), which builds one GpuSolver
// per fixed set of instances and steps it to the end.  Here requests arrive while others relax:
//
//   ws_submit   queues a request (its model and state stay caller-owned until its result is polled; the final
//               state is written into that same ws_state, as GpuSolver.download overwrites the host arrays)
//   ws_pump     one scheduling round on the caller's thread: admit queued requests in arrival order while a slot
//               is free and their device bytes fit the HBM budget; relax every resident instance one chunk (their
//               launches shared, each instance stopping at its own iteration limit or convergence); retire the
//               instances that finished, writing their state back and freeing slot and memory
//   ws_poll     returns finished results (status, iterations, residual, submit / admit / done times)
//
// Resident instances share the batch's launches, so an instance's arithmetic is the same alone and in any mix:
// every reduction runs in a fixed order inside the instance, and no instance reads another's memory.
#include "batch.hpp"

namespace {

double now_ms(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct ws_server {
  struct Req {
    int64_t id = -1;
    const ws_model* md = nullptr;
    ws_state* st = nullptr;
    int32_t max_iter = 0;
    int32_t it0 = 0;
    size_t bytes = 0;
    double t_submit = 0, t_admit = 0;
  };
  ws_ctx* ctx = nullptr;
  std::mutex mu;                  // serialises the calls on this handle (submit / poll from one thread during a pump)
  ws_server_opts opts{};
  std::unique_ptr<BatchBase> batch;
  std::deque<Req> queue;
  std::vector<Req> slot;          // resident request per slot (id < 0: free)
  std::vector<ws_result> finished;
  size_t bytes_in_use = 0;
  int64_t next_id = 1;
  std::chrono::steady_clock::time_point t0;

  ws_result result_of(const Req& r, int status, const ws_control& c) {
    ws_result o{};
    o.id = r.id; o.status = status; o.iterations = c.iter; o.residual = c.resid;
    o.t_submit_ms = r.t_submit; o.t_admit_ms = r.t_admit; o.t_done_ms = now_ms(t0);
    o.n = r.md->n; o.K = r.md->par.contacts ? r.md->K : 0;
    return o;
  }

  // write the state back, free the slot and its memory
  void retire(int q, int status, const ws_control& c) {
    Req& r = slot[q];
    batch->download(q, *r.st);
    batch->remove_instance(q);
    bytes_in_use -= r.bytes;
    finished.push_back(result_of(r, status, c));
    r = Req{};
  }

  int active() const { int a = 0; for (auto& r : slot) a += r.id >= 0; return a; }

  void pump(int32_t* n_active, int32_t* n_queued) {
    bool changed = false;
    // admission: arrival order, while a slot is free and the request's bytes fit the budget
    while (!queue.empty()) {
      Req& r = queue.front();
      int q = -1;
      for (int s = 0; s < int(slot.size()); s++) if (slot[s].id < 0) { q = s; break; }
      if (q < 0 || bytes_in_use + r.bytes > size_t(opts.hbm_budget)) break;
      batch->add_instance(q, *r.md, *r.st);
      r.t_admit = now_ms(t0);
      r.it0 = r.st->iteration;
      bytes_in_use += r.bytes;
      slot[q] = r;
      queue.pop_front();
      changed = true;
    }
    if (changed) batch->relist();
    // one chunk for every resident instance, each up to its own iteration limit
    std::vector<int32_t> sel, cnt;
    std::vector<ws_control> ctl(slot.size());
    batch->read_control(ctl.data());
    for (int q = 0; q < int(slot.size()); q++) {
      if (slot[q].id < 0) continue;
      int left = slot[q].max_iter - ctl[q].iter;
      if (left > 0) { sel.push_back(q); cnt.push_back(std::min(left, opts.chunk)); }
    }
    if (!sel.empty()) batch->step(0, sel.data(), int(sel.size()), 0, nullptr, cnt.data());
    // retirement
    batch->read_control(ctl.data());
    bool removed = false;
    for (int q = 0; q < int(slot.size()); q++) {
      if (slot[q].id < 0) continue;
      const ws_control& c = ctl[q];
      int status = -1;
      if (!std::isfinite(c.ke) || std::isnan(c.resid)) status = 3;
      else if (c.done) status = 0;
      else if (c.iter >= slot[q].max_iter) status = 1;
      if (status >= 0) { retire(q, status, c); removed = true; }
    }
    if (removed) batch->relist();
    if (n_active) *n_active = active();
    if (n_queued) *n_queued = int32_t(queue.size());
  }
};

extern "C" {

WS_API int ws_server_create(ws_ctx* ctx, const ws_server_opts* opts, ws_server** out) {
  if (!ctx || !out) return ctx ? (ctx->err = "ws_server_create: bad arguments", WS_ERR_ARG) : WS_ERR_ARG;
  try {
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<ws_server> s(new ws_server());
    s->ctx = ctx;
    s->opts = opts ? *opts : ws_server_opts{};
    ws_server_opts& o = s->opts;
    if (o.chunk <= 0) o.chunk = 64;
    if (o.max_slots <= 0) o.max_slots = 256;
    if (o.tile == 0) o.tile = 64;
    if (o.tile != 64 && o.tile != 128 && o.tile != 256) throw Fail{WS_ERR_ARG, "server tile must be 64, 128 or 256"};
    if (o.hbm_budget <= 0) {
      // default: 80 % of the memory free at creation
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      o.hbm_budget = int64_t(0.8 * double(fr));
    }
    if (o.precision == WS_FP64) s->batch.reset(new Batch<double>()); else s->batch.reset(new Batch<float>());
    s->batch->ctx = ctx; s->batch->precision = o.precision; s->batch->tile = o.tile; s->batch->per_submit = o.chunk;
    s->batch->eager_graphs = false;   // relisted as requests come and go: each length is captured when a round needs it
    s->batch->create_slots(o.max_slots);
    s->slot.resize(o.max_slots);
    s->t0 = std::chrono::steady_clock::now();
    *out = s.release();
    return WS_OK;
  } catch (const Fail& f) { return fail(ctx, f); }
}

WS_API int ws_server_destroy(ws_server* srv) {
  if (!srv) return WS_OK;
  { std::lock_guard<std::mutex> lk(srv->mu); }   // a call running on another thread finishes first
  delete srv;
  return WS_OK;
}

WS_API int ws_submit(ws_server* srv, const ws_model* model, ws_state* state, int32_t max_iter, int64_t* id_out) {
  if (!srv || !model || !state) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  try {
    ws_server::Req r;
    r.md = model; r.st = state;
    r.max_iter = max_iter > 0 ? max_iter : model->par.max_iter;
    // validates the model and prices it: the admission test compares these bytes with the budget
    r.bytes = srv->opts.precision == WS_FP64 ? instance_bytes<double>(*model, *state, srv->opts.tile, srv->batch->hist_len)
                                             : instance_bytes<float>(*model, *state, srv->opts.tile, srv->batch->hist_len);
    if (r.bytes > size_t(srv->opts.hbm_budget)) throw Fail{WS_ERR_ARG, "request larger than the server's HBM budget"};
    r.id = srv->next_id++;
    r.t_submit = now_ms(srv->t0);
    srv->queue.push_back(r);
    if (id_out) *id_out = r.id;
    return WS_OK;
  } catch (const Fail& f) { return fail(srv->ctx, f); }
}

WS_API int ws_cancel(ws_server* srv, int64_t id) {
  if (!srv) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  try {
    for (auto it = srv->queue.begin(); it != srv->queue.end(); ++it) {
      if (it->id != id) continue;
      ws_control c{};
      c.iter = it->st->iteration; c.resid = it->st->residual;
      srv->finished.push_back(srv->result_of(*it, 2, c));
      srv->queue.erase(it);
      return WS_OK;
    }
    for (int q = 0; q < int(srv->slot.size()); q++) {
      if (srv->slot[q].id != id) continue;
      std::vector<ws_control> ctl(srv->slot.size());
      srv->batch->read_control(ctl.data());
      srv->retire(q, 2, ctl[q]);
      srv->batch->relist();
      return WS_OK;
    }
    throw Fail{WS_ERR_ARG, "ws_cancel: no queued or resident request with this id"};
  } catch (const Fail& f) { return fail(srv->ctx, f); }
}

WS_API int ws_pump(ws_server* srv, int32_t* n_active, int32_t* n_queued) {
  if (!srv) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  try {
    HIPCHK(hipSetDevice(srv->ctx->device));
    srv->pump(n_active, n_queued);
    return WS_OK;
  } catch (const Fail& f) { return fail(srv->ctx, f); }
}

WS_API int ws_poll(ws_server* srv, ws_result* out, int32_t cap, int32_t* n_out) {
  if (!srv || (cap > 0 && !out)) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  int n = std::min<int>(cap, int(srv->finished.size()));
  for (int q = 0; q < n; q++) out[q] = srv->finished[q];
  srv->finished.erase(srv->finished.begin(), srv->finished.begin() + n);
  if (n_out) *n_out = n;
  return WS_OK;
}

WS_API int ws_fetch(ws_server* srv, int64_t id, ws_state* out) {
  if (!srv || !out) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  try {
    for (int q = 0; q < int(srv->slot.size()); q++)
      if (srv->slot[q].id == id) { srv->batch->download(q, *out); return WS_OK; }
    throw Fail{WS_ERR_STATE, "ws_fetch: the request is not resident (queued or finished)"};
  } catch (const Fail& f) { return fail(srv->ctx, f); }
}

WS_API double ws_server_clock_ms(ws_server* srv) { return srv ? now_ms(srv->t0) : 0.0; }

WS_API int ws_server_stats(ws_server* srv, int32_t* n_active, int32_t* n_queued, int64_t* bytes_in_use, int64_t* hbm_budget) {
  if (!srv) return WS_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (n_active) *n_active = srv->active();
  if (n_queued) *n_queued = int32_t(srv->queue.size());
  if (bytes_in_use) *bytes_in_use = int64_t(srv->bytes_in_use);
  if (hbm_budget) *hbm_budget = srv->opts.hbm_budget;
  return WS_OK;
}

}  // extern "C"
