"""The device solver behind the C ABI: `GpuSolver` mirrors the reference's batched solver class
(web/js/solver_gpu.js:62-523) method for method; `Server` is the continuous-batching request loop.

Every call goes to libweave_hip.so (hand-written HIP for gfx950).  There is no CPU fallback: constructing a
solver raises when the library is missing or no GPU is visible.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .model import StateBuffers

PRECISIONS = {"fp32": A.WS_FP32_FP64, "fp32_fp64": A.WS_FP32_FP64, "fp64": A.WS_FP64}


class DeviceError(RuntimeError):
    pass


class Context:
    """getDevice() (solver_gpu.js:17-29): one HIP device."""

    def __init__(self, device=0):
        self.lib = A.hip_lib()
        self.h = C.c_void_p()
        rc = self.lib.ws_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise DeviceError(f"ws_ctx_create failed ({rc}): no HIP device")

    def check(self, rc):
        if rc != 0:
            raise DeviceError(f"status {rc}: {self.lib.ws_last_error(self.h).decode()}")

    def info(self):
        name = C.create_string_buffer(256)
        hbm, ncu = C.c_int64(), C.c_int32()
        self.check(self.lib.ws_device_info(self.h, name, 256, C.byref(hbm), C.byref(ncu)))
        return {"name": name.value.decode(), "hbm_bytes": hbm.value, "cu": ncu.value}

    def copy_probe(self, nbytes, reps=20):
        """mean ms of a plain copy kernel over nbytes (read once, written once)"""
        ms = C.c_double()
        self.check(self.lib.ws_copy_probe(self.h, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                self.lib.ws_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _ctl_dict(c):
    def val(v):
        return list(v) if isinstance(v, C.Array) else v
    return {f: val(getattr(c, f)) for f, _ in A.ws_control._fields_ if not f.startswith("pad")}


def autotune(models, precision="fp32", iters=256, ctx=None):
    """the autotuner's pick (tile, per_submit, cross_lanes) for these instances and every candidate's rate"""
    ctx = ctx or default_context()
    lib = ctx.lib
    B = len(models)
    mp = (C.POINTER(A.ws_model) * B)(*[C.pointer(m.md) for m in models])
    sp = (C.POINTER(A.ws_state) * B)(*[C.pointer(m.st) for m in models])
    best = A.ws_batch_opts()
    table = (A.ws_tune_entry * 16)()
    n = C.c_int32()
    ctx.check(lib.ws_autotune(ctx.h, mp, sp, B, PRECISIONS[precision], int(iters), C.byref(best), table, 16, C.byref(n)))
    rows = [{"tile": t.tile, "per_submit": t.per_submit, "cross_lanes": t.cross_lanes, "ms_per_iter": t.ms_per_iter,
             "node_iters_per_s": t.node_iters_per_s} for t in table[:n.value]]
    return {"tile": best.tile, "per_submit": best.per_submit, "cross_lanes": best.cross_lanes}, rows


class GpuSolver:
    """B instances (each its own topology allowed) relaxed in the same launches."""

    def __init__(self, models, precision="fp32", tile=0, per_submit=0, history=4096, ctx=None, states=None, ctrl_chunk=0,
                 cross_lanes=0, ctrl_mode=0, streams=0):
        """tile: 64 / 128 / 256 lanes per strand tile, 0 = the default (64), "autotune" (or -1) = the autotuner's pick;
        per_submit: iterations per captured graph (0 = 64, or the autotuner's pick);
        cross_lanes: lanes per crossing in the contact kernel, 1 or 2 (0 = by size, or the autotuner's pick);
        ctrl_mode: 0 = a control launch after the strand launch, 1 = the strand launch runs the control step;
        streams: 1 or 2 streams of launches (two groups of instances overlapping), 0 = by size"""
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.models = list(models)
        self.B = len(self.models)
        self.precision = precision
        mp = (C.POINTER(A.ws_model) * self.B)(*[C.pointer(m.md) for m in self.models])
        sts = states or [m for m in self.models]
        sp = (C.POINTER(A.ws_state) * self.B)(*[C.pointer(s.st) for s in sts])
        if tile == "autotune":
            tile = A.WS_TILE_AUTOTUNE
        opts = A.ws_batch_opts(PRECISIONS[precision], per_submit, tile, history, ctrl_chunk, cross_lanes, ctrl_mode,
                               streams)
        self.h = C.c_void_p()
        self.ctx.check(self.lib.ws_batch_create(self.ctx.h, mp, sp, self.B, C.byref(opts), C.byref(self.h)))
        self._ctl = (A.ws_control * self.B)()

    @property
    def opts(self):
        """options in effect: precision, per_submit, tile and cross_lanes (as resolved by default, size or autotuner),
        history"""
        o = A.ws_batch_opts()
        self.ctx.check(self.lib.ws_batch_get_opts(self.h, C.byref(o)))
        return {"precision": o.precision, "per_submit": o.per_submit, "tile": o.tile, "history": o.history,
                "ctrl_chunk": o.ctrl_chunk, "cross_lanes": o.cross_lanes, "ctrl_mode": o.ctrl_mode,
                "streams": o.streams}

    @property
    def tile(self):
        return self.opts["tile"]

    def close(self):
        if getattr(self, "h", None):
            self.lib.ws_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n_iter, instances=None, force=True, control=True):
        """run up to n_iter iterations on the selected instances; returns the control records (None with
        control=False: nothing is read back)"""
        if instances is not None:
            sel = (C.c_int32 * len(instances))(*instances)
            ns = len(instances)
        else:
            sel, ns = None, 0
        self.ctx.check(self.lib.ws_step(self.h, int(n_iter), sel, ns, 1 if force else 0, self._ctl if control else None))
        return [_ctl_dict(c) for c in self._ctl] if control else None

    def read_control(self):
        self.ctx.check(self.lib.ws_read_control(self.h, self._ctl))
        return [_ctl_dict(c) for c in self._ctl]

    def download(self, b, into=None, noalloc=False):
        """state of instance b into `into` (StateBuffers) or into the model's own arrays (GpuSolver.download);
        noalloc: ws_download_noalloc (no device allocation, the library's fixed staging region)"""
        tgt = into if into is not None else self.models[b]
        fn = self.lib.ws_download_noalloc if noalloc else self.lib.ws_download
        self.ctx.check(fn(self.h, b, C.byref(tgt.st)))
        return tgt

    def state(self, b):
        buf = StateBuffers(self.models[b].md)
        return self.download(b, buf)

    def upload(self, b, model=None, state=None):
        m = model or self.models[b]
        st = state or m
        self.ctx.check(self.lib.ws_upload(self.h, b, C.byref(m.md), C.byref(st.st)))

    def update_params(self, b):
        self.ctx.check(self.lib.ws_update_params(self.h, b, C.byref(self.models[b].md)))

    def internalize(self, b):
        self.ctx.check(self.lib.ws_internalize(self.h, b))

    def evaluate(self):
        """generalised forces of every instance at its current state (ws_evaluate): lists of F (n × 3), body-axis
        torques T (m × 3) and frame inertia I3 (m × 3) per instance; the state is unchanged"""
        ns = [m.md.n for m in self.models]
        ms = [m.md.m for m in self.models]
        F = np.zeros(3 * sum(ns))
        T = np.zeros(3 * sum(ms))
        I3 = np.zeros(3 * sum(ms))
        self.ctx.check(self.lib.ws_evaluate(self.h, F.ctypes.data_as(A.P_d), T.ctypes.data_as(A.P_d),
                                            I3.ctypes.data_as(A.P_d)))
        fo = np.concatenate([[0], np.cumsum(ns)]) * 3
        to = np.concatenate([[0], np.cumsum(ms)]) * 3
        return ([F[fo[b]:fo[b + 1]].reshape(-1, 3) for b in range(self.B)],
                [T[to[b]:to[b + 1]].reshape(-1, 3) for b in range(self.B)],
                [I3[to[b]:to[b + 1]].reshape(-1, 3) for b in range(self.B)])

    def set_positions(self, XQ):
        """positions and frames [(X, Q), …] of instances 0, 1, … (both buffers), the rest of their state kept"""
        X = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x, _ in XQ]))
        Q = np.ascontiguousarray(np.concatenate([np.asarray(q, dtype=np.float64).reshape(-1) for _, q in XQ]))
        self.ctx.check(self.lib.ws_set_positions(self.h, len(XQ), X.ctypes.data_as(A.P_d), Q.ctypes.data_as(A.P_d)))

    def read_outputs(self, b, cap=4096):
        """replay outputs of instance b recorded so far: rows of (travel, reaction, value, iteration)"""
        out = np.zeros((max(cap, 1), 4))
        n = C.c_int32()
        self.ctx.check(self.lib.ws_read_outputs(self.h, b, out.ctypes.data_as(A.P_d), cap, C.byref(n)))
        return out[:n.value]

    def read_history(self, b, cap=4096):
        out = np.zeros((cap, 9))
        n = C.c_int32()
        self.ctx.check(self.lib.ws_read_history(self.h, b, out.ctypes.data_as(A.P_d), cap, C.byref(n)))
        return out[:n.value]

    def set_profile(self, on):
        self.ctx.check(self.lib.ws_batch_profile(self.h, 1 if on else 0))

    def timing(self):
        a, b, c, e = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self.ctx.check(self.lib.ws_batch_timing(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(e)))
        return {"ms_last_step": a.value, "ms_cross": b.value, "ms_strand": c.value, "ms_ctrl": e.value}

    def kernel_bench(self, reps=50):
        """mean ms per launch of each kernel alone (back-to-back launches between two hipEvents); advances the state"""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.ctx.check(self.lib.ws_batch_kernel_bench(self.h, reps, C.byref(a), C.byref(b), C.byref(c)))
        return {"ms_cross": a.value, "ms_strand": b.value, "ms_ctrl": c.value}

    def bytes(self):
        a, b = C.c_double(), C.c_double()
        self.ctx.check(self.lib.ws_batch_bytes(self.h, C.byref(a), C.byref(b)))
        return {"alg_bytes_per_iter": a.value, "device_bytes": b.value}


STATUS = {0: "converged", 1: "max_iter", 2: "cancelled", 3: "non_finite"}


class Server:
    """Continuous-batching request server (ws_server_*): submit design-variant solves as they arrive, pump, poll.

    A request's model and state buffers are kept referenced here until its result is polled; the final state is
    written into them (the model's own arrays unless `into` StateBuffers are given)."""

    def __init__(self, precision="fp32", chunk=64, tile=64, max_slots=256, hbm_budget=0, ctx=None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        o = A.ws_server_opts(PRECISIONS[precision], chunk, tile, max_slots, int(hbm_budget))
        self.h = C.c_void_p()
        self.ctx.check(self.lib.ws_server_create(self.ctx.h, C.byref(o), C.byref(self.h)))
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.ws_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, model, max_iter=0, into=None):
        tgt = into if into is not None else model
        rid = C.c_int64()
        self.ctx.check(self.lib.ws_submit(self.h, C.byref(model.md), C.byref(tgt.st), int(max_iter), C.byref(rid)))
        self._keep[rid.value] = (model, tgt)
        return rid.value

    def cancel(self, rid):
        self.ctx.check(self.lib.ws_cancel(self.h, int(rid)))

    def pump(self):
        """one round: admit, relax one chunk, retire; returns (active, queued)"""
        a, q = C.c_int32(), C.c_int32()
        self.ctx.check(self.lib.ws_pump(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, cap=1024):
        out = (A.ws_result * cap)()
        n = C.c_int32()
        self.ctx.check(self.lib.ws_poll(self.h, out, cap, C.byref(n)))
        res = []
        for r in out[:n.value]:
            model, tgt = self._keep.pop(r.id, (None, None))
            res.append({"id": r.id, "status": STATUS.get(r.status, r.status), "iterations": r.iterations,
                        "residual": r.residual, "t_submit_ms": r.t_submit_ms, "t_admit_ms": r.t_admit_ms,
                        "t_done_ms": r.t_done_ms, "n": r.n, "K": r.K, "model": model, "state": tgt})
        return res

    def fetch(self, rid, into):
        self.ctx.check(self.lib.ws_fetch(self.h, int(rid), C.byref(into.st)))
        return into

    def clock_ms(self):
        return self.lib.ws_server_clock_ms(self.h)

    def stats(self):
        a, q, u, b = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        self.ctx.check(self.lib.ws_server_stats(self.h, C.byref(a), C.byref(q), C.byref(u), C.byref(b)))
        return {"active": a.value, "queued": q.value, "bytes_in_use": u.value, "hbm_budget": b.value}

    def drain(self, max_rounds=1 << 30):
        """pump until nothing is resident or queued; returns every result"""
        out = []
        for _ in range(max_rounds):
            a, q = self.pump()
            out += self.poll()
            if a == 0 and q == 0:
                break
        return out
