"""ctypes binding of the f64 oracle (oracle/build/liboracle.so) — test infrastructure.

Only tests, the smoke check and the benchmark's CPU-baseline leg use this module.
"""
import ctypes as C
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
A = importlib.import_module("weave-sim_amd._abi")
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "build", "liboracle.so")
        if not os.path.exists(path):
            raise OSError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oracle_create.argtypes = [C.POINTER(A.ws_model), C.POINTER(A.ws_state), C.c_int32]
        L.oracle_create.restype = C.c_void_p
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_free.restype = None
        for f in ("oracle_step",):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int32]
        for f in ("oracle_converged", "oracle_iteration"):
            getattr(L, f).argtypes = [C.c_void_p]
        for f in ("oracle_residual", "oracle_ke"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_double
        L.oracle_get.argtypes = [C.c_void_p, C.POINTER(A.ws_state)]
        L.oracle_get.restype = None
        L.oracle_inertia.argtypes = [C.c_void_p, A.P_d]
        L.oracle_inertia.restype = None
        L.oracle_floor.argtypes = [C.c_void_p, A.P_d, C.POINTER(C.c_uint8), A.P_d]
        L.oracle_floor.restype = C.c_int32
        L.oracle_history.argtypes = [C.c_void_p, A.P_d, A.P_d, A.P_d, C.c_int32]
        L.oracle_rod_apply.argtypes = [C.c_void_p, A.P_d, A.P_d, A.P_d, A.P_d]
        L.oracle_rod_apply.restype = C.c_double
        L.oracle_evaluate.argtypes = [C.c_void_p, A.P_d, A.P_d, A.P_d, A.P_d, A.P_d]
        L.oracle_evaluate.restype = None
        L.oracle_replay.argtypes = [C.c_void_p, C.c_int32, C.c_int32, A.P_d, A.P_d, C.c_int32, C.c_int32, C.c_int32, A.P_d]
        L.oracle_replay.restype = None
        L.oracle_pairwise_sum.argtypes = [A.P_d, C.c_int64]
        L.oracle_pairwise_sum.restype = C.c_double
        L.oracle_probe_classify.argtypes = [C.c_void_p, C.c_int32, A.P_d, A.P_d, A.P_d, C.c_double, A.P_d]
        L.oracle_probe_classify.restype = None
        L.oracle_probe_support.argtypes = [C.c_void_p, C.c_int32, A.P_d, A.P_d, C.c_double, C.c_int32, A.P_d]
        L.oracle_probe_support.restype = C.c_double
        L.oracle_probe_seat.argtypes = [C.c_void_p, C.c_int32, A.P_d, A.P_d, A.P_d, A.P_d, A.P_d, A.P_d]
        L.oracle_probe_seat.restype = C.c_double
        L.oracle_probe_rot.argtypes = [C.c_void_p, C.c_int32, A.P_d, A.P_d, A.P_d, C.c_double, C.c_double, C.c_int32,
                                       C.c_double, A.P_d]
        L.oracle_probe_rot.restype = None
        L.oracle_probe_al_value.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_double]
        L.oracle_probe_al_value.restype = C.c_double
        _lib = L
    return _lib


class Oracle:
    """CpuSolver restated in f64 on a copy of a Model's description and state."""

    def __init__(self, model, history=False, state=None):
        self.model = model
        self.h = lib().oracle_create(C.byref(model.md), C.byref(state.st if state is not None else model.st),
                                     1 if history else 0)

    def __del__(self):
        try:
            lib().oracle_free(self.h)
        except Exception:
            pass

    def step(self, n):
        return lib().oracle_step(self.h, int(n))

    def evaluate(self, X, Q):
        """generalised forces at positions X and frames Q (the stability check's evaluator): F (n × 3) before the
        anchors, body-axis torques T (m × 3), frame inertia I3 (m × 3); the state is unchanged"""
        n, m = self.model.n, self.model.md.m
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1)
        F, T, I3 = np.zeros(3 * n), np.zeros(3 * m), np.zeros(3 * m)
        lib().oracle_evaluate(self.h, X.ctypes.data_as(A.P_d), Q.ctypes.data_as(A.P_d), F.ctypes.data_as(A.P_d),
                              T.ctypes.data_as(A.P_d), I3.ctypes.data_as(A.P_d))
        return F.reshape(-1, 3), T.reshape(-1, 3), I3.reshape(-1, 3)

    def replay(self, mode, probe, axis, sched, settle, avg):
        """UTM load-test replay (calibrate.js replayForce / replayDisplacement) from the current state: one row per
        schedule value of (travel, reaction, value, iteration)"""
        sch = np.ascontiguousarray(sched, dtype=np.float64)
        ax = np.ascontiguousarray(axis, dtype=np.float64)
        out = np.zeros((len(sch), 4))
        lib().oracle_replay(self.h, {"force": 1, "displacement": 2}[mode], int(probe), ax.ctypes.data_as(A.P_d),
                            sch.ctypes.data_as(A.P_d), len(sch), int(settle), int(avg), out.ctypes.data_as(A.P_d))
        return out

    def run(self, max_iter, chunk=1000):
        done = 0
        while done < max_iter and not self.converged:
            done += self.step(min(chunk, max_iter - done))
        return done

    @property
    def converged(self): return bool(lib().oracle_converged(self.h))

    @property
    def iteration(self): return lib().oracle_iteration(self.h)

    @property
    def residual(self): return lib().oracle_residual(self.h)

    @property
    def ke(self): return lib().oracle_ke(self.h)

    def state(self):
        buf = self.model.buffers()
        lib().oracle_get(self.h, C.byref(buf.st))
        return buf

    def inertia(self):
        out = np.zeros(3 * self.model.m)
        lib().oracle_inertia(self.h, out.ctypes.data_as(A.P_d))
        return out

    def floor(self):
        """the ground's per-node forces (n × 3), state (0 off, 1 stick, 2 slide) and the rigid-body residual part"""
        F = np.zeros(3 * self.model.n)
        st = np.zeros(self.model.n, dtype=np.uint8)
        rr = C.c_double()
        if not lib().oracle_floor(self.h, F.ctypes.data_as(A.P_d), st.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(rr)):
            return None
        return F.reshape(-1, 3), st, rr.value

    def history(self, cap=10 ** 7):
        n = min(cap, self.iteration)
        ke, strain, resid = np.zeros(n), np.zeros(n), np.zeros(n)
        k = lib().oracle_history(self.h, ke.ctypes.data_as(A.P_d), strain.ctypes.data_as(A.P_d),
                                 resid.ctypes.data_as(A.P_d), n)
        return ke[:k], strain[:k], resid[:k]

    def rod_apply(self, X=None, Q=None):
        m = self.model
        F, T = np.zeros(3 * m.n), np.zeros(3 * m.m)
        Xp = np.ascontiguousarray(X, dtype=np.float64) if X is not None else None
        Qp = np.ascontiguousarray(Q, dtype=np.float64) if Q is not None else None
        e = lib().oracle_rod_apply(self.h, Xp.ctypes.data_as(A.P_d) if Xp is not None else None,
                                   Qp.ctypes.data_as(A.P_d) if Qp is not None else None,
                                   F.ctypes.data_as(A.P_d), T.ctypes.data_as(A.P_d))
        return e, F, T


def _vec(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))


class ContactProbe:
    """single contact laws of the oracle on constructed geometry: frames RA, RB (rows x, y, t of each side), the
    contact normal n, for crossing c of an oracle built on `model`"""

    def __init__(self, model, c=0):
        self.o = Oracle(model)
        self.c = c

    def classify(self, RA, RB, n, fn):
        out = np.zeros(10)
        RA, RB, n = _vec(RA), _vec(RB), _vec(n)
        lib().oracle_probe_classify(self.o.h, self.c, RA.ctypes.data_as(A.P_d), RB.ctypes.data_as(A.P_d),
                                    n.ctypes.data_as(A.P_d), float(fn), out.ctypes.data_as(A.P_d))
        return {"ctype": int(out[0]), "face_side": int(out[1]), "theta": out[2], "phi": (out[3], out[4]),
                "Anom": out[5], "pressure": out[6], "util": out[7], "cface": (out[8], out[9])}

    def support(self, R, n, sign, side):
        r = np.zeros(3)
        R, n = _vec(R), _vec(n)
        h = lib().oracle_probe_support(self.o.h, self.c, R.ctypes.data_as(A.P_d), n.ctypes.data_as(A.P_d), float(sign),
                                       int(side), r.ctypes.data_as(A.P_d))
        return h, r

    def seat(self, RA, RB, n, St, dents):
        tp = np.zeros(2)
        RA, RB, n, St, dents = _vec(RA), _vec(RB), _vec(n), _vec(St), _vec(dents)
        mu = lib().oracle_probe_seat(self.o.h, self.c, RA.ctypes.data_as(A.P_d), RB.ctypes.data_as(A.P_d),
                                     n.ctypes.data_as(A.P_d), St.ctypes.data_as(A.P_d), dents.ctypes.data_as(A.P_d),
                                     tp.ctypes.data_as(A.P_d))
        return mu, tp

    def rot(self, RA, RB, n, fn, theta0, active, kt):
        out = np.zeros(3)
        RA, RB, n = _vec(RA), _vec(RB), _vec(n)
        lib().oracle_probe_rot(self.o.h, self.c, RA.ctypes.data_as(A.P_d), RB.ctypes.data_as(A.P_d),
                               n.ctypes.data_as(A.P_d), float(fn), float(theta0), 1 if active else 0, float(kt),
                               out.ctypes.data_as(A.P_d))
        return {"M": out[0], "theta0": out[1], "rot_slip": out[2]}


def al_value(lam, pen_mean, kn, locked, enabled, pinned, cap):
    """PolylineContact.alValue (solver_cpu.js:492-497) as the oracle computes it"""
    return lib().oracle_probe_al_value(float(lam), float(pen_mean), float(kn), int(locked), int(enabled), int(pinned),
                                       float(cap))


def pairwise_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().oracle_pairwise_sum(a.ctypes.data_as(A.P_d), len(a))
