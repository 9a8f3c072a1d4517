"""Host models: scenes built by the C++ model builder (libweave_model.so), exposed as ws_model / ws_state.

A `Model` plays the role of the reference's host `CpuSolver` object that `GpuSolver` uploads from
(web/js/solver_gpu.js:62-123): it owns the f64 topology, goals and state; the device solver copies them in and
downloads back into a `StateBuffers` (or into the model's own state arrays, as `download` overwrites the
CpuSolver's typed arrays, solver_gpu.js:431-503).
"""
import ctypes as C

import numpy as np

from . import _abi as A


def _lengths(md):
    return {"n": md.n, "3n": 3 * md.n, "2n": 2 * md.n, "m": md.m, "2m": 2 * md.m, "3m": 3 * md.m, "4m": 4 * md.m,
            "3p": 3 * md.p, "K": md.K, "2K": 2 * md.K, "3K": 3 * md.K, "2C": 2 * md.C}


def _view(ptr, count, dtype):
    if count == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,))


class Scene:
    """Scene parameters (params.js defaults, mm·N); attributes map onto wsm_scene / ws_params."""

    def __init__(self, **kw):
        self.c = A.wsm_scene()
        A.model_lib().wsm_scene_defaults(C.byref(self.c))
        for k, v in kw.items():
            self.set(k, v)

    def set(self, key, value):
        for obj in (self.c.par, self.c.grp, self.c):
            if hasattr(obj, key) and key not in ("par", "grp"):
                if key in ("gravity_dir", "udl_q"):
                    getattr(obj, key)[:] = list(value)
                else:
                    setattr(obj, key, value)
                return self
        raise KeyError(key)

    def get(self, key):
        for obj in (self.c.par, self.c.grp, self.c):
            if hasattr(obj, key):
                return getattr(obj, key)
        raise KeyError(key)


class StateBuffers:
    """Caller-owned ws_state buffers (numpy) of one instance, the target of a download."""

    def __init__(self, md):
        self.st = A.ws_state()
        self.arrays = {}
        L = _lengths(md)
        for name, ptype, key in A.STATE_ARRAYS:
            dt = {A.P_d: np.float64, A.P_u8: np.uint8, A.P_i32: np.int32}[ptype]
            arr = np.zeros(L[key], dtype=dt)
            self.arrays[name] = arr
            setattr(self.st, name, arr.ctypes.data_as(ptype) if arr.size else ptype())

    def __getitem__(self, k):
        return self.arrays[k]

    def copy_from(self, other):
        """copy the arrays and scalars of another StateBuffers / Model state"""
        for name in self.arrays:
            self.arrays[name][:] = other[name]
        arrays = {name for name, _, _ in A.STATE_ARRAYS}
        for f, _ in A.ws_state._fields_:
            if f not in arrays:
                setattr(self.st, f, getattr(other.st, f))


class Model:
    """One instance built on the host (the reference's CpuSolver after _buildGoals)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.md = A.ws_model()
        self.st = A.ws_state()
        A.model_lib().wsm_get(self.h, C.byref(self.md), C.byref(self.st))
        self.info = A.wsm_info()
        A.model_lib().wsm_info_get(self.h, C.byref(self.info))
        self._L = _lengths(self.md)

    def __del__(self):
        try:
            if self.h:
                A.model_lib().wsm_free(self.h)
                self.h = None
        except Exception:
            pass

    # ------------------------------------------------------------------ construction
    @staticmethod
    def _check(rc, out):
        if rc != 0:
            raise ValueError(A.model_lib().wsm_last_error().decode())
        return Model(out.value)

    @classmethod
    def hex_dome(cls, span, spacing, segment, contraction=0.15, noise=1.0, seed=1, scene=None):
        """Kagome weave of a hexagonal patch raised onto a spherical cap (the benchmark scenes)."""
        sc = scene or Scene()
        spec = A.wsm_hex_spec(span, spacing, segment, contraction, noise, seed, 0)
        out = C.c_void_p()
        return cls._check(A.model_lib().wsm_hex_dome(C.byref(spec), C.byref(sc.c), C.byref(out)), out)

    @classmethod
    def from_mesh(cls, V, F, scene=None, anchors=(), loads=(), flips=(), noise=0.0, seed=1, weld=True, wavelength=0.0,
                  groups=None, assign=(), joints=None, pieces=()):
        """Kagome weave of a triangle mesh.  A hybrid weave passes `groups` (wsm_group of lath groups 2, 3, …) and
        `assign` = [(strand, group), ...]; strands left out are in group 1, the scene's own (scene.js:50-62).
        Groups of pieces of a discretisation: `joints` = {strand: [arc lengths]} and `pieces` = [(strand, piece,
        group), ...] (wsm_from_mesh_pieces; scene.js:57-90)."""
        sc = scene or Scene()
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1)
        an = (A.wsm_anchor * max(len(anchors), 1))(*[_anchor(a) for a in anchors])
        ld = (A.wsm_point_load * max(len(loads), 1))(*[_load(x) for x in loads])
        fl = np.ascontiguousarray(np.array(flips, dtype=np.int32).reshape(-1))
        prep = A.wsm_prep(noise, wavelength, seed, 0)
        out = C.c_void_p()
        head = (V.ctypes.data_as(A.P_d), len(V) // 3, F.ctypes.data_as(A.P_i32), len(F) // 3, 1 if weld else 0,
                C.byref(sc.c), an, len(anchors), ld, len(loads), fl.ctypes.data_as(A.P_i32), len(fl) // 2, C.byref(prep))
        if groups is None:
            return cls._check(A.model_lib().wsm_from_mesh(*head, C.byref(out)), out)
        gr = (A.wsm_group * max(len(groups), 1))(*groups)
        asg = np.ascontiguousarray(np.array(list(assign), dtype=np.int32).reshape(-1))
        if not pieces:
            rc = A.model_lib().wsm_from_mesh_groups(*head, gr, len(groups), asg.ctypes.data_as(A.P_i32), len(asg) // 2,
                                                    C.byref(out))
            return cls._check(rc, out)
        js = [(int(s), float(v)) for s, vs in (joints or {}).items() for v in vs]
        jst = np.ascontiguousarray(np.array([s for s, _ in js], dtype=np.int32))
        jsv = np.ascontiguousarray(np.array([v for _, v in js], dtype=np.float64))
        pg = np.ascontiguousarray(np.array(list(pieces), dtype=np.int32).reshape(-1))
        rc = A.model_lib().wsm_from_mesh_pieces(*head, gr, len(groups), asg.ctypes.data_as(A.P_i32), len(asg) // 2,
                                                jst.ctypes.data_as(A.P_i32), jsv.ctypes.data_as(A.P_d), len(js),
                                                pg.ctypes.data_as(A.P_i32), len(pg) // 3, C.byref(out))
        return cls._check(rc, out)

    @classmethod
    def from_polylines(cls, polylines, scene=None, knot_rows=None, normals=None, knots=(), anchors=(), loads=()):
        """Stand-alone laths (topologyFromPolylines) with explicit crossings knots = [(sA, kA, sB, kB), ...]."""
        sc = scene or Scene()
        P = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1) for p in polylines]))
        counts = np.array([len(np.asarray(p).reshape(-1)) // 3 for p in polylines], dtype=np.int32)
        if knot_rows is None:
            knot_rows = [[0, c - 1] for c in counts]
        kr = np.array([r for rows in knot_rows for r in rows], dtype=np.int32)
        kc = np.array([len(rows) for rows in knot_rows], dtype=np.int32)
        nrm = None
        if normals is not None:
            nrm = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in normals]))
        kn = np.array(knots, dtype=np.int32).reshape(-1)
        an = (A.wsm_anchor * max(len(anchors), 1))(*[_anchor(a) for a in anchors])
        ld = (A.wsm_point_load * max(len(loads), 1))(*[_load(x) for x in loads])
        out = C.c_void_p()
        rc = A.model_lib().wsm_from_polylines(
            P.ctypes.data_as(A.P_d), counts.ctypes.data_as(A.P_i32), len(counts), kr.ctypes.data_as(A.P_i32),
            kc.ctypes.data_as(A.P_i32), nrm.ctypes.data_as(A.P_d) if nrm is not None else None,
            kn.ctypes.data_as(A.P_i32), len(kn) // 4, C.byref(sc.c), an, len(anchors), ld, len(loads), C.byref(out))
        m = cls._check(rc, out)
        m._keep = (P, counts, kr, kc, nrm, kn)
        return m

    def clone(self):
        out = C.c_void_p()
        A.model_lib().wsm_clone(self.h, C.byref(out))
        return Model(out.value)

    def set_protocol(self, mode, probe, axis, sched, settle, avg):
        """UTM load-test replay on the device (md.protocol of the reference's GPU solver): mode "force" (the probe
        node carries axis·value) or "displacement" (the probe pinned at P0 + value·axis); per schedule value at most
        `settle` iterations (ending at the residual), then the probe's travel and reaction along the axis averaged over
        `avg` iterations (GpuSolver.read_outputs).  mode None switches it off."""
        if mode is None:
            self.md.proto_mode = 0
            self._proto = None
            return self
        codes = {"force": 1, "displacement": 2}
        sch = np.ascontiguousarray(sched, dtype=np.float64)
        ax = np.asarray(axis, dtype=np.float64)
        self.md.proto_mode = codes[mode]
        self.md.proto_probe = int(probe)
        self.md.proto_n = len(sch)
        self.md.proto_settle = int(settle)
        self.md.proto_avg = int(avg)
        for k in range(3):
            self.md.proto_axis[k] = float(ax[k])
        self.md.proto_sched = sch.ctypes.data_as(A.P_d)
        self._proto = sch                      # the schedule stays alive with the model
        return self

    # ------------------------------------------------------------------ access
    @property
    def n(self): return self.md.n

    @property
    def m(self): return self.md.m

    @property
    def p(self): return self.md.p

    @property
    def K(self): return self.md.K

    def __getitem__(self, name):
        """numpy view of a state array (writes go to the model)"""
        for nm, ptype, key in A.STATE_ARRAYS:
            if nm == name:
                dt = {A.P_d: np.float64, A.P_u8: np.uint8, A.P_i32: np.int32}[ptype]
                return _view(getattr(self.st, nm), self._L[key], dt)
        raise KeyError(name)

    def arr(self, name, count=None):
        """numpy view of a model (ws_model) array"""
        sizes = {"node_offset": self.md.S + 1, "seg_offset": self.md.S + 1, "joint_offset": self.md.S + 1,
                 "strand_closed": self.md.S, "seg_i0": self.m, "seg_i1": self.m, "j_seg0": self.p, "j_seg1": self.p,
                 "j_node": self.p, "l0": self.m, "j_len": self.p, "EA": self.m, "GA": self.m, "EIx": self.m,
                 "EIy": self.m, "GJ": self.m, "omega0": 3 * self.p, "mass": self.n, "krot_static": 3 * self.m,
                 "load_const": 3 * self.n, "node_free": 3 * self.n, "k_strand": 2 * self.K, "k_seg_off": 2 * self.K,
                 "k_nseg": 2 * self.K, "k_wk": 2 * self.K, "k_hk": 2 * self.K, "k_w": self.K, "k_h": self.K,
                 "k_enabled": self.K, "k_locked": self.K, "c_seg": self.md.C, "c_len": self.md.C, "c_B": 3 * self.md.C,
                 "c_Q": 4 * self.md.C, "k_fc90": 2 * self.K, "k_E90": 2 * self.K, "k_Hs": 2 * self.K,
                 "node_w": self.n, "node_h": self.n}
        ptr = getattr(self.md, name)
        t = type(ptr)
        dt = {A.P_d: np.float64, A.P_u8: np.uint8, A.P_i32: np.int32}[t]
        return _view(ptr, count if count is not None else sizes[name], dt)

    def knots(self):
        """per crossing: mesh edge id, strand_a, k_a, strand_b, k_b, over strand (K × 6; wsm_knots)"""
        K = C.c_int32()
        A.model_lib().wsm_knots(self.h, None, 0, C.byref(K))
        out = np.zeros((K.value, 6), dtype=np.int32)
        if K.value:
            A.model_lib().wsm_knots(self.h, out.ctypes.data_as(A.P_i32), K.value, C.byref(K))
        return out

    def frames(self):
        """the model's frames: rows (strand, frame k, node) in frame-key order (st.nodeFrameIds; wsm_frames)"""
        n = C.c_int32()
        A.model_lib().wsm_frames(self.h, None, 0, C.byref(n))
        out = np.zeros((n.value, 3), dtype=np.int32)
        if n.value:
            A.model_lib().wsm_frames(self.h, out.ctypes.data_as(A.P_i32), n.value, C.byref(n))
        return out

    @property
    def knot_edges(self):
        """the mesh edge id of each crossing (st.knotEdges)"""
        return self.knots()[:, 0].copy()

    def lock_contacts(self, edges=None, lock=True):
        """CpuSolver.lockContacts(edges) / unlockContacts (lock=False): the crossings on the given mesh edges, all
        when edges is None; GpuSolver.update_params takes the locks to the device"""
        e = None if edges is None else np.ascontiguousarray(np.asarray(edges, dtype=np.int32).reshape(-1))
        rc = A.model_lib().wsm_lock_contacts(self.h, e.ctypes.data_as(A.P_i32) if e is not None else None,
                                             0 if e is None else len(e), 1 if lock else 0)
        if rc:
            raise ValueError("wsm_lock_contacts failed")
        return self

    def unlock_contacts(self, edges=None):
        return self.lock_contacts(edges, lock=False)

    def add_load(self, node, F):
        f = (C.c_double * 3)(*F)
        A.model_lib().wsm_add_load(self.h, int(node), f)

    def buffers(self):
        return StateBuffers(self.md)

    @property
    def span(self):
        return self.info.span_plan


def _anchor(a):
    """(type, s, k, fix[6][, radius]) or dict(type=..., s=, k=, fix=[, radius=]); radius: the region selector's"""
    if isinstance(a, dict):
        t, s, k, fix, r = a["type"], a.get("s", 0), a.get("k", 0), a["fix"], a.get("radius", 0.0)
    else:
        t, s, k, fix = a[:4]
        r = a[4] if len(a) > 4 else 0.0
    return A.wsm_anchor(t, s, k, (C.c_int32 * 6)(*[1 if f else 0 for f in fix]), 0, float(r))


def _load(x):
    """(s, k, F[3], local)"""
    s, k, F = x[0], x[1], x[2]
    local = x[3] if len(x) > 3 else 0
    return A.wsm_point_load(s, k, 1 if local else 0, 0, (C.c_double * 3)(*F))
