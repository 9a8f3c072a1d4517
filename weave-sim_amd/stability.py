"""Stability of a state (the reference's stability check, web/js/stability.js): the eigenvalue nearest zero of the
mass-scaled tangent stiffness A = M^-1/2 ½(K + Kᵀ) M^-1/2 and the number of its negative eigenvalues, K = −∂g/∂q the
tangent of the generalised forces g (node forces on the free translations, body-axis torques on the 3 rotations of
every segment frame).

K is assembled from central differences of g, column groups chosen by a distance-2 colouring of the structural stencil
(a strand's DOFs within ±3 nodes / segments, both strands' DOFs within ±5 of every crossing), so that one perturbed
state per colour and sign gives all the columns of that colour (stability.js:55-95, 160-185).  The perturbed states
are evaluated on the device in batches (GpuSolver.evaluate, the FP64 form: central differences at 1e-6 of the model
size need f64 forces) — the force kernels reused, one launch pair for a whole batch of perturbed copies.  The factor
is a sparse LDLᵀ with a symmetric fill-reducing ordering (SuperLU in symmetric mode, diagonal pivots): the number of
negative pivots is the number of negative eigenvalues (Sylvester's law of inertia), and inverse iteration on the factor
gives the eigenvalue nearest zero (Rayleigh quotient), as the reference does with its own LDLᵀ
(stability.js:98-225)."""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _abi as A


def _qexp(v):
    """exp of rotation vectors v [k, 3] → unit quaternions [k, 4] (w x y z)"""
    a = np.linalg.norm(v, axis=1)
    half = 0.5 * a
    s = np.where(a > 1e-12, np.sin(half) / np.where(a > 1e-12, a, 1.0), 0.5)
    return np.column_stack([np.cos(half), v[:, 0] * s, v[:, 1] * s, v[:, 2] * s])


def _qmul(a, b):
    """products of quaternion rows a, b [k, 4]"""
    aw, ax, ay, az = a.T
    bw, bx, by, bz = b.T
    return np.column_stack([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                            aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


class _Layout:
    """DOFs, stencil rows and colouring of a model (dofList, stencil, colour; stability.js:17-23, 57-108), built by the
    host model library (wsm_stability_layout)"""

    def __init__(self, model, u, W=3):
        L = A.model_lib()
        u = np.ascontiguousarray(u, dtype=np.float64)
        nd, nnz, ncol = C.c_int32(0), C.c_int64(0), C.c_int32(0)
        pu = u.ctypes.data_as(A.P_d) if u.size else None
        rc = L.wsm_stability_layout(C.byref(model.md), pu, W, C.byref(nd), C.byref(nnz), C.byref(ncol),
                                    None, None, None, None, None, None)
        if rc:
            raise ValueError(L.wsm_last_error().decode())
        kind, idx, comp = (np.zeros(max(nd.value, 1), dtype=np.int32) for _ in range(3))
        row_ptr = np.zeros(nd.value + 1, dtype=np.int64)
        cols = np.zeros(max(nnz.value, 1), dtype=np.int32)
        col = np.zeros(max(nd.value, 1), dtype=np.int32)
        rc = L.wsm_stability_layout(C.byref(model.md), pu, W, C.byref(nd), C.byref(nnz), C.byref(ncol),
                                    kind.ctypes.data_as(A.P_i32), idx.ctypes.data_as(A.P_i32),
                                    comp.ctypes.data_as(A.P_i32), row_ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                    cols.ctypes.data_as(A.P_i32), col.ctypes.data_as(A.P_i32))
        if rc:
            raise ValueError(L.wsm_last_error().decode())
        n = nd.value
        self.kind, self.index, self.comp = kind[:n].astype(np.int64), idx[:n].astype(np.int64), comp[:n].astype(np.int64)
        self.row_ptr, self.cols = row_ptr, cols[:nnz.value].astype(np.int64)
        self.col, self.ncol = col[:n].astype(np.int64), ncol.value

    def __len__(self):
        return len(self.kind)

    def row(self, d):
        return self.cols[self.row_ptr[d]:self.row_ptr[d + 1]]


def device_evaluator(W, model, base=None, batch=256):
    """g-evaluations on the device: a function [(X, Q), …] → [(F, T, I3), …].  One batch of `batch` copies of the
    model is built once (FP64 form); each round sets the copies' positions and frames and evaluates them all in one
    launch pair.  base: the state the copies carry besides X and Q (contact parameters, stick flags, normals,
    iteration …; the model's own by default)"""
    src = model if base is None else base
    solver = {}

    def evaluate(states):
        out = []
        nb = max(1, min(batch, len(states)))
        if solver.get("n") != nb:
            if "g" in solver:
                solver["g"].close()
            bufs = []
            for _ in range(nb):
                b = W.StateBuffers(model.md)
                b.copy_from(src)
                bufs.append(b)
            solver["g"] = W.GpuSolver([model] * nb, precision="fp64", states=bufs)
            solver["n"] = nb
        g = solver["g"]
        for s0 in range(0, len(states), nb):
            chunk = states[s0:s0 + nb]
            g.set_positions(chunk)
            F, T, I3 = g.evaluate()
            out += list(zip(F, T, I3))[:len(chunk)]
        return out
    return evaluate


def stability(model, evaluate, X=None, Q=None, u=None, iters=40, tol=1e-9):
    """stability of the state (X, Q, u) of `model` (its own arrays by default) with the evaluator `evaluate`
    ([(X, Q), …] → [(F, T, I3), …]) → {lam, neg, ndof, colours, evaluations, nnz, ms, vec}"""
    t0 = time.perf_counter()
    X0 = np.array(model["X"] if X is None else X, dtype=np.float64).reshape(-1)
    Q0 = np.array(model["Q"] if Q is None else Q, dtype=np.float64).reshape(-1)
    U = np.array(model["u"] if u is None else u, dtype=np.float64)
    L = _Layout(model, U)
    nd = len(L)
    if nd == 0:
        return {"lam": float("nan"), "neg": 0, "ndof": 0}
    P = X0.reshape(-1, 3)
    size = max(float(np.linalg.norm(P.max(0) - P.min(0))), 1e-9)
    hT, hR = 1e-6 * size, 1e-6
    kind, idx_i, idx_k = L.kind, L.index, L.comp
    order = np.argsort(L.col, kind="stable")
    by_col = np.split(order, np.cumsum(np.bincount(L.col, minlength=L.ncol))[:-1])

    def perturbed(ds, sgn):
        # a colour holds each node component and each segment at most once (they share stencil rows)
        X, Q = X0.copy(), Q0.copy()
        t = ds[kind[ds] == 0]
        X[3 * idx_i[t] + idx_k[t]] += sgn * hT
        r = ds[kind[ds] == 1]
        if r.size:
            j = idx_i[r]
            v = np.zeros((r.size, 3))
            v[np.arange(r.size), idx_k[r]] = sgn * hR
            Qr = Q.reshape(-1, 4)
            q = _qmul(Qr[j], _qexp(v))
            Qr[j] = q / np.linalg.norm(q, axis=1)[:, None]
        return X, Q

    states = [(X0, Q0)]
    for c in range(L.ncol):
        states += [perturbed(by_col[c], +1), perturbed(by_col[c], -1)]
    evals = evaluate(states)

    tr, ro = kind == 0, kind == 1
    flat_t, flat_r = 3 * idx_i[tr] + idx_k[tr], 3 * idx_i[ro] + idx_k[ro]

    def gvec(F, T):
        g = np.empty(nd)
        g[tr] = F.reshape(-1)[flat_t]
        g[ro] = T.reshape(-1)[flat_r]
        return g

    _, _, I3b = evals[0]
    mass = np.ctypeslib.as_array(model.md.mass, shape=(model.md.n,))
    mdof = np.empty(nd)
    mdof[tr] = mass[idx_i[tr]]
    mdof[ro] = I3b.reshape(-1)[flat_r]
    rows, cols, vals = [], [], []
    for c in range(L.ncol):
        gp = gvec(*evals[1 + 2 * c][:2])
        gm = gvec(*evals[2 + 2 * c][:2])
        ds = by_col[c]
        start, lens = L.row_ptr[ds], L.row_ptr[ds + 1] - L.row_ptr[ds]
        # every stencil row of the colour's columns, column by column (CSR slices gathered at once)
        pos = np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens) + np.repeat(start, lens)
        r, d = L.cols[pos], np.repeat(ds, lens)
        v = -(gp[r] - gm[r]) / (2 * np.where(kind[d] == 0, hT, hR))
        nz = v != 0
        rows.append(r[nz]); cols.append(d[nz]); vals.append(v[nz])
    rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    sc = 1.0 / np.sqrt(np.maximum(mdof, 1e-300))
    K = sp.coo_matrix((vals * sc[rows] * sc[cols], (rows, cols)), shape=(nd, nd)).tocsc()
    A = (0.5 * (K + K.T)).tocsc()
    lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0, options={"SymmetricMode": True})
    if not np.array_equal(lu.perm_r, lu.perm_c):
        raise RuntimeError("the factorisation pivoted off the diagonal: the inertia count does not hold")
    neg = int((lu.U.diagonal() < 0).sum())
    rng = np.random.default_rng(12345)
    x = rng.uniform(-0.5, 0.5, nd)
    lam = float("nan")
    for _ in range(iters):
        y = lu.solve(x)
        x = y / np.linalg.norm(y)
        rq = float(x @ (A @ x))
        if np.isfinite(lam) and abs(rq - lam) <= tol * max(abs(rq), 1e-300):
            lam = rq
            break
        lam = rq
    return {"lam": lam, "neg": neg, "ndof": nd, "colours": L.ncol, "evaluations": len(states),
            "nnz": int(lu.L.nnz + lu.U.nnz), "ms": (time.perf_counter() - t0) * 1e3, "vec": x}
