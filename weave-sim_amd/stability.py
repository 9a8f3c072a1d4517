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
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _qexp(v):
    a = np.linalg.norm(v)
    half = 0.5 * a
    s = np.sin(half) / a if a > 1e-12 else 0.5
    return np.array([np.cos(half), v[0] * s, v[1] * s, v[2] * s])


def _qmul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


class _Layout:
    """DOFs and the stencil of a model (dofList, stencil; stability.js:17-23, 57-94)"""

    def __init__(self, model, u, W=3):
        md = model.md
        n, m, S = md.n, md.m, md.S
        free = np.ctypeslib.as_array(md.node_free, shape=(3 * n,)).astype(bool)
        self.dofs = [(0, i, k) for i in range(n) for k in range(3) if free[3 * i + k]] + \
                    [(1, j, k) for j in range(m) for k in range(3)]
        nd = len(self.dofs)
        node_dofs = [[] for _ in range(n)]
        seg_dofs = [[] for _ in range(m)]
        for d, (kind, i, _) in enumerate(self.dofs):
            (node_dofs if kind == 0 else seg_dofs)[i].append(d)
        noff = np.ctypeslib.as_array(md.node_offset, shape=(S + 1,))
        soff = np.ctypeslib.as_array(md.seg_offset, shape=(S + 1,))
        node_strand = np.repeat(np.arange(S), np.diff(noff))
        seg_strand = np.repeat(np.arange(S), np.diff(soff))
        closed = (np.diff(soff) == np.diff(noff))

        def window(s, loc, w):
            out = []
            nn, ns = noff[s + 1] - noff[s], soff[s + 1] - soff[s]
            for o in range(-w, w + 1):
                a = loc + o
                if closed[s]:
                    a %= nn
                if 0 <= a < nn:
                    out += node_dofs[noff[s] + a]
                b = loc + o
                if closed[s] and ns:
                    b %= ns
                if 0 <= b < ns:
                    out += seg_dofs[soff[s] + b]
            return out

        N = [set() for _ in range(nd)]
        for d, (kind, i, _) in enumerate(self.dofs):
            s = node_strand[i] if kind == 0 else seg_strand[i]
            loc = i - (noff[s] if kind == 0 else soff[s])
            N[d].update(window(s, loc, W))
        if md.par.contacts and md.K:
            ks = np.ctypeslib.as_array(md.k_strand, shape=(2 * md.K,))
            for c in range(md.K):
                st = window(ks[2 * c], int(np.floor(u[2 * c])), W + 2) + window(ks[2 * c + 1], int(np.floor(u[2 * c + 1])), W + 2)
                for a in st:
                    N[a].update(st)
        self.N = [np.array(sorted(x), dtype=np.int64) for x in N]
        # greedy distance-2 colouring: columns of one colour never touch a common row (colour, stability.js:97-108)
        col = np.full(nd, -1, dtype=np.int64)
        mark = np.full(nd + 1, -1, dtype=np.int64)
        ncol = 0
        for d in range(nd):
            for r in self.N[d]:
                cs = col[self.N[r]]
                mark[cs[cs >= 0]] = d
            c = 0
            while mark[c] == d:
                c += 1
            col[d] = c
            ncol = max(ncol, c + 1)
        self.col, self.ncol = col, ncol


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
    dofs, nd = L.dofs, len(L.dofs)
    if nd == 0:
        return {"lam": float("nan"), "neg": 0, "ndof": 0}
    P = X0.reshape(-1, 3)
    size = max(float(np.linalg.norm(P.max(0) - P.min(0))), 1e-9)
    hT, hR = 1e-6 * size, 1e-6
    by_col = [np.flatnonzero(L.col == c) for c in range(L.ncol)]

    def perturbed(ds, sgn):
        X, Q = X0.copy(), Q0.copy()
        for d in ds:
            kind, i, k = dofs[d]
            if kind == 0:
                X[3 * i + k] += sgn * hT
            else:
                v = np.zeros(3)
                v[k] = sgn * hR
                q = _qmul(Q[4 * i:4 * i + 4], _qexp(v))
                Q[4 * i:4 * i + 4] = q / np.linalg.norm(q)
        return X, Q

    states = [(X0, Q0)]
    for c in range(L.ncol):
        states += [perturbed(by_col[c], +1), perturbed(by_col[c], -1)]
    evals = evaluate(states)
    kind = np.array([d[0] for d in dofs])
    idx_i = np.array([d[1] for d in dofs])
    idx_k = np.array([d[2] for d in dofs])

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
        for d in by_col[c]:
            h = hT if kind[d] == 0 else hR
            r = L.N[d]
            v = -(gp[r] - gm[r]) / (2 * h)
            nz = v != 0
            rows.append(r[nz]); cols.append(np.full(int(nz.sum()), d)); vals.append(v[nz])
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
