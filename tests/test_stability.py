"""The stability check (stability.js) over the oracle's evaluator, pinned by the reference's own known-answer cases
(stability_check.mjs): a pinned–pinned Euler column whose lowest eigenvalue falls linearly to zero at P_e = π²EI/L²
(one negative mode beyond), and lateral–torsional buckling of a 60 × 6 mm cantilever strip under an edgewise tip load,
whose eigenvalue crosses zero at P_cr = 4.013 √(EI GJ)/L² (with the pre-buckling correction; the reference reports
21.0 N against 21.1 N)."""
import importlib
import math

import numpy as np
import pytest

import cases
from oracle.bind import Oracle

W = cases.W
A = W._abi
ST = importlib.import_module("weave-sim_amd.stability")
E, G = 1e4, 625.0


def column(P, w=60.0, h=30.0, nseg=20, anchors=None, load_axis=0):
    sc = W.Scene(E=E, G=G, width=w, height=h, gravity=0, contacts=0, tol=1e-10)
    m = W.Model.from_polylines([cases.straight(1000.0, nseg)], scene=sc, anchors=anchors)
    F = [0.0, 0.0, 0.0]
    F[load_axis] = -P
    m.add_load(nseg, F)
    return m


def relaxed(m, max_iter=400000):
    o = Oracle(m)
    o.run(max_iter, chunk=5000)
    assert o.converged
    st = o.state()
    return o, st["X"].copy(), st["Q"].copy(), st["u"].copy()


def oracle_stability(m):
    o, X, Q, u = relaxed(m)
    return ST.stability(m, lambda states: [o.evaluate(x, q) for x, q in states], X=X, Q=Q, u=u)


PINNED = [(A.WSM_FRAME, 0, 0, [1, 1, 1, 0, 0, 1]), (A.WSM_FRAME, 0, 1, [0, 1, 1, 0, 0, 0])]


def test_euler_column_eigenvalue_falls_linearly():
    EI = E * 60 * 30 ** 3 / 12
    Pe = math.pi ** 2 * EI / 1e6
    s0 = oracle_stability(column(0.0, anchors=PINNED))
    s5 = oracle_stability(column(0.5 * Pe, anchors=PINNED))
    s11 = oracle_stability(column(1.1 * Pe, anchors=PINNED))
    assert s0["neg"] == 0 and s5["neg"] == 0 and s11["neg"] >= 1
    assert abs(s5["lam"] / s0["lam"] - 0.5) < 0.03                    # λ/λ0 = 1 − P/Pe with the Euler mode
    assert s11["lam"] < 0


CLAMP = [(A.WSM_FRAME, 0, 0, [1, 1, 1, 1, 1, 1])]


def ltb_pcr():
    m0 = column(0.0, w=60.0, h=6.0, nseg=40, anchors=CLAMP, load_axis=1)
    EIa, EIb, GJ = m0.arr("EIx")[0], m0.arr("EIy")[0], m0.arr("GJ")[0]
    EIz, EIy = min(EIa, EIb), max(EIa, EIb)
    return 4.013 * math.sqrt(EIz * GJ) / 1e6 / math.sqrt((1 - EIz / EIy) * (1 - GJ / EIy))


def test_lateral_torsional_buckling():
    Pcr = ltb_pcr()
    assert 20.0 < Pcr < 22.5                                          # the reference's 21.1 N case
    pts = []
    for r in (0.95, 1.05):
        s = oracle_stability(column(r * Pcr, w=60.0, h=6.0, nseg=40, anchors=CLAMP, load_axis=1))
        pts.append((r * Pcr, s["lam"], s["neg"]))
    (p0, l0, n0), (p1, l1, n1) = pts
    assert n0 == 0 and l0 > 0 and n1 >= 1 and l1 < 0
    cross = p0 + (p1 - p0) * l0 / (l0 - l1)                           # λ = 0 by linear interpolation
    assert abs(cross / Pcr - 1) < 0.01, (cross, Pcr)


def test_layout_stencil_and_distance2_colouring():
    """the DOF layout of a woven dome (its crossings join the stencils of both laths): free node components first, then
    three rotations per segment; symmetric stencil rows that hold the DOF itself; no two columns of one colour share a
    row; and each DOF takes the smallest colour left free by the earlier DOFs within distance 2 (stability.js:97-108)"""
    import time
    m = cases.config_small()
    u = m["u"].copy()
    t = time.perf_counter()
    L = ST._Layout(m, u)
    dt = time.perf_counter() - t
    assert m.K > 0
    free = m.arr("node_free").astype(bool)
    nt = int(free.sum())
    assert len(L) == nt + 3 * m.m
    assert (L.kind[:nt] == 0).all() and (L.kind[nt:] == 1).all()
    assert np.array_equal(3 * L.index[:nt] + L.comp[:nt], np.flatnonzero(free))
    rows = [set(L.row(d).tolist()) for d in range(len(L))]
    for d, r in enumerate(rows):
        assert d in r
        for a in r:
            assert d in rows[a], (d, a)
    for r in rows:
        cs = L.col[list(r)]
        assert len(set(cs.tolist())) == len(cs)                 # distance-2: one colour per row at most once
    for d in range(len(L)):
        used = {int(L.col[q]) for a in rows[d] for q in rows[a] if q < d}
        c = 0
        while c in used:
            c += 1
        assert L.col[d] == c, d
    assert L.ncol == L.col.max() + 1 and dt < 2.0, dt
