"""The host model builder on the reference's own example meshes (examples/meshes, copied under tests/golden/meshes as
fixtures): the genus-1 torus of the reference's self-crossing check (selfcross_check.mjs) — laths that wind round
and cross themselves, one over pass per knot, the passes h apart along the surface normal, over/under alternating
along every lath including the wrap of closed laths, and a flip that swaps both passes of a self-crossing; the
v5 dome STL builds with one knot per interior mesh edge.  And the builder's speed at the largest configured size."""
import os
import time

import numpy as np

import cases

W = cases.W
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def load_obj(path):
    V, F = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "v":
                V.append([float(x) for x in t[1:4]])
            elif t and t[0] == "f":
                F.append([int(x.split("/")[0]) - 1 for x in t[1:4]])
    return np.array(V), np.array(F)


def load_ascii_stl(path):
    P = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "vertex":
                P.append([float(x) for x in t[1:4]])
    P = np.array(P)
    return P, np.arange(len(P)).reshape(-1, 3)          # unwelded triangle soup: the builder welds it


def torus(flips=()):
    V, F = load_obj(os.path.join(HERE, "genus1_torus_v6_696v.obj"))
    # selfcross_check.mjs:12 — subdivisions 4, no overhang, half-height offsets, h = 0.8
    sc = W.Scene(subdivisions=4, overhang=0.0, offset_half_height=1, height=0.8, width=3.0, gravity=0)
    return W.Model.from_mesh(V, F, scene=sc, flips=flips), V, F


def knot_passes(m, V, F):
    """per crossing side: its knot node, and whether it passes over (along the normal of the knot's mesh edge)"""
    X = m["X"].reshape(-1, 3)
    noff, ks = m.arr("node_offset"), m.arr("k_strand")
    nn = np.diff(noff)
    u0 = np.round(m["u"]).astype(int)
    node = noff[ks] + np.mod(u0, nn[ks])
    # face normals → edge normals at the edge midpoints; a knot sits at the midpoint of its edge
    fn = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    edges = {}
    for f, tri in enumerate(F):
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            edges.setdefault((min(a, b), max(a, b)), []).append(f)
    keys = list(edges)
    mid = np.array([0.5 * (V[a] + V[b]) for a, b in keys])
    nrm = np.array([fn[edges[k]].sum(0) for k in keys])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    A, B = X[node[0::2]], X[node[1::2]]
    e = np.argmin(np.linalg.norm((0.5 * (A + B))[:, None, :] - mid[None, :, :], axis=2), axis=1)
    d = np.einsum("ij,ij->i", A - B, nrm[e])
    over = np.empty(2 * m.K, dtype=bool)
    over[0::2], over[1::2] = d > 0, d < 0
    return node, over, d, ks


def test_self_crossing_torus():
    m, V, F = torus()
    h = 0.8
    node, over, d, ks = knot_passes(m, V, F)
    self_x = np.flatnonzero(ks[0::2] == ks[1::2])
    closed = m.arr("strand_closed")
    assert len(self_x) > 0 and closed.all() and m.K == 3 * len(V)     # every edge of the genus-1 mesh a knot
    assert m.info.parity_conflicts == 0
    assert (np.abs(d) >= 0.99 * h).all() and (np.abs(d) <= 1.01 * h).all()   # ±h/2 on each side of the surface
    u = m["u"]
    for s in range(m.md.S):                                              # alternation, the wrap included
        q = np.flatnonzero(ks == s)
        flags = over[q[np.argsort(u[q])]]
        assert (flags != np.roll(flags, 1)).all(), s


def test_flip_swaps_both_passes_of_a_self_crossing():
    m, V, F = torus()
    _, over, _, ks = knot_passes(m, V, F)
    c = int(np.flatnonzero(ks[0::2] == ks[1::2])[0])
    s, frame = int(ks[2 * c]), int(round(m["u"][2 * c])) // 4
    m1, _, _ = torus(flips=[(s, frame)])
    _, over1, _, _ = knot_passes(m1, V, F)
    assert over1[2 * c] != over[2 * c] and over1[2 * c + 1] != over[2 * c + 1]


def test_v5dome_stl_builds():
    V, F = load_ascii_stl(os.path.join(HERE, "v5dome.stl"))
    m = W.Model.from_mesh(V, F, scene=W.Scene(subdivisions=4, gravity=0))
    nE, nV = m.info.nE, m.info.nV
    assert nV < len(V) and m.K == nE and m.K > 0       # welded; one knot per mesh edge
    assert np.isfinite(m["X"]).all() and m.info.parity_conflicts >= 0


def test_builder_speed_at_two_million_nodes():
    t = time.perf_counter()
    m = cases.dome(30000.0, 200.0, 6.0, 0.18, 2.0, 5)
    dt = time.perf_counter() - t
    assert m.n > 1_900_000 and m.K == 50850
    assert dt < 60.0, dt
