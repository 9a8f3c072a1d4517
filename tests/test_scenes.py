"""The reference's example scenes (scene.json + mesh, examples/scenes and examples/meshes, kept as input fixtures under
tests/golden) read by the scene reader and built by the host model builder.

What pins the reader (no golden outputs exist for these scenes): the selector semantics (all_start / all_end fix one
frame per open strand end, 'knots' the listed frames, united per frame, model.js:148-157), load totals that follow in
closed form from the scene's numbers (self-weight = ρ·A·g per length × Σ l0; a UDL per length × Σ l0; per area × the
mesh area; world point loads as given, model.js:191-231), the unit conversions of units.js (a cm scene's E90 default
is 800 N/mm² × 100), binary and ASCII STL giving the same triangles, and the f64 oracle relaxing every scene."""
import importlib
import json
import os

import numpy as np
import pytest

from oracle.bind import Oracle

W = importlib.import_module("weave-sim_amd")
SC = importlib.import_module("weave-sim_amd.scene")
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
SCENES = ["dome_basic", "v5dome_cm", "tree_cm", "dualv5dome_cm"]


def scene_json(name):
    with open(os.path.join(HERE, name + ".json")) as f:
        return json.load(f)


def mesh_area(name):
    d = scene_json(name)
    V, F = SC.read_mesh(os.path.join(HERE, d["mesh"]["path"]), d["mesh"]["scale"])
    T = V.reshape(-1, 3)[F.reshape(-1, 3)]
    return 0.5 * np.linalg.norm(np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]), axis=1).sum()


@pytest.mark.parametrize("name", SCENES)
def test_scene_builds_with_its_anchors(name):
    d = scene_json(name)
    m = SC.load_scene(os.path.join(HERE, name + ".json"))
    assert m.K == m.info.nE and m.K > 0 and m.info.parity_conflicts == 0     # one knot per mesh edge
    fixed = (m.arr("node_free").reshape(-1, 3) == 0).any(1).sum()
    sel = [a["select"] for a in d["anchors"]]
    closed = m.arr("strand_closed")
    if all(s["type"] in ("all_start", "all_end") for s in sel):
        assert fixed == len(sel) * int((closed == 0).sum())
    else:
        frames = {tuple(i) for s in sel for i in s["ids"]}
        assert 0 < fixed <= len(frames)


@pytest.mark.parametrize("name", SCENES)
def test_scene_load_totals_in_closed_form(name):
    d = scene_json(name)
    m = SC.load_scene(os.path.join(HERE, name + ".json"))
    L = SC.LENGTH_TO_M[d["units"]["length"]]
    F = SC.FORCE_TO_N[d["units"]["force"]]
    sec, mat = d["section"], d["material"]
    qline = mat["density"] * (sec["width"] * sec["height"] * L * L) * 9.80665 / F * L      # units.js:21-25
    l0 = m.arr("l0").sum()
    g = np.array(d["loads"]["gravity"]["direction"], dtype=float)
    expect = qline * l0 * g / np.linalg.norm(g)
    udl = d["loads"]["udl"]
    if udl["enabled"]:
        expect += np.array(udl["q"]) * (l0 if udl["per"] == "length" else mesh_area(name))
    for p in d["loads"]["point"]:
        if p["space"] == "world":
            expect += np.array(p["F"])
    got = m.arr("load_const").reshape(-1, 3).sum(0)
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max()), (got, expect)
    assert m.md.n_point_local == sum(1 for p in d["loads"]["point"] if p["space"] == "local")


def test_scene_parameters_and_units():
    sc, anchors, loads, flips, mesh = SC.scene_from_dict(scene_json("dualv5dome_cm"))
    p, c = sc.c.par, sc.c
    assert c.length_to_m == 1e-2 and c.force_to_n == 1.0
    assert p.rot_friction == 1 and p.seat_law == 1 and p.seat_gap == 1
    assert abs(p.edge_threshold - 0.34) < 1e-7 and abs(p.rot_factor - 0.4) < 1e-7 and abs(p.mu - 0.291) < 1e-7
    assert c.udl == 2 and c.grp.beam_advanced == 0 and abs(c.grp.activation - 0.62) < 1e-7
    assert len(anchors) == 16 and all(a[0] == 0 for a in anchors)                # 'knots' → one frame anchor per id
    t, _, _, _, _ = SC.scene_from_dict(scene_json("tree_cm"))
    assert t.c.par.augmented == 1 and t.c.par.al_every == 1200 and t.c.udl == 1
    b, _, loads, flips, _ = SC.scene_from_dict(scene_json("dome_basic"))
    assert b.c.length_to_m == 1e-3 and flips == [(8, 7)] and b.c.overhang == 150.0
    assert [(l[0], l[1], l[3]) for l in loads] == [(18, 6, False), (37, 12, True)]
    # a scene without E90 / fc90 / H takes the N/mm² defaults in its own stress unit (scene.js:144-147)
    d = scene_json("v5dome_cm")
    for k in ("E90", "fc90", "H"):
        del d["material"][k]
    s, _, _, _, _ = SC.scene_from_dict(d)
    assert (s.c.grp.E90, s.c.grp.fc90, s.c.grp.H) == pytest.approx((80000.0, 800.0, 3000.0))   # the file's own values


def _strand_of(m, offsets, count):
    off = m.arr(offsets)
    return np.repeat(np.arange(len(off) - 1), np.diff(off))[:count]


def _build(d, name="dualv5dome_cm_hybrid"):
    return SC.model_from_scene(d, base_dir=HERE)


def test_hybrid_groups_per_strand_in_closed_form():
    """dualv5dome_cm_hybrid: six strands in group 2 (E 5e6, 6 × 2 section) and the rest in the scene's own (E 1e6,
    4 × 0.8).  Each segment's stiffness is its strand's group's section stiffness × its beam factors, each crossing
    side's section and material its strand's group's (a group's E90 is taken as given, the scene's own default is
    rescaled to its stress unit), each crossing's section the mean of its sides, each node's floor section its strand's
    group's, and the self-weight Σ over segments of ρ·A·g of its group × l0 (model.js:203-252, solver_cpu.js:66-78,
    scene.js:57-61)."""
    d = scene_json("dualv5dome_cm_hybrid")
    m = _build(d)
    g1, g2 = d, d["hybrid"]["groups"][0]
    grp = {int(s): int(g) for s, g in d["hybrid"]["assign"].items()}
    S = m.md.S
    gs = np.array([grp.get(s, 1) for s in range(S)])
    assert (gs == 2).sum() == 6
    E = np.where(gs == 2, g2["material"]["E"], g1["material"]["E"])
    w = np.where(gs == 2, g2["section"]["width"], g1["section"]["width"])
    h = np.where(gs == 2, g2["section"]["height"], g1["section"]["height"])
    rho = np.where(gs == 2, g2["material"]["density"], g1["material"]["density"])
    k = np.where(gs == 2, g2["beam"]["stiffness"], g1["beam"]["stiffness"])
    E90 = np.where(gs == 2, 800.0, 80000.0)          # group 2 leaves E90 out: 800 as given; the scene's own: 800 N/mm² in N/cm²
    sj = _strand_of(m, "seg_offset", m.m)
    assert np.allclose(m.arr("EA"), (E * w * h)[sj], rtol=1e-14)
    assert np.allclose(m.arr("EIx"), (k * E * w * h ** 3 / 12)[sj], rtol=1e-14)
    assert np.allclose(m.arr("EIy"), (k * E * h * w ** 3 / 12)[sj], rtol=1e-14)
    ks = m.arr("k_strand")
    assert np.array_equal(m.arr("k_wk"), w[ks]) and np.array_equal(m.arr("k_hk"), h[ks])
    assert np.array_equal(m.arr("k_E90"), E90[ks])
    assert np.array_equal(m.arr("k_w"), 0.5 * (w[ks[0::2]] + w[ks[1::2]]))
    assert np.array_equal(m.arr("k_h"), 0.5 * (h[ks[0::2]] + h[ks[1::2]]))
    assert len(set(zip(gs[ks[0::2]], gs[ks[1::2]]))) >= 3           # crossings of 1 × 1, 1 × 2 and 2 × 2 laths
    si = _strand_of(m, "node_offset", m.n)
    assert np.array_equal(m.arr("node_w"), w[si]) and np.array_equal(m.arr("node_h"), h[si])
    L = SC.LENGTH_TO_M[d["units"]["length"]]
    qline = rho * (w * h * L * L) * 9.80665 * L                     # units.js:21-25, force unit N
    gdir = np.array(d["loads"]["gravity"]["direction"], dtype=float)
    expect = (m.arr("l0") * qline[sj]).sum() * gdir / np.linalg.norm(gdir)
    udl = d["loads"]["udl"]
    if udl["enabled"]:
        expect += np.array(udl["q"]) * (m.arr("l0").sum() if udl["per"] == "length" else mesh_area("dualv5dome_cm_hybrid"))
    got = m.arr("load_const").reshape(-1, 3).sum(0)
    assert np.allclose(got, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max()), (got, expect)


def test_hybrid_group_assigned_everywhere_equals_a_scene_of_that_group():
    """every strand in group 2 builds the very model of a scene whose own material / section / beam are group 2's
    (the nodes between knots are lifted by the scene's own half-height, topology.js:236-249, so that scene keeps
    group 2's height), and the knot lift of a crossing is the mean of its laths' group heights (topology.js:225): a
    crossing of two group-1 laths sits where it does with no groups, one of two group-2 laths where it does with
    every lath in group 2, and a mixed one halfway between"""
    d = scene_json("dualv5dome_cm_hybrid")
    g2 = d["hybrid"]["groups"][0]
    for g in (d, g2):
        g["material"].update({"E90": 800.0, "fc90": 8.0, "H": 30.0})
    S = SC.model_from_scene(dict(d, hybrid={"enabled": False}), base_dir=HERE).md.S
    every = _build(dict(d, section=dict(d["section"], height=g2["section"]["height"]),
                        hybrid=dict(d["hybrid"], assign={str(s): 2 for s in range(S)})))
    own = _build(dict(d, material=g2["material"], section=g2["section"], beam=g2["beam"], hybrid={"enabled": False}))
    names = ["l0", "j_len", "EA", "GA", "EIx", "EIy", "GJ", "omega0", "load_const", "node_free", "k_wk", "k_hk",
             "k_w", "k_h", "k_fc90", "k_E90", "k_Hs", "node_w", "node_h", "mass"]
    for nm in names:
        assert np.array_equal(every.arr(nm), own.arr(nm)), nm
    assert np.array_equal(every["X"], own["X"]) and np.array_equal(every["Q"], own["Q"])
    plain = _build(dict(d, hybrid={"enabled": False}))
    mixed = _build(d)
    gs = np.ones(S, dtype=int)
    for s, g in d["hybrid"]["assign"].items():
        gs[int(s)] = g
    ks = mixed.arr("k_strand")
    both = gs[ks[0::2]] + gs[ks[1::2]]                              # 2, 3 or 4

    def knots(x):
        """each side's knot node: u starts at the knot row, projected onto the lath to within a fraction of a segment"""
        assert np.array_equal(x.arr("k_strand"), ks)
        row = np.rint(x["u"])
        assert np.abs(x["u"] - row).max() < 0.4
        return x["X"].reshape(-1, 3)[x.arr("node_offset")[ks] + row.astype(int)].reshape(-1, 2, 3)
    Xp, Xe, Xm = knots(plain), knots(every), knots(mixed)
    span = plain.span
    assert np.array_equal(Xm[both == 2], Xp[both == 2]) and np.array_equal(Xm[both == 4], Xe[both == 4])
    assert np.abs(Xm[both == 3] - 0.5 * (Xp[both == 3] + Xe[both == 3])).max() < 1e-14 * span
    assert np.abs(Xp[both == 3] - Xe[both == 3]).max() > 1e-3 * span


def _piece_scene():
    """dualv5dome_cm_hybrid with joints on two strands and groups assigned to pieces: strand s0 (group 1) gets joints
    at 65 % and 30 % of its length (unsorted) and its middle piece in group 2; strand s2 (group 1, no joints) its only
    piece in group 2; keys that are not in pieceKey's exact form, or name a group outside 2 … 1 + groups, apply to
    nothing"""
    d = scene_json("dualv5dome_cm_hybrid")
    base = _build(d)
    S = base.md.S
    grp = {int(s): int(g) for s, g in d["hybrid"]["assign"].items()}
    no = base.arr("node_offset")
    free = [s for s in range(S) if grp.get(s, 1) == 1 and no[s + 1] - no[s] > 20 and not base.arr("strand_closed")[s]]
    s0, s2 = free[0], free[1]
    X = base["X"].reshape(-1, 3)
    P = X[no[s0]:no[s0 + 1]]
    Sa = np.r_[0.0, np.cumsum(np.linalg.norm(np.diff(P, axis=0), axis=1))]
    d["joints"] = {str(s0): [0.65 * Sa[-1], 0.3 * Sa[-1]]}
    d["hybrid"]["pieces"] = {f"{s0}:1": 2, f"{s2}:0": 2, f"{s0}:02": 2, f"{s0}:2": 7, f"{s0}:0": 1}
    return d, base, s0, s2, Sa


def test_hybrid_piece_groups_by_arc_length():
    """groups assigned to pieces of a discretisation (Scene.hybrid.pieces with Scene.joints; scene.js:57-90,
    discretize.js:334-358): a segment takes the group of the piece that holds its midpoint and a node the piece that
    holds it, by arc length along the strand on the built positions, for its stiffness, self-weight, floor section and
    crossing-side sections; every other strand is bitwise as in the hybrid model without pieces, and so is the
    geometry (the knot lift follows the strands' groups, topology.js:225)"""
    d, base, s0, s2, Sa = _piece_scene()
    m = _build(d)
    g1, g2 = d, d["hybrid"]["groups"][0]
    EA = {1: g1["material"]["E"] * g1["section"]["width"] * g1["section"]["height"],
          2: g2["material"]["E"] * g2["section"]["width"] * g2["section"]["height"]}
    W = {1: g1["section"]["width"], 2: g2["section"]["width"]}
    assert np.array_equal(m["X"], base["X"]) and np.array_equal(m.arr("l0"), base.arr("l0"))
    no, so = m.arr("node_offset"), m.arr("seg_offset")
    js = np.sort(np.array(d["joints"][str(s0)]))
    X = m["X"].reshape(-1, 3)
    i0, j0, j1 = no[s0], so[s0], so[s0 + 1]
    a, b = m.arr("seg_i0")[j0:j1], m.arr("seg_i1")[j0:j1]
    mid = Sa[a - i0] + 0.5 * np.linalg.norm(X[b] - X[a], axis=1)
    gseg = np.where(np.searchsorted(js, mid, "right") == 1, 2, 1)
    gnode = np.where(np.searchsorted(js, Sa, "right") == 1, 2, 1)
    assert set(gseg) == {1, 2} and set(gnode) == {1, 2}
    assert np.allclose(m.arr("EA")[j0:j1], [EA[g] for g in gseg], rtol=1e-14)
    assert np.array_equal(m.arr("node_w")[i0:no[s0 + 1]], [W[g] for g in gnode])
    # the whole joint-less strand s2 in group 2, every other strand as before
    j2 = slice(so[s2], so[s2 + 1])
    assert np.allclose(m.arr("EA")[j2], EA[2], rtol=1e-14) and np.allclose(base.arr("EA")[j2], EA[1], rtol=1e-14)
    rest = np.ones(m.m, bool); rest[j0:j1] = False; rest[j2] = False
    assert np.array_equal(m.arr("EA")[rest], base.arr("EA")[rest])
    # crossing sides take the group of the piece at their knot node
    ks, fr = m.arr("k_strand"), {(s, k): nd for s, k, nd in m.frames()}
    kn = m.knots()
    node = np.array([[fr[(r[1], r[2])], fr[(r[3], r[4])]] for r in kn]).reshape(-1)
    on = ks == s0
    assert np.array_equal(m.arr("k_wk")[on], [W[g] for g in gnode[node[on] - i0]])
    assert np.array_equal(m.arr("k_wk")[ks == s2], np.full((ks == s2).sum(), W[2]))
    other = (ks != s0) & (ks != s2)
    assert np.array_equal(m.arr("k_wk")[other], base.arr("k_wk")[other])
    # self-weight: ρ·A·g of each segment's group × l0, half to each end
    L = SC.LENGTH_TO_M[d["units"]["length"]]
    q = {g: gd["material"]["density"] * gd["section"]["width"] * gd["section"]["height"] * L * L * 9.80665 * L for g, gd in ((1, g1), (2, g2))}
    dl = (m.arr("load_const") - base.arr("load_const")).reshape(-1, 3).sum(0)
    l0 = m.arr("l0")
    gdir = np.array(d["loads"]["gravity"]["direction"], dtype=float)
    changed = l0[j0:j1][gseg == 2].sum() + l0[j2].sum()
    assert np.allclose(dl, changed * (q[2] - q[1]) * gdir / np.linalg.norm(gdir), rtol=1e-9, atol=1e-9 * q[2] * changed)


def test_hybrid_pieces_of_every_strand_equal_strand_groups():
    """every piece of every strand (joint-less: piece 0) assigned as the strands are builds the hybrid model itself"""
    d = scene_json("dualv5dome_cm_hybrid")
    base = _build(d)
    d["hybrid"]["pieces"] = {f"{s}:0": g for s, g in d["hybrid"]["assign"].items()}
    m = _build(d)
    for nm in ["EA", "GA", "EIx", "EIy", "GJ", "omega0", "load_const", "k_wk", "k_hk", "k_E90", "node_w", "node_h", "mass"]:
        assert np.array_equal(m.arr(nm), base.arr(nm)), nm


def test_scene_version_and_group_defaults():
    with pytest.raises(ValueError):
        SC.scene_from_dict({"version": 2})
    # a group that leaves a key out takes its default, not the scene's value
    d = scene_json("v5dome_cm")
    del d["hybrid"]["groups"][0]["section"]["width"]
    sc, _, _, _, _ = SC.scene_from_dict(d)
    assert sc.groups[0].width == 60.0 and sc.c.grp.width == 2.0 and len(sc.assign) > 0


def test_binary_and_ascii_stl_agree():
    path = os.path.join(HERE, "..", "meshes", "dome.stl")
    with open(path, "rb") as f:
        buf = f.read()
    Vb, Fb = SC.parse_stl(buf)
    T = Vb.reshape(-1, 3, 3)
    lines = ["solid s"]
    for t in T:
        lines += ["facet normal 0 0 0", "outer loop"] + [f"vertex {float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in t] + ["endloop", "endfacet"]
    Va, Fa = SC.parse_stl(("\n".join(lines + ["endsolid s"])).encode())
    assert np.array_equal(Va, Vb) and np.array_equal(Fa, Fb) and len(Fb) == 3 * len(T)
    Vo, Fo = SC.parse_obj(open(os.path.join(HERE, "..", "meshes", "genus1_torus_v6_696v.obj")).read())
    assert len(Vo) == 3 * 696 and Fo.min() == 0 and Fo.max() == 695


@pytest.mark.parametrize("name", SCENES + ["dualv5dome_cm_hybrid"])
def test_oracle_relaxes_every_scene(name):
    m = SC.load_scene(os.path.join(HERE, name + ".json"))
    o = Oracle(m)
    o.step(400)
    st = o.state()
    assert np.isfinite(st["X"]).all() and np.isfinite(o.residual) and o.residual > 0
    assert np.abs(st["X"] - m["X"]).max() > 0                # it moved once the warm-up ended


def _mesh_graph(V, F):
    """the builder's welded mesh and its edges, restated with numpy: vertices quantised at 1e-9 of the model size and
    deduplicated in lexicographic order (weld, mesh.js:106-129), edges the sorted unique vertex pairs (edgeTopology,
    mesh.js:133-150; edge id = row)"""
    V = V.reshape(-1, 3)
    q = max(np.ptp(V, axis=0).max(), 1.0) * 1e-9
    _, first, inv = np.unique(np.round(V / q), axis=0, return_index=True, return_inverse=True)
    Vw = V[first]
    Fw = inv.reshape(-1)[F.reshape(-1, 3)]
    Fw = Fw[(Fw[:, 0] != Fw[:, 1]) & (Fw[:, 1] != Fw[:, 2]) & (Fw[:, 0] != Fw[:, 2])]
    rows = np.concatenate([np.sort(Fw[:, [(i + 1) % 3, (i + 2) % 3]], axis=1) for i in range(3)])
    return Vw, np.unique(rows, axis=0)


@pytest.mark.parametrize("radius", [0.0, 450.0, 1200.0, 1e12])
def test_region_selector_by_mesh_graph_distance(radius):
    """the 'region' anchor (regionFrames, model.js:127-145): the frames of every knot whose mesh edge's two vertices
    lie within `radius` on average, by shortest paths along the mesh edges from the seed knot's edge.  Held to a
    restatement with SciPy's Dijkstra on the numpy-welded mesh; radius 0 anchors the seed knot alone and an
    unbounded radius every knot"""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import dijkstra
    d = scene_json("dome_basic")
    V, F = SC.read_mesh(os.path.join(HERE, d["mesh"]["path"]), d["mesh"]["scale"])
    Vw, E = _mesh_graph(V, F)
    probe = SC.load_scene(os.path.join(HERE, "dome_basic.json"))
    kn = probe.knots()
    assert len(E) == probe.info.nE
    seed = kn[len(kn) // 2]                          # a knot in the middle of the list: frame (strand_a, k_a)
    d["anchors"] = [{"select": {"type": "region", "id": [int(seed[1]), int(seed[2])], "radius": radius},
                     "fix": [1, 1, 1, 0, 0, 0]}]
    m = SC.model_from_scene(d, base_dir=HERE)
    fixed = set(np.nonzero((m.arr("node_free").reshape(-1, 3) == 0).all(1))[0].tolist())
    # the restatement
    L = np.linalg.norm(Vw[E[:, 0]] - Vw[E[:, 1]], axis=1)
    G = coo_matrix((np.r_[L, L], (np.r_[E[:, 0], E[:, 1]], np.r_[E[:, 1], E[:, 0]])), shape=(len(Vw), len(Vw))).tocsr()
    dist = dijkstra(G, indices=[E[seed[0], 0], E[seed[0], 1]], min_only=True)
    near = 0.5 * (dist[E[kn[:, 0], 0]] + dist[E[kn[:, 0], 1]]) <= radius
    node = {(s, k): nd for s, k, nd in m.frames()}
    want = {node[(a, b)] for e, a, b, _, _, _ in kn[near]} | {node[(a, b)] for e, _, _, a, b, _ in kn[near]}
    print(f"\nradius {radius}: {int(near.sum())} of {len(kn)} knots, {len(fixed)} nodes anchored")
    assert fixed == want
    if radius == 0.0:
        assert near.sum() == 1
    if radius > 1e9:
        assert near.all()
    if 0.0 < radius < 1e9:
        assert 1 < near.sum() < len(kn)
