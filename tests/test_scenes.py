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


def test_hybrid_groups_that_differ_are_refused_and_equal_ones_accepted():
    with pytest.raises(NotImplementedError):
        SC.scene_from_dict(scene_json("dualv5dome_cm_hybrid"))
    SC.scene_from_dict(scene_json("v5dome_cm"))              # its second group equals the first
    d = scene_json("dome_basic")
    d["anchors"].append({"select": {"type": "region", "id": [0, 0], "radius": 10.0}, "fix": [1, 1, 1, 0, 0, 0]})
    with pytest.raises(NotImplementedError):
        SC.scene_from_dict(d)
    with pytest.raises(ValueError):
        SC.scene_from_dict({"version": 2})
    # a group that leaves a key out takes its default, not the scene's value; an unassigned group is not used
    d = scene_json("v5dome_cm")
    del d["hybrid"]["groups"][0]["section"]["width"]                  # default 60 ≠ the scene's 2.0
    with pytest.raises(NotImplementedError):
        SC.scene_from_dict(d)
    d["hybrid"]["assign"] = {}
    SC.scene_from_dict(d)


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


@pytest.mark.parametrize("name", SCENES)
def test_oracle_relaxes_every_scene(name):
    m = SC.load_scene(os.path.join(HERE, name + ".json"))
    o = Oracle(m)
    o.step(400)
    st = o.state()
    assert np.isfinite(st["X"]).all() and np.isfinite(o.residual) and o.residual > 0
    assert np.abs(st["X"] - m["X"]).max() > 0                # it moved once the warm-up ended
