"""Scenes as the reference stores them: a scene.json (version 1) and the mesh it names, read into a host model.

The reader follows Scene.fromJSON (web/js/scene.js:136-187): every section starts from the params.js defaults and
takes the keys the file holds; older files without E90 / fc90 / H get the N/mm² defaults in the scene's stress unit
(scene.js:144-147).  The mesh follows Mesh.fromBytes (web/js/mesh.js:24-31): STL (binary when the size says so,
else ASCII, parseSTL :62-86) or OBJ (fan-triangulated polygons, parseOBJ :88-101), scaled, then welded by the model
builder.  The instance is then built by wsm_from_mesh — the reference's buildTopology / buildState / Anchors / Loads
/ CpuSolver._buildGoals pipeline — exactly as for a mesh given directly.

A hybrid weave's lath groups (Scene.hybrid, scene.js:50-62) go to wsm_from_mesh_groups with the strand assignments:
each strand's segments, crossing sides, nodes and knot lifts take its group's material, section and beam.

Anchors take every selector of resolveSelector (model.js:115-125), the 'region' one (graph distance along the mesh
edges from a seed knot, model.js:127-145) included.

Groups assigned to pieces of a discretisation (Scene.hybrid.pieces, with the joints of Scene.joints) go to
wsm_from_mesh_pieces: a segment takes the group of the piece that holds its midpoint, a node the piece that holds
it, by arc length along the strand (discretize.js:334-358).
"""
import json
import os
import re
import struct

import numpy as np

from . import _abi as A
from .model import Model, Scene

SCENE_VERSION = 1                                   # scene.js:13
LENGTH_TO_M = {"mm": 1e-3, "cm": 1e-2, "m": 1.0}    # units.js:2
FORCE_TO_N = {"N": 1.0, "kN": 1e3}                  # units.js:3
FLOOR_MU = 0.4                                      # scene.js:14
_MATERIAL_MM_N = {"E90": 800.0, "fc90": 8.0, "H": 30.0}   # params.js:6, N/mm²
_SELECTOR = {"frame": 0, "all_start": 1, "all_end": 2, "all_ends": 3, "strand": 4, "boundary": 5, "region": 6}   # wsm_selector


# ------------------------------------------------------------------------------------------------ meshes
def parse_stl(buf):
    """triangle soup (V [9t], F [3t]) of an STL file's bytes: binary when 84 + 50·t bytes, else ASCII vertex lines"""
    if len(buf) >= 84:
        t = struct.unpack_from("<I", buf, 80)[0]
        if 84 + 50 * t == len(buf):
            rec = np.frombuffer(buf, dtype=np.uint8, offset=84).reshape(t, 50)
            V = rec[:, 12:48].copy().view("<f4").reshape(-1).astype(np.float64)
            return V, np.arange(3 * t, dtype=np.int32)
    text = buf.decode("utf-8", errors="replace")
    num = r"([-+0-9.eE]+)"
    vals = [float(x) for m in re.finditer(r"vertex\s+" + num + r"\s+" + num + r"\s+" + num, text) for x in m.groups()]
    t = len(vals) // 9
    if t == 0:
        raise ValueError("no triangles found (binary or ASCII STL expected)")
    return np.array(vals[:9 * t], dtype=np.float64), np.arange(3 * t, dtype=np.int32)


def parse_obj(text):
    """vertices and fan-triangulated faces of an OBJ file's text (1-based or negative indices)"""
    V, F = [], []
    for raw in text.splitlines():
        line = raw.strip()
        if line.startswith("v "):
            p = line.split()
            V.extend((float(p[1]), float(p[2]), float(p[3])))
        elif line.startswith("f "):
            nv = len(V) // 3
            idx = []
            for tok in line.split()[1:]:
                i = int(tok.split("/")[0])
                idx.append(nv + i if i < 0 else i - 1)
            for k in range(1, len(idx) - 1):
                F.extend((idx[0], idx[k], idx[k + 1]))
    if not F:
        raise ValueError("no faces in the OBJ file")
    return np.array(V, dtype=np.float64), np.array(F, dtype=np.int32)


def read_mesh(path, scale=1.0):
    """(V, F) of an STL or OBJ file, scaled (before welding, as Mesh.fromBytes)"""
    with open(path, "rb") as f:
        buf = f.read()
    V, F = parse_obj(buf.decode("utf-8", errors="replace")) if path.lower().endswith(".obj") else parse_stl(buf)
    return V * float(scale), F


# ------------------------------------------------------------------------------------------------ scene.json
def _flag(v):
    return 1 if v else 0


def _sec(d, key):
    """a section of the file as a dictionary: anything else reads as empty (a property of a non-object is undefined in
    the reference's reader)"""
    v = d.get(key) if isinstance(d, dict) else None
    return v if isinstance(v, dict) else {}


# params.js:6-19: what a section of a scene or of a lath group holds when the file leaves a key out
_DEFAULTS = {
    "material": {"E": 10000.0, "G": 625.0, "density": 700.0, "E90": 800.0, "fc90": 8.0, "H": 30.0},
    "section": {"width": 60.0, "height": 30.0, "Ix": None, "Iy": None, "J": None},
    "beam": {"mode": "simple", "activation": 1.0, "stiffness": 1.0, "activation_flat": 1.0, "activation_edge": 1.0,
             "activation_twist": 1.0, "stiffness_flat": 1.0, "stiffness_edge": 1.0, "stiffness_twist": 1.0,
             "stiffness_axial": 1.0, "stiffness_shear": 1.0},
}


def _int0(v):
    """JavaScript's `v | 0` of a key or value as the reader takes it (scene.js:168): 0 where it is not a number"""
    try:
        return int(float(v))
    except (TypeError, ValueError):
        return 0


def _fill_group(g, mat, sec, beam):
    """a wsm_group from the material / section / beam dictionaries, over what g holds already"""
    for k in ("E", "G", "density", "E90", "fc90", "H"):
        if k in mat and mat[k] is not None:
            setattr(g, k, float(mat[k]))
    for k in ("width", "height"):
        if k in sec and sec[k] is not None:
            setattr(g, k, float(sec[k]))
    for k in ("Ix", "Iy", "J"):
        setattr(g, k, float(sec[k]) if sec.get(k) is not None else 0.0)
    g.beam_advanced = _flag(beam.get("mode", "simple") == "advanced")
    for src, dst in (("activation", "activation"), ("stiffness", "stiffness"), ("activation_flat", "act_flat"),
                     ("activation_edge", "act_edge"), ("activation_twist", "act_twist"), ("stiffness_axial", "k_axial"),
                     ("stiffness_flat", "k_flat"), ("stiffness_edge", "k_edge"), ("stiffness_twist", "k_twist"),
                     ("stiffness_shear", "k_shear")):
        if src in beam:
            setattr(g, dst, float(beam[src]))
    return g


def _hybrid_groups(d):
    """(groups 2, 3, … as wsm_group, [(strand, group)]) of an enabled Scene.hybrid, or None: each group starts from
    the params.js defaults (fromDict, scene.js:167), its E90 / fc90 / H in N/mm² as given (not rescaled)"""
    h = _sec(d, "hybrid")
    if not h.get("enabled"):
        return None
    groups = []
    for gd in h.get("groups") or []:
        g = _fill_group(A.wsm_group(), _DEFAULTS["material"], _DEFAULTS["section"], _DEFAULTS["beam"])
        groups.append(_fill_group(g, _sec(gd, "material"), _sec(gd, "section"), _sec(gd, "beam")))
    assign = [(_int0(k), _int0(v)) for k, v in (h.get("assign") or {}).items()]
    return groups, assign


def _piece_groups(d):
    """(joints {strand: [arc lengths]}, [(strand, piece, group)]) of an enabled Scene.hybrid that assigns pieces of a
    discretisation (scene.js:57-90).  A piece is looked up by its key "s:p" (pieceKey, scene.js:10), so only keys in
    that exact form can apply; joints are Scene.joints (scene.js:164), arc lengths in scene units along each strand"""
    h = _sec(d, "hybrid")
    if not h.get("enabled"):
        return None, []
    pieces = []
    for k, v in (h.get("pieces") or {}).items():
        m = re.match(r"^(\d+):(\d+)$", str(k))
        if m and str(k) == f"{int(m.group(1))}:{int(m.group(2))}":
            pieces.append((int(m.group(1)), int(m.group(2)), _int0(v)))
    joints = {}
    for k, vs in (d.get("joints") or {}).items():
        if re.match(r"^\d+$", str(k)) and vs:
            joints[int(k)] = [float(x) for x in vs]
    return joints, pieces


def scene_from_dict(d):
    """(Scene, anchors, point loads, flips, mesh {path, scale}) of a scene.json dictionary"""
    d = d or {}
    v = d.get("version", SCENE_VERSION)
    if v > SCENE_VERSION:
        raise ValueError(f"scene version {v} newer than supported {SCENE_VERSION}")
    units = _sec(d, "units")
    lu, fu = units.get("length", "mm"), units.get("force", "N")
    if lu not in LENGTH_TO_M or fu not in FORCE_TO_N:
        raise ValueError(f"unknown units {lu}, {fu}")
    L, Fn = LENGTH_TO_M[lu], FORCE_TO_N[fu]
    sc = Scene()
    c = sc.c
    c.length_to_m, c.force_to_n = L, Fn

    stress = Fn / (L * L) / 1e6                    # scene stress unit in N/mm²
    sc.groups, sc.assign = _hybrid_groups(d) or (None, [])
    sc.joints, sc.pieces = _piece_groups(d)

    mat = _sec(d, "material")
    g = _fill_group(c.grp, mat, _sec(d, "section"), _sec(d, "beam"))
    for k, v0 in _MATERIAL_MM_N.items():
        if k not in mat:
            setattr(g, k, v0 / stress)

    p = c.par
    sol = _sec(d, "solver")
    for k in ("mass_scale", "damping", "tol"):
        if k in sol:
            setattr(p, k, float(sol[k]))
    for k in ("warmup_iter", "max_iter"):
        if k in sol:
            setattr(p, k, int(sol[k]))
    if "kinetic_damping" in sol:
        p.kinetic_damping = _flag(sol["kinetic_damping"])
    c.prestress_asbuilt = _flag(sol.get("prestress", "flat") == "asbuilt")
    con = _sec(d, "contact")
    for k in ("mu", "strength", "stick_strength", "torsion_factor", "epsilon", "slip_rate", "rot_factor", "dent_rate",
              "edge_threshold"):
        if k in con:
            setattr(p, k, float(con[k]))
    for k in ("ramp_iter", "al_every"):
        if k in con:
            setattr(p, k, int(con[k]))
    for k in ("rot_friction", "seat_gap", "augmented"):
        if k in con:
            setattr(p, k, _flag(con[k]))
    if "seat_law" in con:
        p.seat_law = _flag(con["seat_law"] != "off")
    if "dent_after_iter" in con:
        p.dent_after = int(con["dent_after_iter"])
    c.locked = _flag(con.get("locked", False))
    fl = _sec(d, "floor")
    fz = fl.get("z")
    fmu = fl.get("mu")
    fmu = float(fmu) if fmu is not None and float(fmu) >= 0 else FLOOR_MU
    p.floor_on = _flag(fl.get("enabled") and fz is not None and np.isfinite(float(fz)))
    p.floor_z = float(fz) if fz is not None else 0.0
    p.floor_mu = fmu

    wv = _sec(d, "weave")
    if "subdivisions" in wv:
        c.subdivisions = int(wv["subdivisions"])
    if "overhang" in wv:
        c.overhang = float(wv["overhang"])
    c.offset_half_height = _flag(wv.get("offset_mode", "half_height") == "half_height")
    if "seed_strand_phase" in wv:
        c.seed_strand_phase = int(wv["seed_strand_phase"])
    flips = [(int(f[0]), int(f[1])) for f in wv.get("flip_knots") or []]

    lo = _sec(d, "loads")
    gr = _sec(lo, "gravity")
    c.gravity = _flag(gr.get("enabled", True))
    c.gravity_dir[:] = [float(x) for x in gr.get("direction", [0, 0, -1])]
    udl = _sec(lo, "udl")
    c.udl = 0 if not udl.get("enabled") else (1 if udl.get("per", "length") == "length" else 2)
    c.udl_q[:] = [float(x) for x in udl.get("q", [0, 0, 0])]
    loads = [(int(pl["frame"][0]), int(pl["frame"][1]), [float(x) for x in pl["F"]], pl.get("space", "world") == "local")
             for pl in lo.get("point") or []]

    anchors = []
    for a in d.get("anchors") or []:
        sel = a.get("select") or {}
        t = sel.get("type", "frame")
        fix = [bool(f) for f in a.get("fix", [1, 1, 1, 0, 0, 0])]
        if t == "frame":
            anchors.append((_SELECTOR["frame"], int(sel["id"][0]), int(sel["id"][1]), fix))
        elif t == "knots":
            anchors.extend((_SELECTOR["frame"], int(i[0]), int(i[1]), fix) for i in sel.get("ids") or [])
        elif t == "strand":
            anchors.append((_SELECTOR["strand"], int(sel.get("strand", 0)), 0, fix))
        elif t == "region":     # regionFrames (model.js:127-145): seed frame id, radius in scene length units
            anchors.append((_SELECTOR["region"], int(sel["id"][0]), int(sel["id"][1]), fix, float(sel.get("radius") or 0)))
        elif t in _SELECTOR:
            anchors.append((_SELECTOR[t], 0, 0, fix))
        else:
            raise ValueError(f"unknown selector type {t}")
    mesh = _sec(d, "mesh")
    return sc, anchors, loads, flips, {"path": mesh.get("path", ""), "scale": float(mesh.get("scale", 1.0))}


def load_scene(path, mesh=None):
    """the host model of a scene.json file; the mesh path inside the file is relative to the file's own folder
    (scene.js:1-2), or `mesh` = (V, F) given directly (scaled already)"""
    with open(path) as f:
        d = json.load(f)
    return model_from_scene(d, base_dir=os.path.dirname(os.path.abspath(path)), mesh=mesh)


def model_from_scene(d, base_dir=".", mesh=None):
    sc, anchors, loads, flips, m = scene_from_dict(d)
    if mesh is None:
        if not m["path"]:
            raise ValueError("the scene names no mesh")
        V, F = read_mesh(os.path.join(base_dir, m["path"]), m["scale"])
    else:
        V, F = mesh
    return Model.from_mesh(V, F, scene=sc, anchors=anchors, loads=loads, flips=flips, weld=True, groups=sc.groups,
                           assign=sc.assign, joints=sc.joints, pieces=sc.pieces)


__all__ = ["parse_stl", "parse_obj", "read_mesh", "scene_from_dict", "load_scene", "model_from_scene"]
