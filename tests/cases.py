"""Scene fixtures of the reference's test suite (web/tests/cases.mjs:10-66, rod_bench.mjs, rod_check.mjs) and the
benchmark domes, built with the host model builder."""
import importlib
import math

W = importlib.import_module("weave-sim_amd")
A = W._abi
ALL6 = [1, 1, 1, 1, 1, 1]


def straight(L=1000.0, nseg=10):
    return [[L * i / nseg, 0.0, 0.0] for i in range(nseg + 1)]


def cantilever(nseg=10, L=1000.0, F=10.0, edge=False, w=60.0, h=30.0, tol=1e-10, GAfac=1.0):
    """clamped cantilever, tip load F (cases.mjs:32-36; rod_bench.mjs:27-37)"""
    kw = dict(gravity=0, contacts=0, tol=tol, width=w, height=h)
    if GAfac != 1.0:
        kw.update(beam_advanced=1, k_shear=GAfac)
    sc = W.Scene(**kw)
    m = W.Model.from_polylines([straight(L, nseg)], scene=sc, anchors=[(A.WSM_FRAME, 0, 0, ALL6)])
    m.add_load(nseg, [0.0, -F, 0.0] if edge else [0.0, 0.0, -F])
    return m


def torsion(nseg=10, L=1000.0, tol=1e-10):
    """a straight lath whose section normals turn by 90° along it, clamped at its start and held in place and about
    its axis at its end (cases.mjs:38-42)"""
    normals = []
    for i in range(nseg + 1):
        a = math.pi / 2 * i / nseg
        normals.append([0.0, -math.sin(a), math.cos(a)])
    sc = W.Scene(gravity=0, contacts=0, tol=tol)
    return W.Model.from_polylines([straight(L, nseg)], scene=sc, normals=[normals],
                                  anchors=[(A.WSM_FRAME, 0, 0, ALL6), (A.WSM_FRAME, 0, 1, [1, 1, 1, 0, 0, 1])])


def crossing(load=-50.0, pull=0.0, mu=0.35, free_b=True, nseg=10, L=1000.0, h=30.0, locked=False, **scene):
    """two orthogonal laths crossing at their middles, B free in x (cases.mjs:14-28)"""
    P, PB = [], []
    for i in range(nseg + 1):
        x = -L / 2 + L * i / nseg
        P.append([x, 0.0, 0.0])
        PB.append([0.0, x, h])
    mid = nseg // 2
    sc = W.Scene(gravity=0, mu=mu, locked=1 if locked else 0, **scene)
    anchors = []
    for s in (0, 1):
        fix = [0, 1, 1, 0, 0, 1] if (s == 1 and free_b) else [1, 1, 1, 0, 0, 1]
        anchors.append((A.WSM_FRAME, s, 0, fix))
        anchors.append((A.WSM_FRAME, s, 2, fix))
    return W.Model.from_polylines([P, PB], scene=sc, knot_rows=[[0, mid, nseg], [0, mid, nseg]],
                                  knots=[(0, 1, 1, 1)], anchors=anchors, loads=[(1, 1, [pull, 0.0, load])])


CROSS_CASES = {"cross_press": dict(pull=0.0), "cross_stick": dict(pull=5.0), "cross_slip": dict(pull=12.0),
               "cross_locked": dict(pull=12.0, locked=True)}


# the benchmark domes live in the package (bench.py builds the same workloads)
_S = importlib.import_module("weave-sim_amd.scenes")
dome_scene = _S.dome_scene
dome = _S.dome
config_small = _S.small_dome
config1 = _S.dome_3m
config2 = _S.dome_10m
config5 = _S.dome_30m
config3_variant = _S.design_variant


def ground_lath(push=0.0, mu=0.4, anchor=False, tilt=0.05, lift=2.0, push_dir=(1.0, 0.0, 0.0), **scene):
    """one straight lath above the ground, cm·N (floor_check.mjs:12-25): 100 cm, 10 segments, 6 × 3 section,
    E = 1e6, G = 6e4, self-weight, optional sideways push (a UDL along x), optionally pinned at its start"""
    L, nseg = 100.0, 10
    P = [[L * i / nseg - L / 2, 0.0, lift + tilt * (L * i / nseg)] for i in range(nseg + 1)]
    kw = dict(width=6.0, height=3.0, E=1e6, G=6e4, strength=3000.0, stick_strength=1000.0, ramp_iter=0, gravity=1,
              tol=1e-6, contacts=0, floor_on=1, floor_z=0.0, floor_mu=mu, length_to_m=0.01)
    if push:
        kw.update(udl=1, udl_q=[push / L * c for c in push_dir])
    kw.update(scene)
    sc = W.Scene(**kw)
    anchors = [(A.WSM_ALL_START, 0, 0, [1, 1, 1, 0, 0, 0])] if anchor else []
    return W.Model.from_polylines([P], scene=sc, anchors=anchors)
