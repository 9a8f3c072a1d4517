"""Scene fixtures of the reference's test suite (web/tests/cases.mjs:10-66, rod_bench.mjs, rod_check.mjs) and the
benchmark domes, built with the host model builder."""
import importlib
import math

import numpy as np

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


def dome_scene(mu=0.3, thickness=4.0, tol=1e-6, **kw):
    """section 40 × 4 mm timber, E = 10 GPa, G = 0.6 GPa, ρ = 500 kg/m³ (the benchmark workloads), mm·N"""
    return W.Scene(E=10000.0, G=600.0, density=500.0, width=40.0, height=thickness, mu=mu, tol=tol, **kw)


def dome(span=3000.0, spacing=300.0, segment=50.0, contraction=0.15, noise=1.0, seed=1, **kw):
    return W.Model.hex_dome(span, spacing, segment, contraction, noise, seed, dome_scene(**kw))


def config_small(seed=1, **kw):
    """a 1.2 m dome (~230 nodes): converging and chaos windows in seconds on the CPU"""
    return dome(1200.0, 300.0, 50.0, 0.15, 1.0, seed, **kw)


def config1(**kw):
    return dome(3000.0, 300.0, 50.0, 0.15, 1.0, 1, **kw)


def config2(**kw):
    return dome(10000.0, 150.0, 20.0, 0.12, 2.0, 2, **kw)


def config5(**kw):
    return dome(30000.0, 200.0, 15.0, 0.18, 2.0, 5, mu=0.05, **kw)


def config3_variant(i, n_target=None, rng_seed=None):
    """one of the 256 batched design variants: span 2–8 m, spacing 0.2–0.4 m, μ 0.1–0.5, contraction 8–20 %,
    thickness 3–6 mm, node count log-uniform 2k–50k (segment length chosen to approach it)"""
    r = np.random.default_rng(3000 + i if rng_seed is None else rng_seed)
    span = r.uniform(2000, 8000)
    spacing = r.uniform(200, 400)
    mu = r.uniform(0.1, 0.5)
    con = r.uniform(0.08, 0.20)
    th = r.uniform(3.0, 6.0)
    n_t = n_target or math.exp(r.uniform(math.log(2000), math.log(50000)))
    R = max(1, round(span / (2 * spacing)))
    K = 3 * R * (3 * R + 1)
    S = 6 * R
    sub = max(1, round((n_t - S) / max(2 * K - S, 1)))
    a = 2 * spacing / math.sqrt(3)
    seg = (a / 2) / sub
    return dome(span, spacing, seg, con, 2.0, 3000 + i, mu=mu, thickness=th)
