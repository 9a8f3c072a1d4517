"""Seeded kagome gridshell workloads (mm·N), built with the host model builder.

Every scene is a kagome weave of a hexagonal patch of a triangular lattice (edge a = 2·spacing/√3, R =
round(span / (2·spacing)) rings, `subdivisions` = round((a/2) / segment)), raised onto a spherical cap whose
boundary contracts by `contraction`, with all lath ends pinned in XYZ and self-weight applied.  The section is
rectangular timber 40 mm wide (E = 10 GPa, G = 0.6 GPa, ρ = 500 kg/m³); the out-of-plane imperfection is a smooth
seeded field of amplitude `noise` (mm).  The reference builds the same weave from a mesh (web/js/topology.js,
web/js/mesh.js) with the material defaults of web/js/params.js.

    small_dome        1.2 m span, 0.3 m spacing, 50 mm segments (~230 nodes)
    dome_3m           3 m span, 0.3 m spacing, 50 mm segments, 15 %, ±1 mm, μ 0.3 (~1.4k nodes, 240 crossings)
    dome_10m          10 m span, 0.15 m spacing, 20 mm segments, 12 %, ±2 mm, μ 0.3 (78,606 nodes, 9,900 crossings)
    dome_30m          30 m span, 0.2 m spacing, 15 mm segments, 18 %, ±2 mm, μ 0.05 (~0.8M nodes, 51k crossings)
    design_variant(i) span 2–8 m, spacing 0.2–0.4 m, μ 0.1–0.5, contraction 8–20 %, thickness 3–6 mm,
                      node count log-uniform in [2k, 50k] (seed 3000 + i)
    request_stream    Poisson arrivals of design variants with node counts log-uniform in [1k, 200k]
"""
import math

import numpy as np

from .model import Model, Scene


def dome_scene(mu=0.3, thickness=4.0, tol=1e-6, **kw):
    """section 40 × thickness mm timber, E = 10 GPa, G = 0.6 GPa, ρ = 500 kg/m³"""
    return Scene(E=10000.0, G=600.0, density=500.0, width=40.0, height=thickness, mu=mu, tol=tol, **kw)


def dome(span=3000.0, spacing=300.0, segment=50.0, contraction=0.15, noise=1.0, seed=1, **kw):
    return Model.hex_dome(span, spacing, segment, contraction, noise, seed, dome_scene(**kw))


def small_dome(seed=1, **kw):
    return dome(1200.0, 300.0, 50.0, 0.15, 1.0, seed, **kw)


def dome_3m(**kw):
    return dome(3000.0, 300.0, 50.0, 0.15, 1.0, 1, **kw)


def dome_10m(**kw):
    return dome(10000.0, 150.0, 20.0, 0.12, 2.0, 2, **kw)


def dome_30m(**kw):
    return dome(30000.0, 200.0, 15.0, 0.18, 2.0, 5, mu=0.05, **kw)


def variant_params(seed, n_lo, n_hi):
    """span, spacing, μ, contraction, thickness and target node count of one seeded design variant"""
    r = np.random.default_rng(seed)
    span = r.uniform(2000, 8000)
    spacing = r.uniform(200, 400)
    mu = r.uniform(0.1, 0.5)
    con = r.uniform(0.08, 0.20)
    th = r.uniform(3.0, 6.0)
    n_t = math.exp(r.uniform(math.log(n_lo), math.log(n_hi)))
    return span, spacing, mu, con, th, n_t


def _variant(seed, n_lo, n_hi, **kw):
    span, spacing, mu, con, th, n_t = variant_params(seed, n_lo, n_hi)
    R = max(1, round(span / (2 * spacing)))
    K = 3 * R * (3 * R + 1)
    S = 6 * R
    sub = max(1, round((n_t - S) / max(2 * K - S, 1)))
    a = 2 * spacing / math.sqrt(3)
    return dome(span, spacing, (a / 2) / sub, con, 2.0, seed, mu=mu, thickness=th, **kw)


def design_variant(i, **kw):
    """the i-th of the batched design variants (node count log-uniform in [2k, 50k])"""
    return _variant(3000 + i, 2000, 50000, **kw)


def request_variant(i, **kw):
    """the i-th request of the serving stream (node count log-uniform in [1k, 200k])"""
    return _variant(40000 + i, 1000, 200000, **kw)


def request_stream(rate=60.0, duration=120.0, seed=4):
    """Poisson arrival times (s) of the serving stream"""
    r = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += r.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)
