"""Host model builder (libweave_model.so): kagome topology invariants (selfcross_check.mjs style), goal setup re-derived
independently in numpy, and determinism of the seeded scene generator."""
import math

import numpy as np
import pytest

import cases


@pytest.mark.parametrize("R_span", [(5, 3000.0), (4, 2400.0), (2, 1200.0)])
def test_kagome_counts(R_span):
    """one crossing per mesh edge of the hexagonal patch: K = 3R(3R+1), 6R laths, n = (2K − S)·s + S nodes"""
    R, span = R_span
    m = cases.dome(span, 300.0, 50.0, 0.15, 0.0, 1)
    s = m.info.subdivisions
    assert m.info.rings == R
    assert m.K == 3 * R * (3 * R + 1)
    assert m.info.S == 6 * R
    assert m.n == (2 * m.K - m.info.S) * s + m.info.S
    assert m.m == m.n - m.info.S and m.p == m.m - m.info.S
    assert m.info.parity_conflicts == 0


def _sides_by_strand(m):
    ks = m.arr("k_strand")
    u = m["u"]
    out = {}
    for q in range(2 * m.K):
        out.setdefault(int(ks[q]), []).append((u[q], q))
    return {s: sorted(v) for s, v in out.items()}


def test_over_under_alternates_and_offset_is_h():
    """each crossing has one over and one under pass, offset by the lath thickness h along the surface normal, and
    passes alternate over/under along every lath (weave parity)"""
    m = cases.dome(3000.0, 300.0, 50.0, 0.15, 0.0, 1)
    X = m["X"].reshape(-1, 3)
    noff = m.arr("node_offset")
    ks = m.arr("k_strand")
    u0 = np.round(m["u"]).astype(int)   # knot rows (the pairing moves u by less than a segment)
    node = lambda q: noff[ks[q]] + u0[q]
    h = m.arr("k_hk")[0]
    over = {}
    for c in range(m.K):
        a, b = X[node(2 * c)], X[node(2 * c + 1)]
        assert abs(np.linalg.norm(a - b) - h) < 0.05 * h
        over[2 * c], over[2 * c + 1] = a[2] > b[2], b[2] > a[2]
    for s, lst in _sides_by_strand(m).items():
        flags = [over[q] for _, q in lst]
        assert all(flags[i] != flags[i + 1] for i in range(len(flags) - 1)), s


def test_boundary_crossings_disabled_and_ends_anchored():
    m = cases.config1()
    en = m.arr("k_enabled")
    R = m.info.rings
    assert int((en == 0).sum()) == 6 * R            # boundary edges of the hexagon: 6R
    free = m.arr("node_free").reshape(-1, 3)
    noff = m.arr("node_offset")
    ends = set(noff[:-1]) | set(noff[1:] - 1)
    fixed = set(np.where(free.sum(1) == 0)[0])
    assert fixed == ends


def test_masses_and_inertias_rederived():
    """node masses ms·Σk and static rotational stiffness (solver_cpu.js:694-705, rod_cosserat.js:186-209) recomputed in
    numpy from the model's own arrays"""
    m = cases.config1()
    n, M, P, K = m.n, m.m, m.p, m.K
    par = m.md.par
    i0, i1, l0 = m.arr("seg_i0"), m.arr("seg_i1"), m.arr("l0")
    EA, GA = m.arr("EA"), m.arr("GA")
    kN = np.zeros(n)
    np.add.at(kN, i0, (EA + GA) / l0)
    np.add.at(kN, i1, (EA + GA) / l0)
    u = m["u"]
    soff, nseg = m.arr("k_seg_off"), m.arr("k_nseg")
    kc = par.strength + par.stick_strength
    for q in range(2 * K):
        i = int(min(max(math.floor(u[q]), 0), max(nseg[q] - 1, 0)))
        seg = soff[q] + i
        kN[i0[seg]] += kc
        kN[i1[seg]] += kc
    assert np.allclose(m.arr("mass"), par.mass_scale * np.maximum(kN, 1e-12), rtol=1e-13)
    j0, j1, jl = m.arr("j_seg0"), m.arr("j_seg1"), m.arr("j_len")
    kr = np.zeros((M, 3))
    B = np.stack([m.arr("EIx")[j0], m.arr("EIy")[j0], m.arr("GJ")[j0]], 1) / jl[:, None]
    np.add.at(kr, j0, B)
    np.add.at(kr, j1, B)
    X = m["X"].reshape(-1, 3)
    L = np.linalg.norm(X[i1] - X[i0], axis=1)
    kr[:, 0] += GA * L * L / l0
    kr[:, 1] += GA * L * L / l0
    assert np.allclose(m.arr("krot_static").reshape(-1, 3), kr, rtol=1e-12)


def test_self_weight_total():
    """gravity loads sum to ρ g A Σ l0 (units.js:21-25, model.js:200-212)"""
    m = cases.config1()
    q = 500.0 * (40.0 * 4.0 * 1e-6) * 9.80665 * 1e-3      # N per mm
    W = m.arr("load_const").reshape(-1, 3).sum(0)
    assert abs(W[2] + q * m.arr("l0").sum()) < 1e-9 * q * m.arr("l0").sum()
    assert abs(W[0]) < 1e-12 and abs(W[1]) < 1e-12


def test_scene_generator_is_deterministic():
    a, b = cases.config1(), cases.config1()
    for k in ("X", "Q", "u", "n_prev"):
        assert np.array_equal(a[k], b[k])
    c = cases.dome(3000.0, 300.0, 50.0, 0.15, 1.0, 2)
    assert not np.array_equal(a["X"], c["X"])


def test_contact_pairing_reduces_offsets():
    """the Gauss-Newton pairing (PolylineContact.initialize) leaves the two contact points of each crossing at the
    closest points of the centrelines: the tangential offset is small against the segment length"""
    from oracle.bind import Oracle
    m = cases.config1()
    o = Oracle(m)
    o.step(1)
    s = o.state()
    assert np.all(np.abs(s["gap"]) < 10.0)
