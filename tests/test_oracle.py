"""Pins of the f64 oracle (oracle/oracle.cpp) against what fixes it independently of its own formulas:
numpy's summation order, finite differences of the rod energy, the reference's published closed-form errors,
Coulomb-cone invariants of its crossing test cases, equilibrium and determinism."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle, pairwise_sum


@pytest.mark.parametrize("n", [0, 1, 5, 7, 8, 9, 15, 16, 17, 100, 127, 128, 129, 255, 256, 257, 1000, 4099, 12345])
def test_pairwise_sum_is_numpy_order(n):
    """solver_cpu.js:42-54 restates numpy's pairwise summation; np.sum must agree bit for bit"""
    a = np.random.default_rng(n).standard_normal(n) * 10.0 ** np.random.default_rng(n + 1).integers(-8, 8, n)
    assert pairwise_sum(a) == np.sum(a)


def _qexp(v):
    a = np.linalg.norm(v)
    s = np.sin(a / 2) / a if a > 1e-12 else 0.5
    return np.array([np.cos(a / 2), v[0] * s, v[1] * s, v[2] * s])


def _qmul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


def _rod_fixture():
    """rod_check.mjs: a bent, twisted strand with rest curvature (activation 0.4), clamps at both ends"""
    W = cases.W
    A = cases.A
    n = 9
    P = [[1000.0 * i / 8, 40 * np.sin(i), 25 * np.cos(1.3 * i)] for i in range(n)]
    sc = W.Scene(gravity=0, contacts=0, activation=0.4)
    m = W.Model.from_polylines([P], scene=sc, anchors=[(A.WSM_FRAME, 0, 0, [0, 0, 0, 1, 1, 1]),
                                                       (A.WSM_FRAME, 0, 1, [0, 0, 0, 1, 0, 1])])
    return m


def test_rod_forces_are_energy_gradient():
    """CosseratRod.apply forces and torques against central differences of the energy (rod_check.mjs, < 1e-5)"""
    m = _rod_fixture()
    assert m.md.C == 2 and np.abs(m.arr("omega0", 3 * m.p)).max() > 0
    o = Oracle(m)
    rng = np.random.default_rng(11)
    X = m["X"].copy() + 5 * (2 * rng.random(3 * m.n) - 1)
    Q = m["Q"].copy().reshape(-1, 4)
    for j in range(m.m):
        q = _qmul(_qexp(rng.uniform(-1, 1, 3) * [0.2, 0.2, 0.5]), Q[j])
        Q[j] = q / np.linalg.norm(q)
    Q = Q.reshape(-1)
    e0, F, T = o.rod_apply(X, Q)
    assert e0 > 0
    h, ef = 1e-4, 0.0
    for i in range(3 * m.n):
        Xp = X.copy(); Xp[i] += h
        Xm = X.copy(); Xm[i] -= h
        fd = -(o.rod_apply(Xp, Q)[0] - o.rod_apply(Xm, Q)[0]) / (2 * h)
        ef = max(ef, abs(fd - F[i]))
    hr, et = 1e-6, 0.0
    for j in range(m.m):
        for k in range(3):
            v = np.zeros(3); v[k] = hr
            qs = []
            for sgn in (1, -1):
                Qd = Q.copy().reshape(-1, 4)
                q = _qmul(_qexp(sgn * v), Qd[j])
                Qd[j] = q / np.linalg.norm(q)
                qs.append(o.rod_apply(X, Qd.reshape(-1))[0])
            fd = -(qs[0] - qs[1]) / (2 * hr)
            et = max(et, abs(fd - T[3 * j + k]))
    assert ef / np.abs(F).max() < 1e-5
    assert et / np.abs(T).max() < 1e-5


def test_rod_energy_rigid_motion_invariant():
    """without clamps the rod energy is unchanged by a rigid rotation and translation of the whole strand"""
    W = cases.W
    P = [[1000.0 * i / 8, 40 * np.sin(i), 25 * np.cos(1.3 * i)] for i in range(9)]
    m = W.Model.from_polylines([P], scene=W.Scene(gravity=0, contacts=0, activation=0.3))
    o = Oracle(m)
    rng = np.random.default_rng(5)
    X = m["X"].copy() + 3 * rng.standard_normal(3 * m.n)
    Q = m["Q"].copy().reshape(-1, 4)
    for j in range(m.m):
        q = _qmul(_qexp(rng.uniform(-0.3, 0.3, 3)), Q[j]); Q[j] = q / np.linalg.norm(q)
    e0 = o.rod_apply(X, Q.reshape(-1))[0]
    r = _qexp(np.array([0.3, -1.1, 0.7]))
    w, x, y, z = r
    Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                   [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                   [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    X2 = (X.reshape(-1, 3) @ Rm.T + [100.0, -50.0, 7.0]).reshape(-1)
    Q2 = np.array([_qmul(r, q) for q in Q]).reshape(-1)
    e1 = o.rod_apply(X2, Q2)[0]
    assert abs(e1 - e0) <= 1e-9 * abs(e0)


# Timoshenko cantilever δ = FL³/3EI + FL/κGA: the reference's own published errors (web/README.md:71-72)
PUBLISHED = {10: 0.50, 20: 0.12, 40: 0.03}


@pytest.mark.parametrize("nseg", [10, 20, 40])
@pytest.mark.parametrize("edge", [False, True])
def test_cantilever_matches_published_timoshenko_error(nseg, edge):
    E, G, kap = 1e4, 625.0, 5 / 6
    I = 30 * 60 ** 3 / 12 if edge else 60 * 30 ** 3 / 12
    delta_t = 10 * 1e9 / (3 * E * I) + 10 * 1000 / (kap * G * 1800)
    m = cases.cantilever(nseg=nseg, edge=edge)
    o = Oracle(m)
    o.run(600000, 5000)
    X = o.state()["X"]
    tip = int(np.argmax(X[0::3]))
    d = -X[3 * tip + (1 if edge else 2)]
    err = 100 * (d / delta_t - 1)
    assert abs(err - PUBLISHED[nseg]) < 0.015, err
    assert o.residual < 1e-9


def _run_cross(name):
    m = cases.crossing(**cases.CROSS_CASES[name])
    o = Oracle(m)
    o.run(60000 if name != "cross_slip" else 20000, 2000)
    return m, o, o.state()


def test_crossing_press_stick_locked_equilibria():
    """cases.mjs:44-50 — stick: the friction force carries the pull and u stays; locked: no u motion"""
    for name in ("cross_press", "cross_stick", "cross_locked"):
        m, o, s = _run_cross(name)
        assert o.converged, name
        assert s["kstate"][0] == 0 and s["stuck"][0] == 1, name
        pull = cases.CROSS_CASES[name]["pull"]
        assert abs(s["Ft"][0] - pull) < 0.05, (name, s["Ft"][0])
        if name != "cross_locked":
            assert s["Ft"][0] <= 0.35 * s["Fn"][0] + 1e-9
        assert abs(s["u"][1] - 5.0) < 1e-6
        # equilibrium: the reactions balance the applied load
        R = s["reactions"].reshape(-1, 3).sum(0)
        assert np.allclose(R + [pull, 0, -50.0], 0, atol=1e-2 * 50), (name, R)


def test_crossing_slip_sits_on_the_cone():
    """cross_slip: the pull exceeds μ·Fn, the crossing slides and |Ft| = μ·Fn exactly (slip_rate 1)"""
    m, o, s = _run_cross("cross_slip")
    assert s["kstate"][0] == 1 and s["stuck"][0] == 0
    assert abs(s["Ft"][0] - 0.35 * s["Fn"][0]) < 1e-9 * s["Fn"][0]
    assert s["slip"][0] > 0 and s["u"][0] != 5.0


def test_oracle_bit_identical_reruns():
    a = Oracle(cases.config_small())
    b = Oracle(cases.config_small())
    a.step(700)
    b.step(700)
    sa, sb = a.state(), b.state()
    for k in ("X", "Q", "u", "slip", "V", "W"):
        assert np.array_equal(sa[k], sb[k]), k


def test_kinetic_damping_resets_at_energy_peaks():
    """Barnes kinetic damping: between resets the recorded kinetic energy never decreases; a decrease is recorded
    as 0 and restarts from rest (solver_cpu.js:809-817)"""
    o = Oracle(cases.config_small(), history=True)
    o.step(3000)
    ke, _, _ = o.history()
    resets = 0
    for i in range(1, len(ke)):
        if ke[i] == 0.0:
            resets += 1
            continue
        assert ke[i] >= ke[i - 1], i
    assert resets > 0


def test_warmup_and_residual_definition():
    """the residual is infinite through warm-up and finite afterwards (solver_cpu.js:835)"""
    m = cases.config_small()
    o = Oracle(m)
    o.step(m.md.par.warmup_iter)
    assert o.residual == np.inf
    o.step(1)
    assert np.isfinite(o.residual)
