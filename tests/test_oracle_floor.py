"""The oracle's ground (Floor, solver_cpu.js:535-634) pinned by the reference's own checks (floor_check.mjs): a lath
falls and rests (Σ ground force = −Σ loads), a push below μW sticks, a push above μW slides with every node on the
friction cone, μ = 0 carries no friction, and with an anchor the rigid-body freedoms are off."""
import numpy as np

import cases
from oracle.bind import Oracle


def weight(m):
    return m.arr("load_const").reshape(-1, 3).sum(0)


def test_rest():
    m = cases.ground_lath()
    W = weight(m)
    o = Oracle(m)
    o.run(20000)
    F, st, _ = o.floor()
    X = o.state()["X"].reshape(-1, 3)
    assert o.converged
    assert abs(F[:, 2].sum() + W[2]) < 1e-4 * abs(W[2])
    assert (st > 0).all() and abs(X[0, 2] - X[-1, 2]) < 1e-6


def test_stick():
    m = cases.ground_lath(push=3.0)
    W = weight(m)
    o = Oracle(m)
    o.run(20000)
    F, st, _ = o.floor()
    assert o.converged and (st != 2).all()
    assert abs(F[:, 0].sum() + W[0]) < 1e-4 * abs(W[0])
    on = st > 0
    assert (np.hypot(F[on, 0], F[on, 1]) / F[on, 2]).max() <= 0.4 + 1e-12


def test_slide():
    m = cases.ground_lath(push=6.0)
    o = Oracle(m)
    o.step(3000)
    F, st, _ = o.floor()
    X = o.state()["X"].reshape(-1, 3)
    assert (st == 2).all()
    r = np.hypot(F[:, 0], F[:, 1]) / F[:, 2]
    assert np.abs(r - 0.4).max() < 1e-9
    assert X[0, 0] + 50 > 1
    assert abs(F[:, 0].sum() + 0.4 * F[:, 2].sum()) < 1e-6 * F[:, 2].sum()


def test_frictionless():
    m = cases.ground_lath(push=3.0, mu=0.0)
    o = Oracle(m)
    o.step(2000)
    F, st, _ = o.floor()
    assert abs(F[:, 0].sum()) < 1e-12 and o.state()["X"][0] + 50 > 1


def test_pinned():
    m = cases.ground_lath(anchor=True, tilt=0.0, lift=1.5)
    W = weight(m)
    o = Oracle(m)
    o.run(40000)
    F, st, rigid = o.floor()
    R = o.state()["reactions"].reshape(-1, 3)
    assert o.converged and rigid == 0.0
    assert abs(F[:, 2].sum() + R[:, 2].sum() + W[2]) < 1e-3 * abs(W[2])


def test_resume_carries_the_rigid_freedoms():
    """a state exported while the lath slides as a whole and imported into a new oracle continues the same trajectory
    bit for bit: the rigid velocities and the model radius travel with the state (Floor.rv / rw / Rb; the reference's
    upload writes rv and rw, solver_gpu.js:236-237), and the reference point is the centroid of the positions
    (Floor.apply, solver_cpu.js:590).  Imported without the velocities, the same state departs from that trajectory."""
    m = cases.ground_lath(push=6.0)
    a = Oracle(m)
    a.step(700)
    b = Oracle(m)
    b.step(300)
    s = b.state()
    X = s["X"].reshape(-1, 3)
    rv = np.array(s.st.rb_v[:])
    assert abs(rv[0]) > 1e-3                                   # sliding along the push: the rigid translation is live
    assert np.abs(np.array(s.st.rb_c[:]) - X.mean(0)).max() < 1e-12 * np.abs(X).max()
    X0 = m["X"].reshape(-1, 3)
    R = np.sqrt(((X0 - X0.mean(0)) ** 2).sum(1)).max()
    assert abs(s.st.rb_radius - R) < 1e-12 * R                 # Rb of the Floor's construction
    c = Oracle(m, state=s)
    c.step(400)
    assert c.iteration == a.iteration
    assert np.array_equal(c.state()["X"], a.state()["X"])
    for k in range(3):
        s.st.rb_v[k] = 0.0
        s.st.rb_w[k] = 0.0
    d = Oracle(m, state=s)
    d.step(400)
    assert np.abs(d.state()["X"] - a.state()["X"]).max() > 1e-3
