"""The oracle's UTM load-test replays (calibrate.js replayForce / replayDisplacement, :154-215) pinned by closed forms
and by each other: force control on a clamped cantilever gives the Timoshenko tip deflection within the reference's
own 10-segment error (0.50 %, web/README.md:71), linearly and with no hysteresis; displacement control at small travel
gives the reaction the force-control compliance predicts; on a single crossing a pull below μ·N sticks and one above
slides, leaving travel after unloading."""
import numpy as np

import cases
from oracle.bind import Oracle

E, G, W_, H_, L = 10000.0, 600.0, 60.0, 30.0, 1000.0
I_ = W_ * H_ ** 3 / 12
COMPLIANCE = L ** 3 / (3 * E * I_) + L / (5.0 / 6.0 * G * W_ * H_)   # Timoshenko tip compliance


def test_force_replay_cantilever():
    o = Oracle(cases.cantilever(F=0.0, tol=1e-9))
    rows = o.replay("force", 10, [0, 0, -1], [5.0, 10.0, 5.0, 0.0], 20000, 50)
    t = rows[:, 0]
    assert np.array_equal(rows[:, 2], [5.0, 10.0, 5.0, 0.0])
    assert (np.diff(rows[:, 3]) > 0).all()
    assert abs(t[0] / (5.0 * COMPLIANCE) - 1) < 0.006           # the 10-segment rod's 0.5 % (web/README.md:71)
    assert abs(t[1] - 2 * t[0]) < 1e-4 * t[1]                    # linear
    assert abs(t[2] - t[0]) < 1e-6 * t[0] and abs(t[3]) < 1e-6   # elastic: no hysteresis
    assert (rows[:, 1] == 0).all()                               # a free probe has no reaction


def test_displacement_replay_matches_force_compliance():
    o = Oracle(cases.cantilever(F=0.0, tol=1e-9))
    c = o.replay("force", 10, [0, 0, -1], [5.0], 20000, 50)[0, 0] / 5.0
    p = Oracle(cases.cantilever(F=0.0, tol=1e-9))
    rows = p.replay("displacement", 10, [0, 0, -1], [0.01, 0.02], 20000, 50)
    assert np.allclose(rows[:, 0], [0.01, 0.02], rtol=1e-12)      # the pinned probe sits at P0 + d·axis
    assert np.allclose(rows[:, 1], rows[:, 0] / c, rtol=1e-3)     # reaction = d / compliance
    assert abs(rows[1, 1] - 2 * rows[0, 1]) < 1e-4 * rows[1, 1]


def test_crossing_replay_sticks_then_slides():
    m = cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_press"])     # 50 N across the crossing, μ = 0.35
    o = Oracle(m)
    rows = o.replay("force", 16, [1, 0, 0], [5.0, 10.0, 15.0, 10.0, 0.0, 20.0], 6000, 100)
    t = rows[:, 0]
    assert t[0] < 0.05 and t[2] > t[1] > t[0] > 0                # sticks and loads below μ·N = 17.5
    assert t[4] > 1.0                                             # travel left after unloading: it slipped
    assert t[5] > t[4] + 1.0 and o.state()["kstate"][0] == 1      # above μ·N: sliding
