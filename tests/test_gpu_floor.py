"""The ground on the device (anchored structures): the pinned lath of the reference's floor checks
(floor_check.mjs:55-59) against the oracle — trajectory, stick points, reactions and the converged load share — and a
pushed, pinned lath whose free end slides.  A structure on the ground with no anchor (rigid-body freedoms) is refused."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_pinned_lath_trajectory(W, gpu_ctx, precision, tol):
    m = cases.ground_lath(anchor=True, tilt=0.0, lift=1.5)
    o = Oracle(m)
    g = W.GpuSolver([m], precision=precision)
    for n in (50, 400, 3000):
        o.step(n - o.iteration)
        g.step(n - g.read_control()[0]["iter"])
        a, b = o.state(), g.state(0)
        assert np.abs(a["X"] - b["X"]).max() <= tol * 100.0, (n, np.abs(a["X"] - b["X"]).max())
        assert np.abs(a["floor_a"] - b["floor_a"]).max() <= tol * 100.0
    g.close()


def test_pinned_lath_converged_load_share(W, gpu_ctx):
    """relaxed to the 1e-6 residual: anchor and ground carry the weight together, as on the oracle"""
    m = cases.ground_lath(anchor=True, tilt=0.0, lift=1.5)
    Wt = m.arr("load_const").reshape(-1, 3).sum(0)
    o = Oracle(m)
    o.run(40000)
    g = W.GpuSolver([m], precision="fp64")
    for _ in range(20):
        if g.step(5000)[0]["done"]:
            break
    b = g.state(0)
    g.close()
    assert b.st.converged == 1 and o.converged
    a = o.state()
    assert np.abs(a["X"] - b["X"]).max() <= 1e-4 * 100.0
    R = b["reactions"].reshape(-1, 3)
    F, st, _ = o.floor()
    assert abs(F[:, 2].sum() + R[:, 2].sum() + Wt[2]) < 1e-3 * abs(Wt[2])
    assert abs(b["reactions"].reshape(-1, 3)[:, 2].sum() - a["reactions"].reshape(-1, 3)[:, 2].sum()) < 1e-4 * abs(Wt[2])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_pushed_pinned_lath_slides(W, gpu_ctx, precision, tol):
    """a pinned lath lying on the ground, pushed sideways above μ·W: it swings about the pin with every free node
    sliding (state 2 on the oracle); positions and stick points as on the oracle"""
    m = cases.ground_lath(anchor=True, tilt=0.0, lift=1.5, push=12.0, push_dir=(0.0, 1.0, 0.0))
    o = Oracle(m)
    g = W.GpuSolver([m], precision=precision)
    o.step(2000)
    g.step(2000)
    a, b = o.state(), g.state(0)
    assert (o.floor()[1][1:] == 2).all() and b["X"][-2] > 0.1
    assert np.abs(a["X"] - b["X"]).max() <= tol * 100.0
    assert np.abs(a["floor_a"] - b["floor_a"]).max() <= tol * 100.0
    g.close()


def test_ground_without_anchor_is_refused(W, gpu_ctx):
    with pytest.raises(W.DeviceError):
        W.GpuSolver([cases.ground_lath()], precision="fp64")
