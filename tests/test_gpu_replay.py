"""UTM load-test replays on the device (the protocol branch of the reference's control kernel, wgsl.js:1048-1093)
against the oracle's replays (calibrate.js replayForce / replayDisplacement): the recorded travel, reaction, value
and iteration per schedule value; the control record's phase and increment; replays batched with each other and with
a plain relaxation equal each run alone."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

CANT_F = ("force", 10, [0.0, 0.0, -1.0], [5.0, 10.0, 5.0, 0.0], 20000, 50)
CANT_D = ("displacement", 10, [0.0, 0.0, -1.0], [0.5, 1.0, 0.25], 20000, 50)
CROSS_F = ("force", 16, [1.0, 0.0, 0.0], [5.0, 10.0, 15.0, 10.0, 0.0, 20.0], 6000, 100)


def cant():
    return cases.cantilever(F=0.0, tol=1e-9)


def cross():
    return cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_press"])


def run_device(W, models, precision):
    g = W.GpuSolver(models, precision=precision)
    for _ in range(200):
        if all(c["done"] for c in g.step(2000)):
            break
    ctl = g.read_control()
    outs = [g.read_outputs(b) for b in range(len(models))]
    g.close()
    return outs, ctl


@pytest.mark.parametrize("name,build,proto", [("cantilever_force", cant, CANT_F), ("cantilever_displacement", cant, CANT_D),
                                              ("crossing_force", cross, CROSS_F)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_replay_matches_oracle(W, gpu_ctx, name, build, proto, precision):
    mode, probe, axis, sched, settle, avg = proto
    ref = Oracle(build()).replay(mode, probe, axis, sched, settle, avg)
    m = build().set_protocol(mode, probe, axis, sched, settle, avg)
    (out,), (c,) = run_device(W, [m], precision)
    assert c["done"] == 1 and c["phase"] == 2 and c["inc"] == len(sched)
    assert out.shape == ref.shape
    assert np.array_equal(out[:, 2], ref[:, 2])
    span = 1000.0
    tol = 1e-9 if precision == "fp64" else 1e-4
    assert np.abs(out[:, 0] - ref[:, 0]).max() <= tol * span, (out[:, 0], ref[:, 0])
    scale = max(np.abs(ref[:, 1]).max(), 1.0)
    assert np.abs(out[:, 1] - ref[:, 1]).max() <= (1e-6 if precision == "fp64" else 1e-2) * scale, (out[:, 1], ref[:, 1])
    # every value: at most `settle` settling iterations, then exactly `avg` averaged ones.  Where the settle limit ends
    # them (the crossing) the iterations are the oracle's; where the residual does (the cantilever), the iteration the
    # residual first dips below 1e-9 while it hovers there is not a parity quantity (the averaged read-outs are)
    steps = np.diff(np.concatenate([[0.0], out[:, 3]]))
    assert ((steps > avg) & (steps <= settle + avg)).all(), steps
    if (np.diff(np.concatenate([[0.0], ref[:, 3]])) == settle + avg).all():
        assert np.array_equal(out[:, 3], ref[:, 3])


def test_replays_batched_equal_alone(W, gpu_ctx):
    """three replays of different modes and a plain dome relaxation in one batch: every output and state bitwise as
    each run alone"""
    specs = [(cant, CANT_F), (cant, CANT_D), (cross, CROSS_F)]
    alone = []
    for build, proto in specs:
        (out,), _ = run_device(W, [build().set_protocol(*proto)], "fp32")
        alone.append(out)
    dome = cases.config_small()
    models = [build().set_protocol(*proto) for build, proto in specs] + [dome]
    g = W.GpuSolver(models, precision="fp32")
    for _ in range(200):
        c = g.step(2000)
        if all(x["done"] for x in c[:3]):
            break
    for b in range(3):
        assert np.array_equal(g.read_outputs(b), alone[b]), b
    assert g.read_control()[3]["phase"] == 0 and len(g.read_outputs(3)) == 0
    g.close()
