"""Two streams of launches (ws_batch_opts.streams): the instances split into two groups of about equal work, each
replaying its own iteration graphs on its own stream.  No instance reads another's memory and every reduction stays
inside its instance, so each instance's state must be bitwise the one-stream state — through whole graphs, remainder
pieces, per-instance stops, a structure only the ground holds (the rigid step) and a load-test replay."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _models():
    m = [cases.config_small(seed=s) for s in (1, 2, 3)] + [cases.config1(), cases.ground_lath(push=6.0)]
    r = cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_press"])
    r.set_protocol("force", 16, [1.0, 0.0, 0.0], [4.0, 8.0, 16.0], 300, 50)
    return m + [r]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_two_streams_give_the_one_stream_states(W, gpu_ctx, precision):
    out = {}
    for ns in (1, 2):
        models = _models()
        g = W.GpuSolver(models, precision=precision, tile=64, per_submit=32, streams=ns)
        assert g.opts["streams"] == ns
        g.step(300)
        g.step(45)                                   # whole graphs, then the remainder's pieces
        g.step(20, instances=[0, 3, 5])              # a subset
        ctl = g.read_control()
        out[ns] = ([g.state(b) for b in range(len(models))], ctl, g.read_outputs(5))
        g.close()
    (a, ca, ra), (b, cb, rb) = out[1], out[2]
    for q in range(len(a)):
        assert ca[q]["iter"] == cb[q]["iter"] and ca[q]["ke"] == cb[q]["ke"], q
        for k in ("X", "Q", "u", "stuck", "floor_a", "V", "W"):
            assert np.array_equal(a[q][k], b[q][k]), (q, k)
    assert np.array_equal(ra, rb) and len(ra) > 0
    assert all(c["iter"] > 300 for c in ca)


def test_streams_chosen_by_size(W, gpu_ctx):
    """two streams from two instances and 8,192 strand tiles, one below; a single instance is one group"""
    big = cases.config2()                            # 1,350 strand tiles of 64 lanes
    g = W.GpuSolver([big.clone() for _ in range(7)], precision="fp32", tile=64)
    assert g.opts["streams"] == 2
    g.close()
    g = W.GpuSolver([big.clone() for _ in range(5)], precision="fp32", tile=64)
    assert g.opts["streams"] == 1
    g.close()
    g = W.GpuSolver([big], precision="fp32", tile=64, streams=2)   # one instance: one group
    assert g.opts["streams"] == 1
    g.close()
    with pytest.raises(W.DeviceError):
        W.GpuSolver([big], streams=3)
