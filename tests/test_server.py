"""Continuous-batching server (ws_server_*): results identical alone and in the mix, HBM-budgeted admission,
cancellation, per-request iteration limits."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _alone(W, model, iters):
    g = W.GpuSolver([model], precision="fp32", tile=64)
    g.step(iters)
    s = g.state(0)
    g.close()
    return s


def test_results_identical_alone_and_in_the_mix(W, gpu_ctx):
    """requests admitted while others relax finish bitwise equal to the same instance relaxed alone"""
    seeds = [11, 12, 13, 14, 15]
    iters = [700, 450, 700, 300, 520]
    ref = [_alone(W, cases.config_small(seed=s), n) for s, n in zip(seeds, iters)]
    srv = W.Server(precision="fp32", chunk=64, tile=64, max_slots=3)
    models = [cases.config_small(seed=s) for s in seeds]
    ids = [srv.submit(models[0], max_iter=iters[0]), srv.submit(models[1], max_iter=iters[1])]
    srv.pump()
    srv.pump()
    ids += [srv.submit(m, max_iter=n) for m, n in zip(models[2:], iters[2:])]
    res = {r["id"]: r for r in srv.drain()}
    assert sorted(res) == sorted(ids)
    for k, rid in enumerate(ids):
        r = res[rid]
        assert r["status"] == "max_iter" and r["iterations"] == iters[k]
        assert r["t_submit_ms"] <= r["t_admit_ms"] <= r["t_done_ms"]
        m = models[k]
        for name in ("X", "Q", "u", "slip", "stuck", "Fn"):
            assert np.array_equal(m[name], ref[k][name]), (k, name)
    srv.close()


def test_hbm_budget_limits_residency(W, gpu_ctx):
    """with a budget for about two instances, at most two are resident; all finish"""
    # price one instance through a server with an ample budget (admitted, then still resident after one chunk)
    one = W.Server(precision="fp32", max_slots=8)
    one.submit(cases.config_small(seed=20), max_iter=100)
    one.pump()
    per = one.stats()["bytes_in_use"]
    one.drain()
    one.close()
    assert per > 0
    models = [cases.config_small(seed=20 + i) for i in range(5)]   # all priced the same (same sizes)
    srv = W.Server(precision="fp32", max_slots=8, hbm_budget=int(2.5 * per))
    for m in models:
        srv.submit(m, max_iter=200)
    peak, done = 0, []
    for _ in range(100):
        a, q = srv.pump()
        st = srv.stats()
        assert st["bytes_in_use"] <= st["hbm_budget"]
        peak = max(peak, a)
        done += srv.poll()
        if a == 0 and q == 0:
            break
    assert peak == 2 and len(done) == 5
    assert all(r["status"] == "max_iter" and r["iterations"] == 200 for r in done)
    srv.close()


def test_cancel_and_oversized(W, gpu_ctx):
    srv = W.Server(precision="fp32", max_slots=1, hbm_budget=16 << 20)
    a, b = cases.config_small(seed=31), cases.config_small(seed=32)
    ra = srv.submit(a, max_iter=5000)
    rb = srv.submit(b, max_iter=5000)
    srv.pump()                      # a resident, b queued
    srv.cancel(rb)
    srv.cancel(ra)
    res = {r["id"]: r for r in srv.poll()}
    assert res[ra]["status"] == "cancelled" and res[rb]["status"] == "cancelled"
    assert res[ra]["iterations"] == 64 and res[rb]["iterations"] == 0
    assert srv.stats()["active"] == 0
    big = cases.config2()
    with pytest.raises(W.DeviceError):
        srv.submit(big, max_iter=10)
    srv.close()
