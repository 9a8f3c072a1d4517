"""Ragged batches: the autotuner's candidates, per-instance iteration counts, and a batch of design variants of
uneven size against the same instances relaxed alone."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_autotune_table_and_pick(W, gpu_ctx):
    models = [cases.config_small(seed=s) for s in (1, 2, 3)] + [cases.config1()]
    best, rows = W.solver.autotune(models, precision="fp32", iters=128)
    assert len(rows) == 9
    assert {(r["tile"], r["per_submit"]) for r in rows} == {(t, p) for t in (64, 128, 256) for p in (32, 64, 128)}
    assert all(r["node_iters_per_s"] > 0 for r in rows)
    top = max(rows, key=lambda r: r["node_iters_per_s"])
    assert (best["tile"], best["per_submit"]) == (top["tile"], top["per_submit"])
    g = W.GpuSolver(models, precision="fp32", tile="autotune")
    assert g.tile in (64, 128, 256)
    c = g.step(100)
    assert [x["iter"] for x in c] == [100] * 4
    g.close()


@pytest.mark.parametrize("tile", [64, 128, 256])
def test_variant_batch_matches_alone(W, gpu_ctx, tile):
    """three design variants of uneven size in one batch equal each relaxed alone, bitwise (same tile)"""
    ms = [cases.config3_variant(i) for i in (0, 5, 9)]
    g = W.GpuSolver(ms, precision="fp32", tile=tile)
    g.step(300)
    for b, m in enumerate(ms):
        a = W.GpuSolver([m], precision="fp32", tile=tile)
        a.step(300)
        x, y = a.state(0), g.state(b)
        for k in ("X", "Q", "u", "stuck"):
            assert np.array_equal(x[k], y[k]), (b, k)
        a.close()
    g.close()
