"""Ragged batches: the autotuner's candidates, per-instance iteration counts, and a batch of design variants of
uneven size against the same instances relaxed alone."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def test_autotune_table_and_pick(W, gpu_ctx):
    models = [cases.config_small(seed=s) for s in (1, 2, 3)] + [cases.config1()]
    best, rows = W.solver.autotune(models, precision="fp32", iters=128)
    assert len(rows) == 10
    grid, extra = rows[:9], rows[9]
    assert {(r["tile"], r["per_submit"]) for r in grid} == {(t, p) for t in (64, 128, 256) for p in (32, 64, 128)}
    assert all(r["cross_lanes"] == 2 for r in grid)           # 4 small domes: two lanes per crossing by size
    top9 = max(grid, key=lambda r: r["node_iters_per_s"])
    assert (extra["tile"], extra["per_submit"], extra["cross_lanes"]) == (top9["tile"], top9["per_submit"], 1)
    assert all(r["node_iters_per_s"] > 0 for r in rows)
    top = max(rows, key=lambda r: r["node_iters_per_s"])
    assert (best["tile"], best["per_submit"], best["cross_lanes"]) == (top["tile"], top["per_submit"], top["cross_lanes"])
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


CROSS_ONE_LANE = 131072   # crossings in a batch from which one lane takes both sides of a crossing


def test_crossing_layout_by_size_and_bad_option(W, gpu_ctx):
    """two lanes per crossing below CROSS_ONE_LANE crossings in a batch, one from there on; other values are
    rejected"""
    m = cases.config1()
    small = W.GpuSolver([m], precision="fp32")
    assert small.opts["cross_lanes"] == 2
    small.close()
    ms = [m] + [m.clone() for _ in range(-(-CROSS_ONE_LANE // m.md.K) - 1)]
    assert sum(x.md.K for x in ms) >= CROSS_ONE_LANE
    big = W.GpuSolver(ms, precision="fp32")
    assert big.opts["cross_lanes"] == 1
    big.close()
    with pytest.raises(Exception):
        W.GpuSolver([cases.config1()], precision="fp32", cross_lanes=3)
