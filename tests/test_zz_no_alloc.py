"""ws_download_noalloc (the download that allocates no device memory, staging its outputs in the library's fixed
per-GPU region) against ws_download, through the C ABI: every state and output array and every scalar bitwise
equal, on two calls in a row with the second on a smaller instance (the region keeps the first call's values), on
an instance with a pending augmented-Lagrangian update (the multipliers staged too), and an instance too large for
the region refused with WS_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

WS_ERR_CAPACITY = 5
SCALARS = ("iteration", "converged", "pen_n", "ke_prev", "residual", "res_force", "res_torque", "fref", "ke",
           "e_stretch", "e_shear", "e_bend", "e_twist", "e_contact", "res_rigid")


def _pair(W, g, b):
    """the same instance downloaded by ws_download and by ws_download_noalloc"""
    m = g.models[b]
    a, n = m.buffers(), m.buffers()
    assert g.lib.ws_download(g.h, b, C.byref(a.st)) == 0
    rc = g.lib.ws_download_noalloc(g.h, b, C.byref(n.st))
    assert rc == 0, g.lib.ws_last_error(g.ctx.h).decode()
    return a, n


def _same(a, n, tag):
    for name in a.arrays:
        assert np.array_equal(a[name], n[name], equal_nan=True), (tag, name)
    for f in SCALARS:
        assert getattr(a.st, f) == getattr(n.st, f) or (np.isnan(getattr(a.st, f)) and np.isnan(getattr(n.st, f))), (tag, f)


def test_noalloc_download_equals_download_twice_in_a_row_second_smaller(W, gpu_ctx):
    big = W.GpuSolver([cases.config1()], precision="fp32")
    big.step(400)
    a, n = _pair(W, big, 0)
    assert np.abs(a["N"]).max() > 0 and np.abs(a["reactions"]).max() > 0 and np.abs(a["Fn"]).max() > 0
    _same(a, n, "dome_3m")
    small = W.GpuSolver([cases.config_small(seed=3)], precision="fp64")
    small.step(350)
    a2, n2 = _pair(W, small, 0)
    assert a2["N"].size < a["N"].size
    _same(a2, n2, "small dome after dome_3m")
    small.close()
    big.close()


def test_noalloc_download_stages_a_pending_multiplier_update(W, gpu_ctx):
    """augmented Lagrangian every 10 iterations from iteration 0: at iteration 10 the update is pending on the
    device and the download applies it (both entry points through their own staging)"""
    m = cases.crossing(augmented=1, al_every=10, ramp_iter=0, warmup_iter=0)
    g = W.GpuSolver([m], precision="fp64")
    c = g.step(10)[0]
    assert c["al_update"] == 1
    a, n = _pair(W, g, 0)
    assert a["lam"][0] > 0
    _same(a, n, "pending AL update")
    g.close()


def test_noalloc_download_refuses_an_instance_beyond_the_region(W, gpu_ctx):
    """about 72 bytes of staging per node: a 1M-node dome needs more than the 64 MiB region"""
    m = cases.dome(34000.0, 200.0, 15.0, 0.18, 2.0, 5, mu=0.05)
    assert m.n > 1_000_000
    g = W.GpuSolver([m], precision="fp32")
    g.step(2)
    buf = m.buffers()
    rc = g.lib.ws_download_noalloc(g.h, 0, C.byref(buf.st))
    assert rc == WS_ERR_CAPACITY
    assert "more than the" in g.lib.ws_last_error(g.ctx.h).decode()
    assert g.lib.ws_download(g.h, 0, C.byref(buf.st)) == 0       # the allocating download still serves it
    assert buf.st.iteration == 2
    g.close()
