"""The GpuSolver methods the reference's callers use around step(): updateParams (app.js:185,192; protocol_ui.js:67),
internalize (protocol_ui.js:60; calibrate.js:466), the calibration flow relax → download → internalize →
lockContacts → replay (calibrate.js:452-499), and the isolation of a non-finite instance (a NaN in one batch member
or one served request leaves the others bitwise as alone)."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_update_params_takes_effect_like_the_oracle(W, gpu_ctx, precision, tol):
    """400 iterations of the 3 m dome, then μ 0.3 → 0.12 and the self-weight × 1.5 (a scene edit): updateParams on the
    device, and the oracle carried on from its own state with the edited model, 300 iterations more; a rebuild
    resets the kinetic-energy peak (solver_gpu.js:331-332)"""
    m = cases.config1()
    o = Oracle(m)
    o.step(400)
    a400 = o.state()
    g = W.GpuSolver([m], precision=precision)
    g.step(400)
    b400 = g.state(0)
    assert np.abs(a400["X"] - b400["X"]).max() <= tol * m.span
    m.md.par.mu = 0.12
    m.arr("load_const")[:] *= 1.5
    g.update_params(0)
    c = g.step(300)[0]
    b = g.state(0)
    g.close()
    a400.st.ke_prev = 0.0
    o2 = Oracle(m, state=a400)
    o2.step(300)
    a = o2.state()
    assert c["iter"] == 700 == o2.iteration
    err = np.abs(a["X"] - b["X"]).max()
    moved = np.abs(a["X"] - a400["X"]).max()
    print(f"\nupdate_params {precision}: moved {moved:.3e} mm after the edit, |dX| vs oracle {err:.3e} mm")
    assert moved > 1e-3 and err <= tol * m.span
    assert np.array_equal(a["stuck"], b["stuck"])
    # the edit changed the run: without it the state differs
    ref = Oracle(cases.config1())
    ref.step(700)
    assert np.abs(ref.state()["X"] - a["X"]).max() > 100 * err


def test_update_params_refuses_changed_clamps(W, gpu_ctx):
    """the same lath with a different set of rotation clamps: WS_ERR_TOPOLOGY, as the reference throws
    'the anchors changed the rotation clamps: rebuild the GPU solver' (solver_gpu.js:325)"""
    a = cases.cantilever(nseg=10)
    b = cases.torsion(nseg=10)
    assert (a.n, a.m, a.p) == (b.n, b.m, b.p) and a.md.C != b.md.C
    g = W.GpuSolver([a], precision="fp64")
    g.models[0] = b
    with pytest.raises(W.DeviceError) as e:
        g.update_params(0)
    assert "status 4" in str(e.value) and "clamps" in str(e.value)
    g.models[0] = a
    g.update_params(0)                   # the original clamps are accepted
    g.close()


def test_internalize_sets_the_origin_and_clears_slip(W, gpu_ctx):
    """after internalize the displacement origin is the current shape (X0 = X, Q0 = Q) and the slip lengths restart
    (solver_gpu.js:507-520); the dynamic state is untouched"""
    m = cases.crossing(**cases.CROSS_CASES["cross_slip"])
    g = W.GpuSolver([m], precision="fp64")
    g.step(3000)
    before = g.state(0)
    assert before["slip"][0] > 0
    g.internalize(0)
    after = g.state(0)
    assert np.array_equal(after["X0"], after["X"]) and np.array_equal(after["Q0"], after["Q"])
    assert np.array_equal(after["X"], before["X"]) and np.array_equal(after["u"], before["u"])
    assert after["slip"][0] == 0.0
    g.close()


CROSS_PULL = ("force", 16, [1.0, 0.0, 0.0], [4.0, 8.0, 16.0, 8.0, 0.0], 6000, 100)


def _calibration_flow(W, precision, lock):
    """calibrate.js:452-499 on the pressed crossing: relax on the device, download, internalize, optionally lock the
    contacts on the host model, then a force replay of a pull along x in a new batch from the downloaded state"""
    m = cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_press"])
    g = W.GpuSolver([m], precision=precision)
    for _ in range(50):
        if g.step(2000)[0]["done"]:
            break
    g.internalize(0)
    relaxed = g.state(0)
    g.close()
    if lock:
        m.lock_contacts()
    m.set_protocol(*CROSS_PULL)
    r = W.GpuSolver([m], precision=precision, states=[relaxed])
    for _ in range(100):
        if r.step(2000)[0]["done"]:
            break
    out = r.read_outputs(0)
    r.close()
    # the oracle: the same flow on its own relaxed state
    mo = cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_press"])
    o = Oracle(mo)
    o.run(100000, 2000)
    s = o.state()
    s["X0"][:] = s["X"]
    s["Q0"][:] = s["Q"]
    s["slip"][:] = 0.0
    if lock:
        mo.lock_contacts()
    ref = Oracle(mo, state=s).replay(*CROSS_PULL)
    return out, ref


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_relax_internalize_lock_replay_matches_the_oracle(W, gpu_ctx, precision):
    """the calibration batch flow with locked contacts and without: the replay's travel and reaction per schedule
    value against the oracle's same flow; the locked crossing cannot slide, so it travels less under the pull"""
    tol = 1e-9 if precision == "fp64" else 1e-4
    res = {}
    for lock in (True, False):
        out, ref = _calibration_flow(W, precision, lock)
        assert out.shape == ref.shape == (5, 4)
        assert np.array_equal(out[:, 2], ref[:, 2])
        assert np.abs(out[:, 0] - ref[:, 0]).max() <= tol * 1000.0, (lock, out[:, 0], ref[:, 0])
        res[lock] = out
    print(f"\n{precision}: travel locked {res[True][:, 0]}, free {res[False][:, 0]}")
    assert abs(res[True][2, 0]) < abs(res[False][2, 0])


def test_lock_contacts_by_mesh_edge(W, gpu_ctx):
    """CpuSolver.lockContacts(edges): only the crossings on the given mesh edges lock; unknown edges are ignored"""
    m = cases.config1()
    edges = m.knot_edges
    m.lock_contacts([int(edges[3]), int(edges[7]), 10 ** 9])
    locked = m.arr("k_locked").copy()
    assert locked.sum() == 2 and locked[3] == 1 and locked[7] == 1
    m.unlock_contacts([int(edges[3])])
    assert m.arr("k_locked").sum() == 1
    m.unlock_contacts()
    assert m.arr("k_locked").sum() == 0


def _poisoned(seed):
    m = cases.config_small(seed=seed)
    m["X"][3 * 10 + 2] = np.nan
    return m


def test_nan_in_one_batch_member_leaves_the_others_unchanged(W, gpu_ctx):
    """a NaN node position in one instance: that instance's kinetic energy goes non-finite after the warm-up, the
    other instances are bitwise as relaxed alone"""
    seeds = (1, 2, 3)
    alone = {}
    for s in (1, 3):
        g = W.GpuSolver([cases.config_small(seed=s)], precision="fp32", tile=64)
        g.step(400)
        alone[s] = g.state(0)
        g.close()
    models = [cases.config_small(seed=1), _poisoned(2), cases.config_small(seed=3)]
    g = W.GpuSolver(models, precision="fp32", tile=64)
    c = g.step(400)
    assert not np.isfinite(c[1]["ke"])
    for b, s in ((0, 1), (2, 3)):
        st = g.state(b)
        assert np.isfinite(c[b]["ke"])
        for k in ("X", "Q", "u", "stuck", "Fn"):
            assert np.array_equal(st[k], alone[s][k]), (s, k)
    g.close()


def test_nan_in_one_served_request_retires_only_that_one(W, gpu_ctx):
    """three requests through the server, the second poisoned: it retires as non-finite, the others reach their
    iteration limit bitwise as alone"""
    alone = {}
    for s in (1, 3):
        g = W.GpuSolver([cases.config_small(seed=s)], precision="fp32", tile=64)
        g.step(500)
        alone[s] = g.state(0)
        g.close()
    srv = W.Server(precision="fp32", chunk=64, tile=64, max_slots=4)
    models = [cases.config_small(seed=1), _poisoned(2), cases.config_small(seed=3)]
    ids = [srv.submit(m, max_iter=500) for m in models]
    res = {r["id"]: r for r in srv.drain()}
    srv.close()
    assert res[ids[1]]["status"] == "non_finite" and res[ids[1]]["iterations"] < 500
    for rid, s, m in ((ids[0], 1, models[0]), (ids[2], 3, models[2])):
        assert res[rid]["status"] == "max_iter" and res[rid]["iterations"] == 500
        for k in ("X", "Q", "u", "stuck"):
            assert np.array_equal(m[k], alone[s][k]), (s, k)
