"""The ground on the device: the pinned lath of the reference's floor checks (floor_check.mjs:55-59) against the
oracle — trajectory, stick points, reactions and the converged load share — a pushed, pinned lath whose free end
slides, and the unanchored laths (rigid-body freedoms on) that fall and rest, stick, slide and slide without
friction."""
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


# the reference's unanchored cases (floor_check.mjs): the ground alone holds the lath, so its rigid-body freedoms are on
RIGID = {
    "falls_and_rests": dict(),                                  # tilted and lifted: falls, turns flat, rests
    "pushed_sticks": dict(push=3.0),                            # pushed below μ·W
    "pushed_slides": dict(push=6.0),                            # pushed above μ·W: slides as a whole
    "frictionless": dict(push=3.0, mu=0.0),
}


@pytest.mark.parametrize("case", sorted(RIGID))
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_unanchored_lath_matches_oracle(W, gpu_ctx, case, precision, tol):
    """positions, stick points, the rigid part of the residual and the rigid velocities' effect against the oracle
    over the fall / push windows"""
    m = cases.ground_lath(**RIGID[case])
    o = Oracle(m)
    g = W.GpuSolver([m], precision=precision)
    for n in (50, 400, 2000):
        o.step(n - o.iteration)
        c = g.step(n - g.read_control()[0]["iter"])[0]
        a, b = o.state(), g.state(0)
        dX = np.abs(a["X"] - b["X"]).max()
        assert dX <= tol * 100.0, (case, n, dX)
        assert np.abs(a["floor_a"] - b["floor_a"]).max() <= tol * 100.0
        # the rigid part |Q_t|/F_ref, |Q_r|/(F_ref·R): near rest the resultant cancels, and its absolute error follows
        # the position error times the ground stiffness
        rigid = o.floor()[2]
        rtol, atol = (1e-6, 1e-8) if precision == "fp64" else (1e-2, 1e-5)
        assert abs(c["res_rb"] - rigid) <= rtol * rigid + atol, (n, c["res_rb"], rigid)
        assert abs(b.st.res_rigid - c["res_rb"]) == 0.0
    g.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_download_upload_carries_the_rigid_freedoms(W, gpu_ctx, precision, tol):
    """a lath sliding as a whole, downloaded after 300 iterations and uploaded into a new batch, goes on as the oracle
    does without the break: the download carries the rigid velocities, the reference point and the model radius
    (Floor.rv / rw / c / Rb, solver_gpu.js:461) and an upload takes the velocities and the radius back
    (solver_gpu.js:236-237).  Uploaded with the velocities zeroed, the same state leaves the oracle's trajectory."""
    m = cases.ground_lath(push=6.0)
    o = Oracle(m)
    o.step(300)
    a = o.state()
    g = W.GpuSolver([m], precision=precision)
    c = g.step(300)[0]
    s = g.state(0)
    g.close()
    X = s["X"].reshape(-1, 3)
    assert c["rv"] == list(s.st.rb_v) and c["rw"] == list(s.st.rb_w) and c["rc"] == list(s.st.rb_c)   # readControl
    rv, rva = np.array(s.st.rb_v[:]), np.array(a.st.rb_v[:])
    rw, rwa = np.array(s.st.rb_w[:]), np.array(a.st.rb_w[:])
    print(f"{precision}: rb_v {rv} (oracle {rva}), rb_w {rw} (oracle {rwa}), rb_radius {s.st.rb_radius}", flush=True)
    assert abs(rv[0]) > 1e-3                                   # sliding along the push
    assert np.abs(rv - rva).max() <= tol * 10.0 and np.abs(rw - rwa).max() <= tol * 0.1
    assert np.abs(np.array(s.st.rb_c[:]) - X.mean(0)).max() <= 1e-12 * 100.0
    assert s.st.rb_radius == a.st.rb_radius
    o.step(400)
    ref = o.state()["X"]
    h = W.GpuSolver([m], precision=precision)
    h.upload(0, m, s)
    h.step(400)
    b = h.state(0)
    for k in range(3):
        s.st.rb_v[k] = 0.0
        s.st.rb_w[k] = 0.0
    h.upload(0, m, s)
    h.step(400)
    z = h.state(0)
    h.close()
    dX, dZ = np.abs(ref - b["X"]).max(), np.abs(ref - z["X"]).max()
    print(f"{precision}: resumed |dX| vs oracle {dX:.3e}, resumed at rest {dZ:.3e}", flush=True)
    assert b.st.iteration == 700 and dX <= tol * 100.0
    assert dZ > 10.0 * tol * 100.0


def test_unanchored_lath_rests_on_the_ground(W, gpu_ctx):
    """relaxed to the 1e-6 residual with the rigid freedoms on: the ground carries the whole weight (floor_check.mjs
    rest case) and the device stops where the oracle does"""
    m = cases.ground_lath()
    Wt = m.arr("load_const").reshape(-1, 3).sum(0)
    o = Oracle(m)
    o.run(20000)
    g = W.GpuSolver([m], precision="fp64")
    for _ in range(10):
        if g.step(2000)[0]["done"]:
            break
    c = g.read_control()[0]
    b = g.state(0)
    g.close()
    assert o.converged and b.st.converged == 1
    assert abs(c["iter"] - o.iteration) <= 0.02 * o.iteration
    assert np.abs(o.state()["X"] - b["X"]).max() <= 1e-4 * 100.0
    X = b["X"].reshape(-1, 3)
    assert abs(X[0, 2] - X[-1, 2]) < 1e-4


# the ground block of a download (Floor.F / state / energy / FnMax, solver_gpu.js:454-461)
GROUND_OUT = {
    "pinned_ramped": dict(anchor=True, tilt=0.0, lift=1.5, ramp_iter=200),   # read inside the ground's ramp too
    "pinned_pushed_slides": dict(anchor=True, tilt=0.0, lift=1.5, push=12.0, push_dir=(0.0, 1.0, 0.0)),
    "unanchored_pushed_slides": dict(push=6.0),
}


@pytest.mark.parametrize("case", sorted(GROUND_OUT))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ground_block_of_a_download_matches_the_oracle(W, gpu_ctx, case, precision):
    """the ground's force on every node, its stick / slip state, energy and largest Fn after the same iterations as the
    oracle's Floor, inside the ramp and past it.  The force bound follows from the measured position error: ten times
    (k_n + k_t) times it; the contact column of the history holds the ground's energy (solver_cpu.js:828)"""
    m = cases.ground_lath(**GROUND_OUT[case])
    kn, kt = m.md.par.strength, m.md.par.stick_strength
    o = Oracle(m)
    g = W.GpuSolver([m], precision=precision)
    for n in (60, 150, 1500):
        o.step(n - o.iteration)
        c = g.step(n - g.read_control()[0]["iter"])[0]
        a, b = o.state(), g.state(0)
        dX = float(np.abs(a["X"] - b["X"]).max())
        Fmax = float(np.abs(a["floor_F"]).max())
        bound = 10.0 * (kn + kt) * dX + 1e-12 * Fmax
        dF = float(np.abs(a["floor_F"] - b["floor_F"]).max())
        print(f"\n{case} {precision} it {n}: |dX| {dX:.2e}, |dF| {dF:.2e} of {Fmax:.3e} (bound {bound:.2e}), "
              f"on the ground {int((a['floor_state'] > 0).sum())}, slipping {int((a['floor_state'] == 2).sum())}")
        assert dF <= bound, (n, dF, bound)
        mism = int((a["floor_state"] != b["floor_state"]).sum())
        assert mism <= (0 if precision == "fp64" else 1), (n, mism)
        assert abs(b.st.fn_floor - a.st.fn_floor) <= 10.0 * kn * dX + 1e-12 * a.st.fn_floor
        assert abs(b.st.e_floor - a.st.e_floor) <= 10.0 * m.n * Fmax * dX + 1e-9 * a.st.e_floor
        assert b.st.e_floor == c["e_floor"] and b.st.fn_floor == c["fn_floor"]
        h = g.read_history(0)[-1]
        assert h[8] == c["iter"] and h[6] == c["e_cont"] + c["e_floor"]   # the lath may stop at its tolerance first
    assert (b["floor_state"] > 0).sum() > 0
    if "slides" in case:
        assert (b["floor_state"] == 2).sum() > 0
    g.close()


def test_ground_block_in_the_results(W, gpu_ctx):
    """buildResults' floor block and the export's floor section (solver_cpu.js:914-915; app.js:377-382) from a
    download: the summary counts and resultant follow the per-node forces and states"""
    import importlib
    R = importlib.import_module("weave-sim_amd.results")
    m = cases.ground_lath(anchor=True, tilt=0.0, lift=1.5, push=12.0, push_dir=(0.0, 1.0, 0.0))
    g = W.GpuSolver([m], precision="fp64")
    g.step(1500)
    s = g.state(0)
    g.close()
    r = R.build_results(m, s)
    fl = r["floor"]
    on = s["floor_state"] > 0
    assert fl["nodes"] == int(on.sum()) > 0 and fl["slipping"] == int((s["floor_state"] == 2).sum())
    assert np.allclose(fl["sum"], s["floor_F"].reshape(-1, 3)[on].sum(0), rtol=1e-12, atol=1e-12)
    assert fl["FnMax"] == s.st.fn_floor == s["floor_F"].reshape(-1, 3)[:, 2].max()
    assert r["energies"]["floor"] == s.st.e_floor > 0
    doc = R.results_json(m, r, s)
    assert doc["floor"]["nodes"] == fl["nodes"] and len(doc["floor"]["contacts"]) == fl["nodes"]
    first = doc["floor"]["contacts"][0]
    i = int(np.flatnonzero(on)[0])
    assert first["F"] == s["floor_F"][3 * i:3 * i + 3].tolist() and first["sliding"] == bool(s["floor_state"][i] == 2)
