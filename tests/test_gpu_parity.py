"""Device path (libweave_hip.so through the C ABI) against the f64 oracle.

Tolerances (the bar the solver is held to): node positions within 1e-4 × span of the oracle and identical stick /
slip states at crossings; plain FP64 is additionally held to 1e-9 × span on non-chaotic windows.  Woven domes are
chaotic beyond ~1000 iterations (stick/slip toggles and global kinetic-damping resets amplify rounding, as the
reference documents in web/README.md:80-82), so trajectories are compared over windows before that, and converged
states are compared on the cases that converge (rods, single crossings)."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


class Snap:
    """oracle scalars at a checkpoint"""
    def __init__(self, o):
        self.iteration, self.residual, self.converged = o.iteration, o.residual, o.converged


def run_pair(W, model, iters, precision, cross_lanes=0):
    o = Oracle(model)
    g = W.GpuSolver([model], precision=precision, cross_lanes=cross_lanes)
    assert cross_lanes == 0 or g.opts["cross_lanes"] == cross_lanes
    done = 0
    out = []
    for n in iters:
        o.step(n - done)
        g.step(n - done)
        done = n
        out.append((n, o.state(), g.state(0), Snap(o), g.read_control()[0]))
    g.close()
    return out


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_cantilever_trajectory(W, gpu_ctx, precision, tol):
    m = cases.cantilever(nseg=10)
    for n, a, b, o, c in run_pair(W, m, [1, 10, 100, 1000, 3000], precision):
        span = 1000.0
        assert np.abs(a["X"] - b["X"]).max() <= tol * span, (n, precision)
        assert c["iter"] == o.iteration


@pytest.mark.parametrize("case", ["cantilever_edge", "torsion"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_edgewise_and_torsion_trajectories(W, gpu_ctx, case, precision, tol):
    """the reference's other two beam checks (gpu_parity.html, cases.mjs:32-42): the cantilever loaded edgewise and
    the lath with a quarter turn of built-in twist, positions and frames against the oracle along the relaxation"""
    m = cases.cantilever(nseg=10, edge=True) if case == "cantilever_edge" else cases.torsion(nseg=10)
    for n, a, b, o, c in run_pair(W, m, [1, 10, 100, 1000, 3000], precision):
        assert np.abs(a["X"] - b["X"]).max() <= tol * 1000.0, (case, n, np.abs(a["X"] - b["X"]).max())
        sgn = np.sign(np.sum(a["Q"].reshape(-1, 4) * b["Q"].reshape(-1, 4), axis=1))[:, None]    # q and −q: one frame
        assert np.abs(a["Q"].reshape(-1, 4) - sgn * b["Q"].reshape(-1, 4)).max() <= 10 * tol, (case, n)
        # the twisted lath meets the 1e-10 residual stop after the warm-up; FP32 forces leave a residual floor above
        # it, so that form runs on where the oracle stops (its state compared against the oracle's final one)
        assert c["iter"] == o.iteration or (precision == "fp32" and o.converged and c["iter"] > o.iteration)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cantilever_converged_matches_oracle_and_timoshenko(W, gpu_ctx, precision):
    """both run to the residual stop; tip deflections agree to 1e-4·span and sit 0.50 % from Timoshenko"""
    tol = 1e-10 if precision == "fp64" else 1e-6
    m = cases.cantilever(nseg=10, tol=tol)
    o = Oracle(m)
    o.run(200000, 5000)
    g = W.GpuSolver([m], precision=precision)
    for _ in range(60):
        c = g.step(5000)[0]
        if c["done"]:
            break
    s = g.state(0)
    assert s.st.converged == 1 and o.converged
    a = o.state()
    assert np.abs(a["X"] - s["X"]).max() <= 1e-4 * 1000.0
    I = 60 * 30 ** 3 / 12
    dt = 10 * 1e9 / (3 * 1e4 * I) + 10 * 1000 / (5 / 6 * 625 * 1800)
    tip = -s["X"][3 * 10 + 2]
    assert abs(100 * (tip / dt - 1) - 0.50) < 0.02
    g.close()


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("name", ["cross_press", "cross_stick", "cross_locked", "cross_slip"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_crossing_cases_match_oracle(W, gpu_ctx, name, precision, lanes):
    """the reference's single-crossing cases under both crossing layouts (one lane per crossing side, one lane per
    crossing): same stick/slip state, same friction force, positions within 1e-4 span"""
    m = cases.crossing(tol=1e-6, **cases.CROSS_CASES[name])
    its = 20000 if name != "cross_slip" else 3000
    (n, a, b, o, c), = run_pair(W, m, [its], precision, lanes)
    assert a["kstate"][0] == b["kstate"][0] and a["stuck"][0] == b["stuck"][0]
    assert np.abs(a["X"] - b["X"]).max() <= 1e-4 * 1000.0
    assert abs(a["Ft"][0] - b["Ft"][0]) <= 1e-3 * max(1.0, a["Ft"][0])
    assert abs(a["Fn"][0] - b["Fn"][0]) <= 1e-3 * max(1.0, a["Fn"][0])
    if name != "cross_slip":
        assert bool(c["done"]) == o.converged


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("precision,tol,window", [("fp64", 1e-9, [1, 100, 300, 600]), ("fp32", 1e-4, [1, 100, 300, 600])])
def test_dome_trajectory_window(W, gpu_ctx, precision, tol, window, lanes):
    """config-1 dome (1.4k nodes, 240 crossings) before chaos, under both crossing layouts: positions, contact
    parameters, stick/slip"""
    m = cases.config1()
    span = m.span
    for n, a, b, o, c in run_pair(W, m, window, precision, lanes):
        assert np.abs(a["X"] - b["X"]).max() <= tol * span, (n, np.abs(a["X"] - b["X"]).max())
        assert np.array_equal(a["stuck"], b["stuck"]), n
        assert c["iter"] == o.iteration == n
        if n > m.md.par.warmup_iter:
            assert abs(c["resid"] - o.residual) <= 1e-3 * o.residual + (1e-6 if precision == "fp64" else 1e-3)


@pytest.mark.parametrize("tile", [128, 256])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_dome_window_in_wider_strand_tiles(W, gpu_ctx, tile, precision, tol):
    """the single-instance form with strand tiles of two and four wavefronts (the autotuner's pick for one dome):
    the 3 m dome against the oracle through 600 iterations, positions and stick/slip"""
    m = cases.config1()
    o = Oracle(m)
    g = W.GpuSolver([m], precision=precision, tile=tile)
    assert g.opts["tile"] == tile
    for n in (300, 600):
        o.step(n - o.iteration)
        g.step(n - g.read_control()[0]["iter"])
        a, b = o.state(), g.state(0)
        assert np.abs(a["X"] - b["X"]).max() <= tol * m.span, (n, np.abs(a["X"] - b["X"]).max())
        assert np.array_equal(a["stuck"], b["stuck"]), n
    g.close()


@pytest.mark.parametrize("precision,tol_f,tol_c,mism", [("fp64", 1e-6, 1e-6, 0), ("fp32", 1e-3, 1e-2, 0.01)])
def test_dome_classification_window(W, gpu_ctx, precision, tol_f, tol_c, mism):
    """the W1 outputs of every crossing of the 3 m dome (solver_cpu.js:231-271) and its contact layers through the
    first 600 iterations against the oracle: contact type exactly (FP32: at most 1 % of the crossings, where θ or
    the roll sits on an edge threshold); Fn and gap within tol_f of each layer's largest value, Ft within tol_f of the
    largest Fn; θ, roll angles, nominal area, pressure and utilisation within tol_c (the FP32+FP64 form classifies in
    FP32, and an edge contact's area follows its small roll angle: FP32 pressures sit ~1.5e-3 of the largest off)"""
    m = cases.config1()
    for n, a, b, o, c in run_pair(W, m, [100, 300, 600], precision):
        same = a["ctype"] == b["ctype"]
        assert (~same).sum() <= mism * m.K, (n, int((~same).sum()))
        for k in ("theta", "phi", "Anom", "pressure", "util", "Fn", "Ft", "gap"):
            x, y = a[k], b[k]
            if k in ("Anom", "pressure", "util", "phi"):     # set by the contact type
                sel = np.repeat(same, 2) if k == "phi" else same
                x, y = x[sel], y[sel]
            # the friction force against the normal force's scale (|Ft| ≤ μ Fn; tiny while the contacts ramp in)
            scale = max(np.abs(a["Fn"] if k == "Ft" else x).max(), 1e-9)
            tol = tol_f if k in ("Fn", "Ft", "gap") else tol_c
            assert np.abs(x - y).max() <= tol * scale, (n, k, np.abs(x - y).max(), scale)
        assert (a["Fn"] > 0).sum() > 0.25 * m.K            # a third of the crossings press by iteration 100


def test_dome_divergence_is_chaos_not_error(W, gpu_ctx):
    """past the chaotic horizon the device FP64 trajectory stays within the spread of a 1e-12-perturbed oracle run
    (the reference's perturbed-CPU yardstick, gpu_parity.html:5-13) — or both stay within 1e-4 span"""
    m = cases.config_small()
    span = m.span
    o = Oracle(m)
    p = m.clone()
    p["X"][:] += 1e-12 * span * np.random.default_rng(0).standard_normal(3 * m.n)
    op = Oracle(p)
    g = W.GpuSolver([m], precision="fp64")
    its = 3000
    o.step(its); op.step(its); g.step(its)
    a, ap, b = o.state(), op.state(), g.state(0)
    chaos = np.abs(a["X"] - ap["X"]).max()
    err = np.abs(a["X"] - b["X"]).max()
    assert err <= max(20 * chaos, 1e-4 * span), (err, chaos)
    g.close()


def test_bitwise_rerun_and_alone_vs_mix(W, gpu_ctx):
    """no float atomics: identical results run to run, and an instance is bitwise the same alone and in a ragged
    batch of other domes (selftest.html:62-74)"""
    m = cases.config_small()
    others = [cases.config1(), cases.dome(2000.0, 250.0, 40.0, 0.1, 1.0, 7)]
    g1 = W.GpuSolver([m], precision="fp32")
    g2 = W.GpuSolver([m], precision="fp32")
    g3 = W.GpuSolver([others[0], m, others[1]], precision="fp32")
    for g in (g1, g2, g3):
        g.step(700)
    a, b, c = g1.state(0), g2.state(0), g3.state(1)
    for k in ("X", "Q", "V", "W", "u", "slip", "stuck", "Fn"):
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k
    for g in (g1, g2, g3):
        g.close()


def test_download_upload_resume_is_exact(W, gpu_ctx):
    """N iterations straight equal k + (N − k) with a full download → upload round trip in between"""
    m1, m2 = cases.config_small(), cases.config_small()
    g1 = W.GpuSolver([m1], precision="fp64")
    g1.step(500)
    g2 = W.GpuSolver([m2], precision="fp64")
    g2.step(200)
    s = g2.state(0)
    g2.upload(0, model=m2, state=s)
    g2.step(300)
    a, b = g1.state(0), g2.state(0)
    for k in ("X", "Q", "u", "slip", "stuck"):
        assert np.array_equal(a[k], b[k]), k
    assert a.st.iteration == b.st.iteration == 500
    g1.close(); g2.close()


def test_step_subsets_and_stop_at(W, gpu_ctx):
    """step(nIter, {instances}) runs only the selected instances, exactly nIter iterations each"""
    ms = [cases.config_small(seed=s) for s in (1, 2, 3)]
    g = W.GpuSolver(ms, precision="fp32")
    c = g.step(70, instances=[0, 2])
    assert [x["iter"] for x in c] == [70, 0, 70]
    c = g.step(30)
    assert [x["iter"] for x in c] == [100, 30, 100]
    g.close()


def test_graph_build_replay_leaves_every_instance_alone(W, gpu_ctx):
    """building the iteration graphs replays each of them once with no live iteration: after an upload (which
    rebuilds the graphs) a step of one selected instance leaves the other's iteration and state untouched, and the
    stepped instance's state is bitwise that of a batch whose graphs were never rebuilt"""
    ms = [cases.config_small(seed=s) for s in (1, 2)]   # past the 300-iteration contact warm-up: the nodes move
    ref = W.GpuSolver(ms, precision="fp32")
    ref.step(400)
    ref.step(37, instances=[1])
    r0, r1 = ref.state(0), ref.state(1)
    ref.close()
    g = W.GpuSolver(ms, precision="fp32")
    g.step(400)
    s0 = g.state(0)
    g.upload(0, ms[0], s0)                       # the instance's own state back: the graphs are captured again
    c = g.step(37, instances=[1])
    a0, a1 = g.state(0), g.state(1)
    g.close()
    assert [x["iter"] for x in c] == [400, 437]
    for k in ("X", "Q", "V", "u"):
        assert np.array_equal(a0[k], s0[k]), k
        assert np.array_equal(a0[k], r0[k]), k
        assert np.array_equal(a1[k], r1[k]), k


def test_outputs_match_oracle(W, gpu_ctx):
    """section forces, joint moments and reactions of the last evaluated state (download) against the oracle"""
    m = cases.crossing(**cases.CROSS_CASES["cross_stick"])
    (n, a, b, o, c), = run_pair(W, m, [2500], "fp64")
    for k in ("N", "Vsh", "M", "reactions"):
        scale = max(1.0, np.abs(a[k]).max())
        assert np.abs(a[k] - b[k]).max() <= 1e-6 * scale, k


def test_config2_size_properties(W, gpu_ctx):
    """the full-size single instance (78,606 nodes): finite state, unit quaternions, energies consistent"""
    m = cases.config2()
    g = W.GpuSolver([m], precision="fp32")
    c = g.step(640)[0]
    s = g.state(0)
    assert c["iter"] == 640
    assert np.isfinite(s["X"]).all() and np.isfinite(s["Q"]).all()
    q = s["Q"].reshape(-1, 4)
    assert np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-12
    assert c["ke"] >= 0 and np.isfinite(c["resid"])
    assert np.abs(s["X"] - m["X0"]).max() < 0.5 * m.span
    g.close()


BRANCHES = {
    "w2_rotational_friction": dict(rot_friction=1),
    "w3_seat_law": dict(seat_law=1, dent_after=350),
    "al_multipliers": dict(augmented=1, ramp_iter=100, al_every=50),
    "all_three": dict(rot_friction=1, seat_law=1, dent_after=350, augmented=1, ramp_iter=100, al_every=50),
}


@pytest.mark.parametrize("branch,lanes", [(b, 2) for b in sorted(BRANCHES)] + [("all_three", 1)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_default_off_branches_match_oracle(W, gpu_ctx, branch, lanes, precision, tol):
    """W2 rotational friction, W3 seat law and the AL multipliers (default off in every workload) on the sliding
    single crossing and on the small dome's pre-chaos window: positions, contact state and the branch's own state
    (all three together also under one lane per crossing)"""
    kw = BRANCHES[branch]
    m = cases.crossing(tol=1e-9, **dict(cases.CROSS_CASES["cross_slip"], **kw))
    (n, a, b, o, c), = run_pair(W, m, [2500], precision, lanes)
    dX = np.abs(a["X"] - b["X"]).max()
    assert dX <= tol * 1000.0
    assert np.array_equal(a["stuck"], b["stuck"]) and np.array_equal(a["kstate"], b["kstate"])
    for k in ("lam", "theta0", "Mrot", "dent", "climb"):
        scale = max(1.0, np.abs(a[k]).max()) if np.isfinite(a[k]).all() else 1.0
        atol = (1e-7 if precision == "fp64" else 1e-3) * scale
        if k == "lam" and precision == "fp32":
            # λ integrates k_n × penetration: its error is k_n times the position error the FP32 form carries
            atol = max(atol, 10 * m.md.par.strength * dX)
        ok = np.isclose(a[k], b[k], rtol=0, atol=atol, equal_nan=True)
        assert ok.all(), (k, a[k], b[k], dX)
    # the dome: within tolerance before the branches' discrete events (dents set and released, edge contacts)
    # amplify rounding; at 600 iterations within the spread of an oracle run perturbed at the form's rounding level
    # (1e-12 of the span for FP64, 1e-7 for FP32 forces) when that is larger
    d = cases.config_small(**kw)
    p = d.clone()
    eps = 1e-12 if precision == "fp64" else 1e-7
    p["X"][:] += eps * d.span * np.random.default_rng(0).standard_normal(3 * d.n)
    op = Oracle(p)
    for n, a, b, o, c in run_pair(W, d, [100, 400, 600], precision, lanes):
        op.step(n - op.iteration)
        spread = np.abs(a["X"] - op.state()["X"]).max()
        err = np.abs(a["X"] - b["X"]).max()
        if n <= 400:
            assert err <= tol * d.span, (n, err)
            assert np.array_equal(a["stuck"], b["stuck"]), n
        else:
            assert err <= max(tol * d.span, 20 * spread), (n, err, spread)


def test_control_reduction_over_several_workgroups(W, gpu_ctx):
    """the 10 m dome in strand tiles of 64 leaves 1,350 tile partials, reduced here by three control workgroups
    of 512 (the last to finish combines them): kinetic energy, strain energies and the residual parts against the
    oracle"""
    m = cases.config2()
    its = 320                                   # past the 300-iteration warm-up: velocities are live
    o = Oracle(m)
    o.step(its)
    a = o.state()
    g = W.GpuSolver([m], precision="fp64", tile=64, ctrl_chunk=512)
    c = g.step(its)[0]
    g.close()
    assert c["iter"] == its
    assert abs(c["ke"] - o.ke) <= 1e-6 * o.ke
    for k_dev, k_ora in (("e_stretch", "e_stretch"), ("e_bend", "e_bend"), ("e_tors", "e_twist"), ("e_shear", "e_shear"),
                         ("e_cont", "e_contact")):
        ref = getattr(a.st, k_ora)
        assert abs(c[k_dev] - ref) <= 1e-6 * max(abs(ref), 1e-12), (k_dev, c[k_dev], ref)
    assert abs(c["res_f"] - a.st.res_force) <= 1e-6 * a.st.res_force
    assert abs(c["res_t"] - a.st.res_torque) <= 1e-6 * a.st.res_torque


def test_wide_control_reduction_and_mixed_widths(W, gpu_ctx):
    """the 30 m dome in strand tiles of 64 leaves 13,080 tile partials, reduced by 13 control workgroups of 1,024
    (the last to finish combines their partials through LDS): energies, kinetic energy and residual parts against
    the oracle; a small dome batched with it is bitwise the same as alone"""
    big = cases.config5()
    big.md.par.warmup_iter = 0                  # residual parts from the first iteration on
    its = 5                                     # past a kinetic-damping roll-back at iteration 3
    o = Oracle(big)
    o.step(its)
    a = o.state()
    small = cases.config1()
    alone = W.GpuSolver([small.clone()], precision="fp64", tile=64)
    c_alone = alone.step(its)[0]
    s_alone = alone.state(0)
    alone.close()
    g = W.GpuSolver([small, big], precision="fp64", tile=64)
    c_small, c = g.step(its)
    s_small = g.state(0)
    g.close()
    assert c["iter"] == its
    assert o.ke > 0 and abs(c["ke"] - o.ke) <= 1e-9 * o.ke
    for k_dev, k_ora in (("e_stretch", "e_stretch"), ("e_bend", "e_bend"), ("e_tors", "e_twist"), ("e_shear", "e_shear"),
                         ("e_cont", "e_contact")):
        ref = getattr(a.st, k_ora)
        assert abs(c[k_dev] - ref) <= 1e-9 * max(abs(ref), 1e-12), (k_dev, c[k_dev], ref)
    assert a.st.res_force > 0 and a.st.res_torque > 0
    assert abs(c["res_f"] - a.st.res_force) <= 1e-9 * a.st.res_force
    assert abs(c["res_t"] - a.st.res_torque) <= 1e-9 * a.st.res_torque
    for k in ("ke", "e_stretch", "e_bend", "res_f", "res_t"):
        assert c_small[k] == c_alone[k], k
    assert np.array_equal(s_small["X"], s_alone["X"])
