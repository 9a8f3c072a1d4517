"""The forms of the control step (solver_cpu.js:809-837: partial sums reduced, kinetic damping, residual, stop
test, history): a control launch after the strand launch (ctrl_mode 0, the default) and the strand launch's last
tiles of each instance running it (ctrl_mode 1: the partials written through the L2, or fenced on a UTM replay, whose
control step reads the probe's new position).  Each against the oracle, on instances whose strand partials fit one
reduction group and on instances split over several (the 10 m dome: ~1,400 tiles; the fused form's groups are 256
tiles), on a batch of mixed sizes, on the unanchored lath whose rigid-body step the control step also carries, and on
replays batched with a plain relaxation."""
import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
@pytest.mark.parametrize("scene", ["config1", "config2"])
def test_control_forms_match_the_oracle(W, gpu_ctx, scene, precision, tol, mode):
    m = getattr(cases, scene)()
    o = Oracle(m, history=True)
    o.step(600)
    a = o.state()
    ke, _, resid = o.history()
    g = W.GpuSolver([m], precision=precision, tile=64, ctrl_mode=mode)
    assert g.opts["ctrl_mode"] == mode
    c = g.step(600)[0]
    b = g.state(0)
    hist = g.read_history(0)
    g.close()
    err = np.abs(a["X"] - b["X"]).max()
    print(f"\n{scene} {precision} ctrl_mode {mode}: n={m.n} |dX| {err:.3e} mm, ke {c['ke']:.6e}")
    assert c["iter"] == 600
    assert err <= tol * m.span
    assert np.array_equal(a["stuck"], b["stuck"])
    # the history ring the control step writes: kinetic energy, residual (infinite through the warm-up), iteration
    assert hist.shape == (600, 9) and np.array_equal(hist[:, 8], np.arange(1, 601))
    assert np.abs(hist[:, 0] - ke).max() <= (1e-9 if precision == "fp64" else 1e-4) * ke.max()
    assert np.array_equal(np.isinf(hist[:, 7]), np.isinf(resid))
    fin = np.isfinite(resid)
    if precision == "fp64":
        assert np.abs(hist[fin, 7] - resid[fin]).max() <= 1e-6 * resid[fin].max()


def test_control_forms_agree_on_a_mixed_batch(W, gpu_ctx):
    """six small domes and the 3 m dome in one batch, the control launch and a fused form: the same states to the
    reduction-order level, and each instance of the fused form bitwise as that instance alone in the fused form"""
    models = [cases.config_small(seed=s) for s in range(1, 7)] + [cases.config1()]
    st = {}
    fused = 1
    for mode in (0, fused):
        g = W.GpuSolver([mm.clone() for mm in models], precision="fp64", tile=64, ctrl_mode=mode)
        g.step(500)
        st[mode] = [g.state(b) for b in range(len(models))]
        g.close()
    for b, mm in enumerate(models):
        assert np.abs(st[0][b]["X"] - st[fused][b]["X"]).max() <= 1e-9 * mm.span, b
        assert np.array_equal(st[0][b]["stuck"], st[fused][b]["stuck"]), b
    for b in (0, 6):
        g = W.GpuSolver([models[b].clone()], precision="fp64", tile=64, ctrl_mode=fused)
        g.step(500)
        alone = g.state(0)
        g.close()
        for k in ("X", "Q", "u", "stuck"):
            assert np.array_equal(alone[k], st[fused][b][k]), (b, k)


@pytest.mark.parametrize("mode", [0, 1])
def test_control_forms_carry_the_rigid_step(W, gpu_ctx, mode):
    """the lath only the ground holds: the control step integrates its rigid-body freedoms"""
    m = cases.ground_lath(push=6.0)
    o = Oracle(m)
    o.step(1500)
    a = o.state()
    g = W.GpuSolver([m], precision="fp64", ctrl_mode=mode)
    c = g.step(1500)[0]
    b = g.state(0)
    g.close()
    assert np.abs(a["X"] - b["X"]).max() <= 1e-9 * 100.0
    rigid = o.floor()[2]
    assert abs(c["res_rb"] - rigid) <= 1e-6 * rigid + 1e-8


def test_ctrl_mode_out_of_range_is_refused(W, gpu_ctx):
    with pytest.raises(W.DeviceError) as e:
        W.GpuSolver([cases.config_small(seed=1)], ctrl_mode=2)
    assert "ctrl_mode" in str(e.value)


def test_control_forms_run_the_replays(W, gpu_ctx):
    """a force-control and a displacement-control cantilever replay batched with a plain dome, control in the strand
    launch: the replays' read-outs (travel, reaction, value) as with the control launch.  The replay instances take
    the fenced path (their control step reads the probe's new position) and the dome the write-through one, in the
    same launches."""
    mode = 1
    from test_gpu_replay import CANT_D, CANT_F, cant
    outs = {}
    for m_ in (0, mode):
        models = [cant().set_protocol(*CANT_F), cant().set_protocol(*CANT_D), cases.config_small()]
        g = W.GpuSolver(models, precision="fp64", ctrl_mode=m_)
        for _ in range(200):     # every step is forced (clears `done`), so a finished replay is told by its phase
            if all(c["phase"] == 2 for c in g.step(2000)[:2]):
                break
        ctl = g.read_control()
        outs[m_] = [g.read_outputs(b) for b in range(2)]
        g.close()
        assert all(c["phase"] == 2 and c["inc"] == n for c, n in zip(ctl, (len(CANT_F[3]), len(CANT_D[3])))), ctl[:2]
    for a, b in zip(outs[0], outs[mode]):
        assert a.shape == b.shape
        assert np.array_equal(a[:, 2], b[:, 2])
        assert np.abs(a[:, 0] - b[:, 0]).max() <= 1e-9 * 1000.0
        assert np.abs(a[:, 1] - b[:, 1]).max() <= 1e-6 * max(np.abs(a[:, 1]).max(), 1.0)
