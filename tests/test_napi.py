"""The Node.js host: the N-API addon over the C ABI and the JavaScript GpuSolver (the interface of
web/js/solver_gpu.js).  Node runs as a child process; the GPU case compares the JS path with the Python path
through the same library, bitwise."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "weave-sim_amd", "js", "solver_mi355x.js")
ADDON = os.path.join(ROOT, "weave-sim_amd", "lib", "weave_napi.node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the N-API addon is absent")

SCENE = "{E:10000,G:600,density:500,width:40,height:4,mu:0.3,tol:1e-6}"


def run_node(script, timeout=240):
    r = subprocess.run([NODE, "-e", f"const W=require({json.dumps(JS)});\n" + script], capture_output=True, text=True,
                       timeout=timeout, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_addon_exports_and_model_builder():
    """the addon loads in Node, the JS model builder matches the Python one, and the GpuSolver class carries the
    reference's method set"""
    out = run_node(f"""
const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
const names = Object.getOwnPropertyNames(W.GpuSolver.prototype).sort();
console.log(JSON.stringify({{info: m.info, x: Array.from(m.st.X.slice(0, 6)), u: Array.from(m.st.u.slice(0, 4)),
  names, scales: W.GpuSolver.computeScales([m])}}));""")
    d = json.loads(out)
    py = cases.config1()
    assert d["info"]["n"] == py.n and d["info"]["K"] == py.K
    assert np.array_equal(np.array(d["x"]), py["X"][:6]) and np.array_equal(np.array(d["u"]), py["u"][:4])
    for name in ("upload", "updateParams", "step", "readControl", "readHistory", "download", "downloadAll",
                 "internalize", "dispose", "compileInfo", "readOutputs", "readAllOutputs"):
        assert name in d["names"], name
    assert d["scales"] == {"PB": 0, "FB": 0, "TB": 0, "KB": 0}


@pytest.mark.gpu
def test_js_gpu_solver_matches_python_path(W, gpu_ctx):
    """600 iterations of the 3 m dome through the JS GpuSolver (Promises, in-place download) equal the Python path"""
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  if (!dev) throw new Error('no device');
  const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  const c = await g.step(600);
  await g.download(0);
  const n = await g.readHistory(0);
  g.dispose();
  console.log(JSON.stringify({{ctl: c[0], X: Array.from(m.st.X), stuck: Array.from(m.st.stuck), n,
                               it: m.st.iteration, hist: m.st.history.kinetic.length}}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    m = cases.config1()
    g = W.GpuSolver([m], precision="fp32", tile=64)
    c = g.step(600)[0]
    s = g.state(0)
    g.close()
    assert d["ctl"]["iter"] == c["iter"] == 600 and d["it"] == 600
    assert np.array_equal(np.array(d["X"]), s["X"])
    assert np.array_equal(np.array(d["stuck"], dtype=np.uint8), s["stuck"])
    assert d["ctl"]["ke"] == c["ke"] and d["hist"] == d["n"] == 600


@pytest.mark.gpu
def test_js_replay_matches_python_path(W, gpu_ctx):
    """a force-controlled replay on a free node of the 3 m dome (Model.setProtocol, GpuSolver.readOutputs) equals the
    same replay through the Python path, bitwise"""
    py = cases.config1()
    probe = py.n // 2
    assert py.md.node_free[3 * probe] and py.md.node_free[3 * probe + 2]
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
  m.setProtocol({{ mode: 'force', probe: {probe}, axis: [0, 0, -1], sched: [2.0, 4.0], settle: 300, avg: 20 }});
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  let c;
  for (let k = 0; k < 20; k++) {{ c = await g.step(128); if (c[0].done) break; }}
  const rows = await g.readOutputs(0);
  const all = await g.readAllOutputs();
  g.dispose();
  console.log(JSON.stringify({{ rows: Array.from(rows), n: all[0].length, phase: c[0].phase, inc: c[0].inc }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    py.set_protocol("force", probe, [0, 0, -1], [2.0, 4.0], 300, 20)
    g = W.GpuSolver([py], precision="fp32", tile=64, per_submit=64)
    for _ in range(20):
        if g.step(128)[0]["done"]:
            break
    rows = g.read_outputs(0)
    g.close()
    assert d["phase"] == 2 and d["inc"] == 2 and d["n"] == 8
    assert np.array_equal(np.array(d["rows"]).reshape(-1, 4), rows)


@pytest.mark.gpu
def test_js_build_results_match_python(W, gpu_ctx):
    """buildResults / resultsJSON in Node over a downloaded 3 m dome equal the Python results over the same download"""
    import importlib
    R = importlib.import_module("weave-sim_amd.results")
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  await g.step(400);
  await g.download(0);
  g.dispose();
  const r = W.buildResults(m);
  const doc = W.resultsJSON(m, r);
  console.log(JSON.stringify({{ Mx: Array.from(r.segment.Mx), T: Array.from(r.segment.T), N: Array.from(r.segment.N),
    Fn: Array.from(r.knot.Fn), disp: Array.from(r.disp), nreact: r.reactions.size, doc }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    py = cases.config1()
    g = W.GpuSolver([py], precision="fp32", tile=64, per_submit=64)
    g.step(400)
    s = g.state(0)
    g.close()
    r = R.build_results(py, s)
    assert np.array_equal(np.array(d["Mx"]), r["segment"]["Mx"]) and np.array_equal(np.array(d["T"]), r["segment"]["T"])
    assert np.array_equal(np.array(d["N"]), r["segment"]["N"]) and np.array_equal(np.array(d["Fn"]), r["knot"]["Fn"])
    assert np.array_equal(np.array(d["disp"]), r["disp"]) and d["nreact"] == len(r["reactions"])
    doc = R.results_json(py, r)
    assert d["doc"]["segment"]["My"] == doc["segment"]["My"] and d["doc"]["reactions"].keys() == doc["reactions"].keys()


def test_js_scene_reader_matches_python():
    """the reference's example scenes read by the JS reader (scene_mi355x.js: scene.json, STL) and by the Python one
    build the same model: sizes, initial state, anchors, loads and solver parameters, bit for bit"""
    import importlib
    SC = importlib.import_module("weave-sim_amd.scene")
    golden = os.path.join(ROOT, "tests", "golden", "scenes")
    names = ["dome_basic", "v5dome_cm", "tree_cm", "dualv5dome_cm", "dualv5dome_cm_hybrid"]
    scene_js = os.path.join(ROOT, "weave-sim_amd", "js", "scene_mi355x.js")
    out = run_node(f"""
const S = require({json.dumps(scene_js)});
const res = {{}};
for (const name of {json.dumps(names)}) {{
  const m = S.loadScene(require('path').join({json.dumps(golden)}, name + '.json'));
  res[name] = {{ info: {{ n: m.info.n, m: m.info.m, K: m.info.K, S: m.info.S, params: m.info.params }},
    X: Array.from(m.st.X), u: Array.from(m.st.u), loads: Array.from(m.info.loadConst), free: Array.from(m.info.nodeFree) }};
}}
let refused = '';
const hy = JSON.parse(require('fs').readFileSync(require('path').join({json.dumps(golden)}, 'dualv5dome_cm_hybrid.json'), 'utf8'));
hy.hybrid.pieces = {{ '3:1': 2 }};
try {{ S.sceneFromJSON(hy); }} catch (e) {{ refused = e.message; }}
res.refused = refused;
console.log(JSON.stringify(res));""")
    d = json.loads(out)
    assert "pieces" in d["refused"]
    for name in names:
        py = SC.load_scene(os.path.join(golden, name + ".json"))
        js = d[name]
        assert (js["info"]["n"], js["info"]["m"], js["info"]["K"], js["info"]["S"]) == (py.n, py.m, py.K, py.md.S)
        assert np.array_equal(np.array(js["X"]), py["X"]) and np.array_equal(np.array(js["u"]), py["u"])
        assert np.array_equal(np.array(js["loads"]), py.arr("load_const"))
        assert np.array_equal(np.array(js["free"], dtype=np.uint8), py.arr("node_free"))
        for k, v in js["info"]["params"].items():
            assert v == getattr(py.md.par, k), (name, k, v, getattr(py.md.par, k))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tree_cm", "dualv5dome_cm_hybrid"])
def test_js_scene_on_the_device_matches_python_path(W, gpu_ctx, name):
    """the reference's tree scene (scene.json + STL, augmented-Lagrangian contact, UDL per length) and its hybrid dual
    dome (two lath groups: their stiffness, activation and contact sections come through the JS reader too) read by
    the JS reader and relaxed 800 iterations by the JS GpuSolver equal the Python reader + Python path, bitwise"""
    import importlib
    SC = importlib.import_module("weave-sim_amd.scene")
    path = os.path.join(ROOT, "tests", "golden", "scenes", name + ".json")
    scene_js = os.path.join(ROOT, "weave-sim_amd", "js", "scene_mi355x.js")
    out = run_node(f"""
(async () => {{
  const S = require({json.dumps(scene_js)});
  const dev = await W.getDevice();
  if (!dev) throw new Error('no device');
  const m = S.loadScene({json.dumps(path)});
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  const c = await g.step(800);
  await g.download(0);
  g.dispose();
  console.log(JSON.stringify({{ctl: c[0], X: Array.from(m.st.X), lam: Array.from(m.st.lam)}}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    m = SC.load_scene(path)
    g = W.GpuSolver([m], precision="fp32", tile=64)
    c = g.step(800)[0]
    s = g.state(0)
    g.close()
    assert d["ctl"]["iter"] == c["iter"] == 800 and d["ctl"]["ke"] == c["ke"]
    assert np.array_equal(np.array(d["X"]), s["X"]) and np.array_equal(np.array(d["lam"]), s["lam"])


@pytest.mark.gpu
def test_js_server_matches_python_path(W, gpu_ctx):
    """three requests through the JS Server (two slots: the third waits for a free one), each relaxed 300 iterations,
    equal the same models relaxed alone through the Python path, bitwise"""
    out = run_node("""
(async () => {
  const dev = await W.getDevice();
  if (!dev) throw new Error('no device');
  const srv = new W.Server(dev, { chunk: 64, tile: 64, maxSlots: 2 });
  for (const seed of [1, 2, 3]) {
    srv.submit(W.Model.hexDome({ span: 1200, spacing: 300, segment: 50, contraction: 0.15, noise: 1, seed },
                               { E: 10000, G: 600, density: 500, width: 40, height: 4, mu: 0.3, tol: 1e-6 }), 300);
  }
  const res = [];
  for (let r = 0; r < 100 && res.length < 3; r++) { await srv.pump(); for (const x of srv.poll()) res.push(x); }
  srv.dispose();
  console.log(JSON.stringify(res.map((x) => ({ id: x.id, iterations: x.iterations, status: x.status, X: Array.from(x.model.st.X) }))));
})().catch((e) => { console.error(e); process.exit(1); });""")
    d = sorted(json.loads(out), key=lambda x: x["id"])
    assert [x["id"] for x in d] == [1, 2, 3]
    for seed, x in zip((1, 2, 3), d):
        m = cases.config_small(seed=seed)
        g = W.GpuSolver([m], precision="fp32", tile=64)
        c = g.step(300)[0]
        s = g.state(0)
        g.close()
        assert x["iterations"] == c["iter"] == 300
        assert np.array_equal(np.array(x["X"]), s["X"]), seed
