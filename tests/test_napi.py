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
    """600 iterations of the 3 m dome through the JS GpuSolver (Promises, in-place download) equal the Python path;
    the download reads the history rows too (solver_gpu.js:499), so a readHistory after it finds none new"""
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
    assert d["ctl"]["ke"] == c["ke"] and d["hist"] == 600 and d["n"] == 0


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


@pytest.mark.gpu
def test_js_ground_block_matches_python(W, gpu_ctx):
    """the 3 m dome on a ground just above its anchors (22 nodes rest on it): the download's ground arrays (floorF, floorState), its energy
    and largest Fn, buildResults' floor block and resultsJSON's floor section in Node equal the Python path's"""
    import importlib
    R = importlib.import_module("weave-sim_amd.results")
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  const sc = Object.assign({SCENE}, {{ floor_on: 1, floor_z: 63, floor_mu: 0.4 }});
  const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, sc);
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  await g.step(300);
  await g.download(0);
  g.dispose();
  const r = W.buildResults(m);
  const doc = W.resultsJSON(m, r);
  console.log(JSON.stringify({{ F: Array.from(r.floor.F), state: Array.from(r.floor.state), nodes: r.floor.nodes,
    slipping: r.floor.slipping, sum: r.floor.sum, FnMax: r.floor.FnMax, e: r.energies.floor, doc: doc.floor }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    py = cases.config1(floor_on=1, floor_z=63.0, floor_mu=0.4)
    g = W.GpuSolver([py], precision="fp32", tile=64, per_submit=64)
    g.step(300)
    s = g.state(0)
    g.close()
    r = R.build_results(py, s)
    fl = r["floor"]
    assert fl["nodes"] > 0 and d["nodes"] == fl["nodes"] and d["slipping"] == fl["slipping"]
    assert np.array_equal(np.array(d["F"]), fl["F"]) and np.array_equal(np.array(d["state"]), fl["state"])
    assert d["sum"] == fl["sum"] and d["FnMax"] == fl["FnMax"] and d["e"] == r["energies"]["floor"]
    doc = json.loads(json.dumps(R.results_json(py, r, s)))
    assert d["doc"] == doc["floor"]


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
const hy = JSON.parse(require('fs').readFileSync(require('path').join({json.dumps(golden)}, 'dualv5dome_cm_hybrid.json'), 'utf8'));
hy.joints = {{ '5': [120.0, 40.0], '9': [10.5] }};
hy.hybrid.pieces = {{ '5:1': 2, '9:0': 2, '9:1': 2, '5:01': 2, '7:0': 9 }};
const mp = S.modelFromScene(hy, require('fs').readFileSync(require('path').join({json.dumps(golden)}, hy.mesh.path)), hy.mesh.path);
const fp = mp.fields();
res.pieces = {{ EA: Array.from(fp.EA), k_wk: Array.from(fp.k_wk), k_E90: Array.from(fp.k_E90), loads: Array.from(mp.info.loadConst) }};
const rg = JSON.parse(require('fs').readFileSync(require('path').join({json.dumps(golden)}, 'dome_basic.json'), 'utf8'));
rg.anchors = [{{ select: {{ type: 'region', id: [3, 2], radius: 900 }}, fix: [1, 1, 1, 0, 0, 0] }}];
const mr = S.modelFromScene(rg, require('fs').readFileSync(require('path').join({json.dumps(golden)}, rg.mesh.path)), rg.mesh.path);
res.region = Array.from(mr.info.nodeFree);
console.log(JSON.stringify(res));""")
    d = json.loads(out)
    # groups of pieces of a discretisation through both readers (scene.js:57-90)
    hy = json.load(open(os.path.join(golden, "dualv5dome_cm_hybrid.json")))
    hy["joints"] = {"5": [120.0, 40.0], "9": [10.5]}
    hy["hybrid"]["pieces"] = {"5:1": 2, "9:0": 2, "9:1": 2, "5:01": 2, "7:0": 9}
    pp = SC.model_from_scene(hy, base_dir=golden)
    plain = SC.load_scene(os.path.join(golden, "dualv5dome_cm_hybrid.json"))
    for k, nm in (("EA", "EA"), ("k_wk", "k_wk"), ("k_E90", "k_E90"), ("loads", "load_const")):
        assert np.array_equal(np.array(d["pieces"][k]), pp.arr(nm)), k
    assert not np.array_equal(pp.arr("EA"), plain.arr("EA"))
    # the 'region' selector through both readers (model.js:127-145)
    rg = json.load(open(os.path.join(golden, "dome_basic.json")))
    rg["anchors"] = [{"select": {"type": "region", "id": [3, 2], "radius": 900}, "fix": [1, 1, 1, 0, 0, 0]}]
    pr = SC.model_from_scene(rg, base_dir=golden).arr("node_free")
    assert np.array_equal(np.array(d["region"], dtype=np.uint8), pr) and 0 < (pr == 0).sum() < len(pr)
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
  for (let r = 0; r < 100 && res.length < 3; r++) { await srv.pump(); for (const x of await srv.poll()) res.push(x); }
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


SHAPE = os.path.join(ROOT, "tests", "js", "cpu_solver_shape.js")
SCENE_JS = os.path.join(ROOT, "weave-sim_amd", "js", "scene_mi355x.js")
GOLDEN = os.path.join(ROOT, "tests", "golden", "scenes")


def test_cpu_solver_field_map_round_trips():
    """an object with the reference's CpuSolver fields (st.segI0, beam.cSeg, contact.enabled, loads.pointLocal, …),
    assembled from a built model, maps onto ws_model / ws_state exactly: the native model over its arrays gives back
    every array and scalar of the original (the 3 m dome and the reference's five example scenes: clamps, frame-local
    loads, hybrid lath groups)"""
    out = run_node(f"""
const {{ asCpuSolver }} = require({json.dumps(SHAPE)});
const S = require({json.dumps(SCENE_JS)});
const path = require('path');
const models = [['dome_3m', W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE})]];
for (const name of ['dome_basic', 'v5dome_cm', 'tree_cm', 'dualv5dome_cm', 'dualv5dome_cm_hybrid'])
  models.push([name, S.loadScene(path.join({json.dumps(GOLDEN)}, name + '.json'))]);
// dome_basic with its lath ends clamped (rotation anchors → clamp joints)
const file = path.join({json.dumps(GOLDEN)}, 'dome_basic.json'), fs = require('fs');
const d = JSON.parse(fs.readFileSync(file, 'utf8'));
d.anchors.forEach((a) => {{ a.fix = [1, 1, 1, 1, 1, 1]; }});
const mp = path.resolve(path.dirname(file), S.sceneFromJSON(d).mesh.path);
models.push(['dome_basic_clamped', S.modelFromScene(d, fs.readFileSync(mp), mp)]);
const same = (a, b) => a.length === b.length && a.every((x, i) => x === b[i] || (x !== x && b[i] !== b[i]));
const res = {{}};
for (const [name, m] of models) {{
  const f = m.fields();
  const cpu = asCpuSolver(f);
  const g = W._internal.modelFields(W._internal.modelFromFields(W._internal.cpuSolverFields(cpu).fields));
  const bad = [];
  for (const k of Object.keys(f)) {{
    if (ArrayBuffer.isView(f[k])) {{ if (!g[k] || !same(f[k], g[k])) bad.push(k); }}
    else if (k !== 'par' && !(f[k] === g[k] || (f[k] !== f[k] && g[k] !== g[k]))) bad.push(k);
  }}
  for (const k of Object.keys(f.par)) if (!(f.par.floor_on === 0 && k.startsWith('floor_')) && f.par[k] !== g.par[k]) bad.push('par.' + k);
  res[name] = {{ bad, C: f.C, npl: f.n_point_local, K: f.K }};
}}
console.log(JSON.stringify(res));""")
    d = json.loads(out)
    assert d["dome_basic_clamped"]["C"] > 0 and d["dome_basic"]["npl"] > 0 and d["dome_3m"]["K"] > 0
    for name, v in d.items():
        assert v["bad"] == [], (name, v["bad"])


def test_cpu_solver_ground_rigid_freedoms_map_both_ways():
    """a CpuSolver-shaped object with a ground (the reference's Floor: rv, rw, c, Rb) maps its rigid velocities and
    model radius onto ws_state (rb_v, rb_w, rb_radius; read on upload, as _uploadFloor writes rv and rw,
    solver_gpu.js:236-237), and a download's rigid freedoms land in the Floor's own arrays (solver_gpu.js:461)"""
    out = run_node(f"""
const {{ asCpuSolver }} = require({json.dumps(SHAPE)});
const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
const cpu = asCpuSolver(m.fields());
const n = cpu.st.n;
const ax = new Float64Array(n), ay = new Float64Array(n);
for (let i = 0; i < n; i++) {{ ax[i] = cpu.st.X[3 * i]; ay[i] = cpu.st.X[3 * i + 1]; }}
const rv = Float64Array.of(0.25, -0.5, 0.125), rw = Float64Array.of(1e-3, -2e-3, 3e-3), c = new Float64Array(3);
cpu.floor = {{ p: {{ z: 0, mu: 0.4 }}, w: new Float64Array(n).fill(40), h: new Float64Array(n).fill(4), ax, ay,
  F: new Float64Array(3 * n), state: new Uint8Array(n), rv, rw, c, Rb: 1234.5 }};
const g = W._internal.modelFields(W._internal.modelFromFields(W._internal.cpuSolverFields(cpu).fields));
const inst = new W._internal.CpuSolverInstance(cpu);
inst.afterDownload({{ iteration: 7, residual: 1, kePrev: 0, converged: 0, resForce: 0, resTorque: 0, resRigid: 0.5,
  eBend: 0, eTwist: 0, eStretch: 0, eShear: 0, eContact: 0, penN: 0, eFloor: 0, fnFloor: 0,
  rbV: [1, 2, 3], rbW: [4, 5, 6], rbC: [7, 8, 9] }});
console.log(JSON.stringify({{ v: Array.from(g.rb_v), w: Array.from(g.rb_w), R: g.rb_radius, same: cpu.floor.rv === rv && cpu.floor.c === c,
  rv: Array.from(cpu.floor.rv), rw: Array.from(cpu.floor.rw), c: Array.from(cpu.floor.c) }}));""")
    d = json.loads(out)
    assert d["v"] == [0.25, -0.5, 0.125] and d["w"] == [1e-3, -2e-3, 3e-3] and d["R"] == 1234.5
    assert d["same"] and d["rv"] == [1, 2, 3] and d["rw"] == [4, 5, 6] and d["c"] == [7, 8, 9]


def test_js_tables_format_like_python():
    """knotsCSV / segmentsCSV in Node give the Python writers' text byte for byte (floats as Python's repr) over the
    same state, and resultsJSON keys knots by mesh edge id"""
    import importlib
    R = importlib.import_module("weave-sim_amd.results")
    out = run_node(f"""
const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
const r = W.buildResults(m);
console.log(JSON.stringify({{ knots: W.knotsCSV(m, r), segments: W.segmentsCSV(m, r), keys: Object.keys(W.resultsJSON(m, r).knot.Fn).slice(0, 5) }}));""")
    d = json.loads(out)
    py = cases.config1()
    r = R.build_results(py, py)
    assert d["knots"] == R.knots_csv(py, r) and d["segments"] == R.segments_csv(py, r)
    assert d["keys"] == [str(int(e)) for e in py.knot_edges[:5]]


@pytest.mark.gpu
def test_js_cpu_solver_objects_step_like_the_package_models(W, gpu_ctx):
    """new GpuSolver(device, [CpuSolver-shaped object]) (solver_gpu.js:65, app.js:131): 600 iterations of the 3 m dome
    given as the reference's fields equal the same model through the Python path, bitwise, with the download written
    into the object's own arrays (st.X, st.stuck, beam.N, anchors.reactions, contact.Fn); then a scene edit (μ, loads)
    taken by updateParams, and a changed clamp set rejected"""
    out = run_node(f"""
const {{ asCpuSolver }} = require({json.dumps(SHAPE)});
(async () => {{
  const dev = await W.getDevice();
  if (!dev) throw new Error('no device');
  const m = W.Model.hexDome({{span:3000,spacing:300,segment:50,contraction:0.15,noise:1,seed:1}}, {SCENE});
  const cpu = asCpuSolver(m.fields());
  const X = cpu.st.X;
  const g = new W.GpuSolver(dev, [cpu], {{ perSubmit: 64, tile: 64 }});
  const c = await g.step(600);
  await g.download(0);
  const first = {{ ctl: c[0], X: Array.from(cpu.st.X), same: cpu.st.X === X, stuck: Array.from(cpu.st.stuck),
    N: Array.from(cpu.beam.N), R: Array.from(cpu.anchors.reactions), Fn: Array.from(cpu.contact.Fn),
    it: cpu.st.iteration, hist: cpu.st.history.kinetic.length }};
  cpu.scene.contact.mu = 0.12;
  for (let i = 0; i < cpu.loads.const.length; i++) cpu.loads.const[i] *= 1.5;
  await g.updateParams(0);
  await g.step(300);
  await g.download(0);
  const second = {{ X: Array.from(cpu.st.X), it: cpu.st.iteration }};
  cpu.beam.cSeg = [0]; cpu.beam.cNode = [0]; cpu.beam.cLen = [cpu.st.l0[0] / 2]; cpu.beam.cB = [[1, 1, 1]];
  cpu.beam.cQ = cpu.st.Q.slice(0, 4); cpu.beam.cM = new Float64Array(2);
  let refused = '';
  try {{ await g.updateParams(0); }} catch (e) {{ refused = e.message; }}
  await g.dispose();
  console.log(JSON.stringify({{ first, second, refused }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    m = cases.config1()
    g = W.GpuSolver([m], precision="fp32", tile=64, per_submit=64)
    c = g.step(600)[0]
    s = g.state(0)
    f = d["first"]
    assert f["same"] and f["it"] == 600 and f["hist"] == 600 and f["ctl"]["ke"] == c["ke"]
    assert np.array_equal(np.array(f["X"]), s["X"]) and np.array_equal(np.array(f["stuck"], dtype=np.uint8), s["stuck"])
    assert np.array_equal(np.array(f["N"]), s["N"]) and np.array_equal(np.array(f["R"]), s["reactions"])
    assert np.array_equal(np.array(f["Fn"]), s["Fn"])
    m.md.par.mu = 0.12
    m.arr("load_const")[:] *= 1.5
    g.update_params(0)
    g.step(300)
    s2 = g.state(0)
    g.close()
    assert d["second"]["it"] == 900 and np.array_equal(np.array(d["second"]["X"]), s2["X"])
    assert "clamps" in d["refused"]


@pytest.mark.gpu
def test_js_calls_made_while_a_pump_runs_wait_for_it(W, gpu_ctx):
    """a request submitted, results polled and the solver's own calls made while a pump() or step() Promise is pending
    run after it (one chain per handle, and the C ABI's lock): the results equal the sequential order, bitwise"""
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  const dome = (seed) => W.Model.hexDome({{ span: 1200, spacing: 300, segment: 50, contraction: 0.15, noise: 1, seed }}, {SCENE});
  const srv = new W.Server(dev, {{ chunk: 64, tile: 64, maxSlots: 4 }});
  await srv.submit(dome(1), 300);
  const pending = srv.pump();                   // on a worker thread
  const id2 = srv.submit(dome(2), 300);         // made while the pump runs
  const early = srv.poll();
  await pending;
  const res = [...(await early)];
  for (let r = 0; r < 100 && res.length < 2; r++) {{ await srv.pump(); for (const x of await srv.poll()) res.push(x); }}
  await srv.dispose();
  // a solver: step, updateParams, download and dispose issued back to back, none awaited before the next
  const m = dome(3);
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  const s1 = g.step(200); const u = g.updateParams(0); const s2 = g.step(100); const dl = g.download(0); const dp = g.dispose();
  const [c1, , c2] = await Promise.all([s1, u, s2, dl, dp]);
  let after = '';
  try {{ await g.step(1); }} catch (e) {{ after = e.message; }}
  console.log(JSON.stringify({{ id2: await id2, res: res.map((x) => ({{ id: x.id, it: x.iterations, X: Array.from(x.model.st.X) }})),
    c1: c1[0].iter, c2: c2[0].iter, X: Array.from(m.st.X), after }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    res = sorted(d["res"], key=lambda x: x["id"])
    assert [x["id"] for x in res] == [1, 2] and d["id2"] == 2
    for seed, x in zip((1, 2), res):
        g = W.GpuSolver([cases.config_small(seed=seed)], precision="fp32", tile=64)
        g.step(300)
        s = g.state(0)
        g.close()
        assert x["it"] == 300 and np.array_equal(np.array(x["X"]), s["X"]), seed
    m = cases.config_small(seed=3)
    g = W.GpuSolver([m], precision="fp32", tile=64)
    g.step(200)
    g.update_params(0)
    g.step(100)
    s = g.state(0)
    g.close()
    assert d["c1"] == 200 and d["c2"] == 300 and np.array_equal(np.array(d["X"]), s["X"])
    assert "disposed" in d["after"]


@pytest.mark.gpu
def test_js_failed_device_call_rejects_its_promise(W, gpu_ctx):
    """a call the C ABI refuses on the worker thread (an instance index out of range) rejects its Promise with the
    library's message; the solver keeps working"""
    out = run_node(f"""
(async () => {{
  const dev = await W.getDevice();
  const m = W.Model.hexDome({{ span: 1200, spacing: 300, segment: 50, contraction: 0.15, noise: 1, seed: 1 }}, {SCENE});
  const g = new W.GpuSolver(dev, [m], {{ perSubmit: 64, tile: 64 }});
  let msg = '';
  try {{ await g.step(10, {{ instances: [7] }}); }} catch (e) {{ msg = e.message; }}
  const c = await g.step(10);
  await g.dispose();
  console.log(JSON.stringify({{ msg, it: c[0].iter }}));
}})().catch((e) => {{ console.error(e); process.exit(1); }});""")
    d = json.loads(out)
    assert "out of range" in d["msg"] and d["it"] == 10
