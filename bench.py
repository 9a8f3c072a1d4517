"""Benchmark of the MI355X dynamic-relaxation path: node-iterations per second of woven kagome gridshells.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload variants] [--precision fp32|fp64]

A step is one dynamic-relaxation iteration (crossing kernel, strand kernel, control kernel) over the whole input.
The default workload is the batched one: 256 seeded design variants (span 2-8 m, spacing 0.2-0.4 m, mu 0.1-0.5,
contraction 8-20 %, thickness 3-6 mm, 2k-50k target nodes each; 3.74M nodes in all) relaxed in the same launches,
strand tile, graph length and crossing layout picked by the autotuner, FP32 forces with FP64 positions, frames,
kinetic energy and residual.  Other workloads: dome_3m (~1.4k nodes), dome_10m (78,606 nodes), dome_30m (~0.8M
nodes, mu = 0.05).

Prints one JSON line.  `value` is whole-job node-iterations/s with the inputs resident in HBM.  `roofline` is the
strand kernel (the dominant launch): algorithmic bytes per launch over its mean duration inside the iteration
sequence (HIP events between consecutive kernels on the solver's stream) against the 8 TB/s HBM peak, and `traffic`
the HBM bytes per launch read by two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled by what the same
counters read for a plain copy kernel.  On the variants workload the line also carries `solves`: the same batch
relaxed to the stop criterion against the plain single-instance FP64 path.  `cpu_baseline` is the f64 oracle (a
C++ restatement of the reference's CpuSolver) on this host: every host thread (at most 16), one instance each, with
the one-thread figure beside it.
"""
import argparse
import csv
import glob
import importlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20          # the calibration copy: 256 MiB read, 256 MiB written
STREAM_BYTES = 1 << 30          # the plain copy timed beside k_strand: 1 GiB read, 1 GiB written (the batch's size)

WORKLOADS = {
    "dome_3m": "kagome dome 3 m span / 0.3 m spacing / 50 mm segments, 15% contraction, +-1 mm imperfection, mu 0.3",
    "dome_10m": "kagome dome 10 m span / 0.15 m spacing / 20 mm segments, 12% contraction, +-2 mm imperfection, mu 0.3",
    "dome_30m": "kagome dome 30 m span / 0.2 m spacing / 15 mm segments, 18% contraction, +-2 mm imperfection, mu 0.05",
    "variants": "ragged batch of seeded design variants (256 by default; span 2-8 m, spacing 0.2-0.4 m, mu 0.1-0.5, "
                "contraction 8-20%, thickness 3-6 mm, 2k-50k target nodes each, seeds 3000+i)",
}


def log(msg):
    """a progress line on stderr (the JSON line stays alone on stdout)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def scenes():
    return importlib.import_module("weave-sim_amd.scenes")


def build_models(workload, batch):
    S = scenes()
    if workload == "variants":
        return [S.design_variant(i) for i in range(batch)]
    return [getattr(S, workload)()]


def strand_cross_bytes(models, precision):
    """algorithmic HBM bytes per iteration of the strand and crossing kernels: every per-iteration array read once
    and, if it changes, written once; parameters read once (the same accounting as ws_batch_bytes)"""
    R = 4.0 if precision == "fp32" else 8.0
    sb = cb = 0.0
    for m in models:
        K = m.K if m.md.par.contacts else 0
        sb += m.n * (48 + 6 * R + R + 3 * R + 1) + m.m * (64 + 6 * R + 8 + 5 * R + 3 * R) + m.p * (4 * R)
        cb += K * (32 + 6 * R + 16 + R + 1 + 16 + 4 * R + 2 * R + 1)
    return sb, cb


def host_info():
    """the host the CPU baseline ran on: CPU model, the CPUs this process may use (what nproc prints), all CPUs"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": nproc, "cpu_count": os.cpu_count()}


def cpu_baseline(model, budget_s=12.0):
    """the f64 oracle on the same instance from its initial state, one host thread, a bounded sample"""
    from oracle.bind import Oracle
    o = Oracle(model)
    t = time.perf_counter()
    o.step(5)
    dt = max(time.perf_counter() - t, 1e-6) / 5
    its = max(10, int(budget_s / dt))
    t = time.perf_counter()
    o.step(its)
    el = time.perf_counter() - t
    out = {"value": model.n * its / el, "unit": "node-iterations/s", "cores": 1, "kind": "port",
           "sample": f"{its} iterations of the same {model.n}-node instance from its initial state, f64 oracle "
                     f"(CpuSolver restated in C++, -O3, no fast-math), one host thread, {el:.1f} s"}
    out.update(host_info())
    return out


def host_threads():
    """host threads for the all-cores baseline: the CPUs this process may run on, at most 16 (one GPU's share of a
    shared host)"""
    return max(1, min(16, host_info()["nproc"]))


def cpu_solves_baseline(models, iters, budget_s=20.0, per_thread=2):
    """the f64 oracle on all host cores, one instance per thread (a sample of the same variants, evenly spaced over
    the set): each sampled instance runs a bounded number of iterations from its initial state; its solve time is its
    measured time per iteration × the iterations the batched run needed to stop.  All-cores solves/s = threads ÷ the
    mean solve time (every thread busy with one solve after another)"""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.bind import Oracle
    T = host_threads()
    B = len(models)
    k = min(B, T * per_thread)
    idx = sorted(set(int(round(q * (B - 1) / max(k - 1, 1))) for q in range(k)))
    per_inst = budget_s / per_thread

    def one(i):
        o = Oracle(models[i])
        t = time.perf_counter()
        o.step(3)
        dt = max(time.perf_counter() - t, 1e-6) / 3
        its = int(min(max(per_inst / dt, 10), iters[i]))
        t = time.perf_counter()
        o.step(its)
        el = time.perf_counter() - t
        return i, its, el

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=T) as ex:      # the oracle's calls release the interpreter lock
        res = list(ex.map(one, idx))
    wall = time.perf_counter() - t0
    solve_s = [el / its * iters[i] for i, its, el in res]
    mean_solve = sum(solve_s) / len(solve_s)
    out = {"value": T / mean_solve, "unit": "solves/s", "cores": T, "kind": "port",
           "node_iters_per_s": sum(models[i].n * its for i, its, _ in res) / wall,
           "single_core_solves_per_s": 1.0 / mean_solve,
           "sample": f"{len(idx)} of the {B} variants (evenly spaced), {T} host threads, one instance per thread, "
                     f"each run from its initial state for up to {per_inst:.0f} s of f64 oracle iterations (CpuSolver "
                     f"restated in C++, -O3, no fast-math); solve time = time per iteration x the batched run's "
                     f"iterations to the stop; {wall:.1f} s wall"}
    out.update(host_info())
    return out


def cpu_variants_baseline(models, iters):
    """node-iterations/s of the f64 oracle over the batch on every host thread (one instance per thread, a sample of
    the variants) — the baseline of the batched line — with the one-thread figure on the median-size variant and the
    all-thread solves/s beside it"""
    allc = cpu_solves_baseline(models, iters)
    med = sorted(models, key=lambda m: m.n)[len(models) // 2]
    one = cpu_baseline(med, budget_s=8.0)
    out = {"value": allc["node_iters_per_s"], "unit": "node-iterations/s", "cores": allc["cores"], "kind": "port",
           "sample": allc["sample"], "single_thread": {"value": one["value"], "unit": one["unit"], "cores": 1,
                                                       "sample": one["sample"]},
           "solves_per_s": allc["value"], "single_core_solves_per_s": allc["single_core_solves_per_s"]}
    out.update(host_info())
    return out


# ------------------------------------------------------------------------------------------------ counter passes
def solver_opts(args):
    """the batch options of this run: the tile, graph length and crossing layout given on the command line (0 = the
    library's choice by size)"""
    return {"tile": args.tile, "per_submit": args.per_submit, "cross_lanes": args.cross_lanes, "ctrl_mode": args.ctrl_mode,
            "streams": args.streams}


def kernel_only(args):
    """the workload of one counter pass: the calibration copy, then warm-up iterations and strand launches — on one
    stream, so that every dispatch covers the whole batch (per-dispatch counters are then per launch of the line)"""
    W = importlib.import_module("weave-sim_amd")
    ctx = W.solver.default_context()
    ctx.copy_probe(COPY_BYTES, 3)
    models = build_models(args.workload, args.batch)
    g = W.GpuSolver(models, precision=args.precision, **dict(solver_opts(args), streams=1))
    # the iterations launched one by one (no graphs): building the graphs replays each once with no live iteration,
    # and those dispatches would enter the per-dispatch means
    g.set_profile(True)
    g.step(64)
    g.set_profile(False)
    g.kernel_bench(10)
    g.close()


def tune_only(args):
    """the autotuner's pick for the workload, printed as one JSON line (run as a child process, so that the counter
    passes run before the measuring process touches the GPU)"""
    W = importlib.import_module("weave-sim_amd")
    models = build_models(args.workload, args.batch)
    best, rows = W.solver.autotune(models, precision=args.precision, iters=256)
    print(json.dumps({"best": best, "rows": rows}))


def run_child(args, extra, timeout):
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", args.workload, "--precision", args.precision,
           "--batch", str(args.batch)] + extra
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(extra)}: exit {r.returncode}: {r.stderr[-400:]}")
    return r.stdout


def counter_pass(args, counter, outdir):
    """one rocprofv3 pass over kernel_only; returns {kernel: mean counter value per dispatch}"""
    cmd = ["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", outdir,
           "-o", "pass", "--", sys.executable, os.path.abspath(__file__), "--kernel-only", "--workload", args.workload,
           "--precision", args.precision, "--tile", str(args.tile), "--per-submit", str(args.per_submit),
           "--cross-lanes", str(args.cross_lanes), "--ctrl-mode", str(args.ctrl_mode), "--streams", str(args.streams),
           "--batch", str(args.batch)]
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(cmd, cwd=tempfile.gettempdir(), env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 {counter}: exit {r.returncode}, no counter file")
    acc = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"]
            key = next((k for k in ("k_strand", "k_cross", "k_ctrl", "k_copy") if k in name), None)
            if key:
                acc.setdefault(key, []).append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def measure_traffic(args):
    """HBM bytes per strand launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by the factor that
    maps the same counter onto the known bytes of the copy kernel"""
    if shutil.which("rocprofv3") is None:
        return None, {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="ws_traffic_")
    try:
        fetch = counter_pass(args, "FETCH_SIZE", os.path.join(tmp, "fetch"))
        write = counter_pass(args, "WRITE_SIZE", os.path.join(tmp, "write"))
        f_fetch = COPY_BYTES / (fetch["k_copy"] * 1024.0)
        f_write = COPY_BYTES / (write["k_copy"] * 1024.0)
        t = fetch["k_strand"] * 1024.0 * f_fetch + write["k_strand"] * 1024.0 * f_write
        other = {k: fetch[k] * 1024.0 * f_fetch + write[k] * 1024.0 * f_write
                 for k in ("k_cross", "k_ctrl") if k in fetch and k in write}
        return t, {"fetch_kib": fetch["k_strand"], "write_kib": write["k_strand"], "fetch_factor": f_fetch,
                   "write_factor": f_write, "copy_fetch_kib": fetch["k_copy"], "copy_write_kib": write["k_copy"],
                   "other_kernels_bytes_per_launch": other}
    except Exception as e:   # the throughput line stands without the counter passes
        return None, {"error": str(e)[:200]}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ------------------------------------------------------------------------------------------------ solves/s
def solves_compare(W, models, opts, limit, plain_sample):
    """the batch relaxed to the stop criterion (residual <= tol, or `limit` iterations) in the batched FP32+FP64 form
    (all instances in one solver, the given options) against the plain single-instance FP64 path: every sampled
    instance (evenly spaced over the set) run alone from its initial state for as many iterations as it took in the
    batch, one after another.  Both sides are timed with HIP events around the step (graph instantiation excluded),
    the batched wall clock beside it.  Plain solves/s = 1 / the mean plain solve time of the sample."""
    B = len(models)
    g = W.GpuSolver(models, precision="fp32", **opts)
    t0 = time.perf_counter()
    ctl = g.step(limit)
    wall = time.perf_counter() - t0
    dev_s = g.timing()["ms_last_step"] * 1e-3
    iters = [c["iter"] for c in ctl]
    converged = sum(1 for c in ctl if c["done"])
    g.close()
    batched = B / dev_s
    idx = sorted(set(int(round(k * (B - 1) / max(plain_sample - 1, 1))) for k in range(plain_sample)))
    per_solve = []
    for i in idx:
        p = W.GpuSolver([models[i]], precision="fp64", tile=0)
        c = p.step(iters[i])[0]
        if c["iter"] != iters[i]:
            raise RuntimeError(f"variant {i}: the plain path stopped at {c['iter']}, the batch at {iters[i]}")
        per_solve.append(p.timing()["ms_last_step"] * 1e-3)
        p.close()
    plain = 1.0 / (sum(per_solve) / len(per_solve))
    return {"batched_fp32_fp64_solves_per_s": batched, "batched_wall_s": wall, "batched_device_s": dev_s,
            "batched_wall_solves_per_s": B / wall,
            "node_iters_per_s": sum(m.n * k for m, k in zip(models, iters)) / dev_s,
            "iteration_limit": limit, "converged": converged, "iterations_mean": sum(iters) / B,
            "plain_single_instance_fp64_solves_per_s": plain,
            "plain_sample": {"instances": idx, "iterations": "each as many as in the batch (whole solves)",
                             "solve_ms_mean": 1e3 * sum(per_solve) / len(per_solve)},
            "batched_over_plain": batched / plain}, iters


def run_solves(args):
    """design variants relaxed to the stop criterion: the batched FP32+FP64 form (autotuned) against the plain
    single-instance FP64 path (solves_compare)"""
    W = importlib.import_module("weave-sim_amd")
    models = build_models("variants", args.batch)
    B = len(models)
    best, _ = W.solver.autotune(models, precision="fp32", iters=256)
    cmp, iters = solves_compare(W, models, best, args.max_iter, args.plain_sample)
    out = {
        "metric": "solves/s", "value": cmp["batched_fp32_fp64_solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (seeded design variants)",
        "config": {"workload": f"{B} design variants ({WORKLOADS['variants']}), relaxed to residual <= tol or "
                               f"{args.max_iter} iterations", "instances": B, "nodes": sum(m.n for m in models),
                   **best},
    }
    out.update(cmp)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_solves_baseline(models, iters)
    return out


# ------------------------------------------------------------------------------------------------ serving
def run_serve(args):
    """a Poisson stream of design-variant requests (1k–200k nodes) through the continuous-batching server: each
    request relaxed to the stop criterion (residual ≤ tol or the iteration limit); latency = completion − arrival"""
    import queue
    import threading
    W = importlib.import_module("weave-sim_amd")
    S = scenes()
    arrivals = S.request_stream(rate=args.rate, duration=args.duration, seed=4)
    built = {}
    lock = threading.Lock()
    ready = threading.Condition(lock)
    todo = queue.Queue()
    for i in range(len(arrivals)):
        todo.put(i)
    inflight = threading.Semaphore(args.prefetch)

    def builder():
        while True:
            try:
                i = todo.get_nowait()
            except queue.Empty:
                return
            inflight.acquire()
            m = S.request_variant(i)
            with lock:
                built[i] = m
                ready.notify_all()

    workers = [threading.Thread(target=builder, daemon=True) for _ in range(args.builders)]
    for w in workers:
        w.start()
    srv = W.Server(precision="fp32", chunk=args.chunk, tile=64, max_slots=args.slots)
    lat, nodes, next_i, done, results, its_of = [], 0, 0, 0, [], {}
    t0 = time.perf_counter()
    t_report = 30.0
    while done < len(arrivals):
        now = time.perf_counter() - t0
        if now >= t_report:                   # progress of the stream on stderr (the JSON line stays on stdout)
            st = srv.stats()
            print(f"serve: {now:.0f} s, {next_i} submitted, {done} done, {st['active']} resident, {st['queued']} queued",
                  file=sys.stderr, flush=True)
            t_report += 30.0
        while next_i < len(arrivals) and arrivals[next_i] <= now:
            with lock:
                m = built.pop(next_i, None)
                if m is None and srv.stats()["active"] == 0:
                    ready.wait(0.01)          # nothing to relax meanwhile: wait for the scene builders
                    m = built.pop(next_i, None)
            if m is None:
                break
            srv.submit(m, max_iter=args.max_iter)
            next_i += 1
        a, q = srv.pump()
        for r in srv.poll():
            idx = r["id"] - 1
            lat.append((time.perf_counter() - t0 - arrivals[idx]) * 1e3)
            its_of[idx] = r["iterations"]
            nodes += r["n"] * r["iterations"]
            results.append(r["status"])
            done += 1
            inflight.release()
        if a == 0 and q == 0 and next_i < len(arrivals):
            time.sleep(max(0.0, min(arrivals[next_i] - (time.perf_counter() - t0), 0.002)))
    wall = time.perf_counter() - t0
    srv.close()
    lat.sort()
    pct = lambda p: lat[min(len(lat) - 1, int(p / 100.0 * len(lat)))]
    out = {
        "metric": "solves/s", "value": len(lat) / wall, "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded request stream)",
        "config": {"workload": f"Poisson stream at {args.rate}/s over {args.duration} s of design variants "
                               f"(1k-200k nodes), each relaxed to residual <= tol or {args.max_iter} iterations",
                   "requests": len(arrivals), "chunk": args.chunk, "max_slots": args.slots},
        "latency_ms": {"p50": pct(50), "p99": pct(99), "max": lat[-1]},
        "wall_s": wall, "node_iters_per_s": nodes / wall,
        "status": {k: results.count(k) for k in set(results)},
    }
    if not args.no_cpu:   # all host cores on a sample of the same requests, after the stream
        R = len(arrivals)
        k = min(R, 2 * host_threads())
        idx = sorted(set(int(round(q * (R - 1) / max(k - 1, 1))) for q in range(k)))
        out["cpu_baseline"] = cpu_solves_baseline([S.request_variant(i) for i in idx],
                                                  [its_of.get(i, args.max_iter) for i in idx])
        out["cpu_baseline"]["sample"] += f" (requests {idx[0]}..{idx[-1]} of {R})"
    return out


def run_stability(args):
    """the stability check (stability.js) of design variant 6 (7,800 nodes, 156 crossings) after 1,200 iterations:
    the perturbed copies' generalised forces evaluated on the device in batches vs by the oracle on one host thread,
    the same host colouring and factorisation after both"""
    W = importlib.import_module("weave-sim_amd")
    ST = importlib.import_module("weave-sim_amd.stability")
    from oracle.bind import Oracle
    m = importlib.import_module("weave-sim_amd.scenes").design_variant(6)
    o = Oracle(m)
    o.step(1200)
    st = o.state()
    X, Q, u = st["X"].copy(), st["Q"].copy(), st["u"].copy()
    timing = {}

    def timed(name, ev):
        def f(states):
            t = time.perf_counter()
            out = ev(states)
            timing[name] = (time.perf_counter() - t, len(states))
            return out
        return f
    ev = ST.device_evaluator(W, m, base=st, batch=256)
    ST.stability(m, timed("setup", ev), X=X, Q=Q, u=u)     # builds the batch of copies (a one-time upload)
    sd = ST.stability(m, timed("device", ev), X=X, Q=Q, u=u)
    so = ST.stability(m, timed("oracle", lambda states: [o.evaluate(x, q) for x, q in states]), X=X, Q=Q, u=u)
    td, nd = timing["device"]
    tc, nc = timing["oracle"]
    return {"metric": "stability evaluations/s", "value": nd / td, "unit": "evaluations/s", "n_gpus": 1,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded design variant, 1,200 iterations)",
            "config": {"workload": f"stability check of design variant 6 ({m.n} nodes, {m.md.K} crossings) after 1,200 iterations",
                       "dofs": sd["ndof"], "colours": sd["colours"], "evaluations": nd},
            "device": {"lam": sd["lam"], "neg": sd["neg"], "eval_s": td, "total_ms": sd["ms"],
                       "first_call_s_with_upload": timing["setup"][0]},
            "cpu_baseline": {"value": nc / tc, "unit": "evaluations/s", "cores": 1, "kind": "port",
                             "sample": f"the same {nc} perturbed states evaluated by the f64 oracle, one host thread",
                             "lam": so["lam"], "neg": so["neg"], "total_ms": so["ms"]}}


# ------------------------------------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="variants", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=256, help="instances of the variants workload")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--tile", type=int, default=None,
                    help="strand tile (64/128/256); 0 = the default (64); unset: the autotuner's pick (tile, graph "
                         "length and crossing layout)")
    ap.add_argument("--per-submit", type=int, default=0, help="iterations per captured graph (0 = the default, 64)")
    ap.add_argument("--cross-lanes", type=int, default=0, help="lanes per crossing, 1 or 2 (0 = by size)")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams of launches: 1, or 2 (the instances in two overlapping groups); 0 = by size")
    ap.add_argument("--ctrl-mode", type=int, default=0,
                    help="0: a control launch after the strand launch; 1: the strand launch runs the control step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--no-solves", action="store_true", help="leave out the batched-vs-plain solves comparison")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--kernel-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tune-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--mode", default="throughput", choices=["throughput", "solves", "serve", "stability"],
                    help="throughput: node-iterations/s of one workload; solves: design variants to the stop "
                         "criterion, batched vs plain; serve: a Poisson request stream through the server; "
                         "stability: the stability check of a relaxed dome, device-evaluated vs the oracle")
    ap.add_argument("--max-iter", type=int, default=5000, help="iteration limit of a solve (solves, serve)")
    ap.add_argument("--plain-sample", type=int, default=32, help="variants solved on the plain path (solves)")
    ap.add_argument("--rate", type=float, default=60.0, help="requests per second (serve)")
    ap.add_argument("--duration", type=float, default=120.0, help="seconds of arrivals (serve)")
    ap.add_argument("--chunk", type=int, default=64, help="iterations per server round (serve)")
    ap.add_argument("--slots", type=int, default=256, help="resident instances at most (serve)")
    ap.add_argument("--builders", type=int, default=4, help="host threads building request scenes (serve)")
    ap.add_argument("--prefetch", type=int, default=320, help="request scenes held at once: built ahead, queued or resident (serve)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU (got --gpus {args.gpus}): instances are independent solves that would "
              "shard one per GPU, and this build launches none across GPUs", file=sys.stderr)
        sys.exit(2)
    if args.kernel_only:
        kernel_only(args)
        return
    if args.tune_only:
        tune_only(args)
        return
    if args.mode == "solves":
        print(json.dumps(run_solves(args)))
        return
    if args.mode == "serve":
        # switched off: a run of this mode on an earlier build ended in a host segmentation fault, and no cause has
        # been found, neither by reading nor on the CPU (the request loop over a stand-in device runtime under
        # AddressSanitizer, tests/test_server_host.py, runs clean).  The server's own GPU tests cover admission,
        # the HBM budget, cancellation, mixed request sizes and alone-vs-mix equality.
        print("bench.py --mode serve is switched off: a run of it once ended in a host segmentation fault whose "
              "cause is not found; see tests/test_server.py and tests/test_server_host.py for the server",
              file=sys.stderr)
        sys.exit(3)
    if args.mode == "stability":
        print(json.dumps(run_stability(args)))
        return
    # the autotuner and the counter passes run as child processes, before this process touches the GPU
    tune = None
    if args.tile is None:
        log("autotuning (child process)")
        tune = json.loads(run_child(args, ["--tune-only"], 900).strip().splitlines()[-1])
        log(f"autotuner's pick: {tune['best']}")
        args.tile, args.per_submit, args.cross_lanes = (tune["best"]["tile"], tune["best"]["per_submit"],
                                                        tune["best"]["cross_lanes"])
    if not args.no_traffic:
        log("counter passes (FETCH_SIZE, WRITE_SIZE)")
    traffic, traffic_detail = (None, {"skipped": True}) if args.no_traffic else measure_traffic(args)

    log("building the workload")
    W = importlib.import_module("weave-sim_amd")
    models = build_models(args.workload, args.batch)
    n_nodes = sum(m.n for m in models)
    n_cross = sum(m.K for m in models)
    g = W.GpuSolver(models, precision=args.precision, **solver_opts(args))
    opts = g.opts
    dev = g.ctx.info()
    log(f"{len(models)} instances, {n_nodes} nodes on the device; {args.warmup} warm-up and {args.steps} timed steps")
    g.step(args.warmup)
    it0 = sum(c["iter"] for c in g.read_control())
    t0 = time.perf_counter()
    g.step(args.steps, control=False)   # returns after a synchronisation of the solver's stream
    t1 = time.perf_counter()
    it1 = sum(c["iter"] for c in g.read_control())
    wall = t1 - t0
    if it1 - it0 != args.steps * len(models):
        raise SystemExit("an instance stopped inside the timed region; rerun with fewer steps")
    value = n_nodes * args.steps / wall
    step_dev_ms = g.timing()["ms_last_step"]
    # per-kernel durations inside the iteration sequence: the kernels launched one by one, a HIP event between
    # consecutive kernels on the solver's stream
    g.set_profile(True)
    g.step(args.kernel_reps)
    t = g.timing()
    g.set_profile(False)
    kin = {"ms_cross": t["ms_cross"], "ms_strand": t["ms_strand"], "ms_ctrl": t["ms_ctrl"]}
    # and each kernel alone, launched back to back between two events
    kb = g.kernel_bench(args.kernel_reps)
    sb, cb = strand_cross_bytes(models, args.precision)
    achieved = sb / (kin["ms_strand"] * 1e-3) / 1e9
    # what a plain copy kernel (16-byte loads and stores, grid-stride) reaches on this GPU: the practical HBM ceiling
    copy_gbs = 2 * STREAM_BYTES / (g.ctx.copy_probe(STREAM_BYTES, 20) * 1e-3) / 1e9
    alg = g.bytes()["alg_bytes_per_iter"]
    # the host-buffer boundary (not part of `value`): every instance uploaded from its host model and state, and
    # downloaded into host buffers allocated beforehand
    bufs = [W.StateBuffers(m.md) for m in models]
    t2 = time.perf_counter()
    for b in range(len(models)):
        g.upload(b)
    t3 = time.perf_counter()
    for b in range(len(models)):
        g.download(b, bufs[b])
    t4 = time.perf_counter()
    g.close()
    solve_its = 5000
    host = {"upload_s": t3 - t2, "download_s": t4 - t3, "iterations_per_solve": solve_its,
            "node_iters_per_s_with_transfers": n_nodes * solve_its / (t4 - t2 + solve_its * wall / args.steps)}
    line = {
        "metric": "node-iters/s",
        "value": value,
        "unit": "node-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded kagome gridshells from the host model builder)",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "instances": len(models),
                   "nodes": n_nodes, "crossings": n_cross, "precision": args.precision, "strand_tile": opts["tile"],
                   "per_submit": opts["per_submit"], "cross_lanes": opts["cross_lanes"], "ctrl_mode": opts["ctrl_mode"],
                   "streams": opts["streams"],
                   "autotuned": tune is not None, "device": dev["name"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_strand",
                     "alg_bytes_per_launch": sb, "ms_per_launch": kin["ms_strand"],
                     "duration_from": "HIP events between consecutive kernels of the iteration sequence",
                     "copy_kernel_GBps": copy_gbs, "frac_of_copy_kernel": achieved / copy_gbs,
                     "traffic_detail": traffic_detail},
        "ms_per_step_device": step_dev_ms / args.steps,
        "kernels_ms_in_iteration": kin,
        "kernels_ms_in_iteration_sum": kin["ms_cross"] + kin["ms_strand"] + kin["ms_ctrl"],
        "kernels_ms_alone": kb,
        "alg_bytes_per_iter": alg,
        "alg_bytes_per_launch": {"k_strand": sb, "k_cross": cb},
        "alg_bytes_frac_of_peak_per_iteration": alg / (wall / args.steps) / (HBM_PEAK_GBS * 1e9),
        "host_transfers": host,
    }
    if tune is not None:
        line["autotune"] = tune["rows"]
    iters = None
    if args.workload == "variants" and not args.no_solves:
        log(f"solves: the batch to the stop criterion ({args.max_iter} iterations at most), then the plain path")
        cmp, iters = solves_compare(W, models, solver_opts(args), args.max_iter, args.plain_sample)
        line["solves"] = cmp
    if not args.no_cpu:
        log("CPU baseline (f64 oracle on the host)")
        if args.workload == "variants":
            line["cpu_baseline"] = cpu_variants_baseline(models, iters or [args.max_iter] * len(models))
        else:
            line["cpu_baseline"] = cpu_baseline(max(models, key=lambda m: m.n))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
