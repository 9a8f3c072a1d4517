"""Benchmark of the MI355X dynamic-relaxation path: node-iterations per second of woven kagome gridshells.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--precision fp32|fp64] [--tile 0|64|128|256]

A step is one dynamic-relaxation iteration (crossing kernel, strand kernel, control kernel) over the whole input.
Default workload: the single large instance — a kagome dome of 10 m developed width, 0.15 m strand spacing,
0.02 m segments, 12 % contraction, seeded ±2 mm noise (78,606 nodes, 9,900 crossings), FP32 forces with FP64
positions, frames, kinetic energy and residual.  Other configurations: 1 (3 m dome, ~1.4k nodes), 3 (256 ragged
design variants in one batch), 5 (30 m dome, μ = 0.05, ~0.8M nodes).

Prints one JSON line.  `value` is whole-job node-iterations/s with the inputs resident in HBM; `roofline` is the
strand kernel (the dominant launch) against the 8 TB/s HBM peak; `cpu_baseline` is the f64 oracle on this host.
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def strand_cross_bytes(models, precision):
    """algorithmic HBM bytes per iteration split by kernel: every per-iteration array read once and written once
    if it changes (the same accounting as ws_batch_bytes)"""
    R = 4.0 if precision == "fp32" else 8.0
    sb = cb = 0.0
    for m in models:
        K = m.K if m.md.par.contacts else 0
        sb += m.n * (48 + 6 * R + R + 3 * R + 1) + m.m * (64 + 6 * R + 8 + 5 * R + 3 * R) + m.p * (4 * R)
        cb += K * (32 + 6 * R + 16 + R + 1 + 16 + 4 * R + 2 * R + 1)
    return sb, cb


def build_models(config, batch=256):
    import cases
    if config == 1:
        return [cases.config1()], "kagome dome 3 m / 0.3 m spacing / 0.05 m segments, 15% contraction, ±1 mm noise"
    if config == 2:
        return [cases.config2()], "kagome dome 10 m / 0.15 m spacing / 0.02 m segments, 12% contraction, ±2 mm noise"
    if config == 3:
        return [cases.config3_variant(i) for i in range(batch)], f"{batch} ragged kagome design variants, 2k-50k nodes each"
    if config == 5:
        return [cases.config5()], "kagome dome 30 m / 0.2 m spacing / 0.015 m segments, mu 0.05, 18% contraction"
    raise SystemExit(f"unknown config {config}")


def cpu_baseline(model, budget_s=12.0):
    """the f64 oracle (a port of the reference's CpuSolver, 1 thread) on the same instance, a bounded sample"""
    from oracle.bind import Oracle
    o = Oracle(model)
    t = time.perf_counter()
    o.step(5)
    dt = max(time.perf_counter() - t, 1e-6) / 5
    its = max(10, int(budget_s / dt))
    t = time.perf_counter()
    o.step(its)
    el = time.perf_counter() - t
    return {"value": model.n * its / el, "unit": "node-iterations/s", "cores": 1, "kind": "port",
            "sample": f"{its} iterations of the same {model.n}-node instance from its initial state, f64 oracle "
                      f"(CpuSolver restated in C++), one host thread, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3200)
    ap.add_argument("--warmup", type=int, default=320)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--batch", type=int, default=256, help="instances of config 3")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run the copy kernel over this many bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (got --gpus {args.gpus}); instances are independent and would shard "
              "one solve per GPU, which this build does not launch", file=sys.stderr)
        sys.exit(2)
    W = importlib.import_module("weave-sim_amd")
    if args.copy_probe:
        ctx = W.solver.default_context()
        ms = ctx.copy_probe(args.copy_probe, 20)
        print(json.dumps({"copy_bytes": args.copy_probe, "ms": ms, "GBps": 2 * args.copy_probe / (ms * 1e-3) / 1e9}))
        return
    models, desc = build_models(args.config, args.batch)
    n_nodes = sum(m.n for m in models)
    n_cross = sum(m.K for m in models)
    g = W.GpuSolver(models, precision=args.precision, tile=args.tile)
    dev = g.ctx.info()
    # warmup, then exactly K iterations between two synchronisations of the solver's stream
    g.step(args.warmup)
    it0 = sum(c["iter"] for c in g.read_control())
    t0 = time.perf_counter()
    ctl = g.step(args.steps)
    t1 = time.perf_counter()
    it1 = sum(c["iter"] for c in ctl)
    ev_ms = g.timing()["ms_last_step"]
    wall = t1 - t0
    # every instance must have run exactly K iterations (none may stop inside the timed region)
    if it1 - it0 != args.steps * len(models):
        raise SystemExit("an instance stopped inside the timed region; rerun with fewer steps")
    node_iters = n_nodes * args.steps
    value = node_iters / wall
    # dominant kernel: mean duration of a strand-kernel launch on the solver's stream
    kb = g.kernel_bench(args.kernel_reps)
    sb, cb = strand_cross_bytes(models, args.precision)
    achieved = sb / (kb["ms_strand"] * 1e-3) / 1e9
    alg = g.bytes()["alg_bytes_per_iter"]
    g.close()
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
        "data": "synthetic (seeded kagome domes built by the host model builder)",
        "config": {"workload": f"config {args.config}: {desc}", "instances": len(models), "nodes": n_nodes,
                   "crossings": n_cross, "precision": args.precision, "device": dev["name"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_strand",
                     "alg_bytes_per_launch": sb, "ms_per_launch": kb["ms_strand"]},
        "kernels_ms": kb,
        "alg_bytes_per_iter": alg,
        "whole_iteration_hbm_frac": alg / (wall / args.steps) / (HBM_PEAK_GBS * 1e9),
        "event_ms_timed_region": ev_ms,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(models[0] if len(models) == 1 else max(models, key=lambda m: m.n))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
