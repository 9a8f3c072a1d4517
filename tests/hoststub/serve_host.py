"""Driver of the request server's host code over the CPU stand-in runtime (test infrastructure, run as a child
process by tests/test_server_host.py, under AddressSanitizer when the library was built with it).

    serve_host.py LIBDIR bench DURATION RATE MAX_ITER SLOTS
        the serving benchmark's own start-up and loop (bench.run_serve): scene-building threads feeding a Poisson
        stream of design-variant requests while this thread submits, pumps and polls; prints its JSON line
    serve_host.py LIBDIR nan N MAX_ITER SLOTS POISON
        N requests of the stream, request POISON with a NaN in node 0; prints each result's id, status, iterations

LIBDIR holds the stand-in libweave_hip.so and the host model library.
"""
import argparse
import importlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
A = importlib.import_module("weave-sim_amd._abi")
A.LIB_DIR = sys.argv[1]


def bench_mode(duration, rate, max_iter, slots):
    import bench
    args = argparse.Namespace(rate=rate, duration=duration, max_iter=max_iter, chunk=64, slots=slots, builders=4,
                              prefetch=slots + 32, no_cpu=True)
    print(json.dumps(bench.run_serve(args)), flush=True)


def nan_mode(n, max_iter, slots, poison):
    import numpy as np
    W = importlib.import_module("weave-sim_amd")
    S = importlib.import_module("weave-sim_amd.scenes")
    models = [S.request_variant(i) for i in range(n)]
    X = np.ctypeslib.as_array(models[poison].st.X, shape=(3 * models[poison].n,))
    X[0] = float("nan")
    srv = W.Server(precision="fp32", chunk=64, tile=64, max_slots=slots)
    ids = [srv.submit(m, max_iter=max_iter) for m in models]
    res = srv.drain()
    srv.close()
    print(json.dumps({"ids": ids, "results": [{"id": r["id"], "status": r["status"], "iterations": r["iterations"]}
                                              for r in res]}), flush=True)


if __name__ == "__main__":
    mode = sys.argv[2]
    v = sys.argv[3:]
    if mode == "bench":
        bench_mode(float(v[0]), float(v[1]), int(v[2]), int(v[3]))
    else:
        nan_mode(int(v[0]), int(v[1]), int(v[2]), int(v[3]))
