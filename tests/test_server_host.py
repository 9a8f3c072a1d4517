"""The request server's host code on the CPU (no GPU): solver.hip and server.hip compiled with g++ against a CPU
stand-in for the HIP runtime (tests/hoststub: device memory on the host heap, synchronous streams, graph capture
recorded and replayed, kernels emulated as the control records' iteration bookkeeping), under AddressSanitizer and
UBSan when g++ provides them.  Over it run the serving benchmark's own start-up and loop — four scene-building
threads feeding a Poisson stream of 1k-200k-node requests while the main thread creates the server, submits, pumps
and polls — and a stream with one poisoned request.  Admission, relisting, the lazily captured graphs of each
length, retirement and the write-back into the caller's buffers are the product's host code; a heap error anywhere
in them, or in the work lists the kernels would be handed, ends the child process with the sanitizer's report."""
import json
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUB = os.path.join(ROOT, "tests", "hoststub")
MODEL_LIB = os.path.join(ROOT, "weave-sim_amd", "lib", "libweave_model.so")


def _runtime(name):
    r = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def stub_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    if not os.path.exists(MODEL_LIB):
        pytest.skip("libweave_model.so is not built")
    d = tmp_path_factory.mktemp("hoststub")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"] if asan and ubsan else []
    cmd = ["g++", "-std=c++17", "-O0", "-g", "-w", "-fPIC", "-shared", *san, "-I" + STUB, "-I" + os.path.join(ROOT, "include"),
           os.path.join(STUB, "tu_solver.cpp"), os.path.join(STUB, "tu_server.cpp"), "-o", str(d / "libweave_hip.so")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    os.symlink(MODEL_LIB, d / "libweave_model.so")
    env = dict(os.environ)
    if san:
        env["LD_PRELOAD"] = f"{asan}:{ubsan}"
        env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
        env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["HIPSTUB_FREE_MIB"] = "2048"
    return str(d), env, bool(san)


def _run(stub_lib, *args, timeout=600):
    d, env, _ = stub_lib
    r = subprocess.run([sys.executable, os.path.join(STUB, "serve_host.py"), d, *map(str, args)], env=env,
                       capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    assert r.returncode == 0, f"rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.timeout(900)
def test_serving_bench_startup_and_stream_on_the_host(stub_lib):
    """bench.py --mode serve's own code path over 4 s of a 60 req/s stream, 16 slots, a 1.6 GB budget (80 % of the
    stand-in's 2 GiB): every request is admitted in arrival order as slots and bytes free up and retires at its limit"""
    out = _run(stub_lib, "bench", 4.0, 60.0, 320, 16)
    n = out["config"]["requests"]
    print(f"\n{n} requests, {out['wall_s']:.1f} s, statuses {out['status']}, sanitizers {stub_lib[2]}")
    assert n > 150
    assert out["status"] == {"max_iter": n}


def test_a_poisoned_request_retires_alone_on_the_host(stub_lib):
    """the retire-on-non-finite path: 12 requests through 4 slots, the fifth with a NaN in node 0 — it retires as
    non-finite after its first chunk, every other request at its limit"""
    out = _run(stub_lib, "nan", 12, 200, 4, 4)
    st = {r["id"]: (r["status"], r["iterations"]) for r in out["results"]}
    ids = out["ids"]
    assert sorted(st) == sorted(ids)
    for k, rid in enumerate(ids):
        assert st[rid] == (("non_finite", 64) if k == 4 else ("max_iter", 200)), (k, st[rid])
