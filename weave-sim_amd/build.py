"""Build the native parts in-tree: the host model builder (g++), the gfx950 device library (hipcc) and the
Node N-API addon (g++ against the system Node headers).  Outputs go to weave-sim_amd/lib/."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")
HIP_SRC = [os.path.join(HERE, "csrc", "hip", f) for f in ("kernels.hip", "solver.hip", "server.hip")]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               # packed FP32 instructions stay off (v_pk_*_f32 have given silent wrong results on these boxes)
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_model():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libweave_model.so")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-I" + INC,
          os.path.join(HERE, "csrc", "model", "model.cpp"), "-o", out])
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libweave_hip.so")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    _run([hipcc] + HIPCC_FLAGS + ["-I" + INC] + HIP_SRC + ["-o", out])
    return out


def build_napi():
    """weave_napi.node: the thin N-API addon over libweave_hip.so and libweave_model.so"""
    node_inc = "/usr/include/node"
    if not os.path.exists(os.path.join(node_inc, "node_api.h")):
        return None
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "weave_napi.node")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + INC, "-I" + node_inc,
          "-DNAPI_VERSION=8", "-DNODE_GYP_MODULE_NAME=weave_napi", os.path.join(HERE, "csrc", "napi", "addon.cpp"), "-o", out,
          "-L" + LIB, "-lweave_hip", "-lweave_model", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle():
    """the f64 oracle (test infrastructure, oracle/Makefile)"""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "build", "liboracle.so")


def build_all():
    outs = [build_model(), build_hip(), build_oracle()]
    if os.path.exists(os.path.join(HERE, "csrc", "napi", "addon.cpp")):
        n = build_napi()
        if n:
            outs.append(n)
    return outs
