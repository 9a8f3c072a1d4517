"""Results over a device download (buildResults, web/js/solver_cpu.js:878-922) against the same results over the
oracle's state: the relaxed cantilever and a stuck single crossing; and the Node buildResults / resultsJSON against the
Python ones over the same downloaded state."""
import importlib
import json

import numpy as np
import pytest

import cases
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu
R = importlib.import_module("weave-sim_amd.results")


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-4)])
def test_results_match_oracle(W, gpu_ctx, precision, tol):
    for build, its in ((lambda: cases.cantilever(F=10.0, tol=1e-10), 3000),       # before either stops at the residual
                       (lambda: cases.crossing(tol=1e-7, **cases.CROSS_CASES["cross_stick"]), 3000)):
        m = build()
        o = Oracle(m)
        o.step(its)
        g = W.GpuSolver([m], precision=precision)
        g.step(its)
        a, b = R.build_results(m, o.state()), R.build_results(m, g.state(0))
        g.close()
        assert a["iterations"] == b["iterations"] == its
        for k in ("Mx", "My", "T", "N"):
            scale = max(np.abs(a["segment"][k]).max(), 1.0)
            assert np.abs(a["segment"][k] - b["segment"][k]).max() <= tol * 1e3 * scale, k
        assert np.abs(a["disp"] - b["disp"]).max() <= tol * 1000.0
        assert a["reactions"].keys() == b["reactions"].keys()
        for key in a["reactions"]:      # components far below the reaction's size cancel: tolerance on its size
            ra, rb = np.array(a["reactions"][key]), np.array(b["reactions"][key])
            assert np.abs(ra - rb).max() <= tol * 1e3 * max(np.abs(ra).max(), 1.0), (key, ra, rb)
        for k in a["knot"]:
            assert np.allclose(a["knot"][k], b["knot"][k], rtol=tol * 1e3, atol=tol * 10), k
