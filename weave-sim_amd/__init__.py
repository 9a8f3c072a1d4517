"""weave-sim_amd: MI355X dynamic-relaxation solver for woven elastic-rod gridshells.

Python side of the C ABI (include/weave.h, include/weave_model.h).  The device path is libweave_hip.so
(hand-written HIP for gfx950); there is no CPU fallback: the solver classes raise when it is not built.
"""
import importlib

from . import _abi  # noqa: F401
from .model import Model, Scene, StateBuffers  # noqa: F401

__all__ = ["Model", "Scene", "StateBuffers", "GpuSolver", "Server", "Context"]


def __getattr__(name):
    # the device classes load libweave_hip.so lazily, so the host model builder works without a GPU
    if name in ("solver", "GpuSolver", "Server", "Context", "DeviceError", "default_context"):
        mod = importlib.import_module(__name__ + ".solver")
        return mod if name == "solver" else getattr(mod, name)
    raise AttributeError(name)
