"""MI355X (gfx950) implementation of MatMRI's iterative-SENSE operator chain.

The operators and solvers keep the reference's MATLAB interfaces
(``nufftOp``, ``rcvrOp``, ``sampHighOrder``, ``mrSampFunc``, ``cgne``,
``bfista``); the arithmetic runs in hand-written HIP kernels behind the C ABI
of ``include/matmri_hip.h`` (``libmatmri_hip.so`` in this directory).
"""
from ._lib import MMError, OutOfMemory, device_count, lib, LIB_PATH, EXPORTS  # noqa: F401
from .operators import (LinearOp, cartOp, mrSampFunc, mrSampFuncMat, mrSampHandle, nufftOp,  # noqa: F401
                        rcvrOp, sampHighOrder, senseOp, sumDimsOp)
from . import workloads  # noqa: F401
from . import finddel  # noqa: F401
from .finddel import findDelAuto, interpTrajTime, load_traj_mat  # noqa: F401
from .solvers import bfista, cgne, eye, powermethod, set_cg_graph, softthresh  # noqa: F401


def __getattr__(name):
    # the warm-start module is loaded on first use (it is also runnable with
    # ``python -m <package>.warmstart``)
    if name in ("warmstart", "WarmStartNet"):
        import importlib
        mod = importlib.import_module(__name__ + ".warmstart")
        return mod if name == "warmstart" else mod.WarmStartNet
    raise AttributeError(name)

__all__ = ["nufftOp", "rcvrOp", "sampHighOrder", "senseOp", "cartOp", "sumDimsOp", "mrSampFunc", "mrSampHandle", "cgne",
           "bfista", "powermethod", "softthresh", "set_cg_graph", "device_count", "lib", "WarmStartNet",
           "findDelAuto", "interpTrajTime", "load_traj_mat", "eye", "mrSampFuncMat"]
