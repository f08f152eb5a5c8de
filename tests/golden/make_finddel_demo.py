"""Writes finddel_demo_48x4.json: the FP64 oracle's findDelAuto (oracle/finddel.py,
findDel/findDelAuto.m:1-143) on the delay demo's setting (demos/demo_autoDel.m)
reduced to a 48 x 48 image and 4 coils over the reference's spiral R4 record
(traj_spiral_R4.npz): data simulated by the oracle with a 2 us delay, no
noise, maxNit_cgne 20, delJumpFact 3, no coarse search, started at 0.  The
oracle run takes about a minute, so its estimate and steps are kept here for
tests/test_gpu_finddel.py.  Run from the repository root:
    python tests/golden/make_finddel_demo.py
"""
import importlib.util
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle.finddel import findDelAuto, interpTrajTime  # noqa: E402
from oracle.highorder import HighOrderPlan  # noqa: E402
from oracle.mrsamp import mrSampFunc  # noqa: E402
from oracle.rcvr import RcvrPlan  # noqa: E402


def problem(N=48, C=4):
    spec = importlib.util.spec_from_file_location(
        "workloads", os.path.join(ROOT, "uwobrainhack2026-matmri_amd", "workloads.py"))
    w = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(w)
    P = w.delay_demo_problem(np.load(os.path.join(HERE, "traj_spiral_R4.npz")), N=N, C=C)
    a, b = interpTrajTime(P["spha"], P["conc"], P["tdwell"], P["delay"], P["datatime"])
    S = HighOrderPlan(P["b0"], P["datatime"], a.T, b.T, P["grid"])
    y = mrSampFunc(P["phantom"], "notransp", S, RcvrPlan(P["maps"]))
    return P, y


if __name__ == "__main__":
    P, y = problem()
    est, per = findDelAuto(0.0, y, P["tdwell"], P["spha"], P["conc"], P["datatime"], P["grid"], P["b0"],
                           P["maps"], 20, 3, 0)
    out = {"N": 48, "C": 4, "applied_delay_us": P["delay"], "estimate_us": float(est),
           "steps_us": [float(v) for v in per]}
    with open(os.path.join(HERE, "finddel_demo_48x4.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)
