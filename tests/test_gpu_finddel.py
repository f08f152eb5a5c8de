"""interpTrajTime and findDelAuto (findDel/) on the device against the
oracle, on the reference's spiral trajectory (tests/golden/traj_spiral_R4.npz)."""
import os

import numpy as np
import pytest

from conftest import load_pkg
from oracle.finddel import findDelAuto as o_findDelAuto
from oracle.finddel import interpTrajTime as o_interpTrajTime

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "traj_spiral_R4.npz")


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


@pytest.mark.parametrize("delay,start", [(0.0, 50.0), (2.0, 50.0), (-3.7, 0.0), (1.25, 14995.0)])
def test_interp_traj_time_vs_oracle(mm, delay, start):
    """The device makima resampling (FP64) of phs_spha [15000 x 16] and
    phs_conc [15000 x 4] at the acquisition times, delayed: equal to the
    oracle to FP64 rounding, NaN at the same places (times past the end)."""
    print(f"[case] interpTrajTime delay={delay} start={start}", flush=True)
    d = np.load(GOLD)
    dt = d["datatime"].reshape(-1) + start
    a, b = mm.finddel.interpTrajTime(d["phs_spha"], d["phs_conc"], 1.0, delay, dt)
    ao, bo = o_interpTrajTime(d["phs_spha"], d["phs_conc"], 1.0, delay, dt)
    for u, v in ((a, ao), (b, bo)):
        assert u.shape == v.shape
        assert np.array_equal(np.isnan(u), np.isnan(v))
        ok = ~np.isnan(v)
        err = np.max(np.abs(u[ok] - v[ok])) / np.max(np.abs(v[ok])) if ok.any() else 0.0
        print(f"  max rel err {err:.2e}, NaN {int((~ok).sum())}", flush=True)
        assert err < 1e-12


def test_find_del_auto_vs_oracle(mm):
    """findDelAuto.m on a 24x24 / 3-coil high-order problem simulated with a
    2 us delay on 2000 samples of the spiral: the device search (device
    sampHighOrder, cgne, setPhiDiv) lands where the FP64 oracle's does
    (within 0.01 us) and within 0.02 us of the applied delay."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_finddel import _delay_problem
    print("[case] findDelAuto 24x24 C=3 K=2000, delay 2 us", flush=True)
    d, dt, grid, b0, maps, y = _delay_problem()
    est, per = mm.finddel.findDelAuto(0.0, y, 1.0, d["phs_spha"], d["phs_conc"], dt, grid, b0, maps, 12)
    est_o, per_o = o_findDelAuto(0.0, y, 1.0, d["phs_spha"], d["phs_conc"], dt, grid, b0, maps, 12)
    print(f"  device {est:.4f} ({len(per)} steps), oracle {est_o:.4f} ({len(per_o)} steps)", flush=True)
    assert abs(est - est_o) < 0.01 and abs(est - 2.0) < 0.02
    # the data already on the device (this package's torch layout, [C, K])
    # take the same path: the same estimate and steps
    import torch
    yd = torch.as_tensor(np.ascontiguousarray(np.asarray(y).T), dtype=torch.complex64, device="cuda")
    est_d, per_d = mm.finddel.findDelAuto(0.0, yd, 1.0, d["phs_spha"], d["phs_conc"], dt, grid, b0, maps, 12)
    print(f"  device-resident data: {est_d:.4f} ({len(per_d)} steps)", flush=True)
    assert len(per_d) == len(per) and abs(est_d - est) < 1e-9


def test_find_del_auto_demo_record_vs_oracle(mm):
    """findDelAuto in the delay demo's setting (demos/demo_autoDel.m:48-86:
    the reference's spiral R4 record with its 16 harmonic and 4 concomitant
    terms, 5268 samples, delJumpFact 3, no coarse search) reduced to 48 x 48
    and 4 coils, on data the FP64 oracle simulated with a 2 us delay: every
    step of the device search equals the oracle's (1e-4 us; the oracle's run
    is kept in tests/golden/finddel_demo_48x4.json by make_finddel_demo.py)
    and the estimate lands within 0.01 us of the applied delay."""
    import json
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_finddel_demo import problem
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "finddel_demo_48x4.json")))
    print("[case] findDelAuto demo record 48x48 C=4, delay 2 us", flush=True)
    P, y = problem(gold["N"], gold["C"])
    est, per = mm.finddel.findDelAuto(0.0, y, P["tdwell"], P["spha"], P["conc"], P["datatime"], P["grid"],
                                      P["b0"], P["maps"], 20, 3, 0)
    print(f"  device {est:.6f} steps {np.round(per, 4)}; oracle {gold['estimate_us']:.6f}", flush=True)
    assert len(per) == len(gold["steps_us"])
    assert np.max(np.abs(np.asarray(per) - np.asarray(gold["steps_us"]))) < 1e-4
    assert abs(est - gold["applied_delay_us"]) < 0.01

