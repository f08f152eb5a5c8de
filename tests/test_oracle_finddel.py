"""Pins of the trajectory-resampling and delay-search oracle (oracle/finddel.py)."""
import os

import numpy as np
import pytest
from scipy.interpolate import Akima1DInterpolator

from oracle.finddel import _conv_same, findDelAuto, interpTrajTime, makima

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "traj_spiral_R4.npz")


def test_makima_matches_scipy_makima():
    """SciPy's independent modified-Akima implementation on non-uniform data."""
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.random(40) + 0.1)
    y = rng.standard_normal((40, 3))
    xq = np.linspace(x[0], x[-1], 500)
    want = Akima1DInterpolator(x, y, method="makima")(xq)
    assert np.max(np.abs(makima(x, y, xq) - want)) < 1e-12


def test_makima_properties():
    """Exact on linear data; no overshoot across a flat step (the reason
    makima exists); NaN outside the data range; a wrong weight (Akima's
    original, without the |d1 + d2|/2 terms) differs on flat-slope data."""
    x = np.arange(10.0)
    xq = np.linspace(0, 9, 97)
    assert np.allclose(makima(x, 3 - 2 * x, xq), 3 - 2 * xq, atol=1e-13)
    y = np.array([0, 0, 0, 0, 1, 1, 1, 1, 1, 1.0])
    v = makima(x, y, xq)
    assert v.min() >= -1e-15 and v.max() <= 1 + 1e-15
    assert np.all(v[xq <= 3] == 0) and np.all(np.abs(v[xq >= 5] - 1) < 1e-15)
    out = makima(x, y, np.array([-0.5, 9.5]))
    assert np.all(np.isnan(out))
    # classic Akima on y = [1 1 1 2 2 2 ...] style data gives a different curve
    yy = np.array([1, 2, 1, 2, 1, 2, 1, 2, 1, 2.0])
    ak = Akima1DInterpolator(x, yy)(xq)   # method="akima"
    assert np.max(np.abs(makima(x, yy, xq) - ak)) > 1e-3


def test_interp_traj_time_shift_closed_form():
    """A linear ramp sampled every tdwell from time -delTime, read at datatime:
    the value at datatime + delTime (interpTrajTime.m:21-27)."""
    L, tdw, dl = 300, 1.0, 2.37
    tt = tdw * np.arange(L)
    ramp = np.stack([5 + 0.3 * tt, -1 + 2 * tt], axis=1)
    dt = np.linspace(3, 250, 77)
    a, b = interpTrajTime(ramp, ramp[:, :1], tdw, dl, dt)
    assert np.allclose(a[:, 0], 5 + 0.3 * (dt + dl), atol=1e-11)
    assert np.allclose(a[:, 1], -1 + 2 * (dt + dl), atol=1e-11)
    assert np.allclose(b[:, 0], a[:, 0])
    with pytest.raises(ValueError):
        interpTrajTime(None, None, 1, 0, dt)


def test_traj_fixture_and_derivative_kernel():
    """The converted reference trajectory (MAT-v5 data_traj_spiral_R4.mat):
    shapes and dwell; the smoothed derivative kernel of findDelAuto.m:73-75
    is exact on a quadratic away from the ends."""
    d = np.load(GOLD)
    assert d["phs_spha"].shape == (15000, 16) and d["phs_conc"].shape == (15000, 4)
    assert d["datatime"].shape == (1, 5268) and float(d["tdwelltraj"].reshape(-1)[0]) == 1.0
    t = np.arange(50.0)
    q = (0.5 * t ** 2)[:, None]
    der = _conv_same(q, np.array([1 / 8, 1 / 4, 0, -1 / 4, -1 / 8]))
    # (q[r+2] + 2 q[r+1] - 2 q[r-1] - q[r-2]) / 8 = t for q = t^2 / 2
    assert np.allclose(der[2:-2, 0], t[2:-2], rtol=0, atol=1e-12)
    assert not np.allclose(der[:2, 0], t[:2])   # zero padding at the ends ('same')


def _delay_problem(N=24, nsamp=2000, C=3, delay=2.0, seed=7):
    """A small high-order problem on the reference spiral: data simulated with
    the trajectory delayed by `delay` us (dataStartTime 50 us)."""
    from oracle.highorder import HighOrderPlan
    from oracle.mrsamp import mrSampFunc
    from oracle.rcvr import RcvrPlan
    d = np.load(GOLD)
    rng = np.random.default_rng(seed)
    dt = d["datatime"].reshape(-1)[:nsamp] + 50.0
    gx = (np.arange(N) - N // 2) * (0.2 / N)
    X, Y = np.meshgrid(gx, gx, indexing="ij")
    grid = {"x": X, "y": Y, "z": np.full_like(X, 0.01)}
    b0 = 2 * np.pi * 30e-6 * np.outer(np.cos(np.linspace(0, 2, N)), np.sin(np.linspace(0, 3, N)))   # rad/us
    r = np.hypot(X, Y) / 0.1
    maps = np.stack([np.exp(-((X - 0.08 * np.cos(a)) ** 2 + (Y - 0.08 * np.sin(a)) ** 2) / 0.01) *
                     np.exp(1j * a) for a in 2 * np.pi * np.arange(C) / C], axis=-1)
    im = (r < 0.7).astype(float) + 0.5 * (np.hypot(X - 0.02, Y) < 0.03)
    a, b = interpTrajTime(d["phs_spha"], d["phs_conc"], 1.0, delay, dt)
    S = HighOrderPlan(b0, dt, a.T, b.T, grid)
    y = mrSampFunc(im, "notransp", S, RcvrPlan(maps))
    y = y + 1e-4 * np.max(np.abs(y)) * (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))
    return d, dt, grid, b0, maps, y


def test_find_del_auto_recovers_delay():
    """findDelAuto.m on data simulated with a 2 us trajectory delay, started
    at 0, 2000 samples (5 ms of the spiral), 12 cgne iterations per image:
    the estimate converges to the applied delay (within 0.02 us; ~0.002
    measured); an update without the delJumpFact damping or with the wrong
    sign of the derivative kernel would not converge there."""
    d, dt, grid, b0, maps, y = _delay_problem()
    est, per_it = findDelAuto(0.0, y, 1.0, d["phs_spha"], d["phs_conc"], dt, grid, b0, maps, 12)
    assert abs(est - 2.0) < 0.02 and len(per_it) >= 2


def test_load_traj_mat_round_trip(tmp_path):
    """load_traj_mat reads a MAT-v5 file (written here by scipy.io.savemat)
    with the fixture's variable names and shapes."""
    import importlib
    import scipy.io as sio
    fd = importlib.import_module("uwobrainhack2026-matmri_amd.finddel")
    d = np.load(GOLD)
    p = tmp_path / "traj.mat"
    sio.savemat(p, {"phs_spha": d["phs_spha"][:100], "phs_conc": d["phs_conc"][:100],
                    "datatime": d["datatime"][:, :40], "tdwelltraj": np.uint8(1)})
    t = fd.load_traj_mat(str(p))
    assert np.array_equal(t["phs_spha"], d["phs_spha"][:100]) and np.array_equal(t["phs_conc"], d["phs_conc"][:100])
    assert t["datatime"].shape == (40,) and t["tdwelltraj"] == 1.0


def test_delay_demo_problem_conventions():
    """workloads.delay_demo_problem (the setting of demos/demo_autoDel.m:12-59
    on synthetic images): the record's data times shifted to the spiral
    case's 50 us start, a 1.35 mm grid, B0 within +-50 Hz in rad/us, and the
    record's coefficients unchanged."""
    import importlib
    w = importlib.import_module("uwobrainhack2026-matmri_amd.workloads")
    d = np.load(GOLD)
    P = w.delay_demo_problem(d, N=32, C=4)
    assert np.array_equal(P["datatime"], d["datatime"].reshape(-1) + 50.0)
    assert P["tdwell"] == 1.0 and P["delay"] == 2.0
    assert np.array_equal(P["spha"], d["phs_spha"]) and np.array_equal(P["conc"], d["phs_conc"])
    gx = P["grid"]["x"]
    assert gx.shape == (32, 32) and abs((gx[1, 0] - gx[0, 0]) - 0.00135) < 1e-15
    assert np.max(np.abs(P["b0"])) <= 2 * np.pi * 50e-6 * (1 + 1e-12)
    assert P["maps"].shape == (32, 32, 4) and P["phantom"].shape == (32, 32)
    # the record covers every data time for delays within the search's reach
    assert P["datatime"][-1] + 5.0 < (P["spha"].shape[0] - 1) * P["tdwell"]
