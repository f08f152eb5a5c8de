"""Pins of the interpolated high-order oracle (oracle/highorder_interp.py)."""
import numpy as np
import pytest
from scipy.interpolate import PchipInterpolator

from oracle.highorder import HighOrderPlan
from oracle.highorder_interp import InterpPlan, gauss_window, pchip


def projection(Nproj, NperProj):
    """Radial trajectory of the reference's trajectory/projection.m (2-D)."""
    phi0 = np.linspace(0, np.pi, Nproj + 1)[:-1]
    r0 = np.linspace(-0.5, 0.5, NperProj + 1)[:-1]
    r0 = r0 + (r0[1] - r0[0]) / 2
    phi = np.tile(phi0, (NperProj, 1))
    r = np.tile(r0[:, None], (1, Nproj))
    return np.stack([(r * np.cos(phi)).reshape(-1, order="F"), (r * np.sin(phi)).reshape(-1, order="F")], axis=1)


def interp_problem(N0=64, seed=1, orient=1):
    """The setup of @sampHighOrder/tests.m:66-106: radial k-space on terms 2-3,
    random 4th-order polynomials in time on the other terms and on the
    concomitant ones, a 5th-order polynomial B0 map, a shifted grid."""
    rng = np.random.default_rng(seed)
    kloc = projection(N0, N0 - 4)
    t = np.arange(1, kloc.shape[0] + 1, dtype=float)
    x, y = np.meshgrid(np.arange(-N0 / 2, N0 / 2), np.arange(-N0 / 2, N0 / 2))
    dx, dy = 0.8, 0.95
    kloc[:, 0] = 2 * np.pi * kloc[:, 0] / np.max(np.abs(kloc[:, 0])) / 2 / dx
    kloc[:, 1] = 2 * np.pi * kloc[:, 1] / np.max(np.abs(kloc[:, 1])) / 2 / dy
    x, y = (x + 3) * dx, (y - 2) * dy
    scl = 0.5 * np.max(np.abs(x))
    r = rng.standard_normal
    b0 = (scl ** 5 * (r() * x + r() * y) + scl * (r() * x ** 4 + r() * x ** 3 * y + r() * x ** 2 * y ** 2
          + r() * x * y ** 3 + r() * y ** 4) + r() * x ** 5 + r() * x ** 4 * y + r() * x ** 3 * y ** 2
          + r() * x ** 2 * y ** 3 + r() * x * y ** 4 + r() * y ** 5)
    b0 = b0 / np.max(np.abs(b0)) * np.pi / t[-1]
    sclt = 0.5 * t[-1]
    hs = np.zeros((16, len(t)))
    hc = np.zeros((4, len(t)))
    hs[1:3] = kloc.T
    poly = lambda: r() + 2 * r() / sclt * t + r() / sclt ** 2 * t ** 2 + r() / sclt ** 3 * t ** 3 + r() / sclt ** 4 * t ** 4
    for n in [1] + list(range(4, 17)):
        p = poly()
        if 1 < n < 5:
            p = 0.05 * p / scl
        elif 1 < n < 10:
            p = 0.05 * p / scl ** 2
        elif n >= 10:
            p = 0.05 * p / scl ** 3
        hs[n - 1] += p
    for n in range(4):
        hc[n] = 0.05 * poly() / scl ** 2
    gx, gy = {1: (x, y), 2: (x.T, y.T), 3: (-x, -y), 4: (-x.T, -y.T)}[orient]
    grid = {"x": gx, "y": gy, "z": scl * np.ones_like(gx)}
    return b0, t, hs, hc, grid, rng


@pytest.mark.parametrize("orient", [1, 2, 3, 4])
def test_interp_reference_test(orient):
    """@sampHighOrder/tests.m:107-148 (svdThresh 0.01, subFact 5): adjoint
    identity 1e-8 and the distance to the direct operator below 1e-4 (norm
    over numel, the reference's measure); here also rel-L2 < 2e-2."""
    b0, t, hs, hc, grid, rng = interp_problem(orient=orient)
    xv = rng.standard_normal((64, 64)) * (np.hypot(*np.meshgrid(np.arange(64) - 32, np.arange(64) - 32)) < 25)
    S0 = HighOrderPlan(b0, t, hs, hc, grid)
    S1 = InterpPlan(b0, t, hs, hc, grid, svdThresh=0.01, subFact=5)
    yv = rng.standard_normal(len(t)) + 1j * rng.standard_normal(len(t))
    Sx, Sxb, Syb = S0.forward(xv), S1.forward(xv), S1.adjoint(yv)
    d1, d2 = np.vdot(xv.reshape(-1, order="F"), Syb.reshape(-1, order="F")), np.vdot(Sxb, yv)
    assert abs(d1 - d2) / min(abs(d1), abs(d2)) < 1e-8
    cost = np.linalg.norm(Sx - Sxb) / Sx.size
    rel = np.linalg.norm(Sx - Sxb) / np.linalg.norm(Sx)
    assert cost < 1e-4 and rel < 2e-2 and S1.Ns > 1


def test_interp_converges_to_direct():
    """With subFact 1 and a tiny svdThresh the interpolated operator is the
    direct one up to the W = 6 NUFFT's error (~1e-4); a wrong orientation
    sign, axis or phase ramp would leave an O(1) difference."""
    b0, t, hs, hc, grid, rng = interp_problem(N0=32, orient=2)
    xv = rng.standard_normal((32, 32))
    S0 = HighOrderPlan(b0, t, hs, hc, grid)
    S1 = InterpPlan(b0, t, hs, hc, grid, svdThresh=1e-9, subFact=1)
    rel = np.linalg.norm(S0.forward(xv) - S1.forward(xv)) / np.linalg.norm(S0.forward(xv))
    assert rel < 5e-4
    S1.phsShft = np.conj(S1.phsShft)
    assert np.linalg.norm(S0.forward(xv) - S1.forward(xv)) / np.linalg.norm(S0.forward(xv)) > 0.1


def test_pchip_and_window():
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.random(30) + 0.2)
    y = rng.standard_normal((30, 2))
    xq = np.linspace(x[0], x[-1], 400)
    assert np.max(np.abs(pchip(x, y, xq) - PchipInterpolator(x, y)(xq))) < 1e-12
    mono = np.cumsum(rng.random(30))
    v = pchip(x, mono, xq)
    assert np.all(np.diff(v) >= -1e-12)   # shape preserving
    w = gauss_window(8, 0.5)
    sig = 8 * 0.5 / 2.355
    assert np.allclose(w, np.exp(-0.5 * (np.arange(-4, 4) / sig) ** 2))
