"""Pins of oracle/densemat.py (mrSampFuncMat.m)."""
import numpy as np

from oracle.densemat import mrSampFuncMat
from oracle.highorder import HighOrderPlan
from oracle.mrsamp import mrSampFunc
from oracle.rcvr import RcvrPlan


def _problem(N=10, K=150, C=3, seed=2):
    rng = np.random.default_rng(seed)
    gx = (np.arange(N) - N / 2) * 0.01
    X, Y = np.meshgrid(gx, gx, indexing="ij")
    grid = {"x": X, "y": Y, "z": np.full_like(X, 0.01)}
    b0 = rng.standard_normal((N, N)) * 50
    t = np.arange(K) * 1e-5
    hs = rng.standard_normal((9, K)) * np.array([2.0, 300, 300, 300, 3e3, 3e3, 3e3, 3e3, 3e3])[:, None]
    maps = rng.standard_normal((N, N, C)) + 1j * rng.standard_normal((N, N, C))
    y = rng.standard_normal((K, C)) + 1j * rng.standard_normal((K, C))
    return y, maps, b0, t, hs, grid, rng


def test_dense_normal_is_the_normal_operator():
    y, maps, b0, t, hs, grid, rng = _problem()
    AtA, Aty = mrSampFuncMat(y, maps, b0, t, hs, None, grid, None, single_phase=False)
    S, R = HighOrderPlan(b0, t, hs, None, grid), RcvrPlan(maps)
    x = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    nx = mrSampFunc(mrSampFunc(x.reshape(10, 10, order="F"), "notransp", S, R), "transp", S, R)
    assert np.linalg.norm(AtA @ x - nx.reshape(-1, order="F")) / np.linalg.norm(nx) < 1e-10
    assert np.allclose(AtA, AtA.conj().T, atol=1e-12 * np.abs(AtA).max())
    assert np.allclose(Aty, mrSampFunc(y, "transp", S, R).reshape(-1, order="F"))
    # the single rounding of the phase moves it by the phase rounding only
    A1, _ = mrSampFuncMat(y, maps, b0, t, hs, None, grid, None)
    d = np.linalg.norm(A1 - AtA) / np.linalg.norm(AtA)
    assert 1e-10 < d < 1e-5


def test_dense_normal_mask_scale():
    """With a partial mask the reference scales S by 1/sqrt(numel(Aty)) =
    1/sqrt(Vm) (mrSampFuncMat.m:36), the operator by 1/sqrt(V): AtA is
    V/Vm times the normal matrix restricted to the mask."""
    y, maps, b0, t, hs, grid, rng = _problem()
    mask = rng.random((10, 10)) > 0.3
    AtA, Aty = mrSampFuncMat(y, maps, b0, t, hs, None, grid, mask, single_phase=False)
    Full, Atyf = mrSampFuncMat(y, maps, b0, t, hs, None, grid, None, single_phase=False)
    m = mask.reshape(-1, order="F")
    assert np.allclose(AtA, Full[np.ix_(m, m)] * (100 / m.sum()), rtol=1e-10, atol=1e-12 * np.abs(AtA).max())
    assert np.allclose(Aty, Atyf[m])
