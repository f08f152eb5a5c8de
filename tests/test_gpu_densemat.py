"""mrSampFuncMat (operators/mrSampFuncMat.m) on the device against the oracle."""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.densemat import mrSampFuncMat as o_mrSampFuncMat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N,K,C,masked", [(24, 900, 4, True), (32, 1500, 8, False)])
def test_dense_normal_vs_oracle(mm, N, K, C, masked):
    """AtA (single-precision dense encoding matrix, Gram on rocBLAS, coil
    maps folded in) and Aty over an image mask against the oracle (relative
    Frobenius 1e-5; Aty 1e-5); AtA is Hermitian and applied to a vector it
    equals the device normal operator scaled by V/Vm."""
    print(f"[case] mrSampFuncMat N={N} K={K} C={C} masked={masked}", flush=True)
    P = mm.workloads.highorder_problem(seed=5, N=N, K=K, C=C, L=16)
    rng = np.random.default_rng(5)
    y = rng.standard_normal((K, C)) + 1j * rng.standard_normal((K, C))
    X = P["grid"]["x"]
    mask = (np.hypot(X, P["grid"]["y"]) < 0.08) if masked else None
    AtA, Aty = mm.mrSampFuncMat(y, P["maps"], P["b0"], P["t"], P["h"], None, P["grid"], mask)
    Ao, Atyo = o_mrSampFuncMat(y, P["maps"], P["b0"], P["t"], P["h"], None, P["grid"], mask)
    e, ey = _rel(AtA, Ao), _rel(Aty, Atyo)
    herm = np.linalg.norm(AtA - AtA.conj().T) / np.linalg.norm(AtA)
    print(f"  AtA {AtA.shape} rel {e:.2e}; Aty {ey:.2e}; hermitian {herm:.1e}", flush=True)
    assert AtA.shape == Ao.shape and e < 1e-5 and ey < 1e-5 and herm < 1e-6
    At2, _ = mm.mrSampFuncMat(y, P["maps"], P["b0"], P["t"], P["h"], None, P["grid"], mask, on_device=True)
    assert np.array_equal(At2.cpu().numpy().T, AtA)
