"""Pins the FP64 high-order (expanded encoding) oracle and its basis functions."""
import numpy as np
import pytest

from oracle.basis import HARM_ORDER, basis_conc, basis_harm
from oracle.fftnc import fftnc
from oracle.highorder import HighOrderPlan


def _rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# -- basis functions: homogeneity, harmonicity, independence --------------------
@pytest.mark.parametrize("n", list(range(1, 37)))
def test_harmonic_homogeneous(n):
    rng = np.random.default_rng(n)
    x, y, z = rng.standard_normal((3, 50))
    lam = 1.7
    b1 = basis_harm(x, y, z, n)
    b2 = basis_harm(lam * x, lam * y, lam * z, n)
    assert np.allclose(b2, lam ** HARM_ORDER[n - 1] * b1, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("n", list(range(5, 37)))
def test_harmonic_laplacian_zero(n):
    """Solid harmonics of order >= 2 satisfy Laplace's equation; a wrong
    coefficient in a transcribed polynomial breaks it."""
    rng = np.random.default_rng(100 + n)
    p = rng.standard_normal((3, 20))
    h = 1e-2
    lap = -6 * basis_harm(*p, n)
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        lap = lap + basis_harm(*(p + e), n) + basis_harm(*(p - e), n)
    lap /= h * h
    scale = np.max(np.abs(basis_harm(*p, n))) / np.max(np.abs(p)) ** 2 + 1
    assert np.max(np.abs(lap)) < 1e-4 * scale * 100


def test_basis_independent():
    rng = np.random.default_rng(0)
    x, y, z = rng.standard_normal((3, 400))
    B = np.stack([basis_harm(x, y, z, n) for n in range(1, 37)], axis=1)
    assert np.linalg.matrix_rank(B) == 36
    Bc = np.stack([basis_conc(x, y, z, n) for n in range(1, 5)], axis=1)
    # xz = b8, yz = b6 and 2 z^2 - (x^2+y^2) = b7, so the four concomitant terms
    # add exactly one direction (r^2, not harmonic) to the 16 harmonics
    assert np.linalg.matrix_rank(np.concatenate([B[:, :16], Bc], axis=1)) == 17
    assert np.allclose(Bc[:, 2], B[:, 7]) and np.allclose(Bc[:, 3], B[:, 5])
    assert np.allclose(2 * Bc[:, 0] - Bc[:, 1], B[:, 6])
    assert np.allclose(Bc[:, 0] + Bc[:, 1], x * x + y * y + z * z)


def _random_plan(rng, imN, imk, with_conc=True, mask=None):
    b0 = rng.standard_normal(imN) + rng.random()
    t = rng.standard_normal(imk) / np.sqrt(np.prod(imk))
    hs = rng.standard_normal([16] + list(imk))
    hc = rng.standard_normal([4] + list(imk)) if with_conc else None
    grid = {"x": rng.standard_normal(imN), "y": rng.standard_normal(imN), "z": rng.standard_normal(imN)}
    return HighOrderPlan(b0, t, hs, hc, grid, mask)


# -- brute force on tiny sizes ---------------------------------------------------
def test_brute_force_tiny():
    rng = np.random.default_rng(1)
    S = _random_plan(rng, (3, 2), (5,), mask=None)
    x = _rand_c(rng, (3, 2))
    y = S.forward(x)
    xv = x.reshape(-1, order="F")
    for k in range(5):
        acc = 0
        for v in range(6):
            ph = 0.0
            for n in range(1, 17):
                ph += basis_harm(S.x[v], S.y[v], S.z[v], n) * S.phs_spha[n - 1, k]
            for n in range(1, 5):
                ph += basis_conc(S.x[v], S.y[v], S.z[v], n) * S.phs_conc[n - 1, k]
            ph += S.b0[v] * S.sampTimes.reshape(-1)[k]
            acc += np.exp(1j * ph) * xv[v]
        assert abs(acc / np.sqrt(6) - y[k]) < 1e-12


def test_mask_semantics():
    """b0mask multiplies harmonics n > 4, the conc terms and B0 (sampHighOrder.m:564-585)."""
    rng = np.random.default_rng(2)
    mask = (rng.random((4, 3)) > 0.5).astype(float)
    S = _random_plan(rng, (4, 3), (7,), mask=mask)
    ph = S.phase()
    m = mask.reshape(-1, order="F")
    lin = sum(basis_harm(S.x, S.y, S.z, n)[:, None] * S.phs_spha[n - 1][None] for n in range(1, 5))
    assert np.allclose(ph[m == 0], lin[m == 0], atol=1e-12)


# -- @sampHighOrder/tests.m:10-40 adjoint and segmented == direct ---------------
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_adjoint_and_segmented(nd):
    rng = np.random.default_rng(20 + nd)
    N0 = 16
    imN = tuple([int(round(N0 ** (2 / nd)))] * nd)
    imk = tuple([int(round(N0 ** (1 / nd)))] * nd)
    S = _random_plan(rng, imN, imk)
    x = rng.standard_normal(imN + (2,)) + rng.standard_normal()
    y = rng.standard_normal(imk + (2,)) + rng.standard_normal()
    Sx, Sy = S.forward(x), S.adjoint(y)
    d1 = np.vdot(x.reshape(-1, order="F"), Sy.reshape(-1, order="F"))
    d2 = np.vdot(Sx.reshape(-1, order="F"), y.reshape(-1, order="F"))
    assert abs(d1 - d2) / min(abs(d1), abs(d2)) < 1e-8
    Sx2 = S.forward_segmented(x, 3).reshape(-1, order="F")
    Sy2 = S.adjoint_segmented(y, 3).reshape(-1, order="F")
    assert np.linalg.norm(Sx.reshape(-1, order="F") - Sx2) + \
        np.linalg.norm(Sy.reshape(-1, order="F") - Sy2) < 1e-8


def test_phidiv_segmented():
    rng = np.random.default_rng(30)
    S = _random_plan(rng, (8, 8), (16,))
    dh = np.diff(np.concatenate([np.zeros((16, 1)), S.phs_spha], axis=1), axis=1)
    dc = np.diff(np.concatenate([np.zeros((4, 1)), S.phs_conc], axis=1), axis=1)
    S.set_phi_div(dh, dc)
    x = _rand_c(rng, (8, 8, 2))
    a = S.forward(x).reshape(-1, order="F")
    b = S.forward_segmented(x, 2).reshape(-1, order="F")
    assert np.linalg.norm(a - b) < 1e-8


# -- @sampHighOrder/tests.m:265-283 known answer vs Fourier transform ------------
def test_kat_vs_fft():
    N = 16
    gx, gy = np.meshgrid(np.arange(-N / 2, N / 2), np.arange(-N / 2, N / 2))   # meshgrid
    grid = {"x": gx, "y": gy, "z": np.zeros_like(gx)}
    hs = np.zeros((16, N, N))
    hc = np.zeros((4, N, N))
    hs[2] = np.pi * 2 / N * np.tile(np.arange(-N / 2, N / 2), N).reshape(N, N, order="F")
    hs[1] = np.pi * 2 / N * np.tile(np.arange(-N / 2, N / 2), (N, 1))
    t = np.arange(N * N, dtype=float).reshape(N, N, order="F")
    S = HighOrderPlan(np.zeros_like(gx), t, hs, hc, grid)
    rng = np.random.default_rng(4)
    im0 = rng.random((N, N))
    k0 = S.forward(im0)
    k1 = fftnc(im0, 2)
    cost = np.sqrt(np.sum((np.abs(k0) - np.abs(k1)) ** 2)) / k0.size
    assert cost < 1e-10
    im1 = S.adjoint(k0)
    assert np.sqrt(np.sum(np.abs(im1 - im0) ** 2)) / im0.size < 1e-10
