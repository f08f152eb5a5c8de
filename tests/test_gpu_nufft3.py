"""GPU parity of the 3-D NUFFT (nufftOp with a 3-D imN) against the FP64
oracle: forward and adjoint at the per-application bound 1e-5 (even and odd
sides, W 4/6/8, os 1.5/2, dcf, 1 to 20 transforms: partly filled chunks,
grid sides up to 486 along each axis),
the reference's on-grid known answer in 3-D (@nufftOp/tests.m:90-101), and a
3-D SENSE cgne against the FP64 loop.  The oracle's 3-D operator is pinned by
the brute-force NUDFT (tests/test_oracle_nufft.py::test_vs_nudft) and the
odd-side closed form.
"""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.fftnc import fftnc
from oracle.mrsamp import mrSampFunc as o_mrSampFunc
from oracle.nufft import NufftPlan, ToepPlan
from oracle.rcvr import RcvrPlan
from oracle.solvers import cgne as o_cgne

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a).reshape(-1, order="F")
    b = np.asarray(b).reshape(-1, order="F")
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


CASES_3D = [
    # imN, M, kwidth, os, dcf, nrep
    ((12, 16, 20), 2000, 6, 1.5, False, 1),
    ((16, 16, 16), 3000, 4, 1.5, True, 3),
    ((15, 17, 14), 2500, 4, 1.5, False, 2),
    ((24, 20, 18), 4000, 8, 2.0, True, 1),
    ((20, 20, 20), 3000, 4, 1.5, False, 16),
    ((20, 18, 9), 2000, 6, 1.375, True, 20),
    # sides above what the default FFT kernels hold in one LDS buffer: the
    # doubled-register-cap kernels along each axis, 16 and 8 transforms, and
    # a side of 486 points where they run two buffers
    ((176, 16, 12), 3000, 4, 1.5, False, 16),
    ((16, 176, 12), 3000, 4, 1.5, True, 16),
    ((12, 10, 200), 2500, 6, 1.5, False, 3),
    ((200, 12, 10), 2500, 4, 1.5, True, 3),
    ((324, 8, 8), 2000, 4, 1.5, False, 16),
    ((12, 176, 10), 2500, 4, 1.5, False, 8),
]


@pytest.mark.parametrize("imN,M,W,os_,use_dcf,nrep", CASES_3D)
def test_nufft3_vs_oracle(mm, imN, M, W, os_, use_dcf, nrep):
    print(f"[case] 3-D nufft imN={imN} M={M} W={W} os={os_} dcf={use_dcf} nrep={nrep}", flush=True)
    rng = np.random.default_rng(90 + sum(imN) + nrep)
    kloc = rng.random((M, 3)) - 0.5
    dcf = rng.random(M) + 0.5 if use_dcf else None
    So = NufftPlan(list(imN), kloc, dcf, os_, W)
    Sg = mm.nufftOp(list(imN), kloc, dcf, 1, os_, W)
    x = crand(rng, imN + (nrep,))
    y = crand(rng, (M, nrep))
    e_f = rel(Sg * x, So.forward(x))
    e_a = rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


@pytest.mark.parametrize("nrep", [1, 8, 16])
def test_nufft3_dense_centre_vs_oracle(mm, nrep):
    """Samples concentrated at the k-space centre (rows of several hundred
    entries: the gridding's one-wave-per-cell path, 64 / 8 / 4 entry phases)
    against the oracle in both directions (1e-5)."""
    print(f"[case] 3-D nufft dense centre nrep={nrep}", flush=True)
    rng = np.random.default_rng(120 + nrep)
    imN, M = (12, 16, 20), 20000
    kloc = np.clip(rng.normal(0.0, 0.04, (M, 3)), -0.5, 0.4999)
    So = NufftPlan(list(imN), kloc, None, 1.5, 6)
    Sg = mm.nufftOp(list(imN), kloc, None, 1, 1.5, 6)
    x = crand(rng, imN + ((nrep,) if nrep > 1 else ()))
    y = crand(rng, (M, nrep) if nrep > 1 else (M,))
    e_f = rel(Sg * x, So.forward(x))
    e_a = rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_nufft3_on_grid_known_answer(mm):
    """@nufftOp/tests.m:90-101 in 3-D: on-grid kloc makes the forward a
    unitary centred inverse DFT, fftnc(S*im) == im (1e-4, L2 / numel)."""
    print("[case] 3-D nufft on-grid KAT 12x16x20", flush=True)
    imN = [12, 16, 20]
    axes = [(np.arange(n) - n / 2) / n for n in imN]
    g = np.meshgrid(*axes, indexing="ij")
    kloc = np.stack([a.reshape(-1, order="F") for a in g], axis=1)
    Sg = mm.nufftOp(imN, kloc, None, 1, 1.5, 6)
    im = np.random.default_rng(4).random(imN)
    k = (Sg * im).reshape(imN, order="F")
    err = np.sqrt(np.sum(np.abs(fftnc(k, 3) - im) ** 2)) / im.size
    print(f"  KAT err {err:.2e}", flush=True)
    assert err < 1e-4


def test_nufft3_sense_cgne_matches_oracle(mm):
    """A 3-D SENSE operator (rcvrOp over the 3-D NUFFT, mrSampFunc.m:97-107,
    :126-144) and cgne (cgne.m:213-336) for 8 iterations against the FP64
    loop: residual history within 1e-3, iterate NRMSE < 1e-4."""
    print("[case] 3-D sense cgne 16x16x12 C=4", flush=True)
    rng = np.random.default_rng(95)
    imN, M, C = (16, 16, 12), 2500, 4
    kl = rng.random((M, 3)) - 0.5
    maps = crand(rng, imN + (C,))
    So, Ro = NufftPlan(list(imN), kl, None, 1.5, 4), RcvrPlan(maps)
    Sg, Rg = mm.nufftOp(list(imN), kl, None, 1, 1.5, 4), mm.rcvrOp(maps)
    E = mm.senseOp(Sg, Rg)
    x = crand(rng, imN)
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    A = lambda v, tr: o_mrSampFunc(v, tr, So, Ro)
    y = A(crand(rng, imN), "notransp")
    e_a = rel(E.H * y, A(y, "transp"))
    print(f"  sense fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL
    opt = {"resThresh": None, "xChangeThresh": None}
    x0 = np.zeros(imN, complex)
    nit = 8
    xo, ro, *_ = o_cgne(A, y, x0, nit, opt)
    xg, rg, *_ = mm.cgne(E, y, x0, nit, opt)
    ro = np.real(ro)
    d = np.abs(rg - ro) / ro
    nrmse = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"  resSq rel diff max {d.max():.2e}; x nrmse {nrmse:.2e}", flush=True)
    assert np.all(d < 1e-3)
    assert nrmse < 1e-4


@pytest.mark.parametrize("imN,nrep", [((12, 10, 8), 1), ((9, 10, 7), 2)])
def test_nufft3_toeplitz_vs_oracle(mm, imN, nrep):
    """prepToep of a 3-D nufftOp (nufftOp.m:410-432, applied :348-352): both
    directions against the oracle's Toeplitz form (1e-5); for even sides also
    within the reference's own bound (rms < 1e-3) of the accurate normal
    operator S0'*(S0*x) (os 2, W 8, @nufftOp/tests.m:104-124)."""
    print(f"[case] 3-D Toeplitz imN={imN} nrep={nrep}", flush=True)
    rng = np.random.default_rng(97 + sum(imN))
    kloc = rng.random((int(0.5 * np.prod(imN)), 3)) - 0.5
    dcf = rng.random(kloc.shape[0]) + 0.5
    T = mm.nufftOp(list(imN), kloc, dcf, 1, 1.5, 6).prepToep()
    x = crand(rng, imN + ((nrep,) if nrep > 1 else ()))
    want = ToepPlan(list(imN), kloc, dcf, 1.5, 6).apply(x)
    e, e_h = rel(np.asarray(T * x), want), rel(np.asarray(T.H * x), want)
    print(f"  vs oracle {e:.2e} / adjoint {e_h:.2e}", flush=True)
    assert e < TOL and e_h < TOL
    if all(n % 2 == 0 for n in imN):
        S0 = NufftPlan(list(imN), kloc, dcf, 2, 8)
        rms = np.sqrt(np.mean(np.abs(np.asarray(T * x) - S0.adjoint(S0.forward(x))) ** 2))
        print(f"  rms vs S0'S0 {rms:.2e}", flush=True)
        assert rms < 1e-3


def test_nufft3_toeplitz_sense(mm):
    """mrSampFunc on a 3-D prepToep'd nufftOp: forward T(R x) (coil volumes),
    adjoint R'(T y), against the oracle; cgne refuses the operator."""
    print("[case] 3-D Toeplitz SENSE 10x8x8 C=3", flush=True)
    rng = np.random.default_rng(99)
    imN, C = (10, 8, 8), 3
    kl = rng.random((300, 3)) - 0.5
    maps = crand(rng, imN + (C,))
    T = mm.nufftOp(list(imN), kl, None, 1, 1.5, 4).prepToep()
    E = mm.senseOp(T, mm.rcvrOp(maps))
    Tp = ToepPlan(list(imN), kl, None, 1.5, 4)
    x = crand(rng, imN)
    y = crand(rng, imN + (C,))
    want_f = np.stack([Tp.apply(maps[..., c] * x) for c in range(C)], axis=-1)
    want_a = sum(np.conj(maps[..., c]) * Tp.apply(y[..., c]) for c in range(C))
    e_f, e_a = rel(E * x, want_f), rel(E.H * y, want_a)
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL
    with pytest.raises(NotImplementedError):
        mm.cgne(E, y, None, 3, {})


def test_nufft3_rejects_unsupported(mm):
    """Grid sides above 512 and multi-volume 3-D plans raise, not fault."""
    kl = np.zeros((4, 3))
    with pytest.raises(NotImplementedError):
        mm.nufftOp([400, 8, 8], kl, None, 1, 1.5, 4)
    with pytest.raises(NotImplementedError):
        mm.nufftOp.batch([([8, 8, 8], kl, None), ([8, 8, 8], kl, None)], 1.5, 4)
