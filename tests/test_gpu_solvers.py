"""bfista / powermethod with device-resident vectors against the FP64 oracle
(iterativeSolvers/bfista.m:127-231, powermethod.m:38-67)."""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.mrsamp import mrSampFunc as o_mrSampFunc
from oracle.nufft import NufftPlan
from oracle.rcvr import RcvrPlan
from oracle.solvers import bfista as o_bfista
from oracle.solvers import powermethod as o_powermethod
from oracle.solvers import softthresh as o_softthresh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


def _rel(a, b):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _sense(mm, N=64, M=6000, C=4, seed=31):
    rng = np.random.default_rng(seed)
    kl = mm.workloads.spiral(M, N // 2)
    maps = mm.workloads.gaussian_coils(N, N, C, rng)
    ph = mm.workloads.ellipse_phantom(N, N, rng)
    So, Ro = NufftPlan([N, N], kl, None, 1.5, 4), RcvrPlan(maps)
    y = mm.workloads.add_noise(o_mrSampFunc(ph, "notransp", So, Ro), rng)
    E = mm.senseOp(mm.nufftOp([N, N], kl, None, 1, 1.5, 4), mm.rcvrOp(maps))
    return (lambda x, tr: o_mrSampFunc(x, tr, So, Ro)), E, y, rng


def test_bfista_identity_closed_form(mm):
    """A = I, R = I: argmin ||x - b||^2 + lam ||x||_1 = soft(b, lam / 2)."""
    import torch
    print("[case] bfista A = R = I closed form", flush=True)
    rng = np.random.default_rng(9)
    b = rng.standard_normal((40, 30)) + 1j * rng.standard_normal((40, 30))
    eye = lambda v, tr: v
    x, *_ = mm.bfista(eye, b, eye, 0.8, np.zeros_like(b), 300, {"maxEig": 1.0, "resThresh": 0.0})
    want = o_softthresh(b, 0.4)
    e = _rel(x, want)
    print(f"  rel-L2 vs soft(b, lam/2): {e:.2e}", flush=True)
    assert isinstance(x, np.ndarray) and x.shape == b.shape
    assert e < 1e-6
    xt, *_ = mm.bfista(eye, torch.as_tensor(b.T.astype(np.complex64), device="cuda"), eye, 0.8, None, 300,
                       {"maxEig": 1.0, "resThresh": 0.0})
    assert xt.is_cuda and _rel(xt.cpu().numpy().T, want) < 1e-6


@pytest.mark.parametrize("reg", ["identity", "dft", "kmask"])
def test_bfista_sense_vs_oracle(mm, reg):
    """20 bfista iterations on a 64x64 / 4-coil spiral SENSE problem with
    R = I, R = the unitary 2-D DFT, or R = I with a k-space mask (kMask,
    bfista.m:161-169): iterate, resSq and |Rx|_1 histories against the FP64
    oracle (rel-L2 <= 1e-4), maxEig from the device powermethod."""
    import torch
    print(f"[case] bfista SENSE 64x64 C=4, R={reg}, 20 iterations", flush=True)
    A, E, y, rng = _sense(mm)
    N = 64
    opt = {"resThresh": 0.0}
    if reg == "dft":
        Ro = lambda v, tr: np.fft.fft2(v, norm="ortho") if tr == "notransp" else np.fft.ifft2(v, norm="ortho")
        Rd = lambda v, tr: torch.fft.fft2(v, norm="ortho") if tr == "notransp" else torch.fft.ifft2(v, norm="ortho")
    else:
        Ro = Rd = (lambda v, tr: v)
    if reg == "kmask":
        kx = np.fft.fftfreq(N)
        opt["kMask"] = (np.hypot(kx[:, None], kx[None, :]) < 0.35).astype(float)
    x0 = np.zeros((N, N), complex)
    ev = mm.powermethod(E, rng.standard_normal((N, N)))
    opt["maxEig"] = float(np.real(ev))
    lam = 2e-3 * np.max(np.abs(A(y, "transp")))
    xd, rd, Rxd, td, _ = mm.bfista(E, y, Rd, lam, x0, 20, opt)
    xo, ro, Rxo, to, _ = o_bfista(A, y, Ro, lam, x0, 20, opt)
    e, er, eR = _rel(xd, xo), _rel(rd, ro), _rel(Rxd, Rxo)
    print(f"  maxEig {opt['maxEig']:.4e}; x {e:.2e} resSq {er:.2e} Rx {eR:.2e}; iterations {len(rd)}/{len(ro)}",
          flush=True)
    assert len(rd) == len(ro) == 20
    assert e < 1e-4 and er < 1e-4 and eR < 1e-4
    if reg == "kmask":
        assert np.max(np.abs(np.fft.fft2(xd)[opt["kMask"] == 0])) < 1e-4 * np.max(np.abs(np.fft.fft2(xd)))


def test_powermethod_vs_oracle(mm):
    """Largest eigenvalue of E'E of the SENSE operator (powermethod.m:38-67)
    from the same start vector: device vs FP64 oracle within 1e-3."""
    print("[case] powermethod SENSE 64x64 C=4", flush=True)
    A, E, y, rng = _sense(mm, seed=32)
    x0 = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    ed = mm.powermethod(E, x0)
    eo = o_powermethod(A, x0)
    e = abs(ed - eo) / abs(eo)
    print(f"  device {np.real(ed):.6e} oracle {np.real(eo):.6e} rel {e:.2e}", flush=True)
    assert e < 1e-3
