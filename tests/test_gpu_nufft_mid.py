"""GPU parity of the 2-D NUFFT on grid sides of 257..540 points: the row and
column FFT kernels there hold twice the default values per thread so that
their passes fit one LDS buffer (option "nu_fft_mid", on by default), and
take the radix-16/12 factorization wherever it saves a pass (336 = 16 x 3 x
7).  Every transform count the plans use (1, 2, 4, 8, 16 per chunk), odd
sides, unequal grid sides, the sides where even those kernels need two LDS
buffers (420, 500) and the Toeplitz form are held to the FP64 oracle at the
per-application bound 1e-5, with the option on and off; grid sides above
620 points (the large-grid kernels) too.
"""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.mrsamp import mrSampFunc as o_mrSampFunc
from oracle.nufft import NufftPlan, ToepPlan
from oracle.rcvr import RcvrPlan

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


PLAIN_CASES = [
    # imN (grid sides at os 1.5 in the comment), M, kwidth, os, dcf, nrep
    ((250, 250), 6000, 4, 1.5, False, 1),     # 378 x 378, one transform
    ((220, 200), 8000, 6, 1.5, True, 2),      # 336 x 300
    ((201, 187), 6000, 4, 1.5, False, 3),     # 320 x 288, odd sides
    ((217, 176), 7000, 4, 1.5, True, 20),     # 336 x 270, a full and a partial 16-transform chunk
    ((200, 200), 5000, 4, 2.0, False, 5),     # 400 x 400 (os 2)
    ((330, 280), 6000, 4, 1.5, False, 16),    # 500 x 420: two LDS buffers at 16 transforms
    ((90, 300), 4000, 4, 1.5, False, 8),      # 140 x 450: one short side, one long
    ((300,), 2000, 4, 1.5, True, 1),          # 1-D, 450 points
    ((220,), 2000, 6, 1.5, False, 16),        # 1-D, 336 points, 16 transforms
]


@pytest.mark.parametrize("mid", [1, 0])
@pytest.mark.parametrize("imN,M,W,os_,use_dcf,nrep", PLAIN_CASES)
def test_nufft_mid_vs_oracle(mm, imN, M, W, os_, use_dcf, nrep, mid):
    print(f"[case] mid nufft imN={imN} M={M} W={W} os={os_} dcf={use_dcf} nrep={nrep} nu_fft_mid={mid}", flush=True)
    rng = np.random.default_rng(300 + sum(imN) + nrep)
    kloc = rng.random((M, len(imN))) - 0.5
    dcf = rng.random(M) + 0.5 if use_dcf else None
    So = NufftPlan(list(imN), kloc, dcf, os_, W)
    Sg = mm.nufftOp(list(imN), kloc, dcf, 1, os_, W).set_option("nu_fft_mid", mid)
    x = crand(rng, imN + ((nrep,) if nrep > 1 else ()))
    y = crand(rng, (M, nrep) if nrep > 1 else (M,))
    e_f = rel(Sg * x, So.forward(x))
    e_a = rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


SENSE_CASES = [
    # imN, coils, M, kwidth
    ((220, 220), 16, 8000, 4),    # 336 x 336, the high-order problem's image
    ((176, 240), 8, 6000, 6),     # 270 x 360, 8-transform row passes
    ((201, 187), 3, 5000, 4),     # 320 x 288, odd sides, 4-transform chunks
    ((330, 280), 16, 5000, 4),    # 500 x 420, two LDS buffers
]


@pytest.mark.parametrize("imN,C,M,W", SENSE_CASES)
def test_sense_mid_vs_oracle(mm, imN, C, M, W):
    """mrSampFunc (mrSampFunc.m:97-107, :126-144) over the fused SENSE NUFFT
    on mid-size grids, forward, adjoint and the normal operator."""
    print(f"[case] mid sense imN={imN} C={C} M={M} W={W}", flush=True)
    rng = np.random.default_rng(400 + sum(imN) + C)
    kl = rng.random((M, 2)) - 0.5
    maps = crand(rng, imN + (C,))
    So, Ro = NufftPlan(list(imN), kl, None, 1.5, W), RcvrPlan(maps)
    E = mm.senseOp(mm.nufftOp(list(imN), kl, None, 1, 1.5, W), mm.rcvrOp(maps))
    x = crand(rng, imN)
    y = o_mrSampFunc(crand(rng, imN), "notransp", So, Ro)
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    e_a = rel(E.H * y, o_mrSampFunc(y, "transp", So, Ro))
    want_n = o_mrSampFunc(o_mrSampFunc(x, "notransp", So, Ro), "transp", So, Ro)
    e_n = rel(E.normal(x), want_n)
    print(f"  fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL and e_n < TOL


@pytest.mark.parametrize("imN,C", [((140, 150), 1), ((140, 140), 8)])
def test_toeplitz_mid_vs_oracle(mm, imN, C):
    """prepToep (nufftOp.m:410-432, applied :348-352): a 280 x 300 Toeplitz
    grid (the fused column kernel in the mid-size mode) and a 350 x 375
    point-spread plan, against the oracle's Toeplitz form; with coils, the
    SENSE composition T(R x) / R'(T y)."""
    print(f"[case] mid Toeplitz imN={imN} C={C}", flush=True)
    rng = np.random.default_rng(500 + sum(imN) + C)
    kloc = rng.random((20000, 2)) - 0.5
    dcf = rng.random(kloc.shape[0]) + 0.5
    T = mm.nufftOp(list(imN), kloc, dcf, 1, 1.5, 4).prepToep()
    Tp = ToepPlan(list(imN), kloc, dcf, 1.5, 4)
    if C == 1:
        x = crand(rng, imN)
        e = rel(np.asarray(T * x), Tp.apply(x))
        print(f"  vs oracle {e:.2e}", flush=True)
        assert e < TOL
        return
    maps = crand(rng, imN + (C,))
    E = mm.senseOp(T, mm.rcvrOp(maps))
    x = crand(rng, imN)
    y = crand(rng, imN + (C,))
    want_f = np.stack([Tp.apply(maps[..., c] * x) for c in range(C)], axis=-1)
    want_a = sum(np.conj(maps[..., c]) * Tp.apply(y[..., c]) for c in range(C))
    e_f, e_a = rel(E * x, want_f), rel(E.H * y, want_a)
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_fft_mid_option_roundtrip(mm):
    """The option reads back what was set, and a plan whose grids fit the
    default kernels gives the same result with it on or off (bitwise: the
    mode only changes for sides above 256 at 16 transforms)."""
    rng = np.random.default_rng(7)
    kl = rng.random((3000, 2)) - 0.5
    S = mm.nufftOp([128, 128], kl, None, 1, 1.5, 4)
    assert S.get_option("nu_fft_mid") == 1
    x = crand(rng, (128, 128, 16))
    a = np.asarray(S * x)
    S.set_option("nu_fft_mid", 0)
    assert S.get_option("nu_fft_mid") == 0
    b = np.asarray(S * x)
    assert np.array_equal(a, b)


LARGE_CASES = [
    # imN, M, kwidth, dcf, nrep (grid sides at os 1.5 in the comment)
    ((512, 480), 20000, 4, False, 8),    # 768 x 720: large-grid column passes of 8 transforms
    ((512, 512), 20000, 4, True, 16),    # 768 x 768: plain transforms in chunks of 8
    ((451, 300), 12000, 6, False, 3),    # 686 x 450, odd side: two LDS buffers in the mid-size mode
]


@pytest.mark.parametrize("imN,M,W,use_dcf,nrep", LARGE_CASES)
def test_nufft_large_grid_vs_oracle(mm, imN, M, W, use_dcf, nrep):
    """Grid sides above what the small and mid-size FFT kernels can hold in
    LDS at 16 transforms (about 620 points): the large-grid kernels (one
    column of at most 8 transforms per work-group), plain transforms in
    chunks of 8, against the oracle (1e-5)."""
    print(f"[case] large nufft imN={imN} M={M} W={W} dcf={use_dcf} nrep={nrep}", flush=True)
    rng = np.random.default_rng(600 + sum(imN) + nrep)
    kloc = rng.random((M, 2)) - 0.5
    dcf = rng.random(M) + 0.5 if use_dcf else None
    So = NufftPlan(list(imN), kloc, dcf, 1.5, W)
    Sg = mm.nufftOp(list(imN), kloc, dcf, 1, 1.5, W)
    x = crand(rng, imN + ((nrep,) if nrep > 1 else ()))
    y = crand(rng, (M, nrep) if nrep > 1 else (M,))
    e_f = rel(Sg * x, So.forward(x))
    e_a = rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


@pytest.mark.parametrize("C", [8, 12])
def test_sense_large_grid_vs_oracle(mm, C):
    """A 512 x 512 SENSE operator against the oracle in both directions and
    the normal operator: 8 coils as the fused coil chunk on the large-grid
    kernels, 12 coils (a 16-coil chunk the LDS cannot hold at a 768-point
    side) as the composition S(R x) with plain transforms in chunks of 8."""
    print(f"[case] large sense 512x512 C={C}", flush=True)
    rng = np.random.default_rng(650 + C)
    imN, M = (512, 512), 20000
    kl = rng.random((M, 2)) - 0.5
    maps = crand(rng, imN + (C,))
    So, Ro = NufftPlan(list(imN), kl, None, 1.5, 4), RcvrPlan(maps)
    E = mm.senseOp(mm.nufftOp(list(imN), kl, None, 1, 1.5, 4), mm.rcvrOp(maps))
    x = crand(rng, imN)
    y = o_mrSampFunc(crand(rng, imN), "notransp", So, Ro)
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    e_a = rel(E.H * y, o_mrSampFunc(y, "transp", So, Ro))
    e_n = rel(E.normal(x), o_mrSampFunc(o_mrSampFunc(x, "notransp", So, Ro), "transp", So, Ro))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL and e_n < TOL
