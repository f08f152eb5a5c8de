"""GPU parity of odd image sides against the FP64 oracle.

nufftOp accepts odd image sizes (nufftOp.m:171-175 deapodises at the
half-integer positions -N/2..N/2-1, :258-273 builds the k-space ramp with
N/2 half-integer).  The oracle's odd-side forward is pinned by its closed
form (tests/test_oracle_nufft.py::test_odd_size_closed_form); here the device
operator, its SENSE composition, the multi-slice plans, cgne and the
Toeplitz form are held to the oracle at the per-application bound 1e-5.
"""
import numpy as np
import pytest

from conftest import load_pkg
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


ODD_CASES = [
    # imN, M, kwidth, os, dcf, nrep
    ((31, 36), 800, 6, 1.5, False, 1),
    ((36, 31), 800, 4, 1.5, True, 2),
    ((33, 35), 1500, 4, 1.5, True, 3),
    ((15,), 200, 6, 1.5, True, 2),
    ((95, 97), 5000, 8, 2.0, False, 1),
    # 192-point columns of 16 transforms (radix 16 x 12) with an odd N1
    ((63, 128), 4000, 4, 1.5, False, 12),
    ((127, 127), 9000, 4, 1.5, True, 8),
]


@pytest.mark.parametrize("imN,M,W,os_,use_dcf,nrep", ODD_CASES)
def test_odd_nufft_vs_oracle(mm, imN, M, W, os_, use_dcf, nrep):
    print(f"[case] odd nufft imN={imN} M={M} W={W} os={os_} dcf={use_dcf} nrep={nrep}", flush=True)
    rng = np.random.default_rng(70 + sum(imN))
    nd = len(imN)
    kloc = rng.random((M, nd)) - 0.5
    dcf = rng.random(M) + 0.5 if use_dcf else None
    So = NufftPlan(list(imN), kloc, dcf, os_, W)
    Sg = mm.nufftOp(list(imN), kloc, dcf, 1, os_, W)
    x = crand(rng, imN + (nrep,))
    y = crand(rng, (M, nrep))
    e_f = rel(Sg * x, So.forward(x))
    e_a = rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_odd_multislice_sense_vs_oracle(mm):
    """A multi-slice SENSE plan mixing odd and even sides (fused coil
    path, 4 coils): forward, adjoint and the normal operator per slice."""
    print("[case] odd/even multi-slice sense, 4 slices C=4", flush=True)
    rng = np.random.default_rng(77)
    C = 4
    dims = [((47, 47), 2500), ((48, 48), 3000), ((45, 50), 2000), ((50, 45), 2200)]
    sl, maps, orc = [], [], []
    for imN, M in dims:
        kl = rng.random((M, 2)) - 0.5
        mp = crand(rng, imN + (C,))
        sl.append((list(imN), kl, None))
        maps.append(mp)
        orc.append((NufftPlan(list(imN), kl, None, 1.5, 4), RcvrPlan(mp)))
    E = mm.senseOp(mm.nufftOp.batch(sl, 1.5, 4), mm.rcvrOp.batch(maps))
    xs = [crand(rng, imN) for imN, _ in dims]
    ys = [crand(rng, (M, C)) for _, M in dims]
    xcat = np.concatenate([x.reshape(-1, order="F") for x in xs])
    ycat = np.concatenate([y.reshape(-1, order="F") for y in ys])
    yg, xg = np.asarray(E * xcat), np.asarray(E.H * ycat)
    ng = np.asarray(E.normal(xcat)).reshape(-1, order="F")
    oy = ox = 0
    for (imN, M), x, y, (So, Ro) in zip(dims, xs, ys, orc):
        V = imN[0] * imN[1]
        e_f = rel(yg[oy:oy + M * C], o_mrSampFunc(x, "notransp", So, Ro))
        e_a = rel(xg[ox:ox + V], o_mrSampFunc(y, "transp", So, Ro))
        e_n = rel(ng[ox:ox + V], o_mrSampFunc(o_mrSampFunc(x, "notransp", So, Ro), "transp", So, Ro))
        print(f"  imN={imN} fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}", flush=True)
        assert e_f < TOL and e_a < TOL and e_n < TOL
        oy += M * C
        ox += V


def test_odd_cgne_matches_oracle(mm):
    """cgne (cgne.m:213-336) over a SENSE operator with odd sides: residual
    history and iterate follow the FP64 loop for 8 iterations."""
    print("[case] odd cgne imN=(63,65) C=4 zero-init", flush=True)
    rng = np.random.default_rng(78)
    imN, M, C = (63, 65), 3500, 4
    kl = rng.random((M, 2)) - 0.5
    maps = crand(rng, imN + (C,))
    So, Ro = NufftPlan(list(imN), kl, None, 1.5, 4), RcvrPlan(maps)
    Sg, Rg = mm.nufftOp(list(imN), kl, None, 1, 1.5, 4), mm.rcvrOp(maps)
    A = lambda v, tr: o_mrSampFunc(v, tr, So, Ro)
    xt = crand(rng, imN)
    y = A(xt, "notransp")
    opt = {"resThresh": None, "xChangeThresh": None}
    x0 = np.zeros(imN, complex)
    nit = 8
    xo, ro, *_ = o_cgne(A, y, x0, nit, opt)
    xg, rg, *_ = mm.cgne(mm.senseOp(Sg, Rg), y, x0, nit, opt)
    ro = np.real(ro)
    d = np.abs(rg - ro) / ro
    nrmse = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"  resSq rel diff max {d.max():.2e}; x nrmse {nrmse:.2e}", flush=True)
    assert len(rg) == nit + 1
    assert np.all(d < 1e-3)
    assert nrmse < 1e-4


@pytest.mark.parametrize("imN", [(31, 33), (40, 33), (45,)])
def test_odd_toeplitz_vs_oracle(mm, imN):
    """prepToep (nufftOp.m:410-432) on odd sides: the pad/crop offsets are
    floor(npc/2) (padcrop.m), the convolution is shift-equivariant, and the
    device form matches the oracle's."""
    print(f"[case] odd Toeplitz imN={imN}", flush=True)
    rng = np.random.default_rng(79 + sum(imN))
    nd = len(imN)
    kloc = rng.random((int(0.3 * np.prod(imN)), nd)) - 0.5
    dcf = rng.random(kloc.shape[0]) + 0.5
    T = mm.nufftOp(list(imN), kloc, dcf, 1, 1.5, 6).prepToep()
    x = crand(rng, imN)
    want = ToepPlan(list(imN), kloc, dcf, 1.5, 6).apply(x)
    e = rel(np.asarray(T * x), want)
    print(f"  vs oracle {e:.2e}", flush=True)
    assert e < TOL
