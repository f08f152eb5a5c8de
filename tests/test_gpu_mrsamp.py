"""mrSampFunc's Cartesian-mask sampling and opt.sumDims on the device against
the FP64 oracle (operators/mrSampFunc.m:82-159; its unit test :163-216)."""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.fftnc import fftnc, ifftnc
from oracle.mrsamp import mrSampFunc as o_mrSampFunc
from oracle.rcvr import RcvrPlan
from oracle.solvers import cgne as o_cgne

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


def rel(a, b):
    a, b = np.asarray(a).reshape(-1, order="F"), np.asarray(b).reshape(-1, order="F")
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("shape,nrep", [((64, 48), 1), ((45, 33), 3), ((37, 20), 2), ((40, 1), 2), ((7, 96), 1),
                                        ((1000, 1), 4)])
def test_cartesian_mask_vs_oracle(mm, shape, nrep):
    """cartOp(mask): forward mask .* ifftnc(x), adjoint fftnc(mask .* y);
    odd sizes separate fftshift from ifftshift; 33 = 3 x 11 and 37 (prime)
    take the direct DFT; a [N x 1] mask is 1-D."""
    print(f"[case] Cartesian mask {shape} nrep={nrep}", flush=True)
    rng = np.random.default_rng(61)
    mask = (rng.random(shape) > 0.4).astype(float)
    nd = 1 if shape[1] == 1 else 2
    imN = shape[:nd]
    A = mm.cartOp(mask)
    x = crand(rng, imN + ((nrep,) if nrep > 1 else ()))
    mk = mask.reshape(imN + ((1,) if nrep > 1 else ()), order="F")
    want_f = mk * ifftnc(x, nd)
    want_a = fftnc(mk * x, nd)
    e_f, e_a = rel(A * x, want_f), rel(A.H * x, want_a)
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_cartesian_sense_and_cgne(mm):
    """mrSampFunc(x, tr, mask, R) against the oracle both ways, and cgne on it
    (the device solver through the generic SENSE composition) against the
    oracle cgne for 8 iterations."""
    print("[case] Cartesian SENSE 48x40 C=4 + cgne", flush=True)
    rng = np.random.default_rng(62)
    N1, N2, C = 48, 40, 4
    mask = (rng.random((N1, N2)) > 0.5).astype(float)
    maps = crand(rng, (N1, N2, C))
    Rg, Ro = mm.rcvrOp(maps), RcvrPlan(maps)
    x = crand(rng, (N1, N2))
    y = crand(rng, (N1, N2, C))
    e_f = rel(mm.mrSampFunc(x, "notransp", mask, Rg), o_mrSampFunc(x, "notransp", mask, Ro))
    e_a = rel(mm.mrSampFunc(y, "transp", mask, Rg), o_mrSampFunc(y, "transp", mask, Ro))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL
    ph = crand(rng, (N1, N2))
    data = o_mrSampFunc(ph, "notransp", mask, Ro)
    opt = {"resThresh": None, "xChangeThresh": None}
    A = lambda v, tr: o_mrSampFunc(v, tr, mask, Ro)
    xo, ro, *_ = o_cgne(A, data, np.zeros((N1, N2), complex), 8, opt)
    xg, rg, *_ = mm.cgne(mm.mrSampHandle(mask, Rg), data, np.zeros((N1, N2), complex), 8, opt)
    e = rel(xg, xo)
    print(f"  cgne(8) vs oracle {e:.2e}; resSq {rel(rg, np.real(ro)):.2e}", flush=True)
    assert e < 1e-4 and rel(rg, np.real(ro)) < 1e-4


@pytest.mark.parametrize("vec_im,vec_da,tik", [(False, False, False), (True, False, False), (False, True, False),
                                               (True, True, False), (True, True, True)])
def test_mrsampfunc_reference_unit_test(mm, vec_im, vec_da, tik):
    """The reference's own mrSampFunc test (mrSampFunc.m:163-216): a [N x 1]
    Cartesian mask with opt.sumDims = 2 after two receivers, every
    vectorisation / Tikhonov combination: the adjoint identity
    dot(x, F'y) == dot(F x, y) (the reference asserts 1e-8 in FP64; 1e-6
    here in FP32), and both directions against the oracle."""
    print(f"[case] mrSampFunc unit test: sumDims=2, imNFull={vec_im} daNFull={vec_da} tikReg={tik}", flush=True)
    rng = np.random.default_rng(63)
    N, NR = 32, 2
    maps = rng.random((N, N, NR))
    Rg, Ro = mm.rcvrOp(maps), RcvrPlan(maps)
    x = crand(rng, (N, N))
    mask = (rng.random((N, 1)) > 0.5).astype(float)
    y = crand(rng, (N, 1, NR))
    opt = {"sumDims": 2}
    x_a, y_a = x, y
    if vec_im:
        opt["imNFull"] = [N, N]
        x_a = x.reshape(-1, order="F")
    if vec_da:
        opt["daNFull"] = list(y.shape)
        y_a = y.reshape(-1, order="F")
    if tik:
        opt["tikReg"] = 0.01
        y_a = np.concatenate([y_a, x_a])
    Fx = np.asarray(mm.mrSampFunc(x_a, "notransp", mask, Rg, opt))
    Fy = np.asarray(mm.mrSampFunc(y_a, "transp", mask, Rg, opt))
    d1 = np.vdot(x_a.reshape(-1, order="F"), Fy.reshape(-1, order="F"))
    d2 = np.vdot(Fx.reshape(-1, order="F"), y_a.reshape(-1, order="F"))
    adj = abs(d1 - d2) / min(abs(d1), abs(d2))
    e_f = rel(Fx, o_mrSampFunc(x_a, "notransp", mask, Ro, opt))
    e_a = rel(Fy, o_mrSampFunc(y_a, "transp", mask, Ro, opt))
    print(f"  adjoint identity {adj:.2e}; fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert Fx.shape == np.asarray(o_mrSampFunc(x_a, "notransp", mask, Ro, opt)).shape
    assert adj < 1e-6 and e_f < TOL and e_a < TOL


def test_sumdims_on_coils_nufft(mm):
    """sumDims over the receiver dimension with a NUFFT sampling operator:
    S(sum_c R_c x) and its adjoint (broadcast) against the oracle."""
    from oracle.nufft import NufftPlan
    print("[case] sumDims=3 with a NUFFT, 40x40 C=3", flush=True)
    rng = np.random.default_rng(64)
    N, C, M = 40, 3, 1500
    kl = rng.random((M, 2)) - 0.5
    maps = crand(rng, (N, N, C))
    Sg, So = mm.nufftOp([N, N], kl, None, 1, 1.5, 4), NufftPlan([N, N], kl, None, 1.5, 4)
    Rg, Ro = mm.rcvrOp(maps), RcvrPlan(maps)
    opt = {"sumDims": [3]}
    x = crand(rng, (N, N))
    y = crand(rng, (M, 1))
    e_f = rel(mm.mrSampFunc(x, "notransp", Sg, Rg, opt), o_mrSampFunc(x, "notransp", So, Ro, opt))
    e_a = rel(mm.mrSampFunc(y, "transp", Sg, Rg, opt), o_mrSampFunc(y, "transp", So, Ro, opt))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL
