"""The C/OpenMP restatement (oracle/c/mmref.c, the CPU baseline) against the
NumPy oracle: per application to 1e-12, cgne histories and iterates to 1e-8
(the OpenMP reductions sum in a thread-count dependent order, and 12 CG
steps over five decades of residual amplify that)."""
import numpy as np
import pytest

from conftest import load_pkg
from oracle import cref
from oracle.mrsamp import mrSampFunc
from oracle.nufft import NufftPlan
from oracle.rcvr import RcvrPlan
from oracle.solvers import cgne as o_cgne


@pytest.fixture(scope="module")
def built():
    cref.build()
    return cref


def _problem(N, M, C, seed, dcf=False):
    mm = load_pkg()
    rng = np.random.default_rng(seed)
    kloc = mm.workloads.spiral(M, N // 2)
    maps = mm.workloads.gaussian_coils(N, N, C, rng)
    ph = mm.workloads.ellipse_phantom(N, N, rng)
    d = mm.workloads.spiral_dcf(M, N // 2) if dcf else None
    return kloc, maps, ph, d, rng


@pytest.mark.parametrize("N,M,C,dcf", [(32, 1500, 4, False), (48, 3000, 3, True), (40, 2200, 2, False)])
def test_forward_adjoint_match_numpy_oracle(built, N, M, C, dcf):
    kloc, maps, ph, d, rng = _problem(N, M, C, 3, dcf)
    So, Ro = NufftPlan([N, N], kloc, d, 1.5, 4), RcvrPlan(maps)
    E = cref.SenseRef(N, kloc, d, maps)
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    y = rng.standard_normal((M, C)) + 1j * rng.standard_normal((M, C))
    yo = mrSampFunc(x, "notransp", So, Ro).reshape(M, C, order="F")
    xo = mrSampFunc(y, "transp", So, Ro).reshape(N, N, order="F")
    ef = np.linalg.norm(E.forward(x) - yo) / np.linalg.norm(yo)
    ea = np.linalg.norm(E.adjoint(y) - xo) / np.linalg.norm(xo)
    assert ef < 1e-12 and ea < 1e-12, (ef, ea)


def test_cgne_matches_numpy_oracle(built):
    N, M, C = 32, 1500, 4
    kloc, maps, ph, _, rng = _problem(N, M, C, 5)
    So, Ro = NufftPlan([N, N], kloc, None, 1.5, 4), RcvrPlan(maps)
    A = lambda v, tr: mrSampFunc(v, tr, So, Ro)
    y = mrSampFunc(ph, "notransp", So, Ro)
    E = cref.SenseRef(N, kloc, None, maps)
    x0 = np.zeros((N, N), complex)
    xo, ro, *_ = o_cgne(A, y, x0, 12, {"resThresh": None, "xChangeThresh": None, "expResN": 5})
    xc, rc, nit = E.cgne(y.reshape(M, C, order="F"), x0, 12, expResN=5)
    assert nit == 12
    assert np.allclose(rc, np.real(ro), rtol=1e-8, atol=0)
    assert np.linalg.norm(xc - xo) / np.linalg.norm(xo) < 1e-8
    # default x0 = b and a relative threshold stop
    xo2, ro2, *_ = o_cgne(A, y, None, 50, {"resThresh": 1e-6 * np.real(np.vdot(A(y, "transp"), A(y, "transp"))),
                                           "xChangeThresh": None})
    xc2, rc2, nit2 = E.cgne(y.reshape(M, C, order="F"), None, 50, resThreshRel=1e-6)
    assert nit2 == len(ro2) - 1
    assert np.linalg.norm(xc2 - xo2) / np.linalg.norm(xo2) < 1e-8


def _ho_problem(seed=7, N=20, K=300, C=3, mask=False, conc=True):
    mm = load_pkg()
    P = mm.workloads.highorder_problem(seed=seed, N=N, K=K, C=C, f0_hz=70e3, dwell=24e-3 / K)
    rng = P["rng"]
    hc = 1e-3 * rng.standard_normal((4, K)) if conc else None
    m = (rng.uniform(size=(N, N)) > 0.3).astype(float) if mask else None
    return P, hc, m, rng


@pytest.mark.parametrize("mask,conc,cache", [(False, False, False), (True, True, False), (True, True, True)])
def test_highorder_sense_matches_numpy_oracle(built, mask, conc, cache):
    """mmref_hio_forward/adjoint against HighOrderPlan o RcvrPlan through
    mrSampFunc (16 harmonics + concomitant terms + B0 + a 70 kHz term, with
    and without the mask, with and without the exp cache): 1e-12."""
    from oracle.highorder import HighOrderPlan
    P, hc, m, rng = _ho_problem(mask=mask, conc=conc)
    N, K, C = P["N"], P["K"], P["C"]
    So = HighOrderPlan(P["b0"], P["t"], P["h"], hc, P["grid"], m)
    Ro = RcvrPlan(P["maps"])
    E = cref.HighOrderSenseRef(P["b0"], P["t"], P["h"], hc, P["grid"], P["maps"], m, cache_exp=cache)
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    y = rng.standard_normal((K, C)) + 1j * rng.standard_normal((K, C))
    yo = mrSampFunc(x, "notransp", So, Ro).reshape(K, C, order="F")
    xo = mrSampFunc(y, "transp", So, Ro).reshape(N, N, order="F")
    ef = np.linalg.norm(E.forward(x) - yo) / np.linalg.norm(yo)
    ea = np.linalg.norm(E.adjoint(y) - xo) / np.linalg.norm(xo)
    assert np.max(np.abs(So.phase())) > 1e4
    assert ef < 1e-12 and ea < 1e-12, (ef, ea)


def test_highorder_cgne_matches_numpy_oracle(built):
    """mmref_hio_cgne against oracle cgne over the NumPy high-order operator:
    12 iterations (with an explicit residual every 5), histories and iterate
    to 1e-8."""
    from oracle.highorder import HighOrderPlan
    P, hc, m, rng = _ho_problem(seed=8, N=24, K=400, C=4)
    So, Ro = HighOrderPlan(P["b0"], P["t"], P["h"], hc, P["grid"]), RcvrPlan(P["maps"])
    A = lambda v, tr: mrSampFunc(v, tr, So, Ro)
    y = A(P["phantom"], "notransp")
    E = cref.HighOrderSenseRef(P["b0"], P["t"], P["h"], hc, P["grid"], P["maps"], cache_exp=True)
    x0 = np.zeros((24, 24), complex)
    xo, ro, *_ = o_cgne(A, y, x0, 12, {"resThresh": None, "xChangeThresh": None, "expResN": 5})
    xc, rc, nit = E.cgne(y.reshape(400, 4, order="F"), x0, 12, expResN=5)
    assert nit == 12
    assert np.allclose(rc, np.real(ro), rtol=1e-8, atol=0)
    assert np.linalg.norm(xc - xo) / np.linalg.norm(xo) < 1e-8


def test_highorder_round_io_matches_numpy_rounding(built):
    """set_round_io: each application's input and output rounded to
    complex64, as the NumPy oracle wrapped in complex64 casts."""
    from oracle.highorder import HighOrderPlan
    P, hc, m, rng = _ho_problem(seed=9)
    N, K, C = P["N"], P["K"], P["C"]
    So, Ro = HighOrderPlan(P["b0"], P["t"], P["h"], hc, P["grid"]), RcvrPlan(P["maps"])
    E = cref.HighOrderSenseRef(P["b0"], P["t"], P["h"], hc, P["grid"], P["maps"])
    E.set_round_io(True)
    c64 = lambda a: np.asarray(a).astype(np.complex64).astype(np.complex128)
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    y = rng.standard_normal((K, C)) + 1j * rng.standard_normal((K, C))
    yo = c64(mrSampFunc(c64(x), "notransp", So, Ro)).reshape(K, C, order="F")
    xo = c64(mrSampFunc(c64(y), "transp", So, Ro)).reshape(N, N, order="F")
    yc, xc = E.forward(x), E.adjoint(y)
    # the same rounded values up to an ulp of complex64 where the FP64 sums differ in their last bits
    assert np.linalg.norm(yc - yo) / np.linalg.norm(yo) < 1e-7
    assert np.linalg.norm(xc - xo) / np.linalg.norm(xo) < 1e-7
    assert np.array_equal(yc, c64(yc)) and np.array_equal(xc, c64(xc))
    E.set_round_io(False)
    assert not np.array_equal(E.forward(x), c64(E.forward(x)))
