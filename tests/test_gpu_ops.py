"""GPU parity of the device operators against the FP64 oracle.

Tolerance: rel-L2 <= 1e-5 per forward/adjoint application (the per-application
bound the port is held to).  Every case prints its name before it runs.
"""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.highorder import HighOrderPlan
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


NUFFT_CASES = [
    # imN, M, kwidth, os, dcf, nrep
    ((32, 36), 500, 6, 1.5, False, 1),
    ((64, 64), 3000, 4, 1.5, True, 3),
    ((48, 40), 2000, 4, 1.375, False, 20),
    ((64,), 300, 6, 1.5, True, 2),
    ((96, 110), 5000, 8, 2.0, False, 1),
    # 192-point columns of 16 transforms: the radix 16 x 12 column passes
    # (12 reps: a partly filled 16-transform chunk; 8 reps: two columns of 8)
    ((64, 128), 4000, 4, 1.5, True, 12),
    ((48, 128), 3000, 6, 1.5, False, 8),
    # grid sides above 1024 (1080 x 144; 1-D 2100): FFT index arithmetic at large n
    ((704, 96), 6000, 4, 1.5, False, 2),
    ((1400,), 800, 6, 1.5, True, 3),
]


@pytest.mark.parametrize("imN,M,W,os_,use_dcf,nrep", NUFFT_CASES)
def test_nufft_vs_oracle(mm, imN, M, W, os_, use_dcf, nrep):
    print(f"[case] nufft imN={imN} M={M} W={W} os={os_} dcf={use_dcf} nrep={nrep}", flush=True)
    rng = np.random.default_rng(7)
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


def test_nufft_spiral_128_full_size(mm):
    """The 128^2 / 16384-sample / W=4 / os=1.5 spiral operator at full size."""
    print("[case] nufft spiral 128x128 M=16384 W=4 8 reps", flush=True)
    W = mm.workloads.nufft_problem(seed=1)
    So = NufftPlan([128, 128], W["kloc"], None, 1.5, 4)
    Sg = mm.nufftOp([128, 128], W["kloc"], None, 1, 1.5, 4)
    rng = np.random.default_rng(1)
    x = crand(rng, (128, 128, 8))
    y = crand(rng, (16384, 8))
    e_f, e_a = rel(Sg * x, So.forward(x)), rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_nufft_on_grid_known_answer(mm):
    """@nufftOp/tests.m:90-101: on-grid kloc makes the forward a unitary
    centred inverse DFT, fftnc(S*im) == im (tolerance 1e-4, L2/numel)."""
    from oracle.fftnc import fftnc
    print("[case] nufft on-grid KAT", flush=True)
    imN = [32, 36]
    axes = [(np.arange(n) - n / 2) / n for n in imN]
    g = np.meshgrid(*axes, indexing="ij")
    kloc = np.stack([a.reshape(-1, order="F") for a in g], axis=1)
    Sg = mm.nufftOp(imN, kloc, None, 1, 1.5, 6)
    rng = np.random.default_rng(3)
    im = rng.random(imN)
    k = (Sg * im).reshape(imN, order="F")
    err = np.sqrt(np.sum(np.abs(fftnc(k, 2) - im) ** 2)) / im.size
    assert err < 1e-4


def test_sense_fused_vs_oracle(mm):
    print("[case] sense nufft N=64 C=8", flush=True)
    rng = np.random.default_rng(11)
    N, M, C = 64, 4000, 8
    kloc = rng.random((M, 2)) - 0.5
    maps = crand(rng, (N, N, C))
    So, Ro = NufftPlan([N, N], kloc, None, 1.5, 4), RcvrPlan(maps)
    Sg, Rg = mm.nufftOp([N, N], kloc, None, 1, 1.5, 4), mm.rcvrOp(maps)
    E = mm.senseOp(Sg, Rg)
    x = crand(rng, (N, N))
    y = crand(rng, (M, C))
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    e_a = rel(E.H * y, o_mrSampFunc(y, "transp", So, Ro))
    e_h = rel(mm.mrSampFunc(x, "notransp", Sg, Rg), o_mrSampFunc(x, "notransp", So, Ro))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e} handle {e_h:.2e}", flush=True)
    assert max(e_f, e_a, e_h) < TOL
    # rcvrOp alone (rcvrOp.m:145-157)
    assert rel(Rg * x, Ro.forward(x)) < TOL
    assert rel(Rg.H * (Ro.forward(x)), Ro.adjoint(Ro.forward(x))) < TOL


def test_multislice_sense_vs_oracle(mm):
    print("[case] multi-slice sense 3 slices", flush=True)
    rng = np.random.default_rng(12)
    C = 4
    dims = [(48, 3000), (64, 5000), (40, 1500)]
    sl, maps, orc = [], [], []
    for N, M in dims:
        kl = rng.random((M, 2)) - 0.5
        mp = crand(rng, (N, N, C))
        sl.append(([N, N], kl, None))
        maps.append(mp)
        orc.append((NufftPlan([N, N], kl, None, 1.5, 4), RcvrPlan(mp)))
    S = mm.nufftOp.batch(sl, 1.5, 4)
    R = mm.rcvrOp.batch(maps)
    E = mm.senseOp(S, R)
    xs = [crand(rng, (N, N)) for N, _ in dims]
    ys = [crand(rng, (M, C)) for _, M in dims]
    xcat = np.concatenate([x.reshape(-1, order="F") for x in xs])
    ycat = np.concatenate([y.reshape(-1, order="F") for y in ys])
    yg, xg = E * xcat, E.H * ycat
    oy, ox = 0, 0
    for (N, M), x, y, (So, Ro) in zip(dims, xs, ys, orc):
        e_f = rel(yg[oy:oy + M * C], o_mrSampFunc(x, "notransp", So, Ro))
        e_a = rel(xg[ox:ox + N * N], o_mrSampFunc(y, "transp", So, Ro))
        print(f"  N={N} fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
        assert e_f < TOL and e_a < TOL
        oy += M * C
        ox += N * N


def test_uneven_batch_64_slices_16_coils(mm):
    """The uneven batch at full size: 64 slices of N in 96..160 (13 grid sizes
    144..240 at os 1.5, among them 150 = 2*3*5^2, 196 = 2^2*7^2 and
    210 = 2*3*5*7), 6000..20000 spiral samples each, 16 coils, one SENSE
    operator tuned by the autotuner.  E and E' of every slice against the
    oracle (nufftOp.m:146-152 grid sizes, rel-L2 <= 1e-5)."""
    print("[case] uneven batch 64 slices C=16 (autotuned)", flush=True)
    B = mm.workloads.batch_problems(seed=4, nslice=64, C=16)
    grids = sorted({mm.workloads.grid_size(b["N"]) for b in B})
    print(f"  grid sizes {grids}", flush=True)
    assert len(grids) == 13 and {150, 196, 210}.issubset(grids)
    E = mm.senseOp(mm.nufftOp.batch([([b["N"]] * 2, b["kloc"], None) for b in B], 1.5, 4),
                   mm.rcvrOp.batch([b["maps"] for b in B]))
    print(f"  autotune: {E.autotune(2)}", flush=True)
    rng = np.random.default_rng(40)
    xs = [crand(rng, (b["N"], b["N"])) for b in B]
    ys = [crand(rng, (b["M"], 16)) for b in B]
    yg = E * np.concatenate([x.reshape(-1, order="F") for x in xs])
    xg = E.H * np.concatenate([y.reshape(-1, order="F") for y in ys])
    oy = ox = 0
    worst = 0.0
    for b, x, y in zip(B, xs, ys):
        N, M = b["N"], b["M"]
        So, Ro = NufftPlan([N, N], b["kloc"], None, 1.5, 4), RcvrPlan(b["maps"])
        e_f = rel(yg[oy:oy + M * 16], o_mrSampFunc(x, "notransp", So, Ro))
        e_a = rel(xg[ox:ox + N * N], o_mrSampFunc(y, "transp", So, Ro))
        worst = max(worst, e_f, e_a)
        assert e_f < TOL and e_a < TOL, f"slice N={N} M={M}: fwd {e_f:.2e} adj {e_a:.2e}"
        oy += M * 16
        ox += N * N
    print(f"  worst rel-L2 over 64 slices, both directions: {worst:.2e}", flush=True)


def test_multislice_shared_trajectory(mm):
    """Slices with the same trajectory share their plan data (sample weights,
    CSR gridding rows): slices 0 and 2 share one spiral, slice 1 has its own,
    one with a dcf.  Each slice equals the oracle and its single-slice plan."""
    print("[case] multi-slice sense, shared trajectories", flush=True)
    rng = np.random.default_rng(21)
    C, N = 4, 48
    kl_a = mm.workloads.spiral(3000, 24)
    kl_b = rng.random((2500, 2)) - 0.5
    dcf_b = rng.random(2500) + 0.5
    sl = [([N, N], kl_a, None), ([N, N], kl_b, dcf_b), ([N, N], kl_a.copy(), None), ([N, N], kl_b.copy(), dcf_b.copy())]
    maps = [crand(rng, (N, N, C)) for _ in sl]
    E = mm.senseOp(mm.nufftOp.batch(sl, 1.5, 4), mm.rcvrOp.batch(maps))
    xs = [crand(rng, (N, N)) for _ in sl]
    ys = [crand(rng, (s[1].shape[0], C)) for s in sl]
    yg = E * np.concatenate([x.reshape(-1, order="F") for x in xs])
    xg = E.H * np.concatenate([y.reshape(-1, order="F") for y in ys])
    oy = ox = 0
    for (imN, kl, dc), mp, x, y in zip(sl, maps, xs, ys):
        M = kl.shape[0]
        So, Ro = NufftPlan(imN, kl, dc, 1.5, 4), RcvrPlan(mp)
        Es = mm.senseOp(mm.nufftOp(imN, kl, dc, 1, 1.5, 4), mm.rcvrOp(mp))
        e_f = rel(yg[oy:oy + M * C], o_mrSampFunc(x, "notransp", So, Ro))
        e_a = rel(xg[ox:ox + N * N], o_mrSampFunc(y, "transp", So, Ro))
        d_f = rel(yg[oy:oy + M * C], (Es * x).reshape(-1, order="F"))
        d_a = rel(xg[ox:ox + N * N], (Es.H * y).reshape(-1, order="F"))
        print(f"  M={M} fwd {e_f:.2e} adj {e_a:.2e}; vs single-slice plan {d_f:.1e} {d_a:.1e}", flush=True)
        assert e_f < TOL and e_a < TOL
        assert d_f < 1e-6 and d_a < 1e-6
        oy += M * C
        ox += N * N


@pytest.mark.parametrize("kind,nrep", [("uniform", 8), ("uniform", 3), ("spiral", 8), ("spiral", 16),
                                       ("w8", 5)])
def test_grid_staged_vs_global_rows(mm, kind, nrep):
    """2-D gridding from LDS-staged rectangles against the gather from the
    global CSR rows (option nu_grid_stage 0) and the oracle.  Without dense
    rectangles (uniform samples) each cell sums the same products in the same
    order, so the two are bitwise equal; the spiral's dense centre splits
    rows differently (rel. difference <= 1e-6)."""
    print(f"[case] staged gridding {kind} nrep={nrep}", flush=True)
    rng = np.random.default_rng(5)
    if kind == "spiral":
        imN, W, kloc = [128, 128], 4, mm.workloads.spiral(16384, 64)
    elif kind == "w8":
        imN, W, kloc = [40, 56], 8, rng.random((3000, 2)) - 0.5
    else:
        imN, W, kloc = [64, 80], 4, rng.random((6000, 2)) - 0.5
    M = kloc.shape[0]
    y = crand(rng, (M, nrep))
    Sg = mm.nufftOp(imN, kloc, None, 1, 1.5, W)
    a_stage = Sg.H * y
    Sg.set_option("nu_grid_stage", 0)
    a_rows = Sg.H * y
    So = NufftPlan(imN, kloc, None, 1.5, W)
    e_o = rel(a_stage, So.adjoint(y))
    d = rel(a_stage, a_rows)
    print(f"  vs oracle {e_o:.2e}; vs global rows {d:.1e}", flush=True)
    assert e_o < TOL
    if kind == "spiral":
        assert d < 1e-6
    else:
        assert np.array_equal(np.asarray(a_stage), np.asarray(a_rows))


@pytest.mark.parametrize("kind,nrep", [("uniform", 8), ("uniform", 3), ("spiral", 8), ("spiral", 16),
                                       ("w8", 5), ("small", 2)])
def test_degrid_staged_vs_global(mm, kind, nrep):
    """2-D degridding from LDS-staged tiles (with pairs of sparse tiles merged
    into one block, and without) against the gather from the grid in global
    memory (option nu_degrid_stage 0): the same taps, products and order, so
    the forward outputs are bitwise equal; all match the oracle."""
    print(f"[case] staged degridding {kind} nrep={nrep}", flush=True)
    rng = np.random.default_rng(6)
    if kind == "spiral":
        imN, W, kloc = [128, 128], 4, mm.workloads.spiral(16384, 64)
    elif kind == "w8":
        imN, W, kloc = [40, 56], 8, rng.random((3000, 2)) - 0.5
    elif kind == "small":
        imN, W, kloc = [6, 8], 6, rng.random((200, 2)) - 0.5
    else:
        imN, W, kloc = [64, 80], 4, rng.random((6000, 2)) - 0.5
    x = crand(rng, tuple(imN) + (nrep,))
    Sg = mm.nufftOp(imN, kloc, None, 1, 1.5, W)
    f_stage = np.asarray(Sg * x)
    Sg.set_option("nu_degrid_merge", 0)   # every block on one tile
    f_one = np.asarray(Sg * x)
    Sg.set_option("nu_degrid_stage", 0)
    f_glob = np.asarray(Sg * x)
    assert np.array_equal(f_one, f_glob)
    e_o = rel(f_stage, NufftPlan(imN, kloc, None, 1.5, W).forward(x))
    print(f"  vs oracle {e_o:.2e}; equal to the global gather: {np.array_equal(f_stage, f_glob)}", flush=True)
    assert e_o < TOL
    assert np.array_equal(f_stage, f_glob)


@pytest.mark.parametrize("nslice,nrep", [(3, 8), (2, 20), (1, 40), (4, 1)])
def test_grid_pairs_vs_single(mm, nslice, nrep):
    """Staged gridding items that serve two (slice, chunk) units of one
    trajectory (option nu_grid_pairs) against one unit per item: each unit's
    cells sum the same products in the same order, so the adjoints are
    bitwise equal (odd unit counts leave one unit unpaired; nrep > 16 pairs
    the chunks of one slice, the last chunk partly filled)."""
    print(f"[case] paired gridding nslice={nslice} nrep={nrep}", flush=True)
    rng = np.random.default_rng(9)
    imN, kl = [96, 96], mm.workloads.spiral(9000, 40)
    S = mm.nufftOp.batch([(imN, kl, None)] * nslice, 1.5, 4)
    M = kl.shape[0]
    y = crand(rng, (nslice * M * nrep,)).astype(np.complex64)
    S.set_option("nu_grid_pairs", 1)
    assert S.get_option("nu_grid_pairs") == 1
    a_pair = np.asarray(S.H * y).reshape(-1, order="F")
    S.set_option("nu_grid_pairs", 0)
    a_one = np.asarray(S.H * y).reshape(-1, order="F")
    V = imN[0] * imN[1]
    # slice 0 against the oracle: its data are the first M * nrep values
    # (repetition-major), its images the first V * nrep
    y0 = y[: M * nrep].reshape(nrep, M).T
    want = NufftPlan(imN, kl, None, 1.5, 4).adjoint(y0 if nrep > 1 else y0[:, 0])
    e_o = rel(a_pair[: V * nrep], np.asarray(want).reshape(-1, order="F"))
    print(f"  slice 0 vs oracle {e_o:.2e}; equal: {np.array_equal(a_pair, a_one)}", flush=True)
    assert e_o < TOL
    assert np.array_equal(a_pair, a_one)


@pytest.mark.parametrize("nD,with_dcf,nrep", [(2, True, 2), (2, False, 1), (1, True, 1)])
def test_toeplitz_vs_oracle(mm, nD, with_dcf, nrep):
    """prepToep (nufftOp.m:410-432): both directions of the prepared operator
    against the oracle's Toeplitz form, and within the reference's own bound
    of the accurate normal operator S0'*(S0*x) (os 2, W 8: rms < 1e-3,
    @nufftOp/tests.m:104-124)."""
    print(f"[case] Toeplitz nD={nD} dcf={with_dcf} nrep={nrep}", flush=True)
    rng = np.random.default_rng(50 + nD + 2 * nrep)
    imN = [64] * nD
    kloc = rng.random((int(0.25 * np.prod(imN)), nD)) - 0.5
    dcf = rng.random(kloc.shape[0]) if with_dcf else None
    T = mm.nufftOp(imN, kloc, dcf, 1, 1.5, 6).prepToep()
    x = crand(rng, tuple(imN) + ((nrep,) if nrep > 1 else ()))
    got, got_h = np.asarray(T * x), np.asarray(T.H * x)
    want = ToepPlan(imN, kloc, dcf, 1.5, 6).apply(x)
    S0 = NufftPlan(imN, kloc, dcf, 2, 8)
    rms = np.sqrt(np.mean(np.abs(got - S0.adjoint(S0.forward(x))) ** 2))
    e, e_h = rel(got, want), rel(got_h, want)
    print(f"  vs oracle {e:.2e} / adjoint {e_h:.2e}; rms vs S0'S0 {rms:.2e}", flush=True)
    assert e < TOL and e_h < TOL
    assert rms < 1e-3
    # the three-kernel middle stage (forward columns, multiply, adjoint
    # columns) against the fused column kernel
    assert T.get_option("nu_toep_fused") == 1
    T.set_option("nu_toep_fused", 0)
    e_u = rel(np.asarray(T * x), want)
    print(f"  unfused vs oracle {e_u:.2e}", flush=True)
    assert e_u < TOL


@pytest.mark.parametrize("N,C", [(128, 8), (96, 16)])
def test_toeplitz_sense_radix16_columns(mm, N, C):
    """The Toeplitz middle stage on grids whose column FFTs factor into radix
    16 and 12 only (N = 128 -> G = 256 = 16 x 16, N = 96 -> G = 192 = 16 x 12)
    with 8 or 16 transforms per column item (a SENSE operator with 8 or 16
    coils): the fused column kernel and the three-kernel form against the
    oracle's Toeplitz form, E'E x = R'(T(R x))."""
    print(f"[case] Toeplitz SENSE normal N={N} C={C} (radix-16/12 columns)", flush=True)
    rng = np.random.default_rng(58)
    kl = mm.workloads.spiral(int(N * N * 0.6), N // 2)
    dcf = mm.workloads.spiral_dcf(kl.shape[0], N // 2)
    maps = crand(rng, (N, N, C))
    T = mm.nufftOp([N, N], kl, dcf, 1, 1.5, 4).prepToep()
    E = mm.senseOp(T, mm.rcvrOp(maps))
    x = crand(rng, (N, N))
    Tp = ToepPlan([N, N], kl, dcf, 1.5, 4)
    want = sum(np.conj(maps[..., c]) * Tp.apply(maps[..., c] * x) for c in range(C))
    e = rel(E.normal(x), want)
    T.set_option("nu_toep_fused", 0)
    e_u = rel(E.normal(x), want)
    print(f"  fused vs oracle {e:.2e}; unfused {e_u:.2e}", flush=True)
    assert e < TOL and e_u < TOL


@pytest.mark.parametrize("C", [4, 20])
def test_toeplitz_sense_forward_adjoint(mm, C):
    """mrSampFunc on a prepToep'd nufftOp (nufftOp.m:348-352): the forward is
    T(R x) (coil images), the adjoint R'(T y); and cgne refuses the operator,
    also with more than 16 coils (the SENSE operator composed of S and R)."""
    print(f"[case] Toeplitz SENSE forward/adjoint N=48 C={C}", flush=True)
    rng = np.random.default_rng(59 + C)
    N = 48
    kl = rng.random((1200, 2)) - 0.5
    maps = crand(rng, (N, N, C))
    T = mm.nufftOp([N, N], kl, None, 1, 1.5, 4).prepToep()
    E = mm.senseOp(T, mm.rcvrOp(maps))
    Tp = ToepPlan([N, N], kl, None, 1.5, 4)
    x = crand(rng, (N, N))
    y = crand(rng, (N, N, C))
    want_f = np.stack([Tp.apply(maps[..., c] * x) for c in range(C)], axis=-1)
    want_a = sum(np.conj(maps[..., c]) * Tp.apply(y[..., c]) for c in range(C))
    e_f, e_a = rel(E * x, want_f), rel(E.H * y, want_a)
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL
    with pytest.raises(NotImplementedError):
        mm.cgne(E, y, None, 3, {})


def test_toeplitz_sense_normal_multislice(mm):
    """A SENSE operator on a two-slice Toeplitz NUFFT (different trajectories
    and dcf): E'E x = R'(T(R x)) per slice against the oracle."""
    print("[case] Toeplitz SENSE normal, 2 slices 48x48 C=4", flush=True)
    rng = np.random.default_rng(57)
    N, C = 48, 4
    kls = [rng.random((900, 2)) - 0.5 for _ in range(2)]
    dcfs = [0.5 + rng.random(900) for _ in range(2)]
    maps = [crand(rng, (N, N, C)) for _ in range(2)]
    T = mm.nufftOp.batch([([N, N], kl, d) for kl, d in zip(kls, dcfs)], 1.5, 4).prepToep()
    E = mm.senseOp(T, mm.rcvrOp.batch(maps))
    xs = [crand(rng, (N, N)) for _ in range(2)]
    xcat = np.concatenate([x.reshape(-1, order="F") for x in xs])
    got = np.asarray(E.normal(xcat)).reshape(-1, order="F")
    want = []
    for x, kl, d, mp in zip(xs, kls, dcfs, maps):
        Tp = ToepPlan([N, N], kl, d, 1.5, 4)
        want.append(sum(np.conj(mp[..., c]) * Tp.apply(mp[..., c] * x) for c in range(C)).reshape(-1, order="F"))
    e = rel(got, np.concatenate(want))
    print(f"  vs oracle {e:.2e}", flush=True)
    assert e < TOL


def test_nufft_tuning_options_and_autotune(mm):
    """Every staged-partition option value keeps the SENSE operator within the
    per-application bound; the autotuner picks from its candidate sets and
    the tuned operator still matches the oracle."""
    print("[case] tuning options + autotune, 3 spiral slices N=96 C=8", flush=True)
    rng = np.random.default_rng(8)
    N, C = 96, 8
    kl = mm.workloads.spiral(9000, 40)
    maps = [crand(rng, (N, N, C)) for _ in range(3)]
    E = mm.senseOp(mm.nufftOp.batch([([N, N], kl, None)] * 3, 1.5, 4), mm.rcvrOp.batch(maps))
    So = NufftPlan([N, N], kl, None, 1.5, 4)
    xs = [crand(rng, (N, N)) for _ in range(3)]
    xcat = np.concatenate([x.reshape(-1, order="F") for x in xs])
    want = np.concatenate([o_mrSampFunc(o_mrSampFunc(x, "notransp", So, RcvrPlan(mp)), "transp", So, RcvrPlan(mp))
                           .reshape(-1, order="F") for x, mp in zip(xs, maps)])
    for key, vals in [("nu_stage_samples", [16, 96, 256]), ("nu_stage_entries", [64, 768, 2048]),
                      ("nu_degrid_block", [16, 128, 256]), ("nu_pair_cols", [0, 1]), ("nu_xcd_order", [1, 0]),
                      ("nu_grid_pairs", [0, 1]), ("nu_snake", [0, 1]), ("nu_slice_group", [1, 0, 2]),
                      ("nu_degrid_merge", [0, 1])]:
        if key == "nu_xcd_order":
            # the region order of the staged gridding (under the XCD order)
            # only moves work between XCDs: bitwise the same result
            E.set_option("nu_xcd_order", 1)
            E.set_option("nu_grid_region", 0)
            ref = E.normal(xcat)
            E.set_option("nu_grid_region", 1)
            assert E.get_option("nu_grid_region") == 1
            assert np.array_equal(E.normal(xcat), ref)
        for v in vals:
            E.set_option(key, v)
            assert E.get_option(key) == v
            e = rel(E.normal(xcat), want)
            print(f"  {key}={v}: E'E vs oracle {e:.2e}", flush=True)
            assert e < TOL
        E.set_option(key, vals[1])
    with pytest.raises(Exception):
        E.set_option("nu_stage_samples", 4096)
    chosen = E.autotune(3)
    print(f"  autotune chose {chosen}", flush=True)
    assert (chosen["nu_stage_samples"], chosen["nu_stage_entries"]) in [(64, 512), (96, 768), (128, 1024),
                                                                          (192, 1536), (256, 2048)]
    assert chosen["nu_degrid_block"] in (32, 64, 128)
    assert chosen["nu_pair_cols"] in (0, 1) and chosen["nu_xcd_order"] in (0, 1)
    assert chosen["nu_grid_pairs"] in (0, 1) and chosen["nu_snake"] in (0, 1)
    assert chosen["nu_grid_region"] in (0, 1)
    assert chosen["nu_slice_group"] == 0   # three slices: one group
    assert rel(E.normal(xcat), want) < TOL


def _hio_plans(mm, rng, N=12, K=300, conc=True, mask=False, Lh=16):
    gx = (np.arange(N) - N / 2) * 0.01
    X, Y = np.meshgrid(gx, gx, indexing="ij")
    grid = {"x": X, "y": Y, "z": np.full_like(X, 0.02)}
    b0 = rng.standard_normal((N, N)) * 300
    t = np.arange(K) * 1e-5
    hs = rng.standard_normal((Lh, K)) * np.array([5.0] + [300.0] * 3 + [3e3] * 5 + [3e4] * 7 + [1e5] * 20)[:Lh, None]
    hc = rng.standard_normal((4, K)) * 1e3 if conc else None
    mk = (rng.random((N, N)) > 0.3).astype(float) if mask else None
    So = HighOrderPlan(b0, t, hs, hc, grid, mk)
    Sg = mm.sampHighOrder(b0, t, hs, hc, grid, mk)
    return So, Sg


@pytest.mark.parametrize("conc,mask,Lh,variant", [(False, False, 16, 0), (True, True, 16, 0), (True, True, 16, 3), (True, True, 16, 7),
                                                  (True, False, 36, 0)])
def test_highorder_vs_oracle(mm, conc, mask, Lh, variant):
    print(f"[case] highorder conc={conc} mask={mask} Lh={Lh} variant={variant}", flush=True)
    rng = np.random.default_rng(21)
    So, Sg = _hio_plans(mm, rng, conc=conc, mask=mask, Lh=Lh)
    Sg.set_option("hio_variant", variant)
    x = crand(rng, (12, 12, 3))
    y = crand(rng, (300, 3))
    e_f, e_a = rel(Sg * x, So.forward(x)), rel(Sg.H * y, So.adjoint(y))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_highorder_phidiv_vs_oracle(mm):
    print("[case] highorder phiDiv", flush=True)
    rng = np.random.default_rng(22)
    So, Sg = _hio_plans(mm, rng)
    dh = np.diff(np.concatenate([np.zeros((16, 1)), So.phs_spha], axis=1), axis=1)
    dc = np.diff(np.concatenate([np.zeros((4, 1)), So.phs_conc], axis=1), axis=1)
    So.set_phi_div(dh, dc)
    Sg.setPhiDiv(dh, dc)
    x = crand(rng, (12, 12))
    assert rel(Sg * x, So.forward(x)) < TOL


def test_highorder_sense_vs_oracle(mm):
    print("[case] highorder sense C=6", flush=True)
    rng = np.random.default_rng(23)
    So, Sg = _hio_plans(mm, rng, N=16, K=400)
    maps = crand(rng, (16, 16, 6))
    Ro, Rg = RcvrPlan(maps), mm.rcvrOp(maps)
    E = mm.senseOp(Sg, Rg)
    x = crand(rng, (16, 16))
    y = crand(rng, (400, 6))
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    e_a = rel(E.H * y, o_mrSampFunc(y, "transp", So, Ro))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


@pytest.mark.parametrize("f0,variant", [(0.0, 0), (70e3, 0), (70e3, 1), (70e3, 2), (70e3, 3), (70e3, 7)])
def test_highorder_full_grid_precision(mm, f0, variant):
    """220x220 grid, 32 coils, 16 harmonics + B0 of the high-order workload on
    1200 of its 24000 time points (spread over the whole readout, so the late
    samples with the largest phase are included); with the 70 kHz zeroth-order
    term max|Phi| exceeds 1e4 rad.  rel-L2 <= 1e-5 against the FP64 oracle."""
    print(f"[case] highorder 220^2 C=32 K=1200/24000 f0={f0} variant={variant}", flush=True)
    P = mm.workloads.highorder_problem(seed=2, f0_hz=f0)
    sel = np.linspace(0, P["K"] - 1, 1200).round().astype(int)
    t, h = P["t"][sel], P["h"][:, sel]
    So = HighOrderPlan(P["b0"], t, h, None, P["grid"])
    ph = So.phase(tsl=slice(len(sel) - 8, len(sel)))
    maxphi = np.max(np.abs(ph))
    print(f"  max|Phi| over the last samples {maxphi:.3e} rad", flush=True)
    if f0:
        assert maxphi > 1e4
    Sg = mm.sampHighOrder(P["b0"], t, h, None, P["grid"])
    Sg.set_option("hio_variant", variant)
    Ro, Rg = RcvrPlan(P["maps"]), mm.rcvrOp(P["maps"])
    E = mm.senseOp(Sg, Rg)
    rng = np.random.default_rng(3)
    x = crand(rng, (220, 220))
    y = crand(rng, (1200, 32))
    e_f = rel(E * x, o_mrSampFunc(x, "notransp", So, Ro))
    e_a = rel(E.H * y, o_mrSampFunc(y, "transp", So, Ro))
    print(f"  fwd {e_f:.2e} adj {e_a:.2e}", flush=True)
    assert e_f < TOL and e_a < TOL


def test_nufft_1d_gridding_bitwise_reproducible(mm):
    """1-D gridding is the per-cell gather over the CSR rows (no atomics): two
    applications of the adjoint are bitwise equal, and each slice of a
    multi-slice 1-D plan equals its single-slice plan bit for bit."""
    print("[case] 1-D gridding reproducibility, 3 slices", flush=True)
    rng = np.random.default_rng(31)
    sl = [([96], rng.random((700, 1)) - 0.5, None), ([64], rng.random((2000, 1)) - 0.5, rng.random(2000) + 0.5),
          ([150], rng.random((400, 1)) - 0.5, None)]
    S = mm.nufftOp.batch(sl, 1.5, 6)
    ys = [crand(rng, (kl.shape[0],)) for _, kl, _ in sl]
    ycat = np.concatenate(ys)
    a1 = np.asarray(S.H * ycat)
    a2 = np.asarray(S.H * ycat)
    assert np.array_equal(a1, a2)
    off = 0
    for (imN, kl, dc), y in zip(sl, ys):
        single = np.asarray(mm.nufftOp(imN, kl, dc, 1, 1.5, 6).H * y).reshape(-1, order="F")
        want = NufftPlan(imN, kl, dc, 1.5, 6).adjoint(y)
        got = a1.reshape(-1, order="F")[off:off + imN[0]]
        print(f"  N={imN[0]} vs oracle {rel(got, want):.2e}; equal to single-slice: {np.array_equal(got, single)}",
              flush=True)
        assert rel(got, want) < TOL
        assert np.array_equal(got, single)
        off += imN[0]
