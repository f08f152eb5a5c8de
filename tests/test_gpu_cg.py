"""GPU cgne (device loop, FP64 scalars) against the FP64 oracle cgne."""
import numpy as np
import pytest

from conftest import load_pkg
from oracle.highorder import HighOrderPlan
from oracle.mrsamp import mrSampFunc as o_mrSampFunc
from oracle.nufft import NufftPlan
from oracle.rcvr import RcvrPlan
from oracle.solvers import cgne as o_cgne

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mm():
    m = load_pkg()
    if m.device_count() < 1:
        pytest.fail("no GPU visible to libmatmri_hip")
    return m


def _problem(mm, N=64, M=6000, C=4, seed=5):
    rng = np.random.default_rng(seed)
    kloc = mm.workloads.spiral(M, N // 2)
    maps = mm.workloads.gaussian_coils(N, N, C, rng)
    ph = mm.workloads.ellipse_phantom(N, N, rng)
    So, Ro = NufftPlan([N, N], kloc, None, 1.5, 4), RcvrPlan(maps)
    y = mm.workloads.add_noise(o_mrSampFunc(ph, "notransp", So, Ro), rng)
    Sg, Rg = mm.nufftOp([N, N], kloc, None, 1, 1.5, 4), mm.rcvrOp(maps)
    return So, Ro, Sg, Rg, y, ph


def test_cgne_matches_oracle(mm):
    """Residual history and iterate of the device loop follow the FP64 loop
    while the relative residual is above 1e-4 (FP32 vectors resolve it)."""
    print("[case] cgne N=64 C=4 zero-init", flush=True)
    So, Ro, Sg, Rg, y, ph = _problem(mm)
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    opt = {"resThresh": None, "xChangeThresh": None}
    b = A(y, "transp")
    bb = np.real(np.vdot(b, b))
    x0 = np.zeros((64, 64), complex)
    nit = 8
    xo, ro, *_ = o_cgne(A, y, x0, nit, opt)
    xg, rg, _, xng, xdg, _, st = mm.cgne(mm.mrSampHandle(Sg, Rg), y, x0, nit, opt)
    ro = np.real(ro)
    ok = ro > 1e-8 * bb
    d = np.abs(rg - ro) / ro
    print("  rel resid", np.sqrt(ro / bb), flush=True)
    print("  resSq rel diff", d, flush=True)
    assert len(rg) == nit + 1 and st == 0
    assert np.all(d[ok] < 1e-3)
    nrmse = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print("  x nrmse", nrmse, flush=True)
    assert nrmse < 1e-4


def _config1(mm):
    """The 128x128 / 8-coil / 16384-sample spiral problem (KB W=4, os=1.5)
    at full size, data from the FP64 forward operator plus noise (SNR 40)."""
    P = mm.workloads.nufft_problem(seed=1)
    N = P["N"]
    So, Ro = NufftPlan([N, N], P["kloc"], None, 1.5, 4), RcvrPlan(P["maps"])
    y = mm.workloads.add_noise(o_mrSampFunc(P["phantom"], "notransp", So, Ro), P["rng"])
    Sg, Rg = mm.nufftOp([N, N], P["kloc"], None, 1, 1.5, 4), mm.rcvrOp(P["maps"])
    return So, Ro, Sg, Rg, y, N


def _c64(a):
    return np.asarray(a).astype(np.complex64).astype(np.complex128)


def _single_oracle(So, Ro):
    """The oracle's operator evaluated in complex64 (the same steps in single
    precision): what rounding an FP32 implementation of the operator does to
    the FP64 cgne loop."""
    import copy
    S1, R1 = copy.copy(So), copy.copy(Ro)
    S1.single = R1.single = True
    return lambda x, tr: o_mrSampFunc(x, tr, S1, R1)


def test_cgne_config1_full_size_10_iterations(mm):
    """Zero-init cgne on the full-size 128x128 / 8-coil / 16384-sample spiral
    problem against the FP64 oracle loop (cgne.m:213-336) at 10 iterations
    (relative residual ~2e-3): NRMSE <= 1e-4.  resSq within 1e-3 of the
    oracle at every iteration, except where the oracle's own operator
    evaluated in single precision moves it further (from iteration ~9 on,
    CG amplifies per-application rounding ~13x per iteration there): then
    within that."""
    print("[case] cgne 128x128 C=8 M=16384, 10 iterations", flush=True)
    So, Ro, Sg, Rg, y, N = _config1(mm)
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    opt = {"resThresh": None, "xChangeThresh": None}
    x0 = np.zeros((N, N), complex)
    xo, ro, *_ = o_cgne(A, y, x0, 10, opt)
    xs, rs, *_ = o_cgne(_single_oracle(So, Ro), y, x0, 10, opt)
    xg, rg, *_, st = mm.cgne(mm.mrSampHandle(Sg, Rg), y, x0, 10, opt)
    ro, rs = np.real(ro), np.real(rs)
    d, ds = np.abs(np.asarray(rg) - ro) / ro, np.abs(rs - ro) / ro
    nrmse = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    nrmse_s = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
    with np.printoptions(precision=2):
        print(f"  rel resid {np.sqrt(ro[-1] / ro[0]):.2e}; x nrmse {nrmse:.2e} (single-precision oracle "
              f"{nrmse_s:.2e})\n  resSq rel diff device {d}\n  single-precision oracle {ds}", flush=True)
    assert st == 0 and len(rg) == 11
    assert np.all(d <= np.maximum(1e-3, ds))
    assert nrmse < 1e-4


def test_cgne_config1_full_size_20_iterations(mm):
    """The same problem at its own 20 iterations (relative residual ~3e-4).
    There CG amplifies rounding: the FP64 oracle run with only its operator's
    input and output rounded to complex64 moves by d ~ 1e-2 (NRMSE).  The
    device solve (FP32 operator, FP64 scalars) must stay within 3 d of the
    FP64 oracle, d computed here on the same problem; until the histories
    decorrelate (the single-precision oracle within 1e-2 of the FP64 one),
    resSq within max(1e-3, that single-precision oracle's deviation)."""
    print("[case] cgne 128x128 C=8 M=16384, 20 iterations", flush=True)
    So, Ro, Sg, Rg, y, N = _config1(mm)
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    A32 = lambda x, tr: _c64(o_mrSampFunc(_c64(x), tr, So, Ro))
    opt = {"resThresh": None, "xChangeThresh": None}
    x0 = np.zeros((N, N), complex)
    xo, ro, *_ = o_cgne(A, y, x0, 20, opt)
    x3, *_ = o_cgne(A32, y, x0, 20, opt)
    _, rs, *_ = o_cgne(_single_oracle(So, Ro), y, x0, 20, opt)
    d = np.linalg.norm(x3 - xo) / np.linalg.norm(xo)
    xg, rg, *_, st = mm.cgne(mm.mrSampHandle(Sg, Rg), y, x0, 20, opt)
    nrmse = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    ro, rs = np.real(ro), np.real(rs)
    ds, dg = np.abs(rs - ro) / ro, np.abs(np.asarray(rg) - ro) / ro
    tight = np.cumprod(ds < 1e-2).astype(bool)
    print(f"  FP32-I/O oracle drift d {d:.2e}; device nrmse {nrmse:.2e} ({nrmse / d:.2f} d); resSq on the first "
          f"{tight.sum()} iterations: max rel diff {dg[tight].max():.2e}", flush=True)
    assert st == 0 and len(rg) == 21
    assert tight.sum() >= 8 and np.all(dg[tight] <= np.maximum(1e-3, ds[tight]))
    assert nrmse < 3 * d


def test_cgne_relative_threshold(mm):
    """Stop at |r|^2 < 1e-8 |b|^2 (relative residual 1e-4) with cgne's default
    x0 = b.  Near this residual single-precision rounding of the operator
    changes how many iterations reach it: the FP64 oracle loop needs n_o, the
    same loop with its operator evaluated in complex64 (_single_oracle) n_s.
    The device (FP32 operator, FP64 scalars) must stop with status 1 at the
    first iteration below the threshold, after n_o .. n_s + 3 iterations.
    Its iterate is then compared with the FP64 oracle run for the same number
    of iterations n_g: within 3 max(d, ds), where d and ds are the distances
    of the FP32-I/O-rounded and the single-precision oracle runs of n_g
    iterations from that FP64 run (computed here on the same problem).  The
    error against the phantom is printed; the NRMSE bound limits its change
    (triangle inequality)."""
    print("[case] cgne default x0, relative residual 1e-4", flush=True)
    So, Ro, Sg, Rg, y, ph = _problem(mm, seed=6)
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    A32 = lambda x, tr: _c64(o_mrSampFunc(_c64(x), tr, So, Ro))
    As = _single_oracle(So, Ro)
    b = A(y, "transp")
    bb = np.real(np.vdot(b, b))
    thr = {"resThresh": 1e-8 * bb, "xChangeThresh": None}
    _, ro, *_, so = o_cgne(A, y, None, 80, thr)
    _, rs, *_, ss = o_cgne(As, y, None, 80, thr)
    xg, rg, _, _, _, _, st = mm.cgne(mm.mrSampHandle(Sg, Rg), y, None, 80,
                                     {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None})
    rg = np.asarray(rg)
    n_o, n_s, n_g = len(ro) - 1, len(rs) - 1, len(rg) - 1
    fixed = {"resThresh": None, "xChangeThresh": None}
    xo, *_ = o_cgne(A, y, None, n_g, fixed)
    x3, *_ = o_cgne(A32, y, None, n_g, fixed)
    xs, *_ = o_cgne(As, y, None, n_g, fixed)
    dist = lambda u: np.linalg.norm(u - xo) / np.linalg.norm(xo)
    d, ds, nrmse = dist(x3), dist(xs), dist(xg)
    eo = np.linalg.norm(xo - ph) / np.linalg.norm(ph)
    eg = np.linalg.norm(xg - ph) / np.linalg.norm(ph)
    print(f"  iterations: FP64 oracle {n_o}, single-precision oracle {n_s}, device {n_g} (status {st}); at "
          f"{n_g} iterations: drift d {d:.2e} ds {ds:.2e}, device {nrmse:.2e} ({nrmse / max(d, ds):.2f} x); "
          f"error vs phantom {eo:.4e} / {eg:.4e}", flush=True)
    assert so == 1 and ss == 1 and st == 1
    assert rg[-1] < 1e-8 * bb * (1 + 1e-5) and rg[-2] >= 1e-8 * bb * (1 - 1e-5)
    assert n_o <= n_g <= n_s + 3
    assert nrmse < 3 * max(d, ds)


def test_cgne_multislice_equals_single(mm):
    """The batched solve of 3 slices equals 3 single solves: every kernel
    computes a slice of the multi-slice plan with the same products in the
    same order as its single-slice plan (2-D gridding is a fixed-order
    gather, CG partial sums are per slice), so iteration counts, statuses
    and iterates are bitwise equal."""
    print("[case] cgne batched 3 slices vs single", flush=True)
    rng = np.random.default_rng(9)
    C = 4
    dims = [(48, 3000), (64, 5000), (40, 2000)]
    sl, mp, ys, singles = [], [], [], []
    for N, M in dims:
        kl = mm.workloads.spiral(M, N // 2)
        maps = mm.workloads.gaussian_coils(N, N, C, rng)
        ph = mm.workloads.ellipse_phantom(N, N, rng)
        y = mm.workloads.add_noise(o_mrSampFunc(ph, "notransp", NufftPlan([N, N], kl, None, 1.5, 4),
                                                RcvrPlan(maps)), rng)
        sl.append(([N, N], kl, None))
        mp.append(maps)
        ys.append(y)
        singles.append(mm.senseOp(mm.nufftOp([N, N], kl, None, 1, 1.5, 4), mm.rcvrOp(maps)))
    E = mm.senseOp(mm.nufftOp.batch(sl, 1.5, 4), mm.rcvrOp.batch(mp))
    ycat = np.concatenate([y.reshape(-1, order="F") for y in ys])
    opt = {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None}
    xb, rb, _, _, _, _, sb = mm.cgne(E, ycat, None, 60, opt)
    off = 0
    for (N, M), Es, y, r, s in zip(dims, singles, ys, rb, sb):
        xs, rs, *_, ss = mm.cgne(Es, y, None, 60, opt)
        e = np.linalg.norm(xb[off:off + N * N] - xs.reshape(-1, order="F")) / np.linalg.norm(xs)
        same = np.array_equal(np.asarray(xb[off:off + N * N]), np.asarray(xs).reshape(-1, order="F"))
        print(f"  N={N} it {len(r)-1}/{len(rs)-1} status {s}/{ss} x diff {e:.2e} bitwise {same}", flush=True)
        assert s == 1 and ss == 1 and len(r) == len(rs)
        assert np.array_equal(np.asarray(r), np.asarray(rs))
        assert same
        off += N * N


def test_cgne_uneven_batch_slices_equal_single(mm):
    """Five slices of the uneven 16-coil batch (different N, M and grid
    sizes, 150 = 2*3*5^2 and 196 = 2^2*7^2 among them), 25 iterations as one
    batched, graph-captured solve: bitwise equal to five single-slice solves,
    and the first slice within the FP64 oracle's band (NRMSE <= 3 d, d the
    drift of the oracle with complex64-rounded operator I/O)."""
    print("[case] cgne 5 slices of the uneven batch, C=16, 25 iterations", flush=True)
    B = mm.workloads.batch_problems(seed=4, nslice=64, C=16)
    by_grid = {}
    for b in B:
        by_grid.setdefault(mm.workloads.grid_size(b["N"]), b)
    # five distinct grid sizes: 150, 196 and the three smallest others
    pick = [by_grid[150], by_grid[196]] + [by_grid[g] for g in sorted(by_grid) if g not in (150, 196)][:3]
    opt = {"resThresh": None, "xChangeThresh": None}
    ys, orc = [], []
    for b in pick:
        N = b["N"]
        So, Ro = NufftPlan([N, N], b["kloc"], None, 1.5, 4), RcvrPlan(b["maps"])
        ys.append(mm.workloads.add_noise(o_mrSampFunc(b["phantom"], "notransp", So, Ro), b["rng"]))
        orc.append((So, Ro))
    E = mm.senseOp(mm.nufftOp.batch([([b["N"]] * 2, b["kloc"], None) for b in pick], 1.5, 4),
                   mm.rcvrOp.batch([b["maps"] for b in pick]))
    ycat = np.concatenate([y.reshape(-1, order="F") for y in ys])
    xb, rb, *_, sb = mm.cgne(E, ycat, np.zeros(E.image_shape, complex), 25, opt)
    off = 0
    for i, (b, y) in enumerate(zip(pick, ys)):
        N = b["N"]
        Es = mm.senseOp(mm.nufftOp([N, N], b["kloc"], None, 1, 1.5, 4), mm.rcvrOp(b["maps"]))
        xs, rs, *_, ss = mm.cgne(Es, y, np.zeros((N, N), complex), 25, opt)
        xbs = np.asarray(xb[off:off + N * N])
        same = np.array_equal(xbs, np.asarray(xs).reshape(-1, order="F"))
        print(f"  N={N} G={mm.workloads.grid_size(N)} M={b['M']}: it {len(rb[i]) - 1}/{len(rs) - 1} "
              f"bitwise {same}", flush=True)
        assert sb[i] == 0 and ss == 0 and np.array_equal(np.asarray(rb[i]), np.asarray(rs)) and same
        if i == 0:
            So, Ro = orc[0]
            A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
            A32 = lambda x, tr: _c64(o_mrSampFunc(_c64(x), tr, So, Ro))
            x0 = np.zeros((N, N), complex)
            xo, *_ = o_cgne(A, y, x0, 25, opt)
            x3, *_ = o_cgne(A32, y, x0, 25, opt)
            d = np.linalg.norm(x3 - xo) / np.linalg.norm(xo)
            e = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
            print(f"  slice 0 vs FP64 oracle: nrmse {e:.2e}, FP32-I/O drift d {d:.2e}", flush=True)
            assert e < max(3 * d, 1e-4)
        off += N * N


def test_cgne_slice_groups_bitwise(mm):
    """The normal operator run group by group (option nu_slice_group, with the
    alternating slice order and finished slices dropping out) gives the same
    batched solve as one group: the same kernels compute every slice in the
    same way, so iterates, histories and iteration counts are bitwise equal."""
    print("[case] cgne 7 shared-trajectory slices, slice groups 0 / 3 / 2", flush=True)
    rng = np.random.default_rng(12)
    N, M, C, P = 48, 3000, 4, 7
    kl = mm.workloads.spiral(M, N // 2)
    mp, ys = [], []
    for _ in range(P):
        maps = mm.workloads.gaussian_coils(N, N, C, rng)
        ph = mm.workloads.ellipse_phantom(N, N, rng)
        ys.append(mm.workloads.add_noise(o_mrSampFunc(ph, "notransp", NufftPlan([N, N], kl, None, 1.5, 4),
                                                      RcvrPlan(maps)), rng))
        mp.append(maps)
    E = mm.senseOp(mm.nufftOp.batch([([N, N], kl, None)] * P, 1.5, 4), mm.rcvrOp.batch(mp))
    ycat = np.concatenate([y.reshape(-1, order="F") for y in ys])
    opt = {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None}
    runs = []
    for g in (0, 3, 2):
        E.set_option("nu_slice_group", g)
        assert E.get_option("nu_slice_group") == g
        x, r, *_, st = mm.cgne(E, ycat, None, 60, opt)
        runs.append((np.asarray(x), [np.asarray(h) for h in r], list(st)))
        print(f"  group {g}: iterations {[len(h) - 1 for h in r]} status {st}", flush=True)
    x0, r0, s0 = runs[0]
    assert len(set(len(h) for h in r0)) > 1   # slices stop at different iterations
    for x, r, st in runs[1:]:
        assert st == s0 and np.array_equal(x, x0)
        assert all(np.array_equal(a, b) for a, b in zip(r, r0))


def test_cgne_after_inner_option_change(mm):
    """A solve, a tuning option changed on the inner NUFFT handle (which
    rebuilds the work tables a captured iteration graph points at), a second
    solve on the same SENSE operator: equal to a solve with a fresh operator
    configured the same way."""
    print("[case] cgne, set_option on the inner NUFFT between two solves", flush=True)
    So, Ro, Sg, Rg, y, ph = _problem(mm, seed=13)
    opt = {"resThresh": None, "xChangeThresh": None}
    E = mm.senseOp(Sg, Rg)
    x0 = np.zeros((64, 64), complex)
    mm.cgne(E, y, x0, 12, opt)
    Sg.set_option("nu_stage_samples", 64)
    Sg.set_option("nu_stage_entries", 512)
    Sg.set_option("nu_degrid_block", 32)
    x1, r1, *_ = mm.cgne(E, y, x0, 12, opt)
    Sf = mm.nufftOp([64, 64], Sg._kl[0], None, 1, 1.5, 4)
    for k, v in (("nu_stage_samples", 64), ("nu_stage_entries", 512), ("nu_degrid_block", 32)):
        Sf.set_option(k, v)
    x2, r2, *_ = mm.cgne(mm.senseOp(Sf, mm.rcvrOp(Ro.maps)), y, x0, 12, opt)
    print(f"  equal to a fresh operator: x {np.array_equal(x1, x2)} resSq {np.array_equal(r1, r2)}", flush=True)
    assert np.array_equal(x1, x2) and np.array_equal(r1, r2)


def test_cgne_stopping_options_between_solves(mm):
    """Two solves on one operator that differ only in xChangeThresh (the
    iteration graph is captured by the first): the second stops where a
    fresh operator with that threshold stops (status 10)."""
    print("[case] cgne, xChangeThresh changed between two solves", flush=True)
    So, Ro, Sg, Rg, y, ph = _problem(mm, seed=14)
    E = mm.senseOp(Sg, Rg)
    x0 = np.zeros((64, 64), complex)
    _, ra, *_, sa = mm.cgne(E, y, x0, 30, {"resThresh": None, "xChangeThresh": None})
    assert sa == 0 and len(ra) == 31
    _, _, _, _, xd, _, _ = mm.cgne(E, y, x0, 30, {"resThresh": None, "xChangeThresh": None})
    th = float(np.real(xd[12]))   # reached at iteration 12 (xdiff decreases overall)
    xb, rb, *_, sb = mm.cgne(E, y, x0, 30, {"resThresh": None, "xChangeThresh": th * 1.0001})
    Ef = mm.senseOp(mm.nufftOp([64, 64], Sg._kl[0], None, 1, 1.5, 4), mm.rcvrOp(Ro.maps))
    xf, rf, *_, sf = mm.cgne(Ef, y, x0, 30, {"resThresh": None, "xChangeThresh": th * 1.0001})
    print(f"  stop: reused operator it {len(rb) - 1} status {sb}; fresh it {len(rf) - 1} status {sf}", flush=True)
    assert sf == 10 and sb == 10 and len(rb) == len(rf) and len(rb) < 31
    assert np.array_equal(xb, xf)


def test_cgne_highorder_220_32coils_30_iterations(mm):
    """The high-order solve: 220x220, 32 coils, 16 harmonics + B0 + a 70 kHz
    zeroth-order term (max|Phi| > 1e4 rad), on 2400 of the 24000 time points
    spread over the whole readout, 30 zero-init cgne iterations
    (sampHighOrder.m:361-391 inside cgne.m:213-336).  Per iteration the
    device resSq is within 1e-3 of the FP64 oracle wherever complex64
    rounding of the operator I/O alone moves it by < 1e-5; the final NRMSE
    is within 3 d, d that rounding's own drift of the solution."""
    print("[case] cgne high-order 220^2 C=32 K=2400/24000 f0=70 kHz, 30 iterations", flush=True)
    P = mm.workloads.highorder_problem(seed=2, f0_hz=70e3)
    sel = np.linspace(0, P["K"] - 1, 2400).round().astype(int)
    t, h = P["t"][sel], P["h"][:, sel]
    So, Ro = HighOrderPlan(P["b0"], t, h, None, P["grid"]), RcvrPlan(P["maps"])
    So.cache_exp = True
    maxphi = np.max(np.abs(So._kb()))
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    A32 = lambda x, tr: _c64(o_mrSampFunc(_c64(x), tr, So, Ro))
    y = mm.workloads.add_noise(A(P["phantom"], "notransp"), P["rng"])
    E = mm.senseOp(mm.sampHighOrder(P["b0"], t, h, None, P["grid"]), mm.rcvrOp(P["maps"]))
    opt = {"resThresh": None, "xChangeThresh": None}
    x0 = np.zeros((220, 220), complex)
    xg, rg, *_, st = mm.cgne(E, y, x0, 30, opt)
    xo, ro, *_ = o_cgne(A, y, x0, 30, opt)
    x3, r3, *_ = o_cgne(A32, y, x0, 30, opt)
    ro, r3 = np.real(ro), np.real(r3)
    d = np.linalg.norm(x3 - xo) / np.linalg.norm(xo)
    e = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    d3, dg = np.abs(r3 - ro) / ro, np.abs(np.asarray(rg) - ro) / ro
    tight = d3 < 1e-5
    print(f"  max|Phi| {maxphi:.3e} rad; rel resid {np.sqrt(ro[-1] / ro[0]):.2e}; device nrmse {e:.2e}, "
          f"FP32-I/O drift d {d:.2e}; resSq on {tight.sum()} tight iterations: max rel diff "
          f"{dg[tight].max():.2e}", flush=True)
    assert maxphi > 1e4 and st == 0 and len(rg) == 31
    assert tight.sum() >= 10 and np.all(dg[tight] < 1e-3)
    assert e < 3 * d


def test_cgne_tikhonov_terms_and_mse_history(mm):
    """Tikhonov terms in the device solve (cgne.m:171-179, :219-227): gam and
    R as cells with R_1 = identity and R_2 a Cartesian k-space weighting
    (cartOp, a device operator), and the mse history against a ground truth
    (cgne.m:201-204, :271-274); 8 iterations against the FP64 oracle:
    iterate NRMSE <= 1e-4, resSq and mse histories within 1e-3."""
    print("[case] cgne with gam = {eye, Cartesian weighting} and opt.gtruth, N=64 C=4", flush=True)
    So, Ro, Sg, Rg, y, ph = _problem(mm, seed=15)
    kx = np.fft.fftfreq(64)
    mask = (np.hypot(kx[:, None], kx[None, :]) > 0.3).astype(float)   # a high-pass penalty
    A = lambda x, tr: o_mrSampFunc(x, tr, So, Ro)
    gam = [0.02, 0.5]
    eye_o = lambda v, tr: v
    cart_o = lambda v, tr: o_mrSampFunc(v, tr, mask)
    opt = {"resThresh": None, "xChangeThresh": None, "gtruth": ph}
    x0 = np.zeros((64, 64), complex)
    xo, ro, mo, *_ = o_cgne(A, y, x0, 8, opt, gam, [eye_o, cart_o])
    xg, rg, mg, *_, st = mm.cgne(mm.mrSampHandle(Sg, Rg), y, x0, 8, opt, gam, [mm.eye, mm.cartOp(mask)])
    ro, mo = np.real(ro), np.real(mo)
    e = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    dr, dm = np.max(np.abs(rg - ro) / ro), np.max(np.abs(mg - mo) / mo)
    print(f"  x nrmse {e:.2e}; resSq max rel {dr:.2e}; mse max rel {dm:.2e}", flush=True)
    assert st == 0 and e < 1e-4 and dr < 1e-3 and dm < 1e-3
    # without them the solve differs (the terms are applied)
    xn, *_ = mm.cgne(mm.mrSampHandle(Sg, Rg), y, x0, 8, {"resThresh": None, "xChangeThresh": None})
    assert np.linalg.norm(xn - xg) / np.linalg.norm(xg) > 1e-3
