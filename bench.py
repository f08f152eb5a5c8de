#!/usr/bin/env python3
"""CG-SENSE solves/s on one MI355X.

Workload (default, ``--workload warmstart``): the learned warm start against
zero-init CG on 100 held-out 2-D spiral SENSE problems (128x128 complex64
image, 8 Gaussian-ring coils with random phases, 16384-sample variable-density
spiral, Kaiser-Bessel W=4 os=1.5, random 10-ellipse phantom, noise at SNR 40).
One step = the 100 problems solved to relative residual |r|/|b| = 1e-4 in one
multi-slice solve (one launch per stage for all problems, hipGraph per
iteration): the density-compensated adjoint images, the network's x0, and
cgne from x0 with the zero-init fallback.  The zero-init solve of the same
batch is timed the same way and reported beside it.  Network weights are the
ones trained on the device by ``python -m uwobrainhack2026-matmri_amd.warmstart``
(20000 problems whose seeds precede the held-out ones).

``--workload highorder``: the 2-D expanded-encoding problem -- 220x220 complex64 image, 32 Gaussian-ring coils, 24000 spiral
samples at 1 us, 16 real solid harmonics up to 3rd order on a 0.2 m FOV with
a seeded smooth drift, a smooth +-100 Hz B0 map; data = E x_true + complex
Gaussian noise at SNR 40.  One step = one cgne solve from a zero initial
guess that stops at relative residual |r|/|b| = 1e-4 or after 30 iterations;
operators in FP32/complex64, scalars in FP64.  Inputs are resident in HBM
before the timed region.

``--workload nufft``: the 128x128 / 8-coil / 16384-sample spiral problem
(Kaiser-Bessel W=4, os=1.5), at most 20 iterations.

The JSON line carries the roofline of the dominant kernel (timed with HIP
events on the library stream) and a CPU baseline: the FP64 NumPy oracle timed
on a bounded sample of the same operator on this host.
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))


def note(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
sys.path.insert(0, ROOT)

PEAK_FP16_MFMA_TFLOPS = 2500.0   # MI355X dense fp16 matrix peak (the high-order GEMM runs on it)
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=["warmstart", "highorder", "nufft", "batch", "copy", "densemat", "interp",
                                          "finddel", "nufft3"],
                   default="warmstart",
                   help="copy: a plain 1 GiB device copy (HBM counter calibration, not a solve); densemat: "
                        "mrSampFuncMat's dense normal matrix at the high-order problem's full size; interp: the "
                        "high-order solve with sampHighOrder approach 1 (interpolated); finddel: findDelAuto's "
                        "delay estimate in the reference demo's setting; nufft3: a 3-D stack-of-spirals SENSE solve")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--heldout", type=int, default=1000, help="held-out problems of the warm-start workload")
    p.add_argument("--precision", action="store_true",
                   help="highorder: add the full-size precision study against the C/OpenMP FP64 oracle "
                        "(per-application errors and 30-iteration NRMSE; a few CPU minutes)")
    p.add_argument("--no-batch-roofline", action="store_true",
                   help="skip the 64-slice batch roofline of the warm-start line")
    p.add_argument("--hio-variant", type=int, default=None, help="high-order MFMA kernel variant 0..3")
    p.add_argument("--no-autotune", action="store_true",
                   help="skip the untimed autotuner of the NUFFT SENSE operator, or of the interpolated "
                        "operator's NUFFT (mm_op_autotune)")
    p.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                   help="tuning option of the sampling operator (mm_op_set_option), e.g. nu_degrid_t=0")
    return p.parse_args()


def highorder_setup(mm, torch, dev):
    import numpy as np
    P = mm.workloads.highorder_problem(seed=2)
    N, K, C = P["N"], P["K"], P["C"]
    S = mm.sampHighOrder(P["b0"], P["t"], P["h"], None, P["grid"])
    R = mm.rcvrOp(P["maps"])
    E = mm.senseOp(S, R)
    x_true = torch.as_tensor(np.ascontiguousarray(P["phantom"].T), dtype=torch.complex64, device=dev)
    y = E * x_true
    g = torch.Generator(device=dev).manual_seed(2)
    sig = float(torch.sqrt(torch.mean(torch.abs(y) ** 2))) / 40.0
    noise = torch.randn(y.shape, generator=g, device=dev, dtype=torch.float32) + \
        1j * torch.randn(y.shape, generator=g, device=dev, dtype=torch.float32)
    y = (y + sig / np.sqrt(2) * noise).to(torch.complex64).contiguous()
    L = S.Lh + S.Lc + 1
    V = N * N
    flops_fwd = 2.0 * K * V * L + 8.0 * K * V * C
    cfg = {"workload": "highorder-2d: 220x220 c64 image, 32 coils, K=24000 spiral @1us, "
                       "16 SH (<=3rd order, 0.2 m FOV) + B0, cgne to rel. resid 1e-4 (<=30 it), zero init",
           "image": [N, N], "coils": C, "samples": K, "phase_terms": L, "max_iterations": 30,
           "rel_resid": 1e-4}
    return dict(P=P, E=E, S=S, R=R, y=y, x0=torch.zeros(V, dtype=torch.complex64, device=dev), NitMax=30,
                flops=flops_fwd, cfg=cfg, kind="highorder")


def nufft_setup(mm, torch, dev):
    import numpy as np
    P = mm.workloads.nufft_problem(seed=1)
    N, M, C = P["N"], P["M"], P["C"]
    S = mm.nufftOp([N, N], P["kloc"], None, 1, 1.5, 4)
    R = mm.rcvrOp(P["maps"])
    E = mm.senseOp(S, R)
    x_true = torch.as_tensor(np.ascontiguousarray(P["phantom"].T), dtype=torch.complex64, device=dev)
    y = E * x_true
    g = torch.Generator(device=dev).manual_seed(1)
    sig = float(torch.sqrt(torch.mean(torch.abs(y) ** 2))) / 40.0
    noise = torch.randn(y.shape, generator=g, device=dev) + 1j * torch.randn(y.shape, generator=g, device=dev)
    y = (y + sig / np.sqrt(2) * noise).to(torch.complex64).contiguous()
    V, G = N * N, 192
    bytes_app = 8.0 * (V + 2 * V * C + 4 * G * G * C + M + M * C)
    cfg = {"workload": "nufft-2d: 128x128 c64 image, 8 coils, 16384-sample VD spiral, KB W=4 os=1.5, "
                       "cgne to rel. resid 1e-4 (<=20 it), zero init",
           "image": [N, N], "coils": C, "samples": M, "max_iterations": 20, "rel_resid": 1e-4}
    return dict(P=P, E=E, S=S, R=R, y=y, x0=torch.zeros(V, dtype=torch.complex64, device=dev), NitMax=20,
                bytes=bytes_app, cfg=cfg, kind="nufft")


def batch_setup(mm, torch, dev):
    import numpy as np
    Ps = mm.workloads.batch_problems(seed=4, nslice=64, C=16)
    S = mm.nufftOp.batch([([p["N"], p["N"]], p["kloc"], None) for p in Ps], 1.5, 4)
    R = mm.rcvrOp.batch([p["maps"] for p in Ps])
    E = mm.senseOp(S, R)
    xt = np.concatenate([p["phantom"].reshape(-1, order="F") for p in Ps])
    y = E * torch.as_tensor(xt, dtype=torch.complex64, device=dev)
    g = torch.Generator(device=dev).manual_seed(4)
    parts, off = [], 0
    for p in Ps:
        n = p["M"] * 16
        ys = y[off:off + n]
        sig = float(torch.sqrt(torch.mean(torch.abs(ys) ** 2))) / 40.0
        nz = torch.randn(n, generator=g, device=dev) + 1j * torch.randn(n, generator=g, device=dev)
        parts.append(ys + sig / np.sqrt(2) * nz)
        off += n
    y = torch.cat(parts).to(torch.complex64).contiguous()
    V = int(sum(p["N"] ** 2 for p in Ps))
    cfg = {"workload": "batch: 64 uneven 2-D slices (N 96..160 even, M 6000..20000 spiral samples), "
                       "16 coils, KB W=4 os=1.5, one launch per stage for all slices, "
                       "cgne to rel. resid 1e-4 (<=25 it) per slice, zero init",
           "slices": 64, "coils": 16, "voxels_total": V, "samples_total": int(sum(p["M"] for p in Ps)),
           "max_iterations": 25, "rel_resid": 1e-4}
    return dict(P=Ps, E=E, S=S, R=R, y=y, x0=torch.zeros(V, dtype=torch.complex64, device=dev), NitMax=25,
                cfg=cfg, kind="batch", solves_per_step=64)


def cpu_baseline_batch(Ps, seconds):
    """Per-slice FP64 solves (25 iterations) on as many slices as fit the
    budget, scaled to the batch: the C/OpenMP restatement when built, else
    the NumPy oracle E'E scaled to a solve."""
    import numpy as np
    cref = _cref()
    if cref is not None:
        t_tot, n = 0.0, 0
        while t_tot < seconds and n < len(Ps):
            p = Ps[n]
            t0 = time.perf_counter()
            E = cref.SenseRef(p["N"], p["kloc"], None, p["maps"], 1.5, 4)
            y = E.forward(p["phantom"])
            t1 = time.perf_counter()
            E.cgne(y, np.zeros((p["N"], p["N"]), complex), 25, resThreshRel=1e-8)
            t_tot += (t1 - t0) + (time.perf_counter() - t1)
            n += 1
        return {"value": n / t_tot, "unit": "solves/s", "cores": cref.num_threads(), "kind": "port",
                "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c): {n} of the 64 slices solved "
                          f"one after another (25 iterations each, plan included), {t_tot:.1f} s"}
    from oracle.mrsamp import mrSampFunc
    from oracle.nufft import NufftPlan
    from oracle.rcvr import RcvrPlan
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    t_tot, n_done, work_done = 0.0, 0, 0
    s = 0
    while t_tot < seconds and s < len(Ps):
        p = Ps[s]
        So = NufftPlan([p["N"], p["N"]], p["kloc"], None, 1.5, 4)
        Ro = RcvrPlan(p["maps"])
        t0 = time.perf_counter()
        mrSampFunc(mrSampFunc(p["phantom"], "notransp", So, Ro), "transp", So, Ro)
        t_tot += time.perf_counter() - t0
        n_done += 1
        s += 1
    per_slice_app = t_tot / (2 * n_done)
    apps = 2 * 25 + 5
    return {"value": 1.0 / (per_slice_app * apps), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"FP64 NumPy oracle E'E on {n_done} of the 64 slices, scaled to one 25-iteration "
                      f"solve per slice ({apps} applications)"}


def cpu_highorder_per_app(b0, t, h, hc, grid, maps, x, seconds):
    """Seconds per operator application of the C/OpenMP FP64 restatement
    (oracle/c/mmref.c, direct approach: the phase and exp(i Phi) recomputed
    every application, as the reference's direct worker evaluates
    exp(1i*kbase) per call) of a high-order SENSE operator: forward +
    adjoint on a bounded sample of time points, scaled to all of them.
    Returns (seconds, samples timed, seconds measured, threads) or None."""
    cref = _cref()
    if cref is None:
        return None
    K = len(t)
    ks = 256
    while True:
        E = cref.HighOrderSenseRef(b0, t[:ks], h[:, :ks], None if hc is None else hc[:, :ks], grid, maps)
        t0 = time.perf_counter()
        y = E.forward(x)
        E.adjoint(y)
        dt = time.perf_counter() - t0
        if dt > seconds / 3 or ks >= K:
            break
        ks = min(K, int(ks * max(2.0, (seconds / 3) / max(dt, 1e-3))))
    return dt / 2 * K / ks, ks, dt, cref.num_threads()


def cpu_baseline_highorder(P, seconds):
    """The C/OpenMP FP64 restatement of the same SENSE operator
    (cpu_highorder_per_app), scaled to a 30-iteration solve (65
    applications)."""
    r = cpu_highorder_per_app(P["b0"], P["t"], P["h"], None, P["grid"], P["maps"], P["phantom"], seconds)
    if r is None:
        return None
    per_app, ks, dt, thr = r
    K = P["K"]
    apps = 2 * 30 + 5  # cgne with 30 iterations: 32 forward + 33 adjoint applications
    return {"value": 1.0 / (per_app * apps), "unit": "solves/s", "cores": thr, "kind": "port",
            "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c), one forward + one adjoint over {ks} of "
                      f"{K} samples ({dt:.2f} s), scaled to one 30-iteration solve ({apps} applications)"}


def highorder_precision(mm, torch, dev, NitMax=30):
    """The precision study on the high-order operator with a 70 kHz
    zeroth-order term (max |Phi| > 1e4 rad at late samples): per-application
    relative L2 error of the device forward and adjoint, and the NRMSE of a
    30-iteration device cgne against the same 30 iterations of the FP64
    C/OpenMP oracle (exp(i Phi) cached between its applications), from zero
    and with no early stop, at the full 220 x 220 / 24000 / 32-coil size."""
    import numpy as np
    cref = _cref()
    if cref is None:
        return None
    P = mm.workloads.highorder_problem(seed=2, f0_hz=70e3)
    t0 = time.perf_counter()
    Eo = cref.HighOrderSenseRef(P["b0"], P["t"], P["h"], None, P["grid"], P["maps"], cache_exp=True)
    S = mm.sampHighOrder(P["b0"], P["t"], P["h"], None, P["grid"])
    E = mm.senseOp(S, mm.rcvrOp(P["maps"]))
    rng = np.random.default_rng(3)
    N, K, C = P["N"], P["K"], P["C"]
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    yr = rng.standard_normal((K, C)) + 1j * rng.standard_normal((K, C))
    dev_fwd = (E * torch.as_tensor(np.ascontiguousarray(x.T), dtype=torch.complex64, device=dev)).cpu().numpy()
    dev_adj = (E.H * torch.as_tensor(np.ascontiguousarray(yr.T), dtype=torch.complex64, device=dev)).cpu().numpy()
    yo, xo = Eo.forward(x), Eo.adjoint(yr)
    ef = float(np.linalg.norm(dev_fwd.reshape(C, K).T - yo) / np.linalg.norm(yo))
    ea = float(np.linalg.norm(dev_adj.reshape(N, N).T - xo) / np.linalg.norm(xo))
    y = Eo.forward(P["phantom"])
    y = y + np.sqrt(np.mean(np.abs(y) ** 2)) / 40.0 / np.sqrt(2) * (rng.standard_normal(y.shape)
                                                                    + 1j * rng.standard_normal(y.shape))
    note("precision: per-application errors done; oracle cgne")
    cref.set_verbose(True)
    xo30, ro, nit = Eo.cgne(y, np.zeros((N, N), complex), NitMax)
    # the same FP64 solve behind complex64 operator I/O: the drift that
    # single-precision operator inputs and outputs alone cause
    note("precision: oracle cgne with complex64 operator I/O")
    Eo.set_round_io(True)
    x3, r3, _ = Eo.cgne(y, np.zeros((N, N), complex), NitMax)
    Eo.set_round_io(False)
    cref.set_verbose(False)
    yt = torch.as_tensor(np.ascontiguousarray(y.T), dtype=torch.complex64, device=dev)
    xg, rg, *_ = mm.cgne(E, yt, torch.zeros(N * N, dtype=torch.complex64, device=dev), NitMax,
                         {"resThresh": None, "xChangeThresh": None})
    xg = xg.cpu().numpy().reshape(N, N).T
    rg = np.asarray(rg, dtype=np.float64).reshape(-1)
    nrmse = float(np.linalg.norm(xg - xo30) / np.linalg.norm(xo30))
    drift = float(np.linalg.norm(x3 - xo30) / np.linalg.norm(xo30))
    dg = np.abs(rg[: len(ro)] - ro) / ro
    d3 = np.abs(r3[: len(ro)] - ro) / ro
    from oracle.highorder import HighOrderPlan
    late = slice(K - 64, K)
    maxphi = float(np.max(np.abs(HighOrderPlan(P["b0"], P["t"][late], P["h"][:, late], None, P["grid"]).phase())))
    return {"workload": "precision: 220x220, 32 coils, K=24000 @1us, 16 SH + B0 + 70 kHz f0 term",
            "max_abs_phase_rad_last_64_samples": maxphi,
            "forward_rel_l2": ef, "adjoint_rel_l2": ea,
            "cgne_iterations": int(nit), "nrmse_30_iterations": nrmse,
            "nrmse_30_iterations_fp64_with_c64_io": drift,
            "nrmse_ratio_device_over_c64_io": nrmse / drift if drift > 0 else None,
            "resSq_rel_diff_per_iteration": [float(v) for v in dg],
            "resSq_rel_diff_per_iteration_c64_io": [float(v) for v in d3],
            "rel_resid_final": float(np.sqrt(ro[-1] / ro[0])),
            "oracle": f"C/OpenMP FP64 (oracle/c/mmref.c), {cref.num_threads()} threads",
            "seconds": time.perf_counter() - t0}


def cpu_baseline_nufft(P, seconds):
    import numpy as np
    cref = _cref()
    if cref is not None:
        N = P["N"]
        E = cref.SenseRef(N, P["kloc"], None, P["maps"], 1.5, 4)
        y = E.forward(P["phantom"])
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            E.cgne(y, np.zeros((N, N), complex), 20, resThreshRel=1e-8)
            n += 1
        el = time.perf_counter() - t0
        return {"value": n / el, "unit": "solves/s", "cores": cref.num_threads(), "kind": "port",
                "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c): {n} solves of the same problem "
                          f"(zero init, 20 iterations), {el:.1f} s"}
    from oracle.mrsamp import mrSampFunc
    from oracle.nufft import NufftPlan
    from oracle.rcvr import RcvrPlan
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    So = NufftPlan([P["N"], P["N"]], P["kloc"], None, 1.5, 4)
    Ro = RcvrPlan(P["maps"])
    x = P["phantom"]
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        mrSampFunc(mrSampFunc(x, "notransp", So, Ro), "transp", So, Ro)
        n += 1
    per_app = (time.perf_counter() - t0) / (2 * n)
    apps = 2 * 20 + 5
    return {"value": 1.0 / (per_app * apps), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"FP64 NumPy oracle E'E on the full problem, {2 * n} applications, "
                      f"scaled to one 20-iteration solve ({apps} applications)"}


def nufft_stage_bytes(N1, N2, G1, G2, M, C):
    """Algorithmic HBM bytes of each NUFFT SENSE stage for one slice and one
    application: compulsory reads and writes of complex64 vectors (8 bytes a
    value; plan constants such as interpolation weights are not counted).
    The forward row pass writes only the N2 image rows of the padded grid and
    the column passes read N2 rows and write G2 (forward) or the reverse
    (adjoint)."""
    V, G = N1 * N2, G1 * G2
    return {
        "nu_fwd_rows": 8.0 * (V + V * C + G1 * N2 * C),   # image, maps in; N2 grid rows out
        "nu_cols": 8.0 * G1 * (N2 + G2) * C,              # N2 rows in, G2 out (or the reverse)
        "nu_degrid": 8.0 * (G * C + M + M * C),           # grid, k-space locations in; samples out
        "nu_grid": 8.0 * (M * C + M + G * C),             # samples, locations in; grid out
        "nu_adj_rows": 8.0 * (G1 * N2 * C + V * C + V),   # N2 grid rows, maps in; image out
        "nu_transpose_y": 2 * 8.0 * M * C,
    }


def read_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 counter
    summary (profiles/pmc_<kernel>_<workload>.md, line 'bytes_per_dispatch: N'),
    or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.md")
    try:
        for line in open(path):
            if line.startswith("bytes_per_dispatch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def oracle_nrmse_warmstart(fam, host, prob, xw, xz, nprob=3):
    """NRMSE of the device warm-start and zero-init solutions (rel. resid
    1e-4) against the FP64 oracle solution run far past it (untimed)."""
    import numpy as np
    from oracle.mrsamp import mrSampFunc
    from oracle.nufft import NufftPlan
    from oracle.rcvr import RcvrPlan
    from oracle.solvers import cgne as o_cgne
    N, M, C = fam.p["N"], fam.p["M"], fam.p["C"]
    So = NufftPlan([N, N], fam.kloc, None, fam.p["os"], fam.p["kwidth"])
    Y = prob["y"].reshape(fam.batch, C, M)[:nprob].cpu().numpy().astype(np.complex128)
    XW = xw.reshape(fam.batch, N, N)[:nprob].cpu().numpy()
    XZ = xz.reshape(fam.batch, N, N)[:nprob].cpu().numpy()
    ew, ez = [], []
    for n in range(nprob):
        Ro = RcvrPlan(host[0][n])
        A = lambda x, tr: mrSampFunc(x, tr, So, Ro)
        y = Y[n].T
        b = A(y, "transp")
        xo, *_ = o_cgne(A, y, np.zeros((N, N), complex), 150,
                        {"resThresh": 1e-14 * float(np.real(np.vdot(b, b))), "xChangeThresh": None})
        no = np.linalg.norm(xo)
        ew.append(float(np.linalg.norm(XW[n].T - xo) / no))
        ez.append(float(np.linalg.norm(XZ[n].T - xo) / no))
    return {"warm": float(np.mean(ew)), "zero": float(np.mean(ez)), "problems": nprob,
            "reference": "FP64 NumPy oracle cgne, zero init, 150 iterations (|r|/|b| ~ 1e-5)"}


def nufft_roofline(mm, E, dev, slices, workload, iters=20):
    """Dominant kernel of a NUFFT SENSE normal operator (HIP events on the
    library stream, every stage timed) and its HBM roofline: algorithmic
    bytes = per-slice stage formula summed over `slices` [(N, M, C)]."""
    import ctypes as C
    import torch
    ms = C.c_double()
    name = C.create_string_buffer(64)
    ni = int(sum(n * n for n, _, _ in slices))
    xi = torch.zeros(ni, dtype=torch.complex64, device=dev)
    xo = torch.empty_like(xi)
    L = mm.lib()
    rc = L.mm_op_profile(E._plan.h, 2, xi.data_ptr(), xo.data_ptr(), 1, iters, C.byref(ms), name, 64)
    if rc != 0 or ms.value <= 0:
        return None, {}
    stages = {}
    k = 0
    while True:
        sm, sl, sn = C.c_double(), C.c_int(), C.create_string_buffer(64)
        if L.mm_op_profile_stage(E._plan.h, k, C.byref(sm), C.byref(sl), sn, 64) != 0:
            break
        stages[sn.value.decode()] = {"ms_per_launch": sm.value, "launches_per_application": sl.value}
        k += 1
    kname = name.value.decode()
    byts = 0.0
    for n, m, c in slices:
        g = mm.workloads.grid_size(n, 1.5)
        byts += nufft_stage_bytes(n, n, g, g, m, c).get(kname, 0.0)
    if byts <= 0:
        return None, stages
    # the formula covers one pass over all slices (the column FFT makes two
    # passes per application); a launch covers one slice group's share
    passes = 2 if kname == "nu_cols" else 1
    launches = stages.get(kname, {}).get("launches_per_application", passes) or passes
    byts = byts * passes / launches
    ach = byts / (ms.value * 1e-3) / 1e9
    kern_file = {"nu_grid": "nu_grid"}.get(kname, kname)
    return ({"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
             "traffic": read_traffic(kern_file, workload), "kernel": kname, "ms_per_launch": ms.value,
             "algorithmic_bytes_per_launch": byts}, stages)


def copy_gbps(torch, dev, reps=5):
    """Measured device copy rate (1 GiB read + 1 GiB written): the practical
    HBM ceiling the roofline fractions are also quoted against."""
    n = 1 << 28
    x = torch.empty(n, dtype=torch.float32, device=dev).uniform_()
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    del x, y
    torch.cuda.empty_cache()
    return 2 * 4 * n * reps / el / 1e9


def annotate_roofline(roof, copy, llc_mb=256.0):
    """Adds the fraction of the measured copy rate and whether one launch's
    compulsory bytes fit in the 256 MB memory-side cache (then part of the
    input can come from it, and the algorithmic rate is cache-assisted)."""
    if not roof:
        return roof
    roof["copy_GBps"] = copy
    roof["frac_of_copy"] = roof["achieved"] / copy
    mb = roof["algorithmic_bytes_per_launch"] / 1e6
    roof["cache_assisted"] = bool(mb < 2 * llc_mb)
    roof["note"] = (f"{mb:.0f} MB of compulsory traffic per launch; the previous stage's output (about half of "
                    f"it) can still sit in the {llc_mb:.0f} MB memory-side cache, so the algorithmic rate is "
                    "cache-assisted" if roof["cache_assisted"] else
                    f"{mb:.0f} MB per launch, well above the {llc_mb:.0f} MB memory-side cache")
    return roof


def batch_roofline(mm, torch, dev, copy):
    """Dominant-kernel roofline of the uneven 64-slice, 16-coil batch (about
    1.4 GB per normal-operator application): a shape the memory-side cache
    cannot hold."""
    Ps = mm.workloads.batch_problems(seed=4, nslice=64, C=16)
    S = mm.nufftOp.batch([([p["N"], p["N"]], p["kloc"], None) for p in Ps], 1.5, 4)
    E = mm.senseOp(S, mm.rcvrOp.batch([p["maps"] for p in Ps]))
    roof, stages = nufft_roofline(mm, E, dev, [(p["N"], p["M"], 16) for p in Ps], "batch")
    if roof:
        roof = annotate_roofline(roof, copy)
        roof["workload"] = "batch: 64 uneven slices (N 96..160), 16 coils, default tuning options"
        roof["normal_operator_applications_per_s"] = applications_per_s(stages)
    del E, S
    torch.cuda.empty_cache()
    return roof


def toeplitz_normal(mm, fam, prob, dev, slices):
    """The Toeplitz form of the same SENSE normal operator (nufftOp.prepToep,
    nufftOp.m:410-432: row and column FFTs on a grid of at least 2N per side
    around a kernel-spectrum multiply, no gridding): per-stage times and
    applications per second on the workload's slices, and the HBM roofline
    of its dominant stage (algorithmic bytes: the same per-stage formulas on
    the Toeplitz grid, the multiply reading and writing the grid once).  The
    solves use the gridding form: the two differ by the interpolation
    kernel's approximation (the reference's own bound is rms 1e-3)."""
    import ctypes as C
    import torch
    t0 = time.perf_counter()
    T = fam.S.prepToep()
    t_prep = time.perf_counter() - t0
    E = mm.senseOp(T, prob["R"])
    ms = C.c_double()
    name = C.create_string_buffer(64)
    ni = int(sum(n * n for n, _, _ in slices))
    xi = torch.zeros(ni, dtype=torch.complex64, device=dev)
    xo = torch.empty_like(xi)
    L = mm.lib()
    if L.mm_op_profile(E._plan.h, 2, xi.data_ptr(), xo.data_ptr(), 1, 20, C.byref(ms), name, 64) != 0:
        return None
    stages = {}
    k = 0
    while True:
        sm, sl, sn = C.c_double(), C.c_int(), C.create_string_buffer(64)
        if L.mm_op_profile_stage(E._plan.h, k, C.byref(sm), C.byref(sl), sn, 64) != 0:
            break
        stages[sn.value.decode()] = {"ms_per_launch": sm.value, "launches_per_application": sl.value}
        k += 1
    kname = name.value.decode()
    byts = 0.0
    for n, m, c in slices:
        g = mm.workloads.grid_size(n, 2.0)
        f = nufft_stage_bytes(n, n, g, g, m, c)
        f["nu_toep_mul"] = 2.0 * 8 * g * g * c          # grid in and out
        f["nu_cols_toep"] = 2.0 * 8 * g * n * c         # N2 rows in, N2 rows out
        byts += f.get(kname, 0.0)
    passes = 2 if kname == "nu_cols" else 1
    launches = stages.get(kname, {}).get("launches_per_application", passes) or passes
    byts = byts * passes / launches
    ach = byts / (ms.value * 1e-3) / 1e9 if byts > 0 else None
    dom_bytes = byts
    del E, T
    torch.cuda.empty_cache()
    return {"applications_per_s": applications_per_s(stages), "stages": stages, "prep_s": t_prep,
            "grid": [mm.workloads.grid_size(slices[0][0], 2.0)] * 2,
            "dominant": {"kernel": kname, "ms_per_launch": ms.value, "algorithmic_bytes_per_launch": dom_bytes,
                         "achieved": ach, "unit": "GB/s",
                         "frac": ach / PEAK_HBM_GBS if ach else None}}


def cpu_baseline_toeplitz(fam, P, seconds=10.0):
    """The CPU normal operator of the same SENSE slices: forward + adjoint of
    the C/OpenMP FP64 restatement (oracle/c/mmref.c; the gridding form the
    reference's CPU solves apply), timed on as many of the 8-coil 128 x 128
    slices as fit `seconds` and scaled to applications of the P-slice
    batch."""
    import numpy as np
    cref = _cref()
    if cref is None:
        return None
    N, C = fam.p["N"], fam.p["C"]
    rng = np.random.default_rng(0)
    maps = rng.standard_normal((N, N, C)) + 1j * rng.standard_normal((N, N, C))
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    E = cref.SenseRef(N, fam.kloc, None, maps, fam.p["os"], fam.p["kwidth"])
    n, t_tot = 0, 0.0
    while t_tot < seconds:
        t0 = time.perf_counter()
        E.adjoint(E.forward(x))
        t_tot += time.perf_counter() - t0
        n += 1
    return {"value": n / t_tot / P, "unit": "normal-operator applications/s (whole batch)",
            "cores": cref.num_threads(), "kind": "port",
            "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c) E'E of one slice, {n} applications in "
                      f"{t_tot:.1f} s, scaled to the {P}-slice batch"}


def applications_per_s(stages):
    """Normal-operator applications (E'E of every slice of the batch) per
    second from the per-stage launch times of one application."""
    t = sum(v["ms_per_launch"] * v["launches_per_application"] for v in stages.values())
    return 1e3 / t if t > 0 else None


def _cref():
    """The C/OpenMP FP64 restatement (oracle/c/mmref.c), or None if not built."""
    try:
        from oracle import cref
        cref.lib()
        return cref
    except Exception:
        return None


def cpu_baseline_warmstart(fam, host, prob, seconds):
    """Zero-init FP64 cgne to relative residual 1e-4 on as many of the
    held-out problems as fit the time budget (the reference CPU path has no
    learned warm start): the C/OpenMP restatement when it is built, else the
    NumPy oracle."""
    import numpy as np
    N, M, C = fam.p["N"], fam.p["M"], fam.p["C"]
    Y = prob["y"].reshape(fam.batch, C, M).cpu().numpy().astype(np.complex128)
    cref = _cref()
    t_tot, n, its = 0.0, 0, []
    if cref is not None:
        while t_tot < seconds and n < fam.batch:
            t0 = time.perf_counter()
            E = cref.SenseRef(N, fam.kloc, None, host[0][n], fam.p["os"], fam.p["kwidth"])
            x, res, nit = E.cgne(Y[n].T, np.zeros((N, N), complex), 200, resThreshRel=1e-8)
            t_tot += time.perf_counter() - t0
            its.append(nit)
            n += 1
        return {"value": n / t_tot, "unit": "solves/s", "cores": cref.num_threads(), "kind": "port",
                "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c), zero-init cgne to rel. resid 1e-4 "
                          f"on {n} of the held-out problems ({np.mean(its):.1f} iterations on average, plan "
                          f"included), {t_tot:.1f} s"}
    from oracle.mrsamp import mrSampFunc
    from oracle.nufft import NufftPlan
    from oracle.rcvr import RcvrPlan
    from oracle.solvers import cgne as o_cgne
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    So = NufftPlan([N, N], fam.kloc, None, fam.p["os"], fam.p["kwidth"])
    while t_tot < seconds and n < fam.batch:
        Ro = RcvrPlan(host[0][n])
        y = Y[n].T
        A = lambda x, tr: mrSampFunc(x, tr, So, Ro)
        t0 = time.perf_counter()
        b = A(y, "transp")
        x, res, *_ = o_cgne(A, y, np.zeros((N, N), complex), 200,
                            {"resThresh": 1e-8 * float(np.real(np.vdot(b, b))), "xChangeThresh": None})
        t_tot += time.perf_counter() - t0
        its.append(len(res) - 1)
        n += 1
    return {"value": n / t_tot, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"FP64 NumPy oracle zero-init cgne to rel. resid 1e-4 on {n} of the held-out "
                      f"problems ({np.mean(its):.1f} iterations on average), {t_tot:.1f} s"}


def run_warmstart(a, mm, torch, dev):
    import ctypes as C
    import numpy as np
    ws = mm.warmstart
    net = mm.WarmStartNet.load()
    P = 100
    NCH = max(1, a.heldout // P)   # 100-problem chunks of the held-out set
    fam = ws.ProblemFamily(P)
    for kv in a.set:
        k, v = kv.split("=")
        fam.S.set_option(k, int(v))
    N, M, Cc = fam.p["N"], fam.p["M"], fam.p["C"]
    seed0 = ws.TRAIN_SEED0 + ws.HELDOUT_OFFSET
    chunks = []
    for c in range(NCH):
        seeds = list(range(seed0 + c * P, seed0 + (c + 1) * P))
        host = fam.host_batch(seeds)
        chunks.append((fam.make(seeds, host), host))
        note(f"held-out chunk {c + 1}/{NCH} on the device")
    prob, host = chunks[0]
    E, R, y = prob["E"], prob["R"], prob["y"]
    tuned = None if a.no_autotune or a.set else E.autotune(10)   # the NUFFT plan all chunks share
    NitMax, rel = 200, 1e-4
    zeros = torch.zeros((P, N, N), dtype=torch.complex64, device=dev)

    def warm_step(j):
        pr = chunks[j % NCH][0]
        xin = ws.input_image(pr["E"], pr["R"], pr["y"], fam.dcf_t).reshape(P, N, N)
        x0 = ws.warm_x0(net, xin)
        return ws.solve(pr["E"], pr["y"], x0, rel, NitMax, fallback=True)

    def zero_step(j):
        pr = chunks[j % NCH][0]
        return ws.solve(pr["E"], pr["y"], zeros, rel, NitMax)

    # one untimed pass over every held-out problem: iteration counts of the
    # whole held-out set (and every chunk's solver graph captured)
    every = {"zero": [], "warm": [], "fallbacks": 0, "b2": [], "rel_w": []}
    for j in range(NCH):
        _, nz, _, _, hz = zero_step(j)
        _, nw, _, fb, hw = warm_step(j)
        every["zero"].append(nz)
        note(f"priming pass: chunk {j + 1}/{NCH} solved")
        every["warm"].append(nw)
        every["fallbacks"] += int(fb.sum())
        b2 = np.array([h[0] for h in hz])
        every["rel_w"].append(np.sqrt(np.array([h[-1] for h in hw]) / b2))
    it_zero, it_warm = np.concatenate(every["zero"]), np.concatenate(every["warm"])
    res = {}
    note("timed steps")
    for name, fn in (("zero", zero_step), ("warm", warm_step)):
        for j in range(a.warmup):
            fn(j)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nits, fbs = [], 0
        for j in range(a.steps):
            x, nit, st, fb, hist = fn(j)
            nits.append(nit)
            fbs += int(fb.sum())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        nits = np.concatenate(nits)
        res[name] = dict(el=el, nit=nits, fallbacks=fbs, status=np.asarray(st), x=x)
    rel_w = np.concatenate(every["rel_w"])
    # the chunk the last timed step solved, for the oracle comparison
    prob, host = chunks[(a.steps - 1) % NCH]
    E, R, y = prob["E"], prob["R"], prob["y"]
    # network inference alone
    xin = ws.input_image(E, R, y, fam.dcf_t).reshape(P, N, N)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        ws.warm_x0(net, xin)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t0) / 10
    note("roofline")
    roof, stages = nufft_roofline(mm, E, dev, [(N, M, Cc)] * P, "warmstart")
    zw, zz = res["warm"], res["zero"]
    cfg = {"workload": f"warmstart: {NCH * P} held-out 2-D spiral SENSE problems (128x128 c64, 8 coils, "
                       "M=16384 VD spiral, KB W=4 os=1.5, SNR 40), 100 per step (step j solves chunk j mod "
                       f"{NCH}) as one multi-slice solve to rel. resid 1e-4; x0 from the learned "
                       "6x(3x3 conv, 32 ch) warm start with zero-init fallback",
           "problems_per_step": P, "heldout_problems": NCH * P, "image": [N, N], "coils": Cc, "samples": M,
           "rel_resid": rel,
           "max_iterations": NitMax, "network": "6 layers 3x3 conv, 2-32-32-32-32-32-2, residual skip, "
                                                "bf16 weights, fp32 accumulation",
           "autotuned": tuned}
    out = {"metric": "CG-SENSE solves/s at rel. resid 1e-4, learned warm start vs zero-init",
           "value": P * a.steps / zw["el"], "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * zw["el"] / a.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "c64",
           "data": "synthetic (seeded phantoms, coils, trajectory, noise); network trained on the device "
                   "from CG solutions of 20000 synthetic problems",
           "config": cfg,
           "zero_init": {"value": P * a.steps / zz["el"], "unit": "solves/s",
                         "ms_per_step": 1e3 * zz["el"] / a.steps,
                         "iterations_mean": float(zz["nit"].mean()), "iterations_max": int(zz["nit"].max())},
           "warm_start": {"iterations_mean": float(zw["nit"].mean()), "iterations_max": int(zw["nit"].max()),
                          "fallbacks": zw["fallbacks"], "inference_ms_per_step": 1e3 * t_inf},
           "iteration_ratio_zero_over_warm": float(zz["nit"].mean() / zw["nit"].mean()),
           "heldout_all": {"problems": int(it_zero.size), "iterations_mean_zero": float(it_zero.mean()),
                           "iterations_mean_warm": float(it_warm.mean()),
                           "iteration_ratio_zero_over_warm": float(it_zero.mean() / it_warm.mean()),
                           "fallbacks": every["fallbacks"]},
           "speedup_warm_over_zero": float(zz["el"] / zw["el"]),
           "rel_resid_final_max": float(rel_w.max()),
           "status_last": {str(int(k)): int(v) for k, v in zip(*np.unique(zw["status"], return_counts=True))}}
    copy = copy_gbps(torch, dev)
    if roof:
        out["roofline"] = annotate_roofline(roof, copy)
    out["normal_operator_stages"] = stages
    out["normal_operator_applications_per_s"] = applications_per_s(stages)
    if not a.no_batch_roofline:
        note("64-slice batch roofline")
        out["roofline_above_cache"] = batch_roofline(mm, torch, dev, copy)
    note("Toeplitz normal operator")
    out["toeplitz_normal"] = toeplitz_normal(mm, fam, prob, dev, [(N, M, Cc)] * P)
    if out["toeplitz_normal"] and not a.no_cpu:
        out["toeplitz_normal"]["cpu_baseline"] = cpu_baseline_toeplitz(fam, P)
    if not a.no_cpu:
        note("FP64 oracle NRMSE")
        out["nrmse_vs_fp64_oracle"] = oracle_nrmse_warmstart(fam, host, prob, zw["x"], zz["x"])
        note("CPU baseline")
        cb = cpu_baseline_warmstart(fam, host, prob, a.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


def run_copy(a, torch, dev):
    """Plain device-to-device copy of 1 GiB: the known-size reference that
    scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of the profiles."""
    n = 1 << 28
    x = torch.empty(n, dtype=torch.float32, device=dev).uniform_()
    y = torch.empty_like(x)
    for _ in range(max(1, a.warmup)):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nb = 4 * n
    print(json.dumps({"metric": "device copy bandwidth (HBM counter calibration)", "value": 2 * nb * a.steps / el / 1e9,
                      "unit": "GB/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                      "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "copy: 1 GiB device-to-device", "bytes_read": nb,
                                 "bytes_written": nb}}), flush=True)


def interp_stage_bytes(N, G, K, R, Ns):
    """Algorithmic HBM bytes of one forward or adjoint application of the
    interpolated high-order operator (R images of N x N, K samples, Ns SVD
    terms), per kernel: the modulation / spread and combination /
    demodulation kernels and the plain NUFFT stages over B = R Ns transforms
    (complex64 values in and out; plan constants such as interpolation
    weights not counted, the svdSpace / svdTime factors counted)."""
    V, B = N * N, R * Ns
    return {
        "hin_modulate_k": 8.0 * (V * R + V * Ns + V * B),
        "hin_combine_k": 8.0 * (K * B + K * Ns + K + K * R),
        "hin_spread_k": 8.0 * (K * R + K * Ns + K + K * B),
        "hin_demodulate_k": 8.0 * (V * B + V * Ns + V * R),
        "nu_fwd_rows": 8.0 * (V * B + G * N * B),
        "nu_cols": 8.0 * G * (N + G) * B,
        "nu_degrid": 8.0 * (G * G * B + K + K * B),
        "nu_transpose_y": 2 * 8.0 * K * B,
        "nu_grid": 8.0 * (K * B + K + G * G * B),
        "nu_adj_rows": 8.0 * (G * N * B + V * B),
    }


def run_interp(a, mm, torch, dev):
    """The high-order problem (220x220, 32 coils, K = 24000, 16 harmonics +
    B0) solved through ``sampHighOrder`` approach 1 (sampHighOrder.m:314-359,
    :624-902: the linear terms as a NUFFT, the rest as a truncated SVD of
    exp(1i*Phi_res), the reference's default svdThresh 0.05 and subFact 5).
    One step = one cgne solve from zero (<= 30 iterations, rel. resid 1e-4)
    of the data of the direct operator.  Also reported: the plan set-up
    time, the SVD terms, the distance of one forward application from the
    direct operator, normal-operator applications/s of both forms and the
    roofline of the interpolated operator's dominant kernel."""
    import ctypes as C
    import numpy as np
    W = highorder_setup(mm, torch, dev)
    P, Ed, y = W["P"], W["E"], W["y"]
    N, K = P["N"], P["K"]
    note("interp: approach-1 plan (host SVD set-up over the device direct operator)")
    t0 = time.perf_counter()
    Si = mm.sampHighOrder(P["b0"], P["t"], P["h"], None, P["grid"], None, 1, 0, 1, {})
    setup_s = time.perf_counter() - t0
    Ei = mm.senseOp(Si, W["R"])
    # untimed: the NUFFT's options for its batch (32 coils x Ns transforms)
    tuned = None if a.no_autotune else Ei.autotune(5)
    V = N * N
    xr = torch.randn(V, dtype=torch.complex64, device=dev)
    fd, fi = Ed * xr, Ei * xr
    rel_fwd = float(torch.linalg.vector_norm(fi - fd) / torch.linalg.vector_norm(fd))

    def normal_rate(E, reps=5):
        E.normal(xr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            E.normal(xr)
        torch.cuda.synchronize()
        return reps / (time.perf_counter() - t)
    rate_i, rate_d = normal_rate(Ei), normal_rate(Ed)
    opt = {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None}
    x0 = torch.zeros(V, dtype=torch.complex64, device=dev)
    for _ in range(a.warmup):
        mm.cgne(Ei, y, x0, 30, opt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nits = []
    for _ in range(a.steps):
        x, res, *_, st = mm.cgne(Ei, y, x0, 30, opt)
        nits.append(len(res) - 1)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / a.steps
    xd, *_ = mm.cgne(Ed, y, x0, 30, opt)
    nrmse_vs_direct = float(torch.linalg.vector_norm(x - xd) / torch.linalg.vector_norm(xd))
    # dominant kernel of a forward application, HIP events on the library stream
    ms = C.c_double()
    name = C.create_string_buffer(64)
    yo = torch.empty(K * W["R"].C, dtype=torch.complex64, device=dev)
    rc = mm.lib().mm_op_profile(Ei._plan.h, 0, xr.data_ptr(), yo.data_ptr(), 1, 10, C.byref(ms), name, 64)
    roof = None
    if rc == 0 and ms.value > 0:
        kname = name.value.decode()
        G = int(mm.workloads.grid_size(N, 1.5))
        byts = interp_stage_bytes(N, G, K, W["R"].C, Si.Ns).get(kname)
        if byts:
            ach = byts / (ms.value * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                    "traffic": read_traffic(kname, "interp"), "kernel": kname, "ms_per_launch": ms.value,
                    "algorithmic_bytes_per_launch": byts}
    out = {"metric": "CG-SENSE solves/s at rel. resid 1e-4, high-order operator by approach 1 (interpolated)",
           "value": 1.0 / el, "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64", "data": "synthetic (seeded phantom, coils, field coefficients, B0, noise)",
           "config": dict(W["cfg"], workload="interp: " + W["cfg"]["workload"] +
                          "; sampHighOrder approach 1 (svdThresh 0.05, subFact 5, NUFFT W=6 os=1.5)"),
           "iterations": nits, "svd_terms": Si.Ns, "setup_s": setup_s, "autotuned": tuned,
           "forward_rel_l2_vs_direct": rel_fwd, "solution_nrmse_vs_direct_solve": nrmse_vs_direct,
           "normal_operator_applications_per_s": {"interp": rate_i, "direct": rate_d},
           "roofline": roof}
    if not a.no_cpu:
        # the same computation with NumPy on the host: the device plan's SVD
        # factors around the oracle's FP64 NUFFT, one coil image forward and
        # adjoint, scaled to 32 coils and a 30-iteration solve (65 applications)
        from oracle.nufft import NufftPlan
        Sg = Si
        So = NufftPlan([N, N], Sg._nufft._kl[0], None, 1.5, 6)
        sp = Sg._sp.astype(np.complex128)
        stt = Sg._st.astype(np.complex128)
        ph = Sg._ph.astype(np.complex128)
        xc = (P["phantom"] * P["maps"][..., 0]).reshape(-1, order="F")
        t0 = time.perf_counter()
        Yk = np.asarray(So.forward((xc[:, None] * sp).reshape(N, N, -1, order="F"))).reshape(K, -1, order="F")
        yk = ph * np.sum(stt * Yk, axis=1)
        Xa = np.asarray(So.adjoint(np.conj(ph[:, None] * stt) * yk[:, None])).reshape(V, -1, order="F")
        np.sum(np.conj(sp) * Xa, axis=1)
        dt = time.perf_counter() - t0
        apps = 65
        per_solve = dt / 2 * W["R"].C * apps
        out["cpu_baseline"] = {"value": 1.0 / per_solve, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"NumPy FP64 (the oracle's NUFFT with the device plan's {Sg.Ns} SVD "
                                         f"terms): one coil image forward + adjoint in {dt:.1f} s, scaled to "
                                         f"{W['R'].C} coils x {apps} applications"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def finddel_problem(mm, N=160, C=16, seed=5):
    """The reference's delay demo (demos/demo_autoDel.m) on synthetic images
    (workloads.delay_demo_problem) over the reference's own spiral R4
    field-probe record (tests/golden fixture: 15000 x 16 harmonic and 4
    concomitant coefficients at 1 us; 5268 data samples at 2.5 us)."""
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "traj_spiral_R4.npz"))
    return mm.workloads.delay_demo_problem(d, N=N, C=C, seed=seed)


def run_finddel(a, mm, torch, dev):
    """findDelAuto (findDel/findDelAuto.m:1-143) as the reference's demo runs
    it (demos/demo_autoDel.m:48-86): the spiral R4 record, a 160 x 160 image,
    16 coils, data with the trajectory delayed 2 us plus noise; maxNit_cgne =
    the iterations a default cgne of at most 20 takes on the undelayed
    operator, delJumpFact 3, no coarse search, started at 0.  One step = one
    whole delay estimate: every sampHighOrder plan (host set-up and device
    rank compression), cgne solve, setPhiDiv application and update, data
    resident in HBM.  Also reported: the estimate and its steps, the operator
    applications per estimate and the high-order GEMM's roofline."""
    import ctypes as C
    import numpy as np
    P = finddel_problem(mm)
    N, C_, dt, tdw = P["N"], P["C"], P["datatime"], P["tdwell"]
    fd = mm.finddel
    R = mm.rcvrOp(P["maps"])

    def op(delay):
        sa, sc = fd.interpTrajTime(P["spha"], P["conc"], tdw, delay, dt)
        return mm.senseOp(mm.sampHighOrder(P["b0"], dt, sa.T, sc.T, P["grid"]), R)
    x_true = torch.as_tensor(np.ascontiguousarray(P["phantom"].T), dtype=torch.complex64, device=dev)
    y = op(P["delay"]) * x_true
    g = torch.Generator(device=dev).manual_seed(5)
    sig = float(torch.sqrt(torch.mean(torch.abs(y) ** 2))) / 1000.0
    noise = torch.randn(y.shape, generator=g, device=dev) + 1j * torch.randn(y.shape, generator=g, device=dev)
    y = (y + sig / np.sqrt(2) * noise).to(torch.complex64).contiguous()
    # the demo's iteration count: a default cgne of at most 20 on the undelayed operator
    _, res0, *_ = mm.cgne(op(0.0), y, None, 20)
    nit = len(res0) - 1
    args = (0.0, y, tdw, P["spha"], P["conc"], dt, P["grid"], P["b0"], P["maps"], nit, 3, 0)
    # untimed warm-up, counting the operator applications of one estimate
    solves = []
    real_cgne = fd.cgne

    def counting(*c_args, **kw):
        out = real_cgne(*c_args, **kw)
        solves.append(len(out[1]) - 1)
        return out
    fd.cgne = counting
    try:
        est, per = fd.findDelAuto(*args)
    finally:
        fd.cgne = real_cgne
    apps = sum(2 * n + 5 for n in solves) + 2 * (len(per) - 1)
    for _ in range(max(0, a.warmup - 1)):
        fd.findDelAuto(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        est, per = fd.findDelAuto(*args)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / a.steps
    # the operator at the estimate: its dominant kernel (HIP events)
    K = len(dt)
    Ee = op(est)
    ms = C.c_double()
    name = C.create_string_buffer(64)
    xin = torch.zeros(N * N, dtype=torch.complex64, device=dev)
    yout = torch.empty(K * C_, dtype=torch.complex64, device=dev)
    rc = mm.lib().mm_op_profile(Ee._plan.h, 0, xin.data_ptr(), yout.data_ptr(), 1, 10, C.byref(ms), name, 64)
    L = P["spha"].shape[1] + P["conc"].shape[1] + 1
    roof = None
    if rc == 0 and ms.value > 0:
        flops = 2.0 * K * N * N * L + 8.0 * K * N * N * C_
        ach = flops / (ms.value * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_FP16_MFMA_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP16_MFMA_TFLOPS, "traffic": None, "kernel": name.value.decode(),
                "ms_per_launch": ms.value, "mfma_products_per_flop": 3}
    out = {"metric": "findDelAuto delay estimates/s (the reference demo's setting)",
           "value": 1.0 / el, "unit": "estimates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64", "data": "the reference's spiral R4 field-probe record; synthetic phantom, coils, B0, "
                                   "noise; 2 us applied delay",
           "config": {"workload": f"finddel: {N}x{N} c64 image at 1.35 mm, {C_} coils, {K} samples @2.5 us "
                                  "(spiral R4, 16 SH + 4 concomitant terms + B0), findDelAuto from 0, "
                                  f"{nit} cgne it/image, delJumpFact 3, no coarse search",
                      "image": [N, N], "coils": C_, "samples": K, "phase_terms": L,
                      "cgne_iterations": nit, "applied_delay_us": P["delay"]},
           "estimate_us": est, "estimate_steps": [float(v) for v in per], "solves": len(solves),
           "solve_iterations": solves, "operator_applications": apps, "roofline": roof}
    if not a.no_cpu:
        sa, sc = fd.interpTrajTime(P["spha"], P["conc"], tdw, est, dt)
        r = cpu_highorder_per_app(P["b0"], dt, sa.T, sc.T, P["grid"], P["maps"], P["phantom"], a.cpu_seconds)
        if r:
            per_app, ks, dts, thr = r
            out["cpu_baseline"] = {"value": 1.0 / (per_app * apps), "unit": "estimates/s", "cores": thr,
                                   "kind": "port",
                                   "sample": f"C/OpenMP FP64 restatement (oracle/c/mmref.c), one forward + one "
                                             f"adjoint over {ks} of {K} samples ({dts:.2f} s), scaled to the "
                                             f"{apps} operator applications of one estimate (plan set-ups not "
                                             "counted)"}
            out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def nufft3_stage_bytes(N, G, M, T):
    """Algorithmic HBM bytes per launch of each 3-D NUFFT stage over T
    transforms (complex64 values in and out; plan tables not counted)."""
    N1, N2, N3 = N
    G1, G2, G3 = G
    V, Gc = N1 * N2 * N3, G1 * G2 * G3
    r1, r2 = G1 * N2 * N3, G1 * G2 * N3   # after the row pass; after the g2 pass
    return {
        "n3_rows_fwd": 8.0 * T * (V + r1), "n3_axis2_fwd": 8.0 * T * (r1 + r2), "n3_axis3_fwd": 8.0 * T * (r2 + Gc),
        "n3_degrid": 8.0 * (T * Gc + M + M * T), "n3_transpose_y": 2 * 8.0 * M * T,
        "n3_grid_csr": 8.0 * (M * T + M + T * Gc), "n3_axis3_adj": 8.0 * T * (Gc + r2),
        "n3_axis2_adj": 8.0 * T * (r2 + r1), "n3_rows_adj": 8.0 * T * (r1 + V), "n3_odd_ramp": 2 * 8.0 * T * Gc,
    }


def run_nufft3(a, mm, torch, dev):
    """3-D SENSE cgne (nufftOp with a 3-D imN, nufftOp.m:89-274 in three
    dimensions, composed with rcvrOp as mrSampFunc.m:97-107 does): a 128 x
    128 x 64 volume, 8 coils, a stack of 64 golden-angle spirals of 4096
    samples (262144), KB W = 4, os 1.5 (grid 192 x 192 x 96).  One step = one
    cgne solve from zero to rel. resid 1e-4 (<= 20 iterations), data in HBM.
    Roofline: the slowest stage of a forward or adjoint application (HIP
    events after every launch, mm_op_profile)."""
    import ctypes as C
    import numpy as np
    P = mm.workloads.nufft3_problem()
    N, Cc, M = P["N"], P["C"], P["M"]
    V = int(np.prod(N))
    G = tuple(int(mm.workloads.grid_size(n, 1.5)) for n in N)
    note("nufft3: plan (host set-up of the 3-D gridding tables)")
    t0 = time.perf_counter()
    S = mm.nufftOp(list(N), P["kloc"], None, 1, 1.5, 4)
    setup_s = time.perf_counter() - t0
    E = mm.senseOp(S, mm.rcvrOp(P["maps"]))
    x_true = torch.as_tensor(np.ascontiguousarray(P["phantom"].transpose(2, 1, 0)), dtype=torch.complex64,
                             device=dev)
    y = E * x_true
    g = torch.Generator(device=dev).manual_seed(6)
    sig = float(torch.sqrt(torch.mean(torch.abs(y) ** 2))) / 40.0
    noise = torch.randn(y.shape, generator=g, device=dev) + 1j * torch.randn(y.shape, generator=g, device=dev)
    y = (y + sig / np.sqrt(2) * noise).to(torch.complex64).contiguous()
    NitMax = 20
    opt = {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None}
    x0 = torch.zeros(V, dtype=torch.complex64, device=dev)
    for _ in range(a.warmup):
        mm.cgne(E, y, x0, NitMax, opt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nits = []
    for _ in range(a.steps):
        x, res, *_, st = mm.cgne(E, y, x0, NitMax, opt)
        nits.append(len(res) - 1)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / a.steps
    xr = torch.randn(V, dtype=torch.complex64, device=dev)
    E.normal(xr)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        E.normal(xr)
    torch.cuda.synchronize()
    napp = 3 / (time.perf_counter() - t1)
    # slowest stage of either direction (HIP events on the library stream)
    byts = nufft3_stage_bytes(N, G, M, Cc)
    roof, stages = None, {}
    yo = torch.empty(M * Cc, dtype=torch.complex64, device=dev)
    xo = torch.empty(V, dtype=torch.complex64, device=dev)
    for adj, (xi, xo_) in ((0, (xr, yo)), (1, (yo, xo))):
        ms = C.c_double()
        name = C.create_string_buffer(64)
        if mm.lib().mm_op_profile(E._plan.h, adj, xi.data_ptr(), xo_.data_ptr(), 1, 5, C.byref(ms), name, 64) == 0:
            stages[name.value.decode()] = ms.value
    if stages:
        kname, kms = max(stages.items(), key=lambda kv: kv[1])
        b = byts.get(kname)
        if b:
            ach = b / (kms * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
                    "traffic": read_traffic(kname, "nufft3"), "kernel": kname, "ms_per_launch": kms,
                    "algorithmic_bytes_per_launch": b}
    out = {"metric": "CG-SENSE solves/s at rel. resid 1e-4, 3-D NUFFT (stack of spirals)",
           "value": 1.0 / el, "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64", "data": "synthetic (seeded phantom, coils, trajectory, noise)",
           "config": {"workload": "nufft3: 128x128x64 c64 volume, 8 coils, stack of 64 golden-angle spirals x 4096 "
                                  "samples, KB W=4 os=1.5 (grid 192x192x96), cgne to rel. resid 1e-4 (<=20 it), "
                                  "zero init", "image": list(N), "coils": Cc, "samples": M, "grid": list(G),
                      "max_iterations": NitMax, "rel_resid": 1e-4},
           "iterations": nits, "plan_setup_s": setup_s, "normal_operator_applications_per_s": napp,
           "dominant_stage_ms": stages, "roofline": roof}
    if not a.no_cpu:
        # the oracle's FP64 3-D NUFFT and coil operator: one coil image forward
        # and adjoint, scaled to the coils and the solve's applications
        from oracle.nufft import NufftPlan
        So = NufftPlan(list(N), P["kloc"], None, 1.5, 4)
        xc = P["phantom"] * P["maps"][..., 0]
        t0 = time.perf_counter()
        So.adjoint(So.forward(xc))
        dt = time.perf_counter() - t0
        apps = 2 * max(nits) + 5
        out["cpu_baseline"] = {"value": 1.0 / (dt / 2 * Cc * apps), "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"NumPy FP64 oracle 3-D NUFFT: one coil volume forward + adjoint in {dt:.1f} "
                                         f"s, scaled to {Cc} coils x {apps} applications"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


PEAK_FP32_MFMA_TFLOPS = 157.3   # MI355X dense FP32 matrix peak (rocBLAS cherk of the dense normal matrix)


def run_densemat(a, mm, torch, dev):
    """``mrSampFuncMat`` (operators/mrSampFuncMat.m:1-47) at the full size of
    the high-order problem (220x220 image, 24000 samples, 32 coils, no mask):
    the single-precision dense encoding matrix S [24000 x 48400], its Gram
    matrix S'S (rocBLAS cherk, lower triangle, mirrored by a kernel) and the
    coil fold, AtA [48400 x 48400] complex64 left in HBM.  One step = one
    mm_hio_dense_normal call (the workspace allocation included).  The
    roofline is that of the dominant kernel, the cherk, timed with HIP events
    on the library stream on an operand of the same shape."""
    import ctypes as C
    import numpy as np
    P = mm.workloads.highorder_problem(seed=2)
    N, K = P["N"], P["K"]
    V = N * N
    S = mm.sampHighOrder(P["b0"], P["t"], P["h"], None, P["grid"])
    R = mm.rcvrOp(P["maps"])
    L = mm.lib()
    mask = np.ones(V, np.uint8)
    AtA = torch.empty((V, V), dtype=torch.complex64, device=dev)
    vm = C.c_int64()

    def call():
        mm._lib.check(L.mm_hio_dense_normal(S._plan.h, R.C, R._maps.ctypes.data, mask.ctypes.data, AtA.data_ptr(),
                                            C.byref(vm), 1))
    note("densemat: warm-up call")
    torch.cuda.synchronize()
    call()
    for _ in range(max(0, a.warmup - 1)):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        call()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / a.steps
    Vm = vm.value
    # the dominant kernel on an operand of the same shape (column-major
    # [K x Vm] = a row-major (Vm, K) tensor), events on the library stream
    note("densemat: cherk timing")
    rb = C.CDLL("/opt/rocm/lib/librocblas.so")
    h = C.c_void_p()
    assert rb.rocblas_create_handle(C.byref(h)) == 0
    assert rb.rocblas_set_stream(h, C.c_void_p(mm._lib.context_stream())) == 0
    Sx = torch.randn((Vm, K), dtype=torch.complex64, device=dev)
    one, zero = C.c_float(1.0), C.c_float(0.0)

    def herk():
        rc = rb.rocblas_cherk(h, 122, 113, C.c_int(Vm), C.c_int(K), C.byref(one), C.c_void_p(Sx.data_ptr()),
                              C.c_int(K), C.byref(zero), C.c_void_p(AtA.data_ptr()), C.c_int(Vm))
        assert rc == 0, f"rocblas_cherk status {rc}"
    herk()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        herk()
    e1.record()
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / reps
    rb.rocblas_destroy_handle(h)
    del Sx
    flops = 4.0 * K * Vm * (Vm + 1)          # one triangle of the Hermitian Gram matrix
    ach = flops / (kms * 1e-3) / 1e12
    out = {"metric": "dense normal matrices (mrSampFuncMat AtA) per second", "value": 1.0 / el,
           "unit": "matrices/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * el,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
           "data": "synthetic (seeded high-order problem: 16 SH + B0, Gaussian-ring coils)",
           "config": {"workload": "densemat: mrSampFuncMat AtA of the 220x220 / K=24000 / 32-coil high-order "
                                  "operator, no mask (AtA 48400 x 48400 c64 in HBM)",
                      "image": [N, N], "samples": K, "coils": R.C, "masked_voxels": Vm},
           "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                        "frac": ach / PEAK_FP32_MFMA_TFLOPS, "traffic": None, "kernel": "rocblas_cherk",
                        "ms_per_launch": kms, "algorithmic_flops_per_launch": flops,
                        "note": "4 K Vm (Vm + 1) flops: the Hermitian product's lower triangle; the full "
                                "product the reference forms (S'*S) is twice that"}}
    if not a.no_cpu:
        # FP64 Gram matrix of a column subset on the host (NumPy/OpenBLAS,
        # every host thread), scaled by (Vm / Vs)^2 to the full matrix
        vs = 6000
        rng = np.random.default_rng(0)
        Sd = np.exp(1j * rng.uniform(-np.pi, np.pi, (K, vs)))
        t0 = time.perf_counter()
        G = Sd.conj().T @ Sd
        t1 = time.perf_counter() - t0
        del G, Sd
        full = t1 * (Vm / vs) ** 2
        out["cpu_baseline"] = {"value": 1.0 / full, "unit": "matrices/s", "cores": os.cpu_count(), "kind": "port",
                               "sample": f"NumPy FP64 Gram matrix S'*S of {K} x {vs} (the oracle's mrSampFuncMat.m:37 "
                                         f"step) in {t1:.1f} s, scaled by (Vm/{vs})^2 to the {Vm}-voxel matrix"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    mm = importlib.import_module("uwobrainhack2026-matmri_amd")
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    mm._lib.context(0, stream.cuda_stream)
    if a.workload == "copy":
        with torch.cuda.stream(stream):
            run_copy(a, torch, dev)
        return
    if a.workload == "warmstart":
        with torch.cuda.stream(stream):
            run_warmstart(a, mm, torch, dev)
        return
    if a.workload == "densemat":
        with torch.cuda.stream(stream):
            run_densemat(a, mm, torch, dev)
        return
    if a.workload == "interp":
        with torch.cuda.stream(stream):
            run_interp(a, mm, torch, dev)
        return
    if a.workload == "finddel":
        with torch.cuda.stream(stream):
            run_finddel(a, mm, torch, dev)
        return
    if a.workload == "nufft3":
        with torch.cuda.stream(stream):
            run_nufft3(a, mm, torch, dev)
        return
    with torch.cuda.stream(stream):
        W = {"highorder": highorder_setup, "nufft": nufft_setup, "batch": batch_setup}[a.workload](mm, torch, dev)
        E, y, x0, NitMax = W["E"], W["y"], W["x0"], W["NitMax"]
        if a.hio_variant is not None and W["kind"] == "highorder":
            E.set_option("hio_variant", a.hio_variant)
        for kv in a.set:
            k, v = kv.split("=")
            W["S"].set_option(k, int(v))
        if W["kind"] in ("nufft", "batch") and not (a.no_autotune or a.set):
            W["cfg"]["autotuned"] = E.autotune(10)
        opt = {"resThresh": None, "resThreshRel": 1e-8, "xChangeThresh": None}
        b = E.H * y
        bb = float(torch.sum(torch.abs(b.reshape(-1).to(torch.complex128)) ** 2))
        torch.cuda.synchronize()
        nits = []
        for _ in range(a.warmup):
            x, res, *_ , st = mm.cgne(E, y, x0, NitMax, opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            x, res, *_ , st = mm.cgne(E, y, x0, NitMax, opt)
            nits.append(len(res) - 1 if W["kind"] != "batch" else max(len(r) - 1 for r in res))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        # dominant kernel, timed with HIP events on the library stream
        import ctypes as C
        ms = C.c_double()
        name = C.create_string_buffer(64)
        xin = torch.zeros(E.image_shape[0] * (E.image_shape[1] if len(E.image_shape) > 1 else 1),
                          dtype=torch.complex64, device=dev)
        yout = torch.empty_like(y.reshape(-1))
        rc = mm.lib().mm_op_profile(E._plan.h, 0, xin.data_ptr(), yout.data_ptr(), 1, 10, C.byref(ms), name, 64)
        kms = ms.value if rc == 0 else None
    if W["kind"] in ("nufft", "batch"):
        with torch.cuda.stream(stream):
            sl = ([(p["N"], p["M"], 16) for p in W["P"]] if W["kind"] == "batch"
                  else [(W["P"]["N"], W["P"]["M"], W["P"]["C"])])
            roof_nu, stages_nu = nufft_roofline(mm, E, dev, sl, W["kind"])
    sps = W.get("solves_per_step", 1)
    out = {"metric": "CG-SENSE solves/s at rel. resid 1e-4, learned warm start vs zero-init",
           "value": sps * a.steps / el, "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * el / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded phantom, coils, trajectory, noise)",
           "config": W["cfg"], "iterations": nits,
           "rel_resid_final": float(np.sqrt(res[-1] / bb)) if W["kind"] != "batch" else None,
           "status_last": st if W["kind"] != "batch" else {str(k): st.count(k) for k in set(st)}}
    if kms:
        if W["kind"] == "highorder":
            # algorithmic flops of one application (phase contraction + complex
            # MAC); the kernel runs them on fp16 MFMA with split operands, which
            # issues three fp16 products per FP32-accurate product
            ach = W["flops"] / (kms * 1e-3) / 1e12
            out["roofline"] = {"bound": "mfma", "achieved": ach, "peak": PEAK_FP16_MFMA_TFLOPS, "unit": "TFLOP/s",
                               "frac": ach / PEAK_FP16_MFMA_TFLOPS,
                               "traffic": read_traffic(name.value.decode(), "highorder"),
                               "kernel": name.value.decode(), "ms_per_launch": kms,
                               "mfma_products_per_flop": 3}
    if W["kind"] in ("nufft", "batch"):
        if roof_nu:
            out["roofline"] = roof_nu
        out["normal_operator_stages"] = stages_nu
        out["normal_operator_applications_per_s"] = applications_per_s(stages_nu)
    if W["kind"] == "highorder" and a.precision:
        with torch.cuda.stream(stream):
            note("precision study (C/OpenMP FP64 oracle at full size)")
            out["precision"] = highorder_precision(mm, torch, dev)
    if not a.no_cpu:
        cb = {"highorder": cpu_baseline_highorder, "nufft": cpu_baseline_nufft,
              "batch": cpu_baseline_batch}[W["kind"]](W["P"], a.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
