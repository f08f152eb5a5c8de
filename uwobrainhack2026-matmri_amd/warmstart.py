"""Learned warm start for cgne (no counterpart in the reference).

:class:`WarmStartNet` drives the device network of the C ABI
(``mm_wsnet_*``): six 3x3 convolutions (2 -> 32 x5 -> 2 channels, ReLU,
residual skip), bf16 weights, FP32 accumulation, computing
``x0 = x + s * net(x / s)`` with ``s = sqrt(mean |x|^2)`` per image.  Its input
is the density-compensated adjoint image ``R'F'(dcf .* y)`` (the gridding
reconstruction of ``demo_regridding.m:35-39``) divided by the coil
sum of squares ``R'R 1 = sum_c |S_c|^2`` and by one global scale fitted on
the training set (:func:`input_image`); a convolutional network is
translation-equivariant and cannot undo the position-dependent coil gain
itself.  Its output is the ``x0`` of ``cgne(A, b, x0, ...)``
(``iterativeSolvers/cgne.m:1``, ``:143-145``).

Training data comes from the plain CG path itself: spiral CG-SENSE problems
(128x128, 8 coils, 16384-sample spiral) with random phantoms and coil phases,
their density-compensated adjoint images as inputs and the zero-init device
cgne solution run toward relative residual 1e-6 as targets (FP32 CG levels
off near 1e-5; the targets stop at that floor).  Network training runs on
the device (``mm_wsnet_train_step``: forward, backward and Adam kernels).

Fallback: :func:`solve_warm` passes ``x0Fallback`` to the device cgne, which
starts a slice from ``x0 = 0`` when the network's initial normal-equation
residual exceeds that of ``x0 = 0`` (``|b|``).
"""
import ctypes as C
import json
import os
import time
import weakref

import numpy as np

from . import _lib, workloads
from ._lib import check, ptr
from .operators import nufftOp, rcvrOp, senseOp
from .solvers import cgne

try:
    import torch
except Exception:  # pragma: no cover
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(HERE, "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "warmstart_128x128.npy")

# problem family of the warm start: the 2-D spiral CG-SENSE problem
PROBLEM = dict(N=128, M=16384, C=8, n_turns=64, snr=40.0, os=1.5, kwidth=4)
TRAIN_SEED0 = 5_000_000          # problem i uses seed TRAIN_SEED0 + i
HELDOUT_OFFSET = 20_000          # held-out problems follow the training ones


def _destroy_net(h):
    try:
        _lib.lib().mm_wsnet_destroy(C.c_void_p(h))
    except Exception:
        pass


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


class WarmStartNet:
    """The device network for complex N1 x N2 images.

    Images are column-major ``[N1, N2]`` (numpy, optionally with a trailing
    batch dimension) or torch device tensors ``[..., N2, N1]`` (the reversed
    shape of the same memory)."""

    def __init__(self, N1, N2, params=None):
        L = _lib.lib()
        h = C.c_void_p()
        check(L.mm_wsnet_create(_lib.context(), int(N1), int(N2), C.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, _destroy_net, h.value)
        self.N1, self.N2 = int(N1), int(N2)
        self.input_scale = 1.0       # applied by input_image()
        n = C.c_int64()
        check(L.mm_wsnet_num_params(h, C.byref(n)))
        self.num_params = int(n.value)
        if params is not None:
            self.set_params(params)

    # ---- parameters -------------------------------------------------------
    def set_params(self, theta):
        """Flat FP32 parameters (per layer W[cout][cin][3][3], then b[cout]);
        resets the optimiser state."""
        if _is_torch(theta):
            t = theta.to(torch.float32).contiguous()
            if t.numel() != self.num_params:
                raise ValueError(f"expected {self.num_params} parameters, got {t.numel()}")
            self._sync_torch(t)
            check(_lib.lib().mm_wsnet_set_params(self._h, ptr(t), 1))
            self._sync()
            return
        t = np.ascontiguousarray(np.asarray(theta, dtype=np.float32).reshape(-1))
        if t.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {t.size}")
        check(_lib.lib().mm_wsnet_set_params(self._h, t.ctypes.data, 0))

    def get_params(self):
        t = np.empty(self.num_params, dtype=np.float32)
        check(_lib.lib().mm_wsnet_get_params(self._h, t.ctypes.data, 0))
        return t

    # ---- helpers ----------------------------------------------------------
    @staticmethod
    def _sync_torch(t):
        torch.cuda.current_stream(t.device).synchronize()

    @staticmethod
    def _sync():
        check(_lib.lib().mm_ctx_sync(_lib.context()))

    def _batch_of(self, x):
        V = self.N1 * self.N2
        n = x.numel() if _is_torch(x) else np.asarray(x).size
        if n % V or n == 0:
            raise ValueError(f"image input has {n} elements, not a multiple of {self.N1}x{self.N2}")
        return n // V

    def _host_images(self, x):
        xa = np.asarray(x)
        return np.asfortranarray(xa).reshape(-1, order="F").astype(np.complex64), xa.shape

    # ---- network ----------------------------------------------------------
    def forward(self, x):
        """x0 for a batch of images (same shape and kind as x)."""
        L = _lib.lib()
        B = self._batch_of(x)
        if _is_torch(x):
            xc = x.to(torch.complex64).contiguous()
            y = torch.empty_like(xc)
            self._sync_torch(xc)
            check(L.mm_wsnet_forward(self._h, ptr(xc), ptr(y), B, 1))
            self._sync()
            return y
        xf, shp = self._host_images(x)
        y = np.empty_like(xf)
        check(L.mm_wsnet_forward(self._h, xf.ctypes.data, y.ctypes.data, B, 0))
        return y.reshape(shp, order="F")

    __call__ = forward

    def loss_grad(self, x, target):
        """(loss, gradient) of mean((x/s + z - t/s)^2) over the batch."""
        L = _lib.lib()
        B = self._batch_of(x)
        loss = C.c_double()
        g = np.empty(self.num_params, dtype=np.float32)
        if _is_torch(x):
            xc = x.to(torch.complex64).contiguous()
            tc = target.to(torch.complex64).contiguous()
            gd = torch.empty(self.num_params, dtype=torch.float32, device=xc.device)
            self._sync_torch(xc)
            check(L.mm_wsnet_loss_grad(self._h, ptr(xc), ptr(tc), B, C.byref(loss), ptr(gd), 1))
            self._sync()
            return loss.value, gd.cpu().numpy()
        xf, _ = self._host_images(x)
        tf, _ = self._host_images(target)
        check(L.mm_wsnet_loss_grad(self._h, xf.ctypes.data, tf.ctypes.data, B, C.byref(loss),
                                   g.ctypes.data, 0))
        return loss.value, g

    def train_step(self, x, target, lr, want_loss=False):
        """One Adam step on device tensors; returns the loss if asked (which
        synchronises), else None."""
        if not _is_torch(x):
            raise TypeError("train_step takes torch device tensors")
        B = self._batch_of(x)
        loss = C.c_double()
        check(_lib.lib().mm_wsnet_train_step(self._h, ptr(x), ptr(target), B, float(lr),
                                             C.byref(loss) if want_loss else None, 1))
        return loss.value if want_loss else None

    # ---- persistence ------------------------------------------------------
    def save(self, path, meta=None):
        """Parameters as a .npy (FP32) plus a .json with the metadata."""
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.save(path, self.get_params())
        m = dict(meta or {})
        m.update(N1=self.N1, N2=self.N2, num_params=self.num_params, input_scale=self.input_scale)
        with open(os.path.splitext(path)[0] + ".json", "w") as f:
            json.dump(m, f, indent=1, sort_keys=True)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        theta = np.load(path, allow_pickle=False)
        with open(os.path.splitext(path)[0] + ".json") as f:
            meta = json.load(f)
        net = cls(meta["N1"], meta["N2"], theta)
        net.input_scale = float(meta.get("input_scale", 1.0))
        net.meta = meta
        return net


# ---------------------------------------------------------------------------
# problem generation (device)

class ProblemFamily:
    """Spiral CG-SENSE problems with a fixed trajectory: one multi-slice NUFFT
    plan for `batch` problems, per-problem coils (random phases), phantoms
    and noise."""

    def __init__(self, batch, device="cuda", **over):
        p = dict(PROBLEM)
        p.update(over)
        self.p = p
        self.batch = int(batch)
        self.device = device
        N, M = p["N"], p["M"]
        self.kloc = workloads.spiral(M, p["n_turns"])
        self.dcf = workloads.spiral_dcf(M, p["n_turns"])
        self.S = nufftOp.batch([([N, N], self.kloc, None)] * self.batch, p["os"], p["kwidth"])
        self.dcf_t = torch.as_tensor(self.dcf, dtype=torch.float32, device=device)

    def host_batch(self, seeds):
        """Per-problem coil maps ([N, N, C] each) and phantoms ([N2, N1] rows, torch layout)."""
        N, Cc = self.p["N"], self.p["C"]
        maps, ph = [], []
        for sd in seeds:
            rng = np.random.default_rng(int(sd))
            maps.append(workloads.gaussian_coils(N, N, Cc, rng))
            ph.append(workloads.ellipse_phantom(N, N, rng).T)
        return maps, np.stack(ph).astype(np.complex64)

    def make(self, seeds, host=None):
        """Device problem batch: dict(E, x_true [P, N, N], y (flat data),
        xin [P, N, N] = R'F'(dcf .* y))."""
        if len(seeds) != self.batch:
            raise ValueError("seed count must equal the plan's batch")
        maps, ph = host if host is not None else self.host_batch(seeds)
        N, M, Cc = self.p["N"], self.p["M"], self.p["C"]
        P = self.batch
        R = rcvrOp.batch(maps)
        E = senseOp(self.S, R)
        x_true = torch.as_tensor(ph, device=self.device)
        y = (E * x_true.reshape(-1)).reshape(P, Cc, M)
        g = torch.Generator(device=self.device).manual_seed(int(seeds[0]))
        sig = torch.sqrt(torch.mean(torch.abs(y) ** 2, dim=(1, 2), keepdim=True)) / self.p["snr"]
        nz = torch.randn((P, Cc, M), generator=g, device=self.device) + \
            1j * torch.randn((P, Cc, M), generator=g, device=self.device)
        y = (y + sig / np.sqrt(2.0) * nz).to(torch.complex64).contiguous()
        xin = input_image(E, R, y.reshape(-1), self.dcf_t).reshape(P, N, N)
        return dict(E=E, R=R, x_true=x_true, y=y.reshape(-1), xin=xin, seeds=list(seeds))


def input_image(E, R, y, dcf, eps=1e-3):
    """Network input of a SENSE problem: R'F'(dcf .* y) / (ssq + eps max ssq),
    ssq = sum_c |S_c|^2 per pixel (the maximum over every slice), with dcf
    per sample (repeated over coils and slices); device tensors, flat image
    vector out.  The normalisation is one library call
    (mm_rcvr_normalize_ssq)."""
    M = dcf.numel()
    xg = (E.H * (y.reshape(-1, M) * dcf.reshape(1, M)).reshape(-1)).reshape(-1).to(torch.complex64).contiguous()
    out = torch.empty_like(xg)
    torch.cuda.current_stream(xg.device).synchronize()
    check(_lib.lib().mm_rcvr_normalize_ssq(R._plan.h, xg.data_ptr(), out.data_ptr(), float(eps), 1))
    check(_lib.lib().mm_ctx_sync(_lib.context()))
    return out


def cg_opts(rel_resid, NitMax):
    """cgne options for "stop at relative residual |r|/|b| <= rel_resid"
    (resThresh = rel_resid^2 |b|^2, no x-change test)."""
    return dict(resThresh=None, resThreshRel=rel_resid ** 2, xChangeThresh=None), NitMax


def solve(E, y, x0, rel_resid=1e-4, NitMax=200, fallback=False):
    """Device cgne on a (multi-slice) SENSE operator; returns (x, nit per
    slice, status per slice, fellback per slice)."""
    opt, n = cg_opts(rel_resid, NitMax)
    opt["x0Fallback"] = bool(fallback)
    x, res, _, _, _, _, st = cgne(E, y, x0.reshape(-1), n, opt)
    res = res if isinstance(res, list) else [res]
    st = st if isinstance(st, list) else [st]
    nit = np.array([len(r) - 1 for r in res])
    fb = np.zeros(E.nslice, dtype=np.int32)
    check(_lib.lib().mm_cg_last_fallback(E._plan.h, fb.ctypes.data))
    return x, nit, np.array(st), fb, res


def warm_x0(net, xin):
    """x0 from network inputs (input_image), applying the fitted scale."""
    xs = xin if net.input_scale == 1.0 else xin * net.input_scale
    return net(xs)


def solve_warm(net, prob, rel_resid=1e-4, NitMax=200):
    """Warm start + device cgne with the zero-init fallback."""
    x0 = warm_x0(net, prob["xin"])
    return solve(prob["E"], prob["y"], x0, rel_resid, NitMax, fallback=True)


def solve_zero(prob, rel_resid=1e-4, NitMax=200):
    x0 = torch.zeros_like(prob["xin"])
    return solve(prob["E"], prob["y"], x0, rel_resid, NitMax)


# ---------------------------------------------------------------------------
# training

def solve_refined(prob, rel_resid, NitMax, rounds=1):
    """Zero-init cgne followed by restarts from the current iterate (each
    restart recomputes the residual b - E'E x explicitly).  With complex64
    operators the relative residual stalls near 1e-5 and the restarts do
    not lower it below that floor (the explicit residual carries the same
    FP32 operator error), so a 1e-6 target stops at the floor.  Returns (x,
    final relative residual per slice, total iterations per slice)."""
    x0 = torch.zeros_like(prob["xin"])
    tot = None
    rel = None
    for k in range(rounds):
        x, nit, st, _, res = solve(prob["E"], prob["y"], x0, rel_resid, NitMax)
        if k == 0:
            b2 = np.array([r[0] for r in res])      # zero start: resSq[0] = |b|^2
        rel = np.sqrt(np.array([r[-1] for r in res]) / b2)
        tot = nit if tot is None else tot + nit
        if np.all(rel <= rel_resid):
            break
        x0 = x.reshape(prob["xin"].shape)
    return x, rel, tot


def generate(fam, n, seed0, target_rel=1e-6, target_nit=100, rounds=1, log=None):
    """(inputs, targets) device tensors [n, N, N] for problems seed0 .. seed0+n-1
    (n a multiple of the family's batch); targets are cgne solutions at
    relative residual target_rel (zero init plus restarts, at most
    target_nit iterations per round)."""
    P = fam.batch
    if n % P:
        raise ValueError("n must be a multiple of the batch")
    N = fam.p["N"]
    X = torch.empty((n, N, N), dtype=torch.complex64, device=fam.device)
    T = torch.empty_like(X)
    nits, rels = [], []
    import concurrent.futures as cf
    pool = cf.ThreadPoolExecutor(1)
    seeds = [list(range(seed0 + i, seed0 + i + P)) for i in range(0, n, P)]
    fut = pool.submit(fam.host_batch, seeds[0])
    t0 = time.time()
    for k, sd in enumerate(seeds):
        host = fut.result()
        if k + 1 < len(seeds):
            fut = pool.submit(fam.host_batch, seeds[k + 1])
        prob = fam.make(sd, host)
        x, rel, nit = solve_refined(prob, target_rel, target_nit, rounds)
        X[k * P:(k + 1) * P] = prob["xin"]
        T[k * P:(k + 1) * P] = x.reshape(P, N, N)
        nits.extend(nit.tolist())
        rels.extend(rel.tolist())
        if log and (k % 20 == 0 or k + 1 == len(seeds)):
            log(f"generate: {(k + 1) * P}/{n} problems, {time.time() - t0:.1f} s, "
                f"target cg iterations mean {np.mean(nits):.1f} max {np.max(nits)}, "
                f"rel resid max {np.max(rels):.2e}")
    pool.shutdown()
    return X, T, np.array(nits), np.array(rels)


def fit_input_scale(X, T):
    """Least-squares scalar c of T ~ c X over the training set."""
    num = float(torch.sum((X.conj() * T).real))
    den = float(torch.sum(X.real ** 2 + X.imag ** 2))
    return num / den


def train(net, X, T, epochs=30, batch=32, lr=3e-3, lr_final=5e-5, seed=0, log=None):
    """Adam over (X, T) on the device, cosine learning-rate decay; returns
    the per-epoch mean training loss."""
    n = X.shape[0]
    steps_per_epoch = n // batch
    total = epochs * steps_per_epoch
    g = torch.Generator(device=X.device).manual_seed(seed)
    hist = []
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device=X.device)
        losses = []
        for i in range(steps_per_epoch):
            idx = perm[i * batch:(i + 1) * batch]
            # the previous step may still read its minibatch on the library
            # stream; finish it before torch's allocator can reuse that memory
            check(_lib.lib().mm_ctx_sync(_lib.context()))
            xb = X.index_select(0, idx).contiguous()
            tb = T.index_select(0, idx).contiguous()
            torch.cuda.current_stream(X.device).synchronize()
            frac = step / max(1, total - 1)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            want = (i % 50 == 0) or i + 1 == steps_per_epoch
            L = net.train_step(xb, tb, lr_t, want_loss=want)
            if want:
                losses.append(L)
            step += 1
        check(_lib.lib().mm_ctx_sync(_lib.context()))
        hist.append(float(np.mean(losses)))
        if log:
            log(f"train: epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} ({time.time() - t0:.1f} s)")
    return hist


def evaluate(net, fam, n, seed0, rel_resid=1e-4, NitMax=200, log=None):
    """Iterations to rel_resid and wall time, warm start vs zero init, over
    n held-out problems (n a multiple of the family's batch)."""
    P = fam.batch
    out = dict(nit_warm=[], nit_zero=[], t_warm=0.0, t_zero=0.0, fellback=0, diff=[])
    for k in range(0, n, P):
        prob = fam.make(list(range(seed0 + k, seed0 + k + P)))
        torch.cuda.synchronize()
        t = time.perf_counter()
        xz, nz, _, _, _ = solve_zero(prob, rel_resid, NitMax)
        torch.cuda.synchronize()
        out["t_zero"] += time.perf_counter() - t
        t = time.perf_counter()
        xw, nw, _, fb, _ = solve_warm(net, prob, rel_resid, NitMax)
        torch.cuda.synchronize()
        out["t_warm"] += time.perf_counter() - t
        out["nit_warm"].extend(nw.tolist())
        out["nit_zero"].extend(nz.tolist())
        out["fellback"] += int(fb.sum())
        d = (torch.linalg.vector_norm((xw - xz).reshape(P, -1), dim=1) /
             torch.linalg.vector_norm(xz.reshape(P, -1), dim=1)).cpu().numpy()
        out["diff"].extend(d.tolist())
    out["nit_warm"] = np.array(out["nit_warm"])
    out["nit_zero"] = np.array(out["nit_zero"])
    out["diff"] = np.array(out["diff"])
    if log:
        log(f"evaluate: {n} problems, iterations zero {out['nit_zero'].mean():.2f} "
            f"warm {out['nit_warm'].mean():.2f} (ratio {out['nit_zero'].mean() / out['nit_warm'].mean():.3f}), "
            f"time zero {out['t_zero']:.2f} s warm {out['t_warm']:.2f} s, fallbacks {out['fellback']}")
    return out


def init_params(seed, channels=(2, 32, 32, 32, 32, 32, 2)):
    """He-normal weights, zero biases, last layer scaled by 0.1 (so the
    untrained network is close to the identity through its residual skip)."""
    rng = np.random.default_rng(seed)
    parts = []
    L = len(channels) - 1
    for l in range(L):
        cin, cout = channels[l], channels[l + 1]
        w = rng.standard_normal((cout, cin, 3, 3)) * np.sqrt(2.0 / (cin * 9))
        if l == L - 1:
            w *= 0.1
        parts += [w.reshape(-1), np.zeros(cout)]
    return np.concatenate(parts).astype(np.float32)


def main(argv=None):
    """Train the warm-start network from scratch on the device and save it:
    ``python -m uwobrainhack2026-matmri_amd.warmstart`` (defaults: 20000
    training problems, 1000 held-out problems)."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=20000)
    ap.add_argument("--n-eval", type=int, default=1000)
    ap.add_argument("--plan-batch", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--target-rel", type=float, default=1e-6)
    ap.add_argument("--target-nit", type=int, default=100)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    if a.n_train % a.plan_batch or a.n_eval % a.plan_batch:
        raise ValueError("--n-train and --n-eval must be multiples of --plan-batch")
    N = PROBLEM["N"]
    t0 = time.time()
    fam = ProblemFamily(a.plan_batch)
    log(f"plan: {a.plan_batch} slices, {time.time() - t0:.1f} s")
    X, T, nits, rels = generate(fam, a.n_train, TRAIN_SEED0, target_rel=a.target_rel, target_nit=a.target_nit,
                                log=log)
    net = WarmStartNet(N, N, init_params(a.init_seed))
    net.input_scale = fit_input_scale(X, T)
    X *= net.input_scale
    log(f"input scale {net.input_scale:.5f}")
    hist = train(net, X, T, a.epochs, a.batch, a.lr, log=log)
    del X, T
    ev = evaluate(net, fam, a.n_eval, TRAIN_SEED0 + HELDOUT_OFFSET, log=log)
    meta = dict(problem=PROBLEM, n_train=a.n_train, train_seed0=TRAIN_SEED0, target_rel_resid=a.target_rel,
                target_max_iterations=a.target_nit, target_iterations_mean=float(nits.mean()),
                target_iterations_max=int(nits.max()), target_rel_resid_max=float(rels.max()), epochs=a.epochs, batch=a.batch, lr=a.lr,
                init_seed=a.init_seed, loss_per_epoch=hist, n_eval=a.n_eval,
                eval_seed0=TRAIN_SEED0 + HELDOUT_OFFSET,
                eval_iterations_zero=float(ev["nit_zero"].mean()),
                eval_iterations_warm=float(ev["nit_warm"].mean()), eval_fallbacks=ev["fellback"],
                wall_seconds=time.time() - t0)
    net.save(a.out, meta)
    log(f"saved {a.out} ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
