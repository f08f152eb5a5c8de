"""Solvers with the reference's calling conventions.

* :func:`cgne`        -- ``iterativeSolvers/cgne.m`` (signature ``:1``, options ``:21-89``).
  When ``A`` is a device SENSE operator (:class:`senseOp`, or the
  :class:`mrSampHandle` of ``@(x,tr) mrSampFunc(x,tr,S,R)``) the whole solve
  runs on the GPU through ``mm_cgne_solve`` (FP64 scalars, hipGraph per
  iteration).  Tikhonov terms (``gam``/``R``) and ``opt.gtruth`` are not part
  of the device path and raise ``NotImplementedError``.
* :func:`bfista`, :func:`powermethod`, :func:`softthresh` --
  ``iterativeSolvers/bfista.m``, ``powermethod.m``, ``general/softthresh.m``:
  host-side loops that call the device operators (the regulariser ``R`` is
  host code in the reference too).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CgOpts, check, ptr
from .operators import LinearOp, mrSampHandle, senseOp

try:
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


def _cg_opts(opt, NitMax):
    o = CgOpts()
    _lib.lib().mm_cg_opts_default(C.byref(o))
    opt = dict(opt or {})

    def g(name, default):
        v = opt.get(name, default)
        return -1.0 if v is None else float(v)
    o.resThresh = g("resThresh", 1e-10)
    o.resThreshRel = g("resThreshRel", None)
    o.resChangeThresh = g("resChangeThresh", None)
    o.xChangeThresh = g("xChangeThresh", 1e-10)
    o.expResN = int(opt.get("expResN", 100))
    o.stopOnResInc = int(opt.get("stopOnResInc", 0))
    o.resIncThresh = float(opt.get("resIncThresh", 1.0))
    o.stopOnXdifInc = int(opt.get("stopOnXdifInc", 0))
    o.nseExtraIt = int(opt.get("nseExtraIt", 4))
    o.x0Fallback = int(bool(opt.get("x0Fallback", 0)))
    o.stopThresh = -1.0
    return o


def _device_sense(A, W):
    if isinstance(A, senseOp):
        if W is not None:
            return senseOp(A.S, A.R, W)
        return A
    if isinstance(A, mrSampHandle):
        if A.opt.get("tikReg") is not None or A.opt.get("sumDims") is not None or A.opt.get("W") is not None:
            return None
        return A.sense(W)
    return None


class _Eye:
    """The identity as a Tikhonov operator R of :func:`cgne` (``R'R = I``)."""

    def __repr__(self):
        return "eye"


eye = _Eye()


def _regs(gam, R):
    """(gam_n, R_n) pairs of cgne's Tikhonov terms (cgne.m:91-133): scalar gam
    with one R, or lists (MATLAB cells) of both; R_n a device operator or
    ``eye``.  All-zero gam means none (cgne.m:91-93)."""
    if gam is None:
        return []
    gl = list(gam) if isinstance(gam, (list, tuple)) else [gam]
    if not isinstance(gam, (list, tuple)) and np.ndim(gam) > 0:
        raise NotImplementedError("cgne: gam must be a scalar (or a list of scalars with a list of R)")
    if all(float(np.real(g)) == 0.0 for g in gl):
        return []
    if R is None:
        raise ValueError("cgne: Tikhonov terms need R (cgne.m:94-96)")
    rl = list(R) if isinstance(R, (list, tuple)) else [R]
    if len(rl) != len(gl):
        raise ValueError("cgne: gam and R must have the same number of terms")
    out = []
    for g, r in zip(gl, rl):
        if r is eye:
            out.append((float(g), None))
        elif isinstance(r, LinearOp):
            out.append((float(g), r))
        else:
            raise NotImplementedError("cgne: R must be a device operator (LinearOp) or solvers.eye on the device path")
    return out


def cgne(A, b, x0=None, NitMax=None, opt=None, gam=None, R=None, W=None):
    """Returns ``(x, resSqAll, mseAll, xnormAll, xdiffAll, stopThresh, status)``
    like ``cgne.m:1``.  ``x0=None`` is the reference default ``x0 = A'b``.
    Tikhonov terms ``gam``/``R`` (R a device operator or :data:`eye`) and
    ``opt.gtruth`` (the mse history) run inside the device solve."""
    NitMax = 1000 if NitMax is None else int(NitMax)
    opt = dict(opt or {})
    regs = _regs(gam, R)
    E = _device_sense(A, W)
    if E is None:
        raise NotImplementedError("cgne: A must be a device SENSE operator (senseOp / mrSampHandle)")
    o = _cg_opts(opt, NitMax)
    stopThresh = None
    noiseVar = opt.get("noiseVar")
    if noiseVar is not None and noiseVar > 0:          # cgne.m:149-162
        rng = np.random.default_rng(opt.get("noiseSeed", 0))
        shp = E.data_shape
        if np.iscomplexobj(b) or _is_torch(b):
            nse_in = np.sqrt(noiseVar / 2) * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
        else:
            nse_in = np.sqrt(noiseVar) * rng.standard_normal(shp)
        nse = E.H * nse_in
        stopThresh = float(opt.get("nseThreshFact", 1.0)) * float(np.real(np.vdot(nse, nse)))
        o.stopThresh = stopThresh
    ns = E.nslice
    L = _lib.lib()
    ni, nd = E._shapes()
    hres = np.zeros(ns * (NitMax + 1))
    hxn = np.zeros_like(hres)
    hxd = np.zeros_like(hres)
    hmse = np.zeros_like(hres)
    nit = np.zeros(ns, dtype=np.int32)
    status = np.zeros(ns, dtype=np.int32)
    gam_arr = (C.c_double * max(1, len(regs)))(*[g for g, _ in regs])
    R_arr = (C.c_void_p * max(1, len(regs)))(*[None if r is None else r._plan.h.value for _, r in regs])
    gamv = C.cast(gam_arr, C.c_void_p) if regs else None
    Rv = C.cast(R_arr, C.c_void_p) if regs else None
    gt = opt.get("gtruth")
    if _is_torch(b):
        yb = b.to(torch.complex64).contiguous().reshape(-1)
        x0b = None if x0 is None else x0.to(torch.complex64).contiguous().reshape(-1)
        gtb = None if gt is None else (gt if _is_torch(gt) else torch.as_tensor(
            np.ascontiguousarray(np.transpose(np.asarray(gt))), device=b.device)).to(torch.complex64).reshape(-1)
        x = torch.empty(ni, dtype=torch.complex64, device=b.device)
        torch.cuda.current_stream(b.device).synchronize()
        check(L.mm_cgne_solve_ex(E._plan.h, ptr(yb), ptr(x0b), ptr(x), NitMax, C.byref(o), len(regs), gamv, Rv,
                                 ptr(gtb), hres.ctypes.data, hmse.ctypes.data, hxn.ctypes.data, hxd.ctypes.data,
                                 nit.ctypes.data, status.ctypes.data, 1))
        x = x.reshape(tuple(reversed(E.image_shape)))
    else:
        yb = np.asfortranarray(np.asarray(b)).reshape(-1, order="F").astype(np.complex64)
        x0b = None if x0 is None else np.asfortranarray(np.asarray(x0)).reshape(-1, order="F").astype(np.complex64)
        gtb = None if gt is None else np.asfortranarray(np.asarray(gt)).reshape(-1, order="F").astype(np.complex64)
        x = np.empty(ni, dtype=np.complex64)
        check(L.mm_cgne_solve_ex(E._plan.h, yb.ctypes.data, ptr(x0b), x.ctypes.data, NitMax, C.byref(o),
                                 len(regs), gamv, Rv, ptr(gtb), hres.ctypes.data, hmse.ctypes.data,
                                 hxn.ctypes.data, hxd.ctypes.data, nit.ctypes.data, status.ctypes.data, 0))
        x = x.reshape(E.image_shape, order="F")
    if ns == 1:
        n = int(nit[0]) + 1
        return (x, hres[:n], hmse[:n], hxn[:n], hxd[:n], stopThresh, int(status[0]))
    cut = lambda h: [h[s * (NitMax + 1): s * (NitMax + 1) + nit[s] + 1] for s in range(ns)]
    return (x, cut(hres), cut(hmse) if gt is not None else None, cut(hxn), cut(hxd), stopThresh, status.tolist())


def set_cg_graph(A, use_graph):
    """Switch hipGraph capture of the cgne iteration on or off for operator A."""
    E = _device_sense(A, None)
    check(_lib.lib().mm_cg_set_graph(E._plan.h, 1 if use_graph else 0))


def _as_handle(A):
    if isinstance(A, LinearOp):
        return lambda x, tr: A * x if tr == "notransp" else A.H * x
    return A


# -- device-resident vectors for bfista / powermethod ------------------------
# NumPy arrays (MATLAB column-major meaning) become complex64 CUDA tensors
# whose row-major shape is the column-major shape reversed -- the layout the
# device operators read and write in place -- and come back the same way.

def _to_dev(a):
    if _is_torch(a):
        return a.to(torch.complex64)
    a = np.asarray(a)
    return torch.as_tensor(np.ascontiguousarray(np.transpose(a)).astype(np.complex64), device="cuda")


def _to_host(t, like):
    return t if _is_torch(like) else np.transpose(t.cpu().numpy())


def _mask_dev(m):
    if m is None:
        return None
    if _is_torch(m):
        return m.to("cuda")
    return torch.as_tensor(np.ascontiguousarray(np.transpose(np.asarray(m))), device="cuda")


def _dot64(a, b):
    """conj(a) . b summed in FP64 on the device (a host complex)."""
    return torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128)).item()


def _sq64(a):
    return float(torch.sum(torch.abs(a.reshape(-1).to(torch.complex128)) ** 2).item())


def softthresh(w, lam):
    """``general/softthresh.m``: ``sign(w).*max(abs(w)-lam,0)`` (complex sign)."""
    if _is_torch(w):
        a = w.abs()
        return torch.where(a > 0, w / torch.where(a > 0, a, torch.ones_like(a)), torch.zeros_like(w)) * \
            torch.clamp(a - lam, min=0)
    a = np.abs(w)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(a > 0, w / np.where(a > 0, a, 1), 0)
    return s * np.maximum(a - lam, 0)


def powermethod(A, x0, NitMax=50, doAtA=1, xdiffThresh=1e-3):
    """``powermethod.m:14-66``: largest eigenvalue of ``A'A`` (or ``A``), the
    iterate kept on the GPU (norms and the Rayleigh quotient in FP64)."""
    h = _as_handle(A)
    x = _to_dev(x0)
    x_prev = x
    nit = 0
    while True:
        x = h(h(x, "notransp"), "transp") if doAtA else h(x, "notransp")
        x = x / float(np.sqrt(_sq64(x)))
        xdiff = float(np.sqrt(_sq64(x - x_prev)))
        x_prev = x
        nit += 1
        if xdiff < xdiffThresh or nit >= NitMax:
            break
    AtAx = h(h(x, "notransp"), "transp") if doAtA else h(x, "notransp")
    return _dot64(x, AtAx) / _sq64(x)


def bfista(A, b, R, lam, x0=None, NitMax=200, opt=None):
    """``bfista.m:1-246`` with every vector on the GPU: ``A`` runs through the
    device operators (mm_op_apply on device pointers), ``R`` is called on the
    device tensors (an orthonormal transform, ``R'R = I``, as balanced FISTA
    requires), soft-thresholding, the projections and the momentum step are
    device array operations, and the cost terms are summed in FP64.  NumPy
    inputs come back as NumPy arrays.  Options: ``maxEig``, ``resThresh``,
    ``xMask``, ``kMask`` (k-space projection, ``:161-169``), ``nRestart``.
    Returns ``(x, resSqAll, RxAll, testAll, maxEig)``."""
    opt = dict(opt or {})
    resThresh = opt.get("resThresh", 1e-5)
    nRestart = opt.get("nRestart", NitMax + 1)
    kMask = opt.get("kMask")
    if kMask is not None and not (np.ndim(kMask) == 2 and np.shape(kMask)[1] > 1):
        raise ValueError("bfista: opt.kMask must be a 2-D mask (other dimensions are not supported)")
    xMask, kMask = _mask_dev(opt.get("xMask")), _mask_dev(kMask)
    lam = float(lam) if np.ndim(lam) == 0 and not _is_torch(lam) else _mask_dev(lam)   # scalar or per coefficient
    h, Rh = _as_handle(A), _as_handle(R)
    bd = _to_dev(b)
    x0d = h(bd, "transp") if x0 is None else _to_dev(x0)          # :73-75
    maxEig = opt.get("maxEig")
    if maxEig is None:                                             # :81-89
        maxEig = powermethod(h, x0d)
    step = 0.9 / (2 * abs(maxEig))                                 # :90
    resSq, Rx, test = [], [], [[np.nan, np.nan]]
    y = x_prev = x = x0d
    t_prev = 1.0
    since_restart = 1
    for nit in range(1, NitMax + 1):                               # :127-231
        r = h(y, "notransp") - bd
        g = y - (2 * step) * h(r, "transp")
        resSq.append(_sq64(r))
        w = Rh(g, "notransp")
        Rx.append(float(torch.sum(torch.abs(lam * w).to(torch.float64)).item()))
        x = Rh(softthresh(w, lam * step), "transp")
        if xMask is not None:
            x = x * xMask
        if kMask is not None:                                      # uncentred, unscaled pair
            x = torch.fft.ifft2(torch.fft.fft2(x) * kMask)
        dx = x - x_prev
        t = 0.5 * (1 + np.sqrt(1 + 4 * t_prev ** 2))
        y = x + ((t_prev - 1) / t) * dx
        t_prev, x_prev = t, x
        stop = nit >= NitMax
        if nit > 1:                                                # :208-222
            cost = np.asarray(resSq) + np.asarray(Rx)
            tv = np.abs(np.diff(cost)) / cost[-1]
            tval = float(np.mean(tv[max(1, nit - 3) - 1:]))
            test.append([tval, _sq64(dx) / _sq64(x)])
            stop = stop or tval < resThresh
        since_restart += 1
        if since_restart >= nRestart:
            since_restart, t_prev = 1, 1.0
        if stop:
            break
    return (_to_host(x, b), np.asarray(resSq), np.asarray(Rx), np.asarray(test, dtype=float), maxEig)
