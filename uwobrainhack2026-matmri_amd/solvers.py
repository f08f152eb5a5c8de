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


def cgne(A, b, x0=None, NitMax=None, opt=None, gam=None, R=None, W=None):
    """Returns ``(x, resSqAll, mseAll, xnormAll, xdiffAll, stopThresh, status)``
    like ``cgne.m:1``.  ``x0=None`` is the reference default ``x0 = A'b``."""
    NitMax = 1000 if NitMax is None else int(NitMax)
    opt = dict(opt or {})
    if gam is not None and np.any(np.asarray(gam) != 0):
        raise NotImplementedError("cgne: Tikhonov terms are not implemented on the device path")
    if opt.get("gtruth") is not None:
        raise NotImplementedError("cgne: opt.gtruth (mse history) is not implemented on the device path")
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
    nit = np.zeros(ns, dtype=np.int32)
    status = np.zeros(ns, dtype=np.int32)
    if _is_torch(b):
        yb = b.to(torch.complex64).contiguous().reshape(-1)
        x0b = None if x0 is None else x0.to(torch.complex64).contiguous().reshape(-1)
        x = torch.empty(ni, dtype=torch.complex64, device=b.device)
        torch.cuda.current_stream(b.device).synchronize()
        check(L.mm_cgne_solve(E._plan.h, ptr(yb), ptr(x0b), ptr(x), NitMax, C.byref(o), hres.ctypes.data,
                              hxn.ctypes.data, hxd.ctypes.data, nit.ctypes.data, status.ctypes.data, 1))
        x = x.reshape(tuple(reversed(E.image_shape)))
    else:
        yb = np.asfortranarray(np.asarray(b)).reshape(-1, order="F").astype(np.complex64)
        x0b = None if x0 is None else np.asfortranarray(np.asarray(x0)).reshape(-1, order="F").astype(np.complex64)
        x = np.empty(ni, dtype=np.complex64)
        check(L.mm_cgne_solve(E._plan.h, yb.ctypes.data, ptr(x0b), x.ctypes.data, NitMax, C.byref(o),
                              hres.ctypes.data, hxn.ctypes.data, hxd.ctypes.data, nit.ctypes.data,
                              status.ctypes.data, 0))
        x = x.reshape(E.image_shape, order="F")
    if ns == 1:
        n = int(nit[0]) + 1
        return (x, hres[:n], np.zeros(n), hxn[:n], hxd[:n], stopThresh, int(status[0]))
    res = [hres[s * (NitMax + 1): s * (NitMax + 1) + nit[s] + 1] for s in range(ns)]
    xn = [hxn[s * (NitMax + 1): s * (NitMax + 1) + nit[s] + 1] for s in range(ns)]
    xd = [hxd[s * (NitMax + 1): s * (NitMax + 1) + nit[s] + 1] for s in range(ns)]
    return (x, res, None, xn, xd, stopThresh, status.tolist())


def set_cg_graph(A, use_graph):
    """Switch hipGraph capture of the cgne iteration on or off for operator A."""
    E = _device_sense(A, None)
    check(_lib.lib().mm_cg_set_graph(E._plan.h, 1 if use_graph else 0))


def _as_handle(A):
    if isinstance(A, LinearOp):
        return lambda x, tr: A * x if tr == "notransp" else A.H * x
    return A


def _vdot(a, b):
    if _is_torch(a):
        return torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128)).item()
    return np.vdot(np.asarray(a).reshape(-1, order="F"), np.asarray(b).reshape(-1, order="F"))


def _norm(a):
    return float(np.sqrt(np.real(_vdot(a, a))))


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
    """``powermethod.m:14-66``: largest eigenvalue of ``A'A`` (or ``A``)."""
    h = _as_handle(A)
    x = x0
    x_prev = x0
    nit = 0
    while True:
        x = h(h(x, "notransp"), "transp") if doAtA else h(x, "notransp")
        x = x / _norm(x)
        xdiff = _norm(x - x_prev)
        x_prev = x
        nit += 1
        if xdiff < xdiffThresh or nit >= NitMax:
            break
    AtAx = h(h(x, "notransp"), "transp") if doAtA else h(x, "notransp")
    return _vdot(x, AtAx) / _vdot(x, x)


def bfista(A, b, R, lam, x0=None, NitMax=200, opt=None):
    """``bfista.m:1-246``.  Returns ``(x, resSqAll, RxAll, testAll, maxEig)``."""
    opt = dict(opt or {})
    maxEig = opt.get("maxEig")
    resThresh = opt.get("resThresh", 1e-5)
    xMask = opt.get("xMask")
    nRestart = opt.get("nRestart", NitMax + 1)
    h = _as_handle(A)
    Rh = _as_handle(R)
    if x0 is None:
        x0 = h(b, "transp")
    if maxEig is None:
        maxEig = powermethod(h, x0)
    stepSz = 0.9 / (2 * abs(maxEig))
    y = x0
    resSqAll = np.zeros(NitMax + 1)
    RxAll = np.zeros(NitMax + 1)
    testAll = np.zeros((NitMax + 1, 2))
    nit, nitToRestart, t_prev = 1, 1, 1.0
    x_prev = x0
    x = x0
    finished = False
    while not finished:
        residual_y = h(y, "notransp") - b
        grad = 2 * h(residual_y, "transp")
        resSqAll[nit - 1] = np.real(_vdot(residual_y, residual_y))
        g = y - stepSz * grad
        x = Rh(g, "notransp")
        RxAll[nit - 1] = float((lam * x).abs().sum()) if _is_torch(x) else np.sum(np.abs(lam * x))
        x = softthresh(x, lam * stepSz)
        x = Rh(x, "transp")
        if xMask is not None:
            x = x * xMask
        xdiff = x - x_prev
        t = 0.5 * (1 + np.sqrt(1 + 4 * t_prev ** 2))
        y = x + (t_prev - 1) / t * xdiff
        t_prev = t
        x_prev = x
        testVal = None
        if nit >= NitMax:
            finished = True
        if nit > 1:
            numAve = 3
            tv = np.abs(np.diff(resSqAll[:nit] + RxAll[:nit])) / (resSqAll[nit - 1] + RxAll[nit - 1])
            i1 = max(1, nit - numAve)
            testVal = np.mean(tv[i1 - 1:])
            testAll[nit - 1, 0] = testVal
            testAll[nit - 1, 1] = np.real(_vdot(xdiff, xdiff)) / np.real(_vdot(x, x))
        if nit > 1 and testVal < resThresh:
            finished = True
        nit += 1
        nitToRestart += 1
        if nitToRestart >= nRestart:
            nitToRestart = 1
            t_prev = 1.0
    n = nit - 1
    testAll = testAll[:n].copy()
    testAll[0, :] = np.nan
    return x, resSqAll[:n], RxAll[:n], testAll, maxEig
