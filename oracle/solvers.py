"""Iterative solvers, FP64 (oracle, test infrastructure).

* :func:`cgne`        -- ``iterativeSolvers/cgne.m:1-341`` (loop ``:213-336``)
* :func:`bfista`      -- ``iterativeSolvers/bfista.m:1-246``
* :func:`powermethod` -- ``iterativeSolvers/powermethod.m:1-67``
* :func:`softthresh`  -- ``general/softthresh.m:1-7``

``A`` is a callable ``A(x, transp)`` with ``transp`` in ``{'notransp',
'transp'}`` (the reference's handle convention, ``cgne.m:111-115``), or an
object with ``forward``/``adjoint``.

Pinned by (``tests/test_oracle_solvers.py``): small dense systems solved by
``numpy.linalg.lstsq`` (CGNE converges to the least-squares solution, and in
exact arithmetic within n iterations), the reference's restart rule (the
first iteration is a steepest-descent step because ``mod(0, expResN) == 0``),
each stopping status of ``cgne.m:286-335``, ``powermethod`` against
``numpy.linalg.eigvalsh`` (``powermethod.m:77-95``), and ``bfista`` with
``A = I``, ``R = I`` against the closed form ``soft(b, lam/2)``; its k-space
projection ``opt.kMask`` (``bfista.m:161-169``) by the closed form of
``A = I``, ``lam = 0``: the iterates converge to the projection of ``b``,
``ifft2(kMask .* fft2(b))``, and every iterate lies in its range.
"""
import numpy as np

from .fftnc import fftnc, ifftnc


def _as_handle(A):
    if callable(A) and not hasattr(A, "forward"):
        return A

    def h(x, transp):
        return A.forward(x) if transp == "notransp" else A.adjoint(x)
    return h


def _vdot(a, b):
    return np.vdot(np.asarray(a).reshape(-1, order="F"), np.asarray(b).reshape(-1, order="F"))


def cgne(Ain, bin, x0=None, NitMax=None, opt=None, gam=None, Rin=None, W=None, noise_rng=None):
    """Returns ``(x, resSqAll, mseAll, xnormAll, xdiffAll, stopThresh, status)``
    exactly as ``cgne.m:1``; every history is complex/real FP64."""
    opt = {} if opt is None else dict(opt)
    NitMax = 1000 if NitMax is None else int(NitMax)
    resThresh = opt.get("resThresh", 1e-10)
    resChangeThresh = opt.get("resChangeThresh", None)
    xChangeThresh = opt.get("xChangeThresh", 1e-10)
    expResN = opt.get("expResN", 100)
    nseThreshFact = opt.get("nseThreshFact", 1.0)
    nseExtraIt = opt.get("nseExtraIt", 4)
    stopOnResInc = opt.get("stopOnResInc", 0)
    resIncThresh = opt.get("resIncThresh", 1.0)
    stopOnXdifInc = opt.get("stopOnXdifInc", 0)
    noiseVar = opt.get("noiseVar", None)
    gtruth = opt.get("gtruth", None)
    cell = isinstance(gam, (list, tuple))                           # gam/R as MATLAB cells
    if gam is not None and not cell and np.all(np.asarray(gam) == 0):   # :91-93
        gam = None
    if gam is None:
        Rin = None
    A = _as_handle(Ain)
    if Rin is None:
        R = None
    elif cell:                                                      # :117-126
        R = [_as_handle(r) for r in Rin]
    else:
        R = _as_handle(Rin)

    b = A(bin if W is None else W * bin, "transp")                  # :136-140
    if x0 is None:                                                  # :143-145
        x0 = b
    stopThresh = None
    if noiseVar is not None and noiseVar > 0:                       # :149-162
        rng = np.random.default_rng(0) if noise_rng is None else noise_rng
        if np.iscomplexobj(bin):
            nse_in = np.sqrt(noiseVar / 2) * (rng.standard_normal(np.shape(bin))
                                              + 1j * rng.standard_normal(np.shape(bin)))
        else:
            nse_in = np.sqrt(noiseVar) * rng.standard_normal(np.shape(bin))
        nse = A(nse_in, "transp")
        stopThresh = nseThreshFact * np.real(_vdot(nse, nse))

    def normal(v):
        q = A(A(v, "notransp") if W is None else W * A(v, "notransp"), "transp")
        if gam is not None:                                         # :171-179, :219-227
            if cell:
                for g, r in zip(gam, R):
                    q = q + g * r(r(v, "notransp"), "transp")
            else:
                q = q + gam * R(R(v, "notransp"), "transp")
        return q

    xold = x0
    res = b - normal(xold)                                          # :165-179
    p = res
    res_sq_old = _vdot(res, res)                                    # :181
    nit = 0
    resSqAll = np.zeros(NitMax + 1, dtype=np.complex128)
    resSqAll[0] = res_sq_old
    xnormAll = np.zeros(NitMax + 1, dtype=np.complex128)
    xnormAll[0] = _vdot(x0, x0)
    xdiffAll = np.zeros(NitMax + 1, dtype=np.complex128)
    xdiffAll[0] = xnormAll[0]
    mseAll = np.zeros(NitMax + 1, dtype=np.complex128)
    if gtruth is not None:
        t = np.asarray(gtruth).reshape(-1, order="F") - np.asarray(x0).reshape(-1, order="F")
        mseAll[0] = np.vdot(t, t)
    finished = False
    resIncsTotal = 0
    xdifIncsTotal = 0
    lastxdifInc = False
    status = 0
    x = x0
    while not finished:                                             # :213
        Ap = normal(p)
        alph = res_sq_old / _vdot(p, Ap)                            # :228
        x = xold + alph * p
        if nit % expResN == 0:                                      # :230-246
            res = b - normal(x)
            betaFact = 0.0
        else:
            res = res - alph * Ap
            betaFact = 1.0
        res_sq_new = _vdot(res, res)
        resSqAll[nit + 1] = res_sq_new
        xnormAll[nit + 1] = _vdot(x, x)
        xdiff = x - xold
        xdiffAll[nit + 1] = _vdot(xdiff, xdiff)
        if gtruth is not None:
            t = np.asarray(gtruth).reshape(-1, order="F") - np.asarray(x).reshape(-1, order="F")
            mseAll[nit + 1] = np.vdot(t, t)
        if xChangeThresh is not None and np.real(xdiffAll[nit + 1]) < xChangeThresh:   # :286
            finished, status = True, 10
        if resThresh is not None and np.real(res_sq_new) < resThresh:                   # :291
            finished, status = True, 1
        if resChangeThresh is not None and abs((resSqAll[nit + 1] - resSqAll[nit]) / resSqAll[nit]) < resChangeThresh:
            finished, status = True, 2                                                  # :296
        if stopThresh is not None and np.real(res_sq_new) < stopThresh:                # :301
            NitMax = nit + nseExtraIt
            stopThresh = 0
            status = 3
        if stopOnResInc > 0 and np.real(resSqAll[nit + 1] / resSqAll[nit]) > resIncThresh:   # :308
            resIncsTotal += 1
            if resIncsTotal >= stopOnResInc:
                finished, status = True, 4
        if stopOnXdifInc > 0 and np.real(xdiffAll[nit + 1]) > np.real(xdiffAll[nit]):       # :315
            if not lastxdifInc:
                xdifIncsTotal += 1
                lastxdifInc = True
                if xdifIncsTotal >= stopOnXdifInc:
                    finished, status = True, 5
        else:
            lastxdifInc = False
        p = res + betaFact * (res_sq_new / res_sq_old) * p          # :328
        res_sq_old = res_sq_new
        xold = x
        nit += 1
        if nit >= NitMax:                                           # :332-335
            finished, status = True, 0
    n = nit + 1
    return x, resSqAll[:n], mseAll[:n], xnormAll[:n], xdiffAll[:n], stopThresh, status


def softthresh(w, lam):
    """``general/softthresh.m:4``: ``sign(w) .* max(abs(w) - lam, 0)`` with
    MATLAB's complex ``sign(w) = w ./ abs(w)`` (0 at 0)."""
    w = np.asarray(w)
    a = np.abs(w)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(a > 0, w / np.where(a > 0, a, 1), 0)
    return s * np.maximum(a - lam, 0)


def powermethod(Ain, x0, NitMax=50, doAtA=1, xdiffThresh=1e-3):
    """``powermethod.m:14-66``."""
    A = _as_handle(Ain)
    x = x0
    x_prev = x0
    nit = 0
    while True:
        x = A(A(x, "notransp"), "transp") if doAtA else A(x, "notransp")
        x = x / np.linalg.norm(np.asarray(x).reshape(-1))
        xdiff = np.linalg.norm(np.asarray(x).reshape(-1) - np.asarray(x_prev).reshape(-1))
        x_prev = x
        nit += 1
        if xdiff < xdiffThresh or nit >= NitMax:
            break
    AtAx = A(A(x, "notransp"), "transp") if doAtA else A(x, "notransp")
    return _vdot(x, AtAx) / _vdot(x, x)


def bfista(Ain, bin, Rin, lam, x0=None, NitMax=200, opt=None):
    """``bfista.m:1-246``.  Returns ``(x, resSqAll, RxAll, testAll, maxEig)``."""
    opt = {} if opt is None else dict(opt)
    maxEig = opt.get("maxEig", None)
    resThresh = opt.get("resThresh", 1e-5)
    xMask = opt.get("xMask", None)
    kMask = opt.get("kMask", None)
    if kMask is not None and not (np.ndim(kMask) == 2 and np.shape(kMask)[1] > 1):
        raise ValueError("TODO: code for other dims, or require an fftFunc supplied to bfista")   # :167
    nRestart = opt.get("nRestart", NitMax + 1)
    A = _as_handle(Ain)
    R = _as_handle(Rin)
    if x0 is None:                                                  # :73-75
        x0 = A(bin, "transp")
    if maxEig is None:                                              # :81-89
        maxEig = powermethod(A, x0)
    stepSz = 0.9 / (2 * abs(maxEig))                                # :90
    y = x0
    resSqAll = np.zeros(NitMax + 1)
    RxAll = np.zeros(NitMax + 1)
    testAll = np.zeros((NitMax + 1, 2))
    nit = 1
    nitToRestart = 1
    t_prev = 1.0
    x_prev = x0
    x = x0
    finished = False
    while not finished:                                             # :127
        residual_y = A(y, "notransp") - bin
        grad = 2 * A(residual_y, "transp")
        resSqAll[nit - 1] = np.real(_vdot(residual_y, residual_y))
        g = y - stepSz * grad
        x = R(g, "notransp")
        RxAll[nit - 1] = np.sum(np.abs(lam * x))
        x = softthresh(x, lam * stepSz)
        x = R(x, "transp")
        if xMask is not None:
            x = x * xMask
        if kMask is not None:                                       # :161-169
            x = ifftnc(fftnc(x, 2, 0, 0) * kMask, 2, 0, 0)
        xdiff = x - x_prev                                          # :195-201
        t = 0.5 * (1 + np.sqrt(1 + 4 * t_prev ** 2))
        y = x + (t_prev - 1) / t * xdiff
        t_prev = t
        x_prev = x
        testVal = None
        if nit >= NitMax:
            finished = True
        if nit > 1:                                                 # :208-222
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
