"""Trajectory resampling and delay search, FP64 (oracle, test infrastructure).

* :func:`makima` -- MATLAB's ``interp1(x, y, xq, 'makima', nan)``: the
  modified Akima piecewise cubic Hermite interpolant (C. Moler, "Makima
  piecewise cubic interpolation", MathWorks blog, 2019): slopes
  ``delta_i = (y_{i+1} - y_i) / (x_{i+1} - x_i)`` extended past both ends by
  ``delta_{-1} = 2 delta_0 - delta_1``, ``delta_{-2} = 2 delta_{-1} - delta_0``
  (and mirrored at the far end); node derivatives
  ``d_i = (w1 delta_{i-1} + w2 delta_i) / (w1 + w2)`` with
  ``w1 = |delta_{i+1} - delta_i| + |delta_{i+1} + delta_i| / 2`` and
  ``w2 = |delta_{i-1} - delta_{i-2}| + |delta_{i-1} + delta_{i-2}| / 2``
  (``d_i = 0`` when both vanish); NaN outside ``[x_1, x_end]``.  Restated
  from that published method; MATLAB itself is absent.
* :func:`interpTrajTime` -- ``findDel/interpTrajTime.m:1-27``.
* :func:`findDelAuto` -- ``findDel/findDelAuto.m:1-143`` over the oracle's
  ``sampHighOrder`` (direct), ``rcvrOp``, ``mrSampFunc`` and ``cgne``.

Pinned by (``tests/test_oracle_finddel.py``): SciPy's independent
``Akima1DInterpolator(method="makima")`` on random data, exactness on linear
data, no overshoot on flat steps (the property makima exists for), NaN
extrapolation; ``interpTrajTime`` by a shift of a linear ramp (closed form);
``findDelAuto`` by recovering a known delay from data simulated with it.
"""
import numpy as np

from .highorder import HighOrderPlan
from .mrsamp import mrSampFunc
from .rcvr import RcvrPlan
from .solvers import cgne


def _slopes_ext(x, y):
    """delta_{-2} .. delta_{n-1+1}: n-1 interval slopes plus two on each side."""
    d = np.diff(y, axis=0) / np.diff(x)[:, None]
    n = d.shape[0]
    if n == 1:
        return np.concatenate([d, d, d, d, d], axis=0)
    dm1 = 2 * d[0] - d[1]
    dm2 = 2 * dm1 - d[0]
    dp1 = 2 * d[-1] - d[-2]
    dp2 = 2 * dp1 - d[-1]
    return np.concatenate([dm2[None], dm1[None], d, dp1[None], dp2[None]], axis=0)


def makima(x, y, xq):
    """``interp1(x, y, xq, 'makima', nan)`` along the first dimension of y."""
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    y = np.asarray(y, dtype=np.float64)
    vec = y.ndim == 1
    y2 = y.reshape(len(x), -1)
    xq = np.asarray(xq, dtype=np.float64).reshape(-1)
    m = _slopes_ext(x, y2)                    # m[k] = delta_{k-2}
    n = len(x)
    d = np.zeros_like(y2)
    for i in range(n):
        dm2, dm1, d0, dp1 = m[i], m[i + 1], m[i + 2], m[i + 3]
        w1 = np.abs(dp1 - d0) + 0.5 * np.abs(dp1 + d0)
        w2 = np.abs(dm1 - dm2) + 0.5 * np.abs(dm1 + dm2)
        s = w1 + w2
        with np.errstate(invalid="ignore", divide="ignore"):
            d[i] = np.where(s > 0, (w1 * dm1 + w2 * d0) / np.where(s > 0, s, 1), 0.0)
    out = np.full((len(xq), y2.shape[1]), np.nan)
    inside = (xq >= x[0]) & (xq <= x[-1])
    i = np.clip(np.searchsorted(x, xq, side="right") - 1, 0, n - 2)
    h = x[i + 1] - x[i]
    s = (xq - x[i]) / h
    s2, s3 = s * s, s * s * s
    h00, h10 = 2 * s3 - 3 * s2 + 1, s3 - 2 * s2 + s
    h01, h11 = -2 * s3 + 3 * s2, s3 - s2
    v = (h00[:, None] * y2[i] + (h10 * h)[:, None] * d[i] + h01[:, None] * y2[i + 1]
         + (h11 * h)[:, None] * d[i + 1])
    out[inside] = v[inside]
    return out[:, 0] if vec else out.reshape((len(xq),) + y.shape[1:])


def interpTrajTime(phs_spha, phs_conc, tdwell_in, delTime, datatime):
    """``findDel/interpTrajTime.m:1-27``: each array resampled along its first
    dimension from times ``-delTime + tdwell_in*(0:L-1)`` to ``datatime``."""
    empty = lambda a: a is None or np.size(a) == 0
    if empty(phs_spha) and empty(phs_conc):
        raise ValueError("at least one of phs_spha or phs_conc must be non-empty!")
    L = np.shape(phs_conc)[0] if empty(phs_spha) else np.shape(phs_spha)[0]
    trajTim = -delTime + tdwell_in * np.arange(L)
    tq = np.asarray(datatime, dtype=np.float64).reshape(-1)
    a = None if empty(phs_spha) else makima(trajTim, phs_spha, tq)
    b = None if empty(phs_conc) else makima(trajTim, phs_conc, tq)
    return a, b


def _conv_same(a, kern):
    """``conv2(a, kern, 'same')`` for a column kernel along the first dimension."""
    k = len(kern)
    out = np.zeros_like(a)
    c = (k - 1) // 2 if k % 2 else k // 2
    L = a.shape[0]
    for r in range(L):
        for j in range(k):
            src = r + c - j
            if 0 <= src < L:
                out[r] += kern[j] * a[src]
    return out


def findDelAuto(delSk0, data_in, tdwell, phs_spha, phs_conc, datatime, phs_sph_grid, b0map, Crcvr, maxNit_cgne,
                delJumpFact=3, numCoarseSearch=2):
    """``findDel/findDelAuto.m:1-143`` (direct approach, FP64); returns
    ``(delSk, delSk_perIt)``."""
    R = RcvrPlan(Crcvr)
    dts = float(datatime[1] - datatime[0])
    kern = np.array([1 / 8, 1 / 4, 0, -1 / 4, -1 / 8])
    dspha = _conv_same(np.asarray(phs_spha, np.float64), kern) / tdwell           # :73-75
    dconc = None if phs_conc is None else _conv_same(np.asarray(phs_conc, np.float64), kern) / tdwell
    t = np.asarray(datatime, np.float64).reshape(-1)

    def build(dl):
        a, b = interpTrajTime(phs_spha, phs_conc, tdwell, dl, datatime)
        return HighOrderPlan(b0map, t, a.T, None if b is None else b.T, phs_sph_grid)

    delSk = delSk0
    if numCoarseSearch > 0:                                                        # :77-101
        resmin = np.inf
        for ct in range(-numCoarseSearch, numCoarseSearch + 1):
            dl = delSk0 + ct * dts
            S = build(dl)
            A = lambda x, tr, S=S: mrSampFunc(x, tr, S, R)
            _, resvec, *_ = cgne(A, data_in, None, maxNit_cgne, {"stopOnResInc": 0})
            if np.real(resvec[-1]) < resmin:
                delmin, resmin = dl, np.real(resvec[-1])
        delSk = delmin
    ntry = 0
    delChange = np.inf
    delChangePrev = 0.0
    per_it = [delSk]
    nit = maxNit_cgne
    resvec = None
    while abs(delChange) > 0.005:                                                  # :106-141
        S = build(delSk)
        A = lambda x, tr, S=S: mrSampFunc(x, tr, S, R)
        opt = {"stopOnResInc": 1}
        if ntry > 0:
            opt = {"stopOnResInc": 0}
            nit = len(resvec) - 1
        x, resvec, *_ = cgne(A, data_in, None, nit, opt)
        bhat = np.asarray(data_in) - A(x, "notransp")
        a, b = interpTrajTime(dspha, dconc, tdwell, delSk, datatime)
        S.set_phi_div(a.T, None if b is None else b.T)
        bhh = A(x, "notransp")
        bv, hv = bhat.reshape(-1, order="F"), bhh.reshape(-1, order="F")
        delChange = delJumpFact * np.real(np.vdot(hv, bv) / np.vdot(hv, hv))       # bhathat(:) \\ bhat(:)
        delSk = delSk + delChange
        if ntry > 0 and np.sign(delChangePrev) != np.sign(delChange) and delJumpFact > 1:
            delJumpFact = delJumpFact / 2
        delChangePrev = delChange
        ntry += 1
        per_it.append(delSk)
        if ntry > 50:
            break
    return delSk, np.array(per_it)
