"""Interpolated expanded-encoding operator (sampHighOrder approach 1), FP64
(oracle, test infrastructure).

Restates ``operators/@sampHighOrder/sampHighOrder.m`` for axial 2-D grids
(no SMS, ``subFactSpc = 1``):

* ``createNufftForInterp`` (``:624-701``): the linear terms 2 and 3 of
  ``phs_spha`` become NUFFT locations in cycles/pixel, their sign and axis
  from the grid orientation, plus the phase ramp ``phsShft`` of the grid
  centre; a ``nufftOp`` with its default W = 6, os = 1.5;
* ``subSampTime`` (``:959-982``): Gaussian smoothing of the coefficients
  (``filtNd(2*(halfsz+1), 0.5, 1, 'gauss')`` without its first tap,
  normalised), every ``subFact``-th time point plus the last;
* ``prepForInterp`` (``:703-902``): ``b = exp(1i*Phi_res)`` of the remaining
  terms (``[1, 4:Nh]``, concomitant, B0) at the subsampled times, its SVD
  truncated at the first singular value below ``svdThresh`` times the
  largest (inclusive, ``:827-836``), ``svdTime = conj(V*S)`` interpolated back
  to every time point with ``pchip`` (real and imaginary parts separately)
  and ``svdSpace = U``;
* ``useInterpWorker`` (``:314-359``): forward
  ``y = sum_n svdTime_n .* phsShft .* NUFFT(x .* svdSpace_n)``, adjoint its
  Hermitian transpose.

The truncated SVD is taken from the full FP64 SVD (the reference calls
``svds``; the top singular triplets are the same).  ``pchip`` follows
Fritsch and Butland's harmonic-mean derivative with the shape-preserving
three-point ends (the method MATLAB's pchip and SciPy's PchipInterpolator
document).

Pinned by (``tests/test_oracle_highorder_interp.py``): the reference's own
approach-1 test (``@sampHighOrder/tests.m:66-148``: random polynomial
coefficients, four grid orientations, adjoint identity 1e-8, distance to the
direct operator below 1e-4 per sample); convergence to the direct operator
as svdThresh -> 0 with subFact = 1 (up to the NUFFT's W = 6 error, which a
wrong orientation sign or phase ramp misses by O(1)); pchip against SciPy
and on monotone data; the smoothing window's closed form.
"""
import numpy as np

from .highorder import HighOrderPlan
from .nufft import NufftPlan


def normpdf(x, mu, sigma):
    return np.exp(-0.5 * ((np.asarray(x, float) - mu) / sigma) ** 2) / (sigma * np.sqrt(2 * np.pi))


def gauss_window(sz, alpha):
    """``filtNd(sz, alpha, 1, 'gauss')`` (general/filtNd.m:13-56)."""
    if sz % 2 == 0:
        w = normpdf(np.arange(-sz / 2, sz / 2), 0, sz * alpha / 2.355)
    else:
        h = sz // 2
        w = normpdf(np.arange(-h, h + 1), 0, sz * alpha / 2.355)
    return w / w.max()


def pchip(x, y, xq):
    """``interp1(x, y, xq, 'pchip')`` along the first dimension of y; complex
    values are interpolated as their real and imaginary parts."""
    y = np.asarray(y)
    if np.iscomplexobj(y):
        return pchip(x, y.real, xq) + 1j * pchip(x, y.imag, xq)
    x = np.asarray(x, float)
    y2 = y.reshape(len(x), -1)
    h = np.diff(x)
    delta = np.diff(y2, axis=0) / h[:, None]
    n = len(x)
    d = np.zeros_like(y2)
    if n == 2:
        d[:] = delta[0]
    else:
        for k in range(1, n - 1):
            w1, w2 = 2 * h[k] + h[k - 1], h[k] + 2 * h[k - 1]
            a, b = delta[k - 1], delta[k]
            same = (np.sign(a) * np.sign(b)) > 0
            with np.errstate(divide="ignore", invalid="ignore"):
                hm = (w1 + w2) / (w1 / a + w2 / b)
            d[k] = np.where(same, hm, 0)

        def end(h0, h1, del0, del1):
            dd = ((2 * h0 + h1) * del0 - h0 * del1) / (h0 + h1)
            dd = np.where(np.sign(dd) != np.sign(del0), 0, dd)
            dd = np.where((np.sign(del0) != np.sign(del1)) & (np.abs(dd) > np.abs(3 * del0)), 3 * del0, dd)
            return dd
        d[0] = end(h[0], h[1], delta[0], delta[1])
        d[-1] = end(h[-1], h[-2], delta[-1], delta[-2])
    xq = np.asarray(xq, float)
    i = np.clip(np.searchsorted(x, xq, side="right") - 1, 0, n - 2)
    hh = x[i + 1] - x[i]
    s = (xq - x[i]) / hh
    s2, s3 = s * s, s * s * s
    out = ((2 * s3 - 3 * s2 + 1)[:, None] * y2[i] + ((s3 - 2 * s2 + s) * hh)[:, None] * d[i]
           + (-2 * s3 + 3 * s2)[:, None] * y2[i + 1] + ((s3 - s2) * hh)[:, None] * d[i + 1])
    return out.reshape((len(xq),) + y.shape[1:])


def nufft_for_interp(grid, phs_spha):
    """``createNufftForInterp`` (``:624-701``) for a 2-D axial grid: returns
    (kloc [K x 2] cycles/pixel, phsShft [K] complex, image size)."""
    gx, gy, gz = (np.asarray(grid[c], float) for c in ("x", "y", "z"))
    if gz.shape[0] > 1 and gz.shape[1] > 1 and gz[1, 1] - gz[0, 0] != 0:
        raise ValueError("non-axial slices not supported for interpolated approach")
    h2, h3 = phs_spha[1], phs_spha[2]
    N1, N2 = gx.shape
    if abs(gx[0, 1] - gx[0, 0]) > np.finfo(float).eps:
        if gy[0, 1] - gy[0, 0] != 0:
            raise ValueError("non-axial slices not supported for interpolated approach")
        res1, res2 = abs(gy[1, 0] - gy[0, 0]), abs(gx[0, 1] - gx[0, 0])
        k1 = (1 if gy[1, 0] - gy[0, 0] > 0 else -1) * h3 / 2 / np.pi * res1
        k2 = (1 if gx[0, 1] - gx[0, 0] > 0 else -1) * h2 / 2 / np.pi * res2
        shft = gy[N1 // 2, 0] * h3 + gx[0, N2 // 2] * h2
    else:
        res1, res2 = abs(gx[1, 0] - gx[0, 0]), abs(gy[0, 1] - gy[0, 0])
        k1 = (1 if gx[1, 0] - gx[0, 0] > 0 else -1) * h2 / 2 / np.pi * res1
        k2 = (1 if gy[0, 1] - gy[0, 0] > 0 else -1) * h3 / 2 / np.pi * res2
        shft = gx[N1 // 2, 0] * h2 + gy[0, N2 // 2] * h3
    return np.stack([k1, k2], axis=1), np.exp(1j * shft), (N1, N2)


def sub_samp_time(phs_spha, phs_conc, sampTimes, subFact):
    """``subSampTime`` (``:959-982``): (phs_spha_out, phs_conc_out, t_out, inds)
    with 0-based inds."""
    K = len(sampTimes)
    inds = list(range(0, K, subFact))
    if inds[-1] != K - 1:
        inds.append(K - 1)
    inds = np.array(inds)
    sp, sc = np.array(phs_spha, float), np.array(phs_conc, float)
    if subFact > 1:
        half = int(np.ceil(subFact / 2))
        win = gauss_window(2 * (half + 1), 0.5)
        win = win[1:] / win[1:].sum()
        conv = lambda a: np.stack([np.convolve(r, win, mode="same") for r in a]) if a.shape[0] else a
        sp2, sc2 = conv(sp), conv(sc)
        for a2, a in ((sp2, sp), (sc2, sc)):
            if a.shape[0]:
                a2[:, :half] = a[:, :half]
                a2[:, -half:] = a[:, -half:]
        sp, sc = sp2, sc2
    return sp[:, inds], sc[:, inds], np.asarray(sampTimes, float)[inds], inds


class InterpPlan:
    """``sampHighOrder(b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask,
    useGPU, useSingle, 1, opt)`` with ``opt.svdThresh`` / ``opt.subFact``."""

    def __init__(self, b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask=None, svdThresh=0.05, subFact=5):
        t = np.asarray(sampTimes, float).reshape(-1)
        hs = np.asarray(phs_spha, float).reshape(np.shape(phs_spha)[0], -1)
        hc = np.zeros((0, len(t))) if phs_conc is None or np.size(phs_conc) == 0 else \
            np.asarray(phs_conc, float).reshape(np.shape(phs_conc)[0], -1)
        self.kloc, self.phsShft, self.imN = nufft_for_interp(phs_grid, hs)
        self.nufft = NufftPlan(list(self.imN), self.kloc, None, 1.5, 6)
        subFact = max(1, int(subFact))
        sp, sc, ts, inds = sub_samp_time(hs, hc, t, subFact)
        sp = sp.copy()
        sp[1:3] = 0   # terms 2 and 3 are in the NUFFT (prepForDirect with [1, 4:end])
        P = HighOrderPlan(b0, ts, sp, sc if sc.shape[0] else None, phs_grid, b0mask)
        b = np.exp(1j * P.phase())
        mask = None if b0mask is None else np.asarray(b0mask).reshape(-1, order="F") > 0.5
        bm = b if mask is None else b[mask]
        U, s, Vh = np.linalg.svd(bm, full_matrices=False)
        below = np.nonzero(s / s.max() < svdThresh)[0]
        if len(below) == 0:
            raise ValueError("many large singular values. Try providing mask or adjusting svdThresh.")
        Ns = int(below[0]) + 1
        # svdTime = conj(V * S) with V = Vh' (so conj(V) = Vh.')
        st = Vh[:Ns].T * s[:Ns]
        if subFact > 1:
            st = pchip(inds, st, np.arange(len(t)))
        self.svdTime = st
        sv = U[:, :Ns]
        if mask is not None:
            full = np.zeros((b.shape[0], Ns), complex)
            full[mask] = sv
            sv = full
        self.svdSpace = sv
        self.Ns = Ns
        self.V = b.shape[0]
        self.K = len(t)

    def forward(self, x):
        x = np.asarray(x, complex).reshape(self.V, -1, order="F")
        R = x.shape[1]
        imgs = (x[:, :, None] * self.svdSpace[:, None, :]).reshape(self.imN + (R * self.Ns,), order="F")
        Y = np.asarray(self.nufft.forward(imgs)).reshape(self.K, R, self.Ns, order="F")
        y = np.sum(self.phsShft[:, None, None] * Y * self.svdTime[:, None, :], axis=2)
        return y[:, 0] if R == 1 else y

    def adjoint(self, y):
        y = np.asarray(y, complex).reshape(self.K, -1, order="F")
        R = y.shape[1]
        Y = np.conj(self.phsShft)[:, None, None] * y[:, :, None] * np.conj(self.svdTime)[:, None, :]
        X = np.asarray(self.nufft.adjoint(Y.reshape(self.K, R * self.Ns, order="F"))).reshape(self.V, R, self.Ns,
                                                                                              order="F")
        x = np.sum(X * np.conj(self.svdSpace)[:, None, :], axis=2)
        return x[:, 0].reshape(self.imN, order="F") if R == 1 else x
