"""Kaiser-Bessel gridding NUFFT, FP64 (oracle, test infrastructure).

Line-by-line restatement of ``operators/@nufftOp/nufftOp.m`` (MatMRI v1.18):

* plan (``doPrep``)                  -- ``nufftOp.m:138-274``
* forward, Cartesian -> non-Cart.    -- ``nufftOp.m:378-394``
* adjoint, non-Cart. -> Cartesian    -- ``nufftOp.m:354-377``
* repetition columns (trailing dims) -- ``nufftOp.m:301-331``

Arrays are column-major like MATLAB: an image of size ``imN`` is a NumPy array
of shape ``imN + reps`` whose linear index is taken in Fortran order, and the
interpolation matrix columns use the column-major linear index of the
oversampled grid (``sub2ind(osN, ...)``, ``nufftOp.m:243``).

Pinned by (``tests/test_oracle_nufft.py``):

* closed-form plan constants: grid size, KB beta, LUT length (the table of known
  values derived from ``nufftOp.m:146-156``);
* the reference's own known-answer test: on-grid ``kloc`` makes the forward a
  unitary centred inverse DFT, ``fftnc(S*im) == im`` (``@nufftOp/tests.m:90-101``);
* point-spread function flat magnitude / zero phase (``@nufftOp/tests.m:67-81``);
* adjoint identity (``@nufftOp/tests.m:28-37``), repetitions (``:39-54``),
  circular wrap (``:57-65``), scale invariance across os/W (``:82-87``);
* a brute-force non-uniform DFT with the "+" sign and centred image index,
  which bounds the gridding approximation (and fails for a wrong sign, a
  wrong centring or a wrong scale).

:class:`ToepPlan` (``prepToep``, ``nufftOp.m:410-432``, applied at
``:348-352``) is pinned by the reference's own Toeplitz test (against
``S0'*(S0*x)`` with os = 2, W = 8, rms < 1e-3, ``@nufftOp/tests.m:104-124``)
and by the exact normal operator ``E' diag(dcf^2) E`` of the brute-force
NUDFT (fails for dcf applied once, a conjugated kernel or a wrong scale).
"""
import numpy as np
import scipy.sparse as sp
from scipy.special import i0

from .fftnc import fftnc, ifftnc


def _max_prime_factor(n):
    n = int(n)
    best = 1
    p = 2
    while p * p <= n:
        while n % p == 0:
            best = p
            n //= p
        p += 1
    if n > 1:
        best = n
    return best


def grid_size(imN, os):
    """Oversampled grid size, ``nufftOp.m:146-152``."""
    osN = []
    for n in np.atleast_1d(imN):
        g = int(np.ceil(n * os / 2.0) * 2)
        while _max_prime_factor(g) > 7:
            g += 2
        osN.append(g)
    return np.array(osN, dtype=np.int64)


def kb_beta(kwidth, os):
    """Kaiser-Bessel shape parameter, ``nufftOp.m:155``."""
    return np.pi * np.sqrt(kwidth ** 2 / os ** 2 * (os - 0.5) ** 2 - 0.8)


def kb_nkern(kwidth, os, precision=1e-6):
    """LUT length, ``nufftOp.m:154-156``."""
    return int(np.ceil(0.5 * (kwidth + 1) * np.sqrt(0.37 / os ** 2 / precision)) * 2)


def kb_lut(kwidth, os):
    """Normalised KB lookup table with the trailing zero sentinel,
    ``nufftOp.m:157-165``.  Length ``Nkern + 1``."""
    bet = kb_beta(kwidth, os)
    nk = kb_nkern(kwidth, os)
    kx = np.linspace(-1.0, 1.0, nk)
    kern = np.concatenate([i0(bet * np.sqrt(1.0 - kx ** 2)), [0.0]])
    kern = kern / np.sum(kern) * (len(kern) - 2) / kwidth
    return kern


def deapod(imN_d, osN_d, kwidth, os):
    """Separable deapodisation ``comp{nD}``, ``nufftOp.m:171-182``.

    ``x = sqrt(x^2 - beta^2)`` is imaginary for every image index at the
    supported (W, os), and ``x ./ sin(x)`` then equals the real value
    ``a / sinh(a)`` with ``a = sqrt(beta^2 - x^2)``; complex arithmetic is kept
    so the rare real branch behaves like MATLAB too.
    """
    bet = kb_beta(kwidth, os)
    j = np.arange(-imN_d / 2.0, imN_d / 2.0 - 1 + 0.5, 1.0)[: int(imN_d)]
    x = j * np.pi * kwidth / osN_d
    x = np.sqrt((x ** 2 - bet ** 2).astype(np.complex128))
    comp = x / np.sin(x)
    comp = comp / comp[np.argmin(np.abs(comp))]
    comp = comp * np.sqrt(osN_d / imN_d)
    return np.real(comp)


class NufftPlan:
    """``nufftOp(imN, kloc, dcf, useGPU, os, kwidth)`` (``nufftOp.m:89-136``)
    followed by ``doPrep`` (``nufftOp.m:138-274``)."""

    def __init__(self, imN, kloc, dcf=None, os=1.5, kwidth=6):
        self.imN = np.array(np.atleast_1d(imN), dtype=np.int64)
        nd = len(self.imN)
        kloc = np.asarray(kloc, dtype=np.float64).reshape(-1, nd, order="F") \
            if np.ndim(kloc) == 1 else np.asarray(kloc, dtype=np.float64)
        self.kloc = kloc
        self.M = kloc.shape[0]
        self.dcf = None if dcf is None else np.asarray(dcf, dtype=np.float64).reshape(-1)
        self.os = float(os)
        self.kwidth = int(kwidth)
        self.osN = grid_size(self.imN, self.os)
        self.beta = kb_beta(self.kwidth, self.os)
        self.nkern = kb_nkern(self.kwidth, self.os)
        self.kern = kb_lut(self.kwidth, self.os)
        self.comp = [deapod(self.imN[d], self.osN[d], self.kwidth, self.os) for d in range(nd)]
        self._prep_matrix()
        self._prep_shift()
        # single = True evaluates the same steps in complex64 arithmetic (a
        # model of a single-precision implementation of this operator, used
        # only to size the tolerances of iterative-solver tests)
        self.single = False
        self._c32 = None

    # -- nufftOp.m:187-257 ---------------------------------------------------
    def _prep_matrix(self):
        nd = len(self.imN)
        W = self.kwidth
        kloc_a = self.kloc + 0.5                                   # :141
        offs1 = np.arange(-W / 2 + 1, W / 2 + 1)                   # :192-198
        # ndgrid over nd dims, first dim fastest (column-major)
        grids = np.meshgrid(*([offs1] * nd), indexing="ij")
        samp_off = [g.reshape(-1, order="F") for g in grids]
        ntap = W ** nd
        C = np.ones((self.M, ntap))
        Nk = self.nkern
        for d in range(nd):                                         # :216-227
            ka = kloc_a[:, d] * self.osN[d]
            ka = np.floor(ka) - ka
            for s in range(ntap):
                ki = ka + samp_off[d][s]
                ki = ki * (Nk - 1) / W + Nk / 2.0 + 0.5
                frac = ki - np.floor(ki)
                ki = np.floor(ki).astype(np.int64)                  # 1-based LUT index
                C[:, s] *= (1 - frac) * self.kern[ki - 1] + frac * self.kern[ki]
        cart = []
        for d in range(nd):                                         # :230-238
            ci = np.floor(kloc_a[:, d] * self.osN[d])[:, None] + samp_off[d][None, :] \
                + self.osN[d] / 2 + 1
            # the reference wraps once in each direction; mod is identical on
            # every input it accepts
            ci = np.mod(ci - 1, self.osN[d]) + 1
            cart.append(ci.astype(np.int64) - 1)                    # 0-based
        lin = np.zeros_like(cart[0])
        stride = 1
        for d in range(nd):                                         # sub2ind, :239-246
            lin = lin + cart[d] * stride
            stride *= int(self.osN[d])
        rows = np.repeat(np.arange(self.M)[:, None], ntap, axis=1)
        self.convMat = sp.csr_matrix(
            (C.reshape(-1), (rows.reshape(-1), lin.reshape(-1))),
            shape=(self.M, int(np.prod(self.osN))))                 # :257 (sums duplicates)
        self.C = C
        self.cart_lin = lin

    # -- nufftOp.m:258-273 ---------------------------------------------------
    def _prep_shift(self):
        nd = len(self.imN)
        shp = tuple(int(g) for g in self.osN)
        m = np.zeros(shp)
        for d in range(nd):
            g = np.arange(self.osN[d], dtype=np.float64)
            v = -(g - self.osN[d] / 2.0) * self.imN[d] / self.osN[d]
            bshape = [1] * nd
            bshape[d] = int(self.osN[d])
            m = m + v.reshape(bshape)
        self.fftShiftArray = np.exp(1j * np.pi * m)

    # -- helpers -------------------------------------------------------------
    def _comp_nd(self):
        nd = len(self.imN)
        c = np.ones(tuple(int(n) for n in self.imN))
        for d in range(nd):
            bshape = [1] * nd
            bshape[d] = int(self.imN[d])
            c = c * self.comp[d].reshape(bshape)
        return c

    def _single_consts(self):
        if self._c32 is None:
            cm = self.convMat.astype(np.complex64)
            self._c32 = (cm, cm.T.tocsr(), self._comp_nd().astype(np.float32),
                         self.fftShiftArray.astype(np.complex64),
                         None if self.dcf is None else self.dcf.astype(np.float32))
        return self._c32

    def _apply_single(self, x, adjoint):
        """The forward (``:378-394``) or adjoint (``:354-377``) steps in
        complex64: unitary FFTs with the same padding and crop, float32
        deapodisation, interpolation matrix and dcf."""
        nd = len(self.imN)
        cm, cmT, comp, sh, dcf = self._single_consts()
        axes = tuple(range(nd))
        imN = tuple(int(n) for n in self.imN)
        osN = tuple(int(g) for g in self.osN)
        if not adjoint:
            x = np.asarray(x).astype(np.complex64)
            reps = x.shape[nd:]
            nrep = int(np.prod(reps)) if reps else 1
            z = x.reshape(imN + (nrep,), order="F") * comp[..., None]
            z = np.fft.ifftn(z, s=osN, axes=axes, norm="ortho").astype(np.complex64) * sh[..., None]
            y = cm @ z.reshape(-1, nrep, order="F")
            if dcf is not None:
                y = y * dcf[:, None]
            return y.astype(np.complex128).reshape((self.M,) + reps, order="F")
        y = np.asarray(x).astype(np.complex64)
        reps = y.shape[1:]
        nrep = int(np.prod(reps)) if reps else 1
        ya = y.reshape(self.M, nrep, order="F")
        if dcf is not None:
            ya = ya * dcf[:, None]
        z = (cmT @ ya).reshape(osN + (nrep,), order="F") * np.conj(sh)[..., None]
        z = np.fft.fftn(z, axes=axes, norm="ortho").astype(np.complex64)
        xa = z[tuple(slice(0, n) for n in imN)] * comp[..., None]
        return xa.astype(np.complex128).reshape(imN + reps, order="F")

    def forward(self, x):
        """``N*x``: image ``imN + reps`` -> samples ``[M] + reps``
        (``nufftOp.m:378-394``)."""
        if self.single:
            return self._apply_single(x, False)
        nd = len(self.imN)
        x = np.asarray(x, dtype=np.complex128)
        reps = x.shape[nd:]
        nrep = int(np.prod(reps)) if reps else 1
        xa = x.reshape(tuple(int(n) for n in self.imN) + (nrep,), order="F")
        ya = xa * self._comp_nd()[..., None]                          # :380-383
        ya = ifftnc(ya, nd, 0, 1, [int(g) for g in self.osN])         # :384
        ya = ya * self.fftShiftArray[..., None]                       # :385-387
        ya = ya.reshape(-1, nrep, order="F")                          # :388
        y = self.convMat @ ya                                         # :389
        if self.dcf is not None:
            y = y * self.dcf[:, None]                                 # :391-393
        return y.reshape((self.M,) + reps, order="F")

    def adjoint(self, y):
        """``N'*y``: samples ``[M] + reps`` -> image ``imN + reps``
        (``nufftOp.m:354-377``)."""
        if self.single:
            return self._apply_single(y, True)
        nd = len(self.imN)
        y = np.asarray(y, dtype=np.complex128)
        reps = y.shape[1:]
        nrep = int(np.prod(reps)) if reps else 1
        ya = y.reshape(self.M, nrep, order="F")
        if self.dcf is not None:
            ya = ya * self.dcf[:, None]                               # :356-358
        za = self.convMat.T @ ya                                      # :359
        za = za.reshape(tuple(int(g) for g in self.osN) + (nrep,), order="F")
        za = za * np.conj(self.fftShiftArray)[..., None]              # :361-363
        za = fftnc(za, nd, 0)                                         # :364
        sl = tuple(slice(0, int(n)) for n in self.imN)                # :365-373
        xa = za[sl] * self._comp_nd()[..., None]                      # :374-377
        return xa.reshape(tuple(int(n) for n in self.imN) + reps, order="F")


def nudft_centred(imN, kloc, x):
    """Brute-force unitary centred NUDFT with the "+" sign:
    ``y_m = prod(N)^-1/2 sum_i x_i exp(+2 pi i sum_d (i_d - N_d/2) k_md)``.
    Used only to bound the gridding approximation of :class:`NufftPlan`."""
    imN = np.atleast_1d(imN)
    nd = len(imN)
    grids = np.meshgrid(*[np.arange(n) - n / 2.0 for n in imN], indexing="ij")
    pos = np.stack([g.reshape(-1, order="F") for g in grids], axis=1)   # [V, nd]
    ph = 2 * np.pi * (np.asarray(kloc).reshape(-1, nd) @ pos.T)         # [M, V]
    E = np.exp(1j * ph) / np.sqrt(np.prod(imN))
    xa = np.asarray(x, dtype=np.complex128).reshape(int(np.prod(imN)), -1, order="F")
    return E @ xa


def padcrop(x, szIn):
    """``general/padcrop.m:1-93`` over the leading ``len(szIn)`` dimensions:
    a longer size zero-pads with ``floor(npc/2)`` before the data, a shorter
    one keeps the entries from ``floor(npc/2)`` on (``npc`` = size change)."""
    y = np.asarray(x)
    for d, n in enumerate(szIn):
        n = int(n)
        npc = n - y.shape[d]
        if npc > 0:
            pad = [(0, 0)] * y.ndim
            pad[d] = (npc // 2, npc - npc // 2)
            y = np.pad(y, pad)
        elif npc < 0:
            idx = [slice(None)] * y.ndim
            idx[d] = slice((-npc) // 2, (-npc) // 2 + n)
            y = y[tuple(idx)]
    return y


class ToepPlan:
    """``S.prepToep(tos, kwidth)`` (``nufftOp.m:410-432``) and the Toeplitz
    application of ``nufftOp.m:348-352``: ``S * x`` (and ``S' * x``) becomes
    ``padcrop(fftnc(convMat .* ifftnc(padcrop(x, 2N))), N)`` with
    ``convMat = 2^nd * ifftnc(S2' * dcf)`` (``S2' * ones`` without dcf) and
    ``S2 = nufftOp(2N, kloc, dcf, ., tos, kwidth)``; ``tos`` defaults to
    ``0.5 (os - 1) + 1`` and ``kwidth`` to the operator's (``:412-419``)."""

    def __init__(self, imN, kloc, dcf=None, os=1.5, kwidth=6, tos=None, tkwidth=None):
        self.imN = [int(n) for n in np.atleast_1d(imN)]
        nd = len(self.imN)
        if tos is None:
            tos = 0.5 * (os - 1) + 1                                  # :414
        if tkwidth is None:
            tkwidth = kwidth                                          # :416-418
        S = NufftPlan([2 * n for n in self.imN], kloc, dcf, tos, tkwidth)   # :419
        w = np.ones(S.M) if dcf is None else np.asarray(dcf, np.float64).reshape(-1)
        psf = S.adjoint(w)                                            # :420-424 (dcf applied again inside)
        self.convMat = (2 ** nd) * ifftnc(psf, nd)                    # :425

    def apply(self, x):
        """``S * x`` of the prepared operator (``nufftOp.m:348-352``)."""
        nd = len(self.imN)
        x = np.asarray(x, dtype=np.complex128)
        reps = x.shape[nd:]
        nrep = int(np.prod(reps)) if reps else 1
        xa = x.reshape(tuple(self.imN) + (nrep,), order="F")
        z = ifftnc(padcrop(xa, [2 * n for n in self.imN]), nd)
        z = fftnc(self.convMat[..., None] * z, nd)
        return padcrop(z, self.imN).reshape(tuple(self.imN) + reps, order="F")
