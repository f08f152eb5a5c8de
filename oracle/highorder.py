"""Expanded-encoding (high-order) sampling operator, FP64 (oracle).

Restates ``operators/@sampHighOrder/sampHighOrder.m`` for the direct and
segmented approaches (0, 2, 3), which compute the same values:

* phase ``kbase`` (``prepForDirect``, ``:496-588``)::

      Phi[v, t] = sum_n B_n(r_v) mu_n(v) h_n(t)          (harmonics; mu only for n > 4)
                + sum_n Cc_n(r_v) mu(v) g_n(t)           (concomitant terms)
                + mu(v) b0(v) t                          (only if numel(sampTimes) > 1)

  with ``mu = b0mask`` (or 1), summed in that order;
* forward ``y = exp(+1i*kbase.') * x / sqrt(V)`` (``:374-387``);
* adjoint ``x = exp(-1i*kbase) * y / sqrt(V)`` (``:369-372``, ``:387``);
* segmented forward/adjoint (``:393-457``) -- chunks over V / K with partial sums;
* ``phiDiv`` forward (``setPhiDiv`` ``:590-622``, use ``:375-382``):
  ``y = (1i*exp(1i*kbase) .* dPhi).' * x / sqrt(V)`` where ``dPhi`` is
  ``prepForDirect(ksphaDiv, kconcDiv, 1)`` (scalar time: no B0 term).

Pinned by (``tests/test_oracle_highorder.py``): the brute-force double loop on
tiny sizes, the reference's DFT known-answer test (``@sampHighOrder/tests.m:265-283``:
linear phase on an N x N grid gives ``|S im| = |fftnc(im)|`` and
``S'S im = im``, 1e-10), the adjoint identity (``:10-28``), segmented = direct
(``:30-40``) and phiDiv consistency across approaches (``:43-63``).
"""
import numpy as np

from .basis import basis_conc, basis_harm


class HighOrderPlan:
    """``sampHighOrder(b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask)``.

    ``b0`` (rad/s) and ``phs_grid`` fields have the image shape; ``sampTimes``
    (s) has the k-space shape; ``phs_spha`` is ``[Nh, kshape]`` and
    ``phs_conc`` is ``[Nc, kshape]`` or ``None``.
    """

    def __init__(self, b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask=None):
        gx = np.asarray(phs_grid["x"], dtype=np.float64)
        self.imSize = gx.shape
        self.x = gx.reshape(-1, order="F")
        self.y = np.asarray(phs_grid["y"], dtype=np.float64).reshape(-1, order="F")
        self.z = np.asarray(phs_grid["z"], dtype=np.float64).reshape(-1, order="F")
        self.V = self.x.size
        if b0 is None or np.size(b0) == 0:
            self.b0 = np.zeros(self.V)
        else:
            if np.shape(b0) != self.imSize:
                raise ValueError("Size mismatch between b0 and phs_grid.x")
            self.b0 = np.asarray(b0, dtype=np.float64).reshape(-1, order="F")
        self.sampTimes = np.asarray(sampTimes, dtype=np.float64)
        self.kSize = self.sampTimes.shape
        self.K = self.sampTimes.size
        self.phs_spha = np.asarray(phs_spha, dtype=np.float64).reshape(np.shape(phs_spha)[0], -1, order="F")
        if self.phs_spha.shape[1] != self.K:
            raise ValueError("Dimension mismatch between phs_spha and sampTimes")
        if phs_conc is None or np.size(phs_conc) == 0:
            self.phs_conc = np.zeros((0, self.K))
        else:
            self.phs_conc = np.asarray(phs_conc, dtype=np.float64).reshape(np.shape(phs_conc)[0], -1, order="F")
        self.b0mask = None if b0mask is None or np.size(b0mask) == 0 else \
            np.asarray(b0mask, dtype=np.float64).reshape(-1, order="F")
        self.kbase = None
        self.ksphaDiv = None
        self.kconcDiv = None
        # exp(1i*kbase) kept between calls when set (repeated applications in
        # an iterative solve); the values are those of the direct model
        self.cache_exp = False
        self._ekb = None

    # -- sampHighOrder.m:496-588 --------------------------------------------
    def phase(self, phs_spha=None, phs_conc=None, sampTimes=None, vsl=slice(None), tsl=slice(None)):
        """``prepForDirect``: ``[Vsub, Ksub]`` FP64 phase, same summation order."""
        phs_spha = self.phs_spha if phs_spha is None else phs_spha
        phs_conc = self.phs_conc if phs_conc is None else phs_conc
        sampTimes = self.sampTimes.reshape(-1, order="F") if sampTimes is None else np.atleast_1d(sampTimes)
        x, y, z = self.x[vsl], self.y[vsl], self.z[vsl]
        mask = None if self.b0mask is None else self.b0mask[vsl]
        hs = phs_spha[:, tsl]
        hc = phs_conc[:, tsl] if phs_conc.shape[0] else phs_conc
        phs = 0.0
        for n in range(1, hs.shape[0] + 1):
            bf = basis_harm(x, y, z, n)
            if n > 4 and mask is not None:
                bf = bf * mask
            phs = phs + bf[:, None] * hs[n - 1][None, :]
        for n in range(1, hc.shape[0] + 1):
            bf = basis_conc(x, y, z, n)
            if mask is not None:
                bf = bf * mask
            phs = phs + bf[:, None] * hc[n - 1][None, :]
        if sampTimes.size > 1:
            b0a = self.b0[vsl]
            if mask is not None:
                b0a = b0a * mask
            phs = phs + b0a[:, None] * sampTimes[tsl][None, :]
        return phs

    def set_phi_div(self, ksphaDiv, kconcDiv):
        """``setPhiDiv`` (``:590-622``)."""
        self.ksphaDiv = np.asarray(ksphaDiv, dtype=np.float64).reshape(np.shape(ksphaDiv)[0], -1, order="F")
        if kconcDiv is None or np.size(kconcDiv) == 0:
            self.kconcDiv = np.zeros((0, self.K))
        else:
            self.kconcDiv = np.asarray(kconcDiv, dtype=np.float64).reshape(np.shape(kconcDiv)[0], -1, order="F")

    def _kb(self):
        if self.kbase is None:
            self.kbase = self.phase()
        return self.kbase

    def _exp_kb(self):
        """``exp(1i*kbase)``, [V x K]; cached when ``cache_exp`` is set."""
        if self.cache_exp:
            if self._ekb is None:
                self._ekb = np.exp(1j * self._kb())
            return self._ekb
        return np.exp(1j * self._kb())

    # -- sampHighOrder.m:361-391 ---------------------------------------------
    def forward(self, x):
        """``S*x``: ``imSize + reps`` -> ``kSize + reps`` (direct)."""
        x = np.asarray(x, dtype=np.complex128)
        nd = len(self.imSize) if not (len(self.imSize) == 2 and self.imSize[1] == 1) else 1
        reps = x.shape[nd:]
        xa = x.reshape(self.V, -1, order="F")
        kb = self._kb()
        if self.ksphaDiv is None:
            y = self._exp_kb().T @ xa
        else:
            d = self.phase(self.ksphaDiv, self.kconcDiv, np.array([1.0]))
            y = (1j * np.exp(1j * kb) * d).T @ xa
        y = y / np.sqrt(self.V)
        return y.reshape(self.kSize + reps, order="F") if reps else y.reshape(self.kSize, order="F")

    def adjoint(self, y):
        """``S'*y``: ``kSize + reps`` -> ``imSize + reps`` (direct)."""
        y = np.asarray(y, dtype=np.complex128)
        ndk = len(self.kSize) if not (len(self.kSize) == 2 and self.kSize[1] == 1) else 1
        reps = y.shape[ndk:]
        ya = y.reshape(self.K, -1, order="F")
        # conj(E) @ y computed as conj(E @ conj(y)): the same products, with
        # no conjugated copy of the [V x K] matrix
        x = np.conj(self._exp_kb() @ np.conj(ya)) / np.sqrt(self.V)
        return x.reshape(self.imSize + reps, order="F") if reps else x.reshape(self.imSize, order="F")

    # -- sampHighOrder.m:393-457 (approach 3: phase recomputed per segment) ---
    def forward_segmented(self, x, ndiv):
        x = np.asarray(x, dtype=np.complex128)
        xa = x.reshape(self.V, -1, order="F")
        seg = int(np.ceil(self.V / ndiv))
        y = 0
        for nc in range(ndiv):
            s1, s2 = nc * seg, min((nc + 1) * seg, self.V)
            if s1 >= s2:
                continue
            ph = self.phase(vsl=slice(s1, s2))
            if self.ksphaDiv is None:
                y = y + np.exp(1j * ph.T) @ xa[s1:s2]
            else:
                d = self.phase(self.ksphaDiv, self.kconcDiv, np.array([1.0]), vsl=slice(s1, s2))
                y = y + (1j * np.exp(1j * ph) * d).T @ xa[s1:s2]
        return y / np.sqrt(self.V)

    def adjoint_segmented(self, y, ndiv):
        y = np.asarray(y, dtype=np.complex128)
        ya = y.reshape(self.K, -1, order="F")
        seg = int(np.ceil(self.K / ndiv))
        x = 0
        for nc in range(ndiv):
            s1, s2 = nc * seg, min((nc + 1) * seg, self.K)
            if s1 >= s2:
                continue
            ph = self.phase(tsl=slice(s1, s2))
            x = x + np.exp(-1j * ph) @ ya[s1:s2]
        return x / np.sqrt(self.V)
