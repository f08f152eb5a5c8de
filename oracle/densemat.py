"""Dense normal matrix of the high-order SENSE operator, FP64 (oracle, test
infrastructure).

Restates ``operators/mrSampFuncMat.m:1-47``: ``S = sampHighOrder(b0,
sampTimes, phs_spha, phs_conc, phs_grid)``, ``R = rcvrOp(Crcvr)``,
``Aty = R'*(S'*y)`` restricted to ``imMask``, and

    Sd  = exp(1i * single(kbase)).' restricted to the masked voxels / sqrt(Vm)
    AtA = sum_c conj(m_c) .* (Sd' * Sd) .* m_c.'           (m_c: masked maps)

The phase is rounded to single precision as the reference does
(``single(S.kbase)``, :34); everything after that is evaluated here in FP64,
so the oracle is the reference's formula without its single-precision
arithmetic noise.

Pinned by (``tests/test_oracle_densemat.py``): with a full mask and the
phase kept in FP64 (``single_phase=False``) AtA is the matrix of the normal
operator ``R'(S'(S(R x)))`` of the oracle high-order SENSE operator (closed
form, 1e-10); with a partial mask the reference's 1/sqrt(numel(Aty)) scale
makes it (V/Vm) times the masked normal matrix; the single rounding moves it
by the phase rounding error only; Hermitian.
"""
import numpy as np

from .highorder import HighOrderPlan
from .rcvr import RcvrPlan


def mrSampFuncMat(y, Crcvr, b0, sampTimes, phs_spha, phs_conc, phs_grid, imMask=None, single_phase=True):
    S = HighOrderPlan(b0, sampTimes, phs_spha, phs_conc, phs_grid)
    R = RcvrPlan(Crcvr)
    V = S.V
    mask = np.ones(V, bool) if imMask is None else np.asarray(imMask).reshape(-1, order="F").astype(bool)
    Aty = R.adjoint(S.adjoint(y)).reshape(-1, order="F")[mask]             # :29-31
    kb = S.phase()                                                          # [V x K], radians
    if single_phase:
        kb = kb.astype(np.float32).astype(np.float64)                       # single(S.kbase), :34
    Sd = np.exp(1j * kb[mask]).T / np.sqrt(mask.sum())                      # :35-36
    G = Sd.conj().T @ Sd                                                    # :37
    m = R.maps.reshape(V, -1, order="F")[mask]                              # [Vm x C]
    AtA = 0
    for c in range(m.shape[1]):                                             # :39-44
        AtA = AtA + np.conj(m[:, c])[:, None] * G * m[:, c][None, :]
    return AtA, Aty
