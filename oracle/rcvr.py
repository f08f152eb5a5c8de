"""Coil-sensitivity operator, FP64 (oracle, test infrastructure).

Restates the maps path of ``operators/@rcvrOp/rcvrOp.m`` (``isCalDat=false``,
``:90-93``): forward ``y = x .* maps`` (``:152``), adjoint
``x = sum(y .* conj(maps), nDim+1)`` (``:149-150``).  The ``doLoop`` variant
(``:116-143``) computes the same values coil by coil.

Pinned by the adjoint identity and the ``doLoop`` equality of
``@rcvrOp/tests.m:19-31`` (both closed forms of an elementwise operator).
"""
import numpy as np


class RcvrPlan:
    """``rcvrOp(maps, false)``: ``maps`` has shape ``imN + [C]``."""

    def __init__(self, maps):
        self.maps = np.asarray(maps, dtype=np.complex128)
        self.imN = self.maps.shape[:-1]
        self.nDim = len(self.imN)
        self.C = self.maps.shape[-1]
        # complex64 evaluation (see NufftPlan.single)
        self.single = False

    def forward(self, x):
        """``R*x`` (``rcvrOp.m:152``): ``imN`` -> ``imN + [C]``."""
        x = np.asarray(x, dtype=np.complex128).reshape(self.imN, order="F")
        if self.single:
            return (x.astype(np.complex64)[..., None] * self.maps.astype(np.complex64)).astype(np.complex128)
        return x[..., None] * self.maps

    def adjoint(self, y, doLoop=False):
        """``R'*y`` (``rcvrOp.m:149-150``; ``doLoop`` form ``:117-129``)."""
        y = np.asarray(y, dtype=np.complex128)
        if y.size == int(np.prod(self.imN)) * self.C:
            y = y.reshape(self.imN + (self.C,), order="F")
        else:  # MATLAB implicit expansion of singleton dims (mrSampFunc sumDims)
            y = np.broadcast_to(y, self.imN + (self.C,))
        if self.single:
            return np.sum(y.astype(np.complex64) * np.conj(self.maps.astype(np.complex64)),
                          axis=self.nDim).astype(np.complex128)
        if doLoop:
            acc = 0
            for c in range(self.C):
                acc = acc + y[..., c] * np.conj(self.maps[..., c])
            return acc
        return np.sum(y * np.conj(self.maps), axis=self.nDim)
