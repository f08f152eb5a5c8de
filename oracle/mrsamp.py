"""SENSE encoding composer, FP64 (oracle, test infrastructure).

Restates ``operators/mrSampFunc.m:82-159`` for sampling operators that expose
``forward``/``adjoint`` (the oracle NUFFT and high-order plans) or a real
Cartesian mask:

* ``'notransp'``: ``hardReal``, Tikhonov copy, ``reshape(imNFull)``, ``R*x``,
  ``sumDims``, ``S*x`` (or ``mask .* ifftnc(x, nDim)``), vectorise, stack the
  Tikhonov rows (``:83-120``);
* ``'transp'``: split the Tikhonov rows, ``x .* W``, ``reshape(daNFull)``,
  ``S'*x`` (or ``fftnc(mask .* x, nDim)``), ``R'*x``, vectorise, add the
  Tikhonov term, ``hardReal`` (``:121-158``).

Pinned by the adjoint identity over the vectorisation / Tikhonov / sumDims
combinations of ``mrSampFunc.m:163-216``.
"""
import numpy as np

from .fftnc import fftnc, ifftnc


def _mask_ndim(mask):
    nd = mask.ndim
    if nd == 2 and mask.shape[1] == 1:
        nd = 1
    return nd


def mrSampFunc(x, transp, samplingOp, R=None, opt=None):
    opt = {} if opt is None else dict(opt)
    imNFull = opt.get("imNFull")
    daNFull = opt.get("daNFull")
    tikReg = opt.get("tikReg")
    sumDims = opt.get("sumDims")
    hardReal = opt.get("hardReal", 0) or 0
    W = opt.get("W")
    if tikReg is not None and (imNFull is None or daNFull is None):
        raise ValueError("Tikhonov regularization requires opt.imNFull and opt.daNFull")
    x = np.asarray(x, dtype=np.complex128)
    is_op = hasattr(samplingOp, "forward")
    if transp == "notransp":
        if hardReal:
            x = np.real(x).astype(np.complex128)
        if tikReg is not None:
            t = np.atleast_1d(tikReg)
            x_tik = t[0] * np.real(x) + 1j * t[1] * np.imag(x) if t.size == 2 else t[0] * x
        if imNFull is not None:
            x = x.reshape(tuple(imNFull), order="F")
        if R is not None:
            x = R.forward(x)
        if sumDims is not None:
            for d in np.atleast_1d(sumDims):
                x = np.sum(x, axis=d - 1, keepdims=True)
        if is_op:
            x = samplingOp.forward(x)
        else:
            m = np.asarray(samplingOp, dtype=np.float64)
            nd = _mask_ndim(m)
            mm = m.reshape(m.shape[:nd] + (1,) * (x.ndim - nd), order="F") if x.ndim > nd else m
            x = mm * ifftnc(x, nd)
        if daNFull is not None:
            x = x.reshape(-1, order="F")
        if tikReg is not None:
            x = np.concatenate([x.reshape(-1, order="F"), x_tik.reshape(-1, order="F")])
        return x
    if transp == "transp":
        if tikReg is not None:
            nda = int(np.prod(daNFull))
            x_tik = x[nda:]
            x = x[:nda]
        if W is not None:
            x = x * np.asarray(W)
        if daNFull is not None:
            x = x.reshape(tuple(daNFull), order="F")
        if is_op:
            x = samplingOp.adjoint(x)
        else:
            m = np.asarray(samplingOp, dtype=np.float64)
            nd = _mask_ndim(m)
            mm = m.reshape(m.shape[:nd] + (1,) * (x.ndim - nd), order="F") if x.ndim > nd else m
            x = fftnc(mm * x, nd)
        if R is not None:
            x = R.adjoint(x)
        if imNFull is not None:
            x = x.reshape(-1, order="F")
        if tikReg is not None:
            t = np.atleast_1d(tikReg)
            x_tik = t[0] * np.real(x_tik) + 1j * t[1] * np.imag(x_tik) if t.size == 2 else t[0] * x_tik
            x = x + x_tik
        if hardReal:
            x = np.real(x).astype(np.complex128)
        return x
    raise ValueError("transp must be 'notransp' or 'transp'")
