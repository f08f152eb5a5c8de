"""Centred/uncentred unitary N-D FFT helpers (oracle, test infrastructure).

Follows ``general/fftnc.m:1-57`` and ``general/ifftnc.m:1-56``:

* transform over the first ``N`` dimensions, one dimension at a time;
* ``doShift``: ``fftshift`` before and after the 1-D transform (``ifftshift``
  for the inverse);
* optional ``sz``: per-dimension transform length; MATLAB's ``fft(a, n, dim)``
  zero-pads at the END of the dimension (or truncates);
* ``doScale``: divide by ``prod(sqrt(n))`` for ``fftnc`` (unitary forward), and
  MULTIPLY by ``prod(sqrt(n))`` for ``ifftnc`` (MATLAB's ``ifft`` already has
  the ``1/n``), so both are unitary.

Pinned by: Parseval/unitarity and the inverse round trip (closed forms), and by
the on-grid NUFFT known-answer test of ``@nufftOp/tests.m:90-101``.
"""
import numpy as np


def _pad_or_crop(a, n, axis):
    cur = a.shape[axis]
    if n == cur:
        return a
    if n < cur:
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(0, n)
        return a[tuple(sl)]
    pad = [(0, 0)] * a.ndim
    pad[axis] = (0, n - cur)
    return np.pad(a, pad)


def fftnc(a, N=None, doShift=1, doScale=1, sz=None):
    """``general/fftnc.m:1-57``."""
    a = np.asarray(a, dtype=np.complex128)
    if N is None:
        N = a.ndim
    N = min(N, a.ndim)
    scale = 1.0
    for n in range(N):
        if doShift:
            a = np.fft.fftshift(a, axes=n)
        if sz is not None:  # fft(a, sz(n), n): pad/crop at the end, after the shift
            a = _pad_or_crop(a, int(sz[n]), n)
        a = np.fft.fft(a, axis=n)
        if doShift:
            a = np.fft.fftshift(a, axes=n)
        scale *= np.sqrt(a.shape[n])
    if doScale:
        a = a / scale
    return a


def ifftnc(a, N=None, doShift=1, doScale=1, sz=None):
    """``general/ifftnc.m:1-56``."""
    a = np.asarray(a, dtype=np.complex128)
    if N is None:
        N = a.ndim
    N = min(N, a.ndim)
    scale = 1.0
    for n in range(N):
        if doShift:
            a = np.fft.ifftshift(a, axes=n)
        if sz is not None:  # ifft(a, sz(n), n): pad/crop at the end, after the shift
            a = _pad_or_crop(a, int(sz[n]), n)
        a = np.fft.ifft(a, axis=n)
        if doShift:
            a = np.fft.ifftshift(a, axes=n)
        scale *= np.sqrt(a.shape[n])
    if doScale:
        a = a * scale
    return a
