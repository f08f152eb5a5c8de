"""ctypes binding of the C/OpenMP FP64 restatement of the 2-D spiral
CG-SENSE path (``oracle/c/mmref.c``).

TEST INFRASTRUCTURE ONLY: the CPU baseline of ``bench.py`` and the checker
in ``tests/test_oracle_c.py``.  The library is built by :func:`build` (gcc,
OpenMP) into ``oracle/_build/``; ``__graft_entry__.build()`` calls it.
Arrays follow the NumPy oracle: column-major complex128.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "mmref.c")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD_DIR, "libmmref.so")
CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=gnu11"]

_lib = None


def build(verbose=False):
    """Compile the restatement with gcc; returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc", *CFLAGS, SRC, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed on mmref.c:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise ImportError(f"{LIB} is not built (oracle.cref.build())")
    L = C.CDLL(LIB)
    vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
    L.mmref_sense_create.restype = vp
    L.mmref_sense_create.argtypes = [i32, i32, i32, vp, vp, dbl, i32, i32, vp]
    L.mmref_sense_destroy.argtypes = [vp]
    L.mmref_forward.argtypes = [vp, vp, vp]
    L.mmref_adjoint.argtypes = [vp, vp, vp]
    L.mmref_cgne.restype = i32
    L.mmref_cgne.argtypes = [vp, vp, vp, vp, i32, dbl, dbl, dbl, i32, vp]
    L.mmref_num_threads.restype = i32
    L.mmref_set_verbose.argtypes = [i32]
    L.mmref_hio_create.restype = vp
    L.mmref_hio_create.argtypes = [i32, i32, i32, vp, vp, i32, vp, i32]
    L.mmref_hio_destroy.argtypes = [vp]
    L.mmref_hio_set_round_io.argtypes = [vp, i32]
    L.mmref_hio_forward.argtypes = [vp, vp, vp]
    L.mmref_hio_adjoint.argtypes = [vp, vp, vp]
    L.mmref_hio_cgne.restype = i32
    L.mmref_hio_cgne.argtypes = [vp, vp, vp, vp, i32, dbl, dbl, dbl, i32, vp]
    _lib = L
    return L


def set_verbose(v):
    """Per-iteration cgne lines on stderr (long solves report progress)."""
    lib().mmref_set_verbose(int(bool(v)))


def num_threads():
    return int(lib().mmref_num_threads())


class SenseRef:
    """FP64 SENSE operator E = S o R of one 2-D slice (N x N image, M samples,
    C coils) with ``cgne``."""

    def __init__(self, N, kloc, dcf, maps, os=1.5, W=4):
        L = lib()
        self.N = int(N)
        kl = np.asfortranarray(np.asarray(kloc, dtype=np.float64))
        self.M = kl.shape[0]
        self._kl = kl
        self._dcf = None if dcf is None else np.ascontiguousarray(np.asarray(dcf, dtype=np.float64))
        mp = np.asfortranarray(np.asarray(maps, dtype=np.complex128))
        self.C = mp.shape[-1]
        self._maps = mp.reshape(-1, order="F")
        h = L.mmref_sense_create(self.N, self.N, self.M, kl.ctypes.data,
                                 None if self._dcf is None else self._dcf.ctypes.data, float(os), int(W),
                                 self.C, self._maps.ctypes.data)
        if not h:
            raise ValueError("mmref_sense_create failed")
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mmref_sense_destroy(self._h)
        except Exception:
            pass

    def forward(self, x):
        xf = np.asfortranarray(np.asarray(x, dtype=np.complex128)).reshape(-1, order="F")
        y = np.empty(self.M * self.C, dtype=np.complex128)
        lib().mmref_forward(self._h, xf.ctypes.data, y.ctypes.data)
        return y.reshape((self.M, self.C), order="F")

    def adjoint(self, y):
        yf = np.asfortranarray(np.asarray(y, dtype=np.complex128)).reshape(-1, order="F")
        x = np.empty(self.N * self.N, dtype=np.complex128)
        lib().mmref_adjoint(self._h, yf.ctypes.data, x.ctypes.data)
        return x.reshape((self.N, self.N), order="F")

    def cgne(self, y, x0=None, NitMax=100, resThresh=None, resThreshRel=None, xChangeThresh=None,
             expResN=100):
        """Returns (x, resSqAll, nit)."""
        yf = np.asfortranarray(np.asarray(y, dtype=np.complex128)).reshape(-1, order="F")
        x0f = None if x0 is None else np.asfortranarray(np.asarray(x0, dtype=np.complex128)).reshape(-1, order="F")
        x = np.empty(self.N * self.N, dtype=np.complex128)
        hist = np.zeros(NitMax + 1)
        g = lambda v: -1.0 if v is None else float(v)
        nit = lib().mmref_cgne(self._h, yf.ctypes.data, None if x0f is None else x0f.ctypes.data, x.ctypes.data,
                               int(NitMax), g(resThresh), g(resThreshRel), g(xChangeThresh), int(expResN),
                               hist.ctypes.data)
        return x.reshape((self.N, self.N), order="F"), hist[: nit + 1], int(nit)


def _cgne_call(fn, h, y, x0, shape, NitMax, resThresh, resThreshRel, xChangeThresh, expResN):
    yf = np.asfortranarray(np.asarray(y, dtype=np.complex128)).reshape(-1, order="F")
    x0f = None if x0 is None else np.asfortranarray(np.asarray(x0, dtype=np.complex128)).reshape(-1, order="F")
    x = np.empty(int(np.prod(shape)), dtype=np.complex128)
    hist = np.zeros(NitMax + 1)
    g = lambda v: -1.0 if v is None else float(v)
    nit = fn(h, yf.ctypes.data, None if x0f is None else x0f.ctypes.data, x.ctypes.data, int(NitMax), g(resThresh),
             g(resThreshRel), g(xChangeThresh), int(expResN), hist.ctypes.data)
    return x.reshape(shape, order="F"), hist[: nit + 1], int(nit)


class HighOrderSenseRef:
    """FP64 expanded-encoding SENSE operator E = S o R (direct approach) of
    one 2-D slice with ``cgne``; the phase model is the NumPy oracle's
    (``oracle.highorder.HighOrderPlan``: harmonics, concomitant terms and
    B0, with the mask), handed over as basis columns B [V x Lt] and
    coefficient columns H [K x Lt].  ``cache_exp`` keeps exp(i Phi) between
    applications (16 V K bytes)."""

    def __init__(self, b0, sampTimes, phs_spha, phs_conc, phs_grid, maps, b0mask=None, cache_exp=False):
        from .basis import basis_conc, basis_harm
        from .highorder import HighOrderPlan
        P = HighOrderPlan(b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask)
        mask = P.b0mask
        cols_b, cols_h = [], []
        for n in range(1, P.phs_spha.shape[0] + 1):
            bf = basis_harm(P.x, P.y, P.z, n)
            cols_b.append(bf * mask if (n > 4 and mask is not None) else bf)
            cols_h.append(P.phs_spha[n - 1])
        for n in range(1, P.phs_conc.shape[0] + 1):
            bf = basis_conc(P.x, P.y, P.z, n)
            cols_b.append(bf * mask if mask is not None else bf)
            cols_h.append(P.phs_conc[n - 1])
        if P.K > 1:
            cols_b.append(P.b0 * mask if mask is not None else P.b0)
            cols_h.append(P.sampTimes.reshape(-1, order="F"))
        self._B = np.asfortranarray(np.stack(cols_b, axis=1))
        self._H = np.asfortranarray(np.stack(cols_h, axis=1))
        mp = np.asfortranarray(np.asarray(maps, dtype=np.complex128))
        self.imSize, self.V, self.K = P.imSize, P.V, P.K
        self.C = mp.shape[-1]
        self._maps = mp.reshape(-1, order="F")
        h = lib().mmref_hio_create(self.V, self.K, self._B.shape[1], self._B.ctypes.data, self._H.ctypes.data,
                                   self.C, self._maps.ctypes.data, int(bool(cache_exp)))
        if not h:
            raise MemoryError("mmref_hio_create failed")
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mmref_hio_destroy(self._h)
        except Exception:
            pass

    def set_round_io(self, v):
        """Round every application's input and output to complex64 (an FP64
        operator behind single-precision I/O)."""
        lib().mmref_hio_set_round_io(self._h, int(bool(v)))

    def forward(self, x):
        xf = np.asfortranarray(np.asarray(x, dtype=np.complex128)).reshape(-1, order="F")
        y = np.empty(self.K * self.C, dtype=np.complex128)
        lib().mmref_hio_forward(self._h, xf.ctypes.data, y.ctypes.data)
        return y.reshape((self.K, self.C), order="F")

    def adjoint(self, y):
        yf = np.asfortranarray(np.asarray(y, dtype=np.complex128)).reshape(-1, order="F")
        x = np.empty(self.V, dtype=np.complex128)
        lib().mmref_hio_adjoint(self._h, yf.ctypes.data, x.ctypes.data)
        return x.reshape(self.imSize, order="F")

    def cgne(self, y, x0=None, NitMax=100, resThresh=None, resThreshRel=None, xChangeThresh=None,
             expResN=100):
        """Returns (x, resSqAll, nit)."""
        return _cgne_call(lib().mmref_hio_cgne, self._h, y, x0, self.imSize, NitMax, resThresh, resThreshRel,
                          xChangeThresh, expResN)
