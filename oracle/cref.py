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
    _lib = L
    return L


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
