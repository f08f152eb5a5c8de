"""MATLAB-style operator classes over the gfx950 library.

Mirrors the reference's class interfaces (paths relative to the MatMRI
repository root):

* :class:`nufftOp`       -- ``operators/@nufftOp/nufftOp.m`` (ctor ``:89``, ``mtimes`` ``:284``)
* :class:`rcvrOp`        -- ``operators/@rcvrOp/rcvrOp.m`` (ctor ``:47``, ``mtimes`` ``:104``)
* :class:`sampHighOrder` -- ``operators/@sampHighOrder/sampHighOrder.m`` (ctor ``:129``,
  ``mtimes`` ``:293``, ``setPhiDiv`` ``:590``)
* :func:`mrSampFunc`     -- ``operators/mrSampFunc.m``

``A * x`` is ``mtimes``; ``A.H`` (or ``A.ctranspose()``) is MATLAB's ``A'``,
a copy with the adjoint flag toggled that shares the device plan.

Array conventions.  NumPy arrays are read and written with MATLAB's
column-major meaning (``x.reshape(-1, order="F")``), so an image of size
``imN`` is an array of shape ``imN + reps`` and the output lives on the host
like a MATLAB CPU array.  Torch CUDA tensors (complex64) stay in HBM: their
row-major shape is the column-major shape reversed, e.g. an image batch is
``(reps, N2, N1)`` and data is ``(reps, C, M)``.
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import check, ptr

try:  # torch is only needed for device-resident tensors
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


class _Plan:
    """Owns one mm_op handle."""

    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = keep  # plans this one references (kept alive)
        self._fin = weakref.finalize(self, _destroy, handle.value)


def _destroy(h):
    try:
        _lib.lib().mm_op_destroy(C.c_void_p(h))
    except Exception:
        pass


class LinearOp:
    """Common ``mtimes``/``ctranspose`` machinery."""

    adjoint = 0

    def _shapes(self):
        ni, nd = C.c_int64(), C.c_int64()
        check(_lib.lib().mm_op_shape(self._plan.h, C.byref(ni), C.byref(nd)))
        return ni.value, nd.value

    # MATLAB A' (ctranspose) and A.' (transpose, nufftOp.m:279-282)
    def ctranspose(self):
        c = object.__new__(type(self))
        c.__dict__.update(self.__dict__)
        c.adjoint = 1 - self.adjoint
        return c

    transpose = ctranspose

    @property
    def H(self):
        return self.ctranspose()

    def __mul__(self, x):
        return self.mtimes(x)

    def __matmul__(self, x):
        return self.mtimes(x)

    # shapes of one repetition on each side, column-major
    def _in_shape(self):
        return self.data_shape if self.adjoint else self.image_shape

    def _out_shape(self):
        return self.image_shape if self.adjoint else self.data_shape

    def mtimes(self, x):
        return _apply(self._plan.h, self.adjoint, x, self._in_shape(), self._out_shape(), self._shapes())

    def set_option(self, key, value):
        """Kernel tuning knob (see mm_op_set_option)."""
        check(_lib.lib().mm_op_set_option(self._plan.h, key.encode(), int(value)))
        return self

    def get_option(self, key):
        """Current value of a tuning knob (see mm_op_get_option)."""
        v = C.c_int()
        check(_lib.lib().mm_op_get_option(self._plan.h, key.encode(), C.byref(v)))
        return v.value

    TUNED_OPTIONS = ("nu_grid_pairs", "nu_stage_samples", "nu_stage_entries", "nu_degrid_block", "nu_pair_cols",
                     "nu_xcd_order", "nu_grid_region", "nu_snake", "nu_slice_group", "nu_degrid_merge")

    def autotune(self, iters=10):
        """Run the device autotuner of the normal operator (mm_op_autotune)
        and return the options it chose."""
        check(_lib.lib().mm_op_autotune(self._plan.h, int(iters)))
        return {k: self.get_option(k) for k in self.TUNED_OPTIONS}

    def normal(self, x):
        """``A'*(A*x)`` in one call (the operator cgne iterates)."""
        return _apply(self._plan.h, 2, x, self.image_shape, self.image_shape, self._shapes())


def _apply(h, mode, x, in_shape, out_shape, shapes):
    ni, nd = shapes
    nin = nd if mode == 1 else ni
    nout = ni if mode in (1, 2) else nd
    L = _lib.lib()
    if _is_torch(x):
        if not x.is_cuda:
            raise ValueError("torch tensors must be on the GPU")
        xc = x.to(torch.complex64).contiguous()
        if xc.numel() % nin:
            raise ValueError(f"input has {xc.numel()} elements, not a multiple of {nin}")
        nrep = xc.numel() // nin
        y = torch.empty(nrep * nout, dtype=torch.complex64, device=x.device)
        torch.cuda.current_stream(x.device).synchronize()
        if mode == 2:
            check(L.mm_op_apply_normal(h, ptr(xc), ptr(y), nrep, 1))
        else:
            check(L.mm_op_apply(h, mode, ptr(xc), ptr(y), nrep, 1))
        check(L.mm_ctx_sync(_lib.context()))
        lead = tuple(xc.shape[: xc.dim() - len(in_shape)]) if xc.dim() >= len(in_shape) else ()
        shp = lead + tuple(reversed(out_shape)) if lead and int(np.prod(lead)) == nrep else \
            ((nrep,) if nrep > 1 else ()) + tuple(reversed(out_shape))
        return y.reshape(shp)
    xa = np.asarray(x)
    in_dtype = xa.dtype
    xf = np.asfortranarray(xa).reshape(-1, order="F").astype(np.complex64)
    if xf.size % nin:
        raise ValueError(f"input has {xf.size} elements, not a multiple of {nin}")
    nrep = xf.size // nin
    y = np.empty(nrep * nout, dtype=np.complex64)
    if mode == 2:
        check(L.mm_op_apply_normal(h, xf.ctypes.data, y.ctypes.data, nrep, 0))
    else:
        check(L.mm_op_apply(h, mode, xf.ctypes.data, y.ctypes.data, nrep, 0))
    reps = xa.shape[len(in_shape):] if xa.ndim >= len(in_shape) and \
        int(np.prod(xa.shape[:len(in_shape)])) == nin else ((nrep,) if nrep > 1 else ())
    y = y.reshape(tuple(out_shape) + tuple(reps), order="F")
    # revert the class like nufftOp.m:404-406
    if in_dtype == np.complex128 or in_dtype == np.float64:
        y = y.astype(np.complex128)
    return y


class nufftOp(LinearOp):
    """``nufftOp(imN, kloc, dcf, useGPU, os, kwidth)`` (nufftOp.m:89-136).

    ``kloc`` is ``[M x nD]`` in cycles/pixel, ``[-0.5, 0.5)`` (wrapped).
    ``kwidth`` defaults to 6 and ``os`` to 1.5 like the reference properties
    (nufftOp.m:65-66).  Only the GPU path exists (``useGPU`` is accepted and
    ignored); 1-D, 2-D and 3-D images are implemented (3-D: one volume per
    plan, grid sides up to 512).
    """

    def __init__(self, imN, kloc, dcf=None, useGPU=1, os=1.5, kwidth=6):
        imN = [int(n) for n in np.atleast_1d(imN)]
        if len(imN) == 2 and imN[1] == 1:
            imN = imN[:1]
        nd = len(imN)
        kloc = np.asarray(kloc, dtype=np.float64)
        if kloc.ndim == 1:
            kloc = kloc.reshape(-1, 1)
        if kloc.shape[1] != nd:
            raise ValueError("kloc must be [M x numel(imN)]")
        self._build([(imN, kloc, dcf)], os, kwidth)

    @classmethod
    def batch(cls, slices, os=1.5, kwidth=6):
        """Multi-slice plan: ``slices`` = list of ``(imN, kloc, dcf)``; data of
        all slices is concatenated (images then samples, per slice)."""
        obj = object.__new__(cls)
        obj._build([([int(n) for n in np.atleast_1d(s[0])], np.asarray(s[1], np.float64),
                     s[2] if len(s) > 2 else None) for s in slices], os, kwidth)
        return obj

    def _build(self, slices, os, kwidth):
        nd = len(slices[0][0])
        ns = len(slices)
        if nd == 3 and ns != 1:
            raise NotImplementedError("3-D nufftOp: one volume per plan")
        if nd not in (1, 2, 3) or any(len(s[0]) != nd for s in slices):
            raise ValueError("nufftOp: every slice needs the same 1, 2 or 3 image dimensions")
        imN = np.zeros(3 if nd == 3 else 2 * ns, dtype=np.int64)
        Ms = np.zeros(ns, dtype=np.int64)
        self._kl = []
        self._dc = []
        for s, (n, kl, dc) in enumerate(slices):
            if nd == 3:
                imN[:] = n
            else:
                imN[2 * s] = n[0]
                imN[2 * s + 1] = n[1] if nd == 2 else 1
            kl = np.asfortranarray(np.asarray(kl, np.float64).reshape(-1, nd))
            Ms[s] = kl.shape[0]
            self._kl.append(kl)
            self._dc.append(None if dc is None else np.ascontiguousarray(np.asarray(dc, np.float64).reshape(-1)))
        kp = (C.c_void_p * ns)(*[k.ctypes.data for k in self._kl])
        dp = (C.c_void_p * ns)(*[(0 if d is None else d.ctypes.data) for d in self._dc]) \
            if any(d is not None for d in self._dc) else None
        h = C.c_void_p()
        check(_lib.lib().mm_nufft_plan_create(
            _lib.context(), nd, ns, imN.ctypes.data_as(C.POINTER(C.c_int64)),
            Ms.ctypes.data_as(C.POINTER(C.c_int64)), kp, dp, float(os), int(kwidth), C.byref(h)))
        self._plan = _Plan(h)
        self.nslice = ns
        self.imN = [list(s[0]) for s in slices]
        self.M = [int(m) for m in Ms]
        self.os = os
        self.kwidth = kwidth
        if ns == 1:
            self.image_shape = tuple(self.imN[0])
            self.data_shape = (self.M[0],)
        else:
            self.image_shape = (int(sum(np.prod(n) for n in self.imN)),)
            self.data_shape = (int(sum(self.M)),)
        self.isToep = False

    def prepToep(self, tos=None, kwidth=None):
        """``prepToep(tos, kwidth)`` (nufftOp.m:410-432): a copy whose ``*``
        and ``'*`` both apply the normal operator in Toeplitz form
        (nufftOp.m:348-352), ``crop(fftnc(convMat .* ifftnc(pad(x, 2N))))``
        with ``convMat = 2^nd ifftnc(S2' * dcf)`` and ``S2 = nufftOp(2N, kloc,
        dcf, useGPU, tos, kwidth)`` (``tos`` default ``(os - 1) / 2 + 1``,
        ``kwidth`` default this operator's).  The point-spread image
        ``S2' * dcf`` (or ``S2' * ones``) is computed by the device NUFFT; the
        returned operator applies it on a grid of at least 2N per side."""
        if self.isToep:
            return self
        if tos is None:
            tos = 0.5 * (self.os - 1) + 1
        if kwidth is None:
            kwidth = self.kwidth
        slices = list(zip(self.imN, self._kl, self._dc))
        S2 = nufftOp.batch([([2 * n for n in imN], kl, dc) for imN, kl, dc in slices], tos, kwidth)
        w = np.concatenate([np.ones(kl.shape[0]) if dc is None else dc for _, kl, dc in slices])
        psf = np.asarray(S2.H * w.astype(np.complex64).reshape(S2.data_shape, order="F"))
        psf = np.ascontiguousarray(psf.reshape(-1, order="F").astype(np.complex64))
        T = nufftOp.batch(slices, 2.0, self.kwidth)
        check(_lib.lib().mm_nufft_set_toeplitz(T._plan.h, psf.ctypes.data))
        T.data_shape = T.image_shape
        T.isToep = True
        return T


class rcvrOp(LinearOp):
    """``rcvrOp(maps, isCalDat=false)`` (rcvrOp.m:47-94): ``maps`` has shape
    ``imN + [C]``.  The calibration path (``isCalDat=true``) is not on the
    accelerated path."""

    def __init__(self, maps, isCalDat=False, apN=None, kalph=None):
        if isCalDat:
            raise NotImplementedError("rcvrOp calibration from low-resolution data is not implemented")
        maps = np.asarray(maps)
        self.imN = maps.shape[:-1]
        self.C = maps.shape[-1]
        V = int(np.prod(self.imN))
        self._maps = np.asfortranarray(maps.astype(np.complex64)).reshape(-1, order="F")
        h = C.c_void_p()
        check(_lib.lib().mm_rcvr_plan_create(_lib.context(), V, self.C, self._maps.ctypes.data, C.byref(h)))
        self._plan = _Plan(h)
        self.image_shape = tuple(self.imN)
        self.data_shape = tuple(self.imN) + (self.C,)
        self.nslice = 1

    @classmethod
    def batch(cls, maps_list):
        """Per-slice maps (each ``imN_s + [C]``, same C) for a multi-slice SENSE operator."""
        obj = object.__new__(cls)
        Cc = maps_list[0].shape[-1]
        Vs = np.array([int(np.prod(m.shape[:-1])) for m in maps_list], dtype=np.int64)
        obj._maps = np.concatenate([np.asfortranarray(np.asarray(m).astype(np.complex64)).reshape(-1, order="F")
                                    for m in maps_list])
        h = C.c_void_p()
        check(_lib.lib().mm_rcvr_plan_create_multi(_lib.context(), len(maps_list),
                                                   Vs.ctypes.data_as(C.POINTER(C.c_int64)), Cc,
                                                   obj._maps.ctypes.data, C.byref(h)))
        obj._plan = _Plan(h)
        obj.C = Cc
        obj.nslice = len(maps_list)
        obj.image_shape = (int(Vs.sum()),)
        obj.data_shape = (int(Vs.sum()) * Cc,)
        return obj


class sampHighOrder(LinearOp):
    """``sampHighOrder(b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask,
    useGPU, useSingle, approach, opt)`` (sampHighOrder.m:129-291).

    ``b0`` is in rad/s with the image shape (or ``None``); ``sampTimes`` in s
    with the k-space shape; ``phs_spha`` is ``[Nh, kshape]`` and
    ``phs_conc`` ``[Nc, kshape]`` or ``None``; ``phs_grid`` a dict of
    ``x, y, z`` position arrays (m) with the image shape.  Approaches 0, 2
    and 3 give the same values and share one kernel; approach 1
    (interpolation: ``opt.svdThresh``, ``opt.subFact``) runs the truncated
    SVD form over a device NUFFT (``mm_hio_interp_create``).
    """

    def __init__(self, b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask=None, useGPU=1,
                 useSingle=0, approach=0, opt=None):
        self.approach = 0 if approach in (None, 2, 3) else int(approach)
        if self.approach == 1:
            self._sampTimes = sampTimes
            self._opt = dict(opt or {})
            self._init_interp(b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask, self._opt)
            return
        gx = np.asarray(phs_grid["x"], np.float64)
        self.imSize = gx.shape
        V = gx.size
        t = np.asarray(sampTimes, np.float64)
        self.kSize = t.shape if t.ndim else (1,)
        K = t.size
        hs = np.asarray(phs_spha, np.float64)
        Lh = hs.shape[0]
        hs = np.ascontiguousarray(hs.reshape(Lh, -1, order="F").T)  # [K][Lh]
        if hs.shape[0] != K:
            raise ValueError("Dimension mismatch between phs_spha and sampTimes")
        if phs_conc is None or np.size(phs_conc) == 0:
            Lc, hc = 0, None
        else:
            hc = np.asarray(phs_conc, np.float64)
            Lc = hc.shape[0]
            hc = np.ascontiguousarray(hc.reshape(Lc, -1, order="F").T)
            if hc.shape[0] != K:
                raise ValueError("Dimension mismatch between phs_conc and sampTimes")
        f = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, order="F"))
        if b0 is not None and np.size(b0) and np.shape(b0) != self.imSize:
            raise ValueError("Size mismatch between b0 and phs_grid.x")
        self._arr = [f(phs_grid["x"]), f(phs_grid["y"]), f(phs_grid["z"]),
                     f(b0) if b0 is not None and np.size(b0) else None, f(t), hs, hc,
                     f(b0mask) if b0mask is not None and np.size(b0mask) else None]
        a = self._arr
        h = C.c_void_p()
        check(_lib.lib().mm_hio_plan_create(_lib.context(), V, K, Lh, Lc, ptr(a[0]), ptr(a[1]), ptr(a[2]),
                                            ptr(a[3]), ptr(a[4]), ptr(a[5]), ptr(a[6]), ptr(a[7]),
                                            C.byref(h)))
        self._plan = _Plan(h)
        self.Lh, self.Lc = Lh, Lc
        self.image_shape = tuple(self.imSize)
        self.data_shape = tuple(self.kSize)
        self.nslice = 1

    # -- approach 1: interpolated (sampHighOrder.m:314-359, :624-902) -------
    def _init_interp(self, b0, sampTimes, phs_spha, phs_conc, phs_grid, b0mask, opt):
        """Plan of the interpolated operator.  On the host: the NUFFT
        locations and phase ramp of the linear terms (createNufftForInterp,
        :624-701), the smoothing and time subsampling of the coefficients
        (subSampTime, :959-982), the small QR and SVD factorisations.  On the
        device: every product with exp(1i*Phi_res) of the truncated SVD (a
        randomized range finder with two power iterations over the direct
        operator of the remaining terms), the NUFFT and the operator."""
        svdThresh = float(opt.get("svdThresh", 0.05))
        subFact = max(1, int(opt.get("subFact", 5)))
        if int(opt.get("subFactSpc", 1)) > 1:
            raise NotImplementedError("sampHighOrder approach 1: subFactSpc > 1 (not recommended by the reference)")
        ntry = int(opt.get("svdNtry", 30))
        gx, gy, gz = (np.asarray(phs_grid[c], np.float64) for c in ("x", "y", "z"))
        if gx.ndim != 2 or min(gx.shape) < 2:
            raise NotImplementedError("sampHighOrder approach 1: 2-D axial grids only")
        t = np.asarray(sampTimes, np.float64).reshape(-1)
        K = t.size
        hs = np.asarray(phs_spha, np.float64).reshape(np.shape(phs_spha)[0], -1)
        if hs.shape[1] != K:
            raise ValueError("Dimension mismatch between phs_spha and sampTimes")
        hc = None if phs_conc is None or np.size(phs_conc) == 0 else \
            np.asarray(phs_conc, np.float64).reshape(np.shape(phs_conc)[0], -1)
        N1, N2 = gx.shape
        V = N1 * N2
        if gz[1, 1] - gz[0, 0] != 0:
            raise ValueError("non-axial slices not supported for interpolated approach")
        # createNufftForInterp: which grid axis each linear term runs along
        if abs(gx[0, 1] - gx[0, 0]) > np.finfo(np.float64).eps:
            if gy[0, 1] - gy[0, 0] != 0:
                raise ValueError("non-axial slices not supported for interpolated approach")
            kl = np.stack([np.sign(gy[1, 0] - gy[0, 0] or -1) * hs[2] * abs(gy[1, 0] - gy[0, 0]),
                           np.sign(gx[0, 1] - gx[0, 0] or -1) * hs[1] * abs(gx[0, 1] - gx[0, 0])], axis=1)
            shift = gy[N1 // 2, 0] * hs[2] + gx[0, N2 // 2] * hs[1]
        else:
            kl = np.stack([np.sign(gx[1, 0] - gx[0, 0] or -1) * hs[1] * abs(gx[1, 0] - gx[0, 0]),
                           np.sign(gy[0, 1] - gy[0, 0] or -1) * hs[2] * abs(gy[0, 1] - gy[0, 0])], axis=1)
            shift = gx[N1 // 2, 0] * hs[1] + gy[0, N2 // 2] * hs[2]
        kloc = kl / (2 * np.pi)
        phsShft = np.exp(1j * shift)
        nu = nufftOp([N1, N2], kloc, None, 1, 1.5, 6)
        # subSampTime: smoothing window filtNd(2*(h+1), 0.5, 1, 'gauss') without
        # its first tap, every subFact-th sample plus the last
        inds = np.arange(0, K, subFact)
        if inds[-1] != K - 1:
            inds = np.append(inds, K - 1)
        hs_s, hc_s = hs.copy(), None if hc is None else hc.copy()
        if subFact > 1:
            half = int(np.ceil(subFact / 2))
            sz = 2 * (half + 1)
            g = np.exp(-0.5 * (np.arange(-sz // 2, sz // 2) / (sz * 0.5 / 2.355)) ** 2)
            win = g[1:] / g[1:].sum()
            for a, a0 in ((hs_s, hs), (hc_s, hc)):
                if a is None:
                    continue
                for r in range(a.shape[0]):
                    a[r] = np.convolve(a0[r], win, mode="same")
                a[:, :half] = a0[:, :half]
                a[:, -half:] = a0[:, -half:]
        hs_s = hs_s[:, inds]
        hs_s[1:3] = 0.0   # the linear terms are the NUFFT's
        hc_s = None if hc_s is None else hc_s[:, inds]
        B = sampHighOrder(b0, t[inds], hs_s, hc_s, phs_grid, b0mask)
        mask = None if b0mask is None or np.size(b0mask) == 0 else \
            (np.asarray(b0mask).reshape(-1, order="F") > 0.5)
        U, sv, Vc, Ns = _interp_svd(B, V, len(inds), mask, svdThresh, ntry)
        st = (Vc[:, :Ns] * sv[:Ns])            # conj(V*S): Vc holds conj(V)
        if subFact > 1:
            from scipy.interpolate import PchipInterpolator
            tk = np.arange(K)
            st = PchipInterpolator(inds, st.real, axis=0)(tk) + 1j * PchipInterpolator(inds, st.imag, axis=0)(tk)
        sp = U[:, :Ns]
        if mask is not None:
            full = np.zeros((V, Ns), complex)
            full[mask] = sp
            sp = full
        self._sp = np.asfortranarray(sp.astype(np.complex64))
        self._st = np.asfortranarray(st.astype(np.complex64))
        self._ph = np.ascontiguousarray(phsShft.astype(np.complex64))
        h = C.c_void_p()
        check(_lib.lib().mm_hio_interp_create(_lib.context(), nu._plan.h, V, K, Ns, self._sp.ctypes.data,
                                              self._st.ctypes.data, self._ph.ctypes.data, C.byref(h)))
        self._plan = _Plan(h, keep=(nu._plan,))
        self._nufft = nu
        self.Ns = Ns
        self.svdValues = sv
        self.imSize = gx.shape
        self.kSize = (K,)
        self.Lh, self.Lc = hs.shape[0], 0 if hc is None else hc.shape[0]
        self.image_shape = tuple(self.imSize)
        self.data_shape = (K,)
        self.nslice = 1

    def updateObj(self, phs_spha, phs_conc, b0map, b0mask, phs_grid):
        """``updateObj`` (sampHighOrder.m:904-957): new field coefficients, B0
        map, mask and grid for the same sample times and options (the next
        slice or volume of a multi-slice acquisition).  For approach 1 the
        interpolated plan (NUFFT locations, phase ramp, truncated SVD) is
        rebuilt on the device (prepForInterp, :937); other approaches raise
        like the reference (:953-955).  Returns the operator."""
        if self.approach != 1:
            raise NotImplementedError("this case not yet coded")   # :953-955
        self._init_interp(b0map, self._sampTimes, phs_spha, phs_conc, phs_grid, b0mask, self._opt)
        self.adjoint = 0
        return self

    def setPhiDiv(self, ksphaDiv, kconcDiv=None):
        """``setPhiDiv`` (sampHighOrder.m:590-622): the forward becomes
        ``(1i*exp(1i*Phi).*dPhi).' * x / sqrt(V)``.  Returns the operator,
        like the reference's value-class mutator."""
        if self.approach == 1:
            raise NotImplementedError("phiDiv not yet implemented for interpolated approach")   # :335-337
        hd = np.ascontiguousarray(np.asarray(ksphaDiv, np.float64).reshape(self.Lh, -1, order="F").T)
        gd = None
        if self.Lc:
            gd = np.ascontiguousarray(np.asarray(kconcDiv, np.float64).reshape(self.Lc, -1, order="F").T)
        check(_lib.lib().mm_hio_set_phi_div(self._plan.h, ptr(hd), ptr(gd)))
        return self


def _interp_svd(B, V, Ks, mask, svdThresh, ntry):
    """Top singular triplets of b = exp(1i*Phi_res) [V x Ks] (rows of the
    mask only) by a randomized range finder with two power iterations: the
    products b*X and b'*Q run on the device through B, the direct operator
    of the remaining phase terms (B*x = b.'x/sqrt(V), B'*y = conj(b)y/sqrt(V),
    so b*X = conj(B'*conj(X))*sqrt(V) and b'*Q = conj(B*conj(Q))*sqrt(V));
    the QR and the small SVD on the host.  Grows the rank like the
    reference's svds loop (x1.2 until the smallest value found is below
    svdThresh times the largest) and keeps Ns = the first index below it
    (inclusive, sampHighOrder.m:818-836).  Returns (U [Vm x l], s, conj(V)
    [Ks x l], Ns)."""
    sq = np.sqrt(V)
    dev = "cuda"

    def bx(X):      # b X, X host [Ks x l] -> host [Vm x l]
        Xt = torch.as_tensor(np.ascontiguousarray(np.conj(X).T).astype(np.complex64), device=dev)   # (l, Ks)
        Y = (B.H * Xt).cpu().numpy().astype(np.complex128)       # (l, N2, N1): conj(b X) / sqrt(V)
        Y = np.conj(Y.reshape(Y.shape[0], -1).T) * sq              # column-major images as columns
        return Y if mask is None else Y[mask]

    def bhq(Q):     # b' Q, Q host [Vm x l] -> host [Ks x l]
        Qf = Q
        if mask is not None:
            Qf = np.zeros((V, Q.shape[1]), complex)
            Qf[mask] = Q
        Qt = torch.as_tensor(np.ascontiguousarray(_to_rowmajor_images(np.conj(Qf), B.imSize)).astype(np.complex64),
                             device=dev)
        Y = (B * Qt).cpu().numpy().astype(np.complex128)          # (l, Ks)
        return np.conj(Y.T) * sq

    rng = np.random.default_rng(12345)
    while True:
        if ntry > 200:
            raise ValueError("many large singular values. Try providing mask or adjusting svdThresh.")
        l = min(ntry + 10, Ks, V if mask is None else int(mask.sum()))
        Om = rng.standard_normal((Ks, l)) + 1j * rng.standard_normal((Ks, l))
        Q, _ = np.linalg.qr(bx(Om))
        for _ in range(2):
            Z, _ = np.linalg.qr(bhq(Q))
            Q, _ = np.linalg.qr(bx(Z))
        Bs = np.conj(bhq(Q)).T                                       # Q' b  [l x Ks]
        Ub, s, Vh = np.linalg.svd(Bs, full_matrices=False)
        U = Q @ Ub
        below = np.nonzero(s / s.max() < svdThresh)[0]
        if len(below) and (below[0] < ntry or l >= min(Ks, Q.shape[0])):
            return U, s, Vh.T, int(below[0]) + 1
        ntry = int(np.ceil(1.2 * ntry))


def _to_rowmajor_images(Q, imSize):
    """[V x l] column-major images -> (l, N2, N1) row-major tensor layout."""
    l = Q.shape[1]
    return Q.T.reshape((l,) + tuple(reversed(imSize)))


class senseOp(LinearOp):
    """E = S o R (mrSampFunc.m:97-107; adjoint :126-144) as one device
    operator; W (data weights) is used by :meth:`normal` and by cgne."""

    def __init__(self, S, R, W=None):
        self.S, self.R = S, R
        self._W = None if W is None else np.ascontiguousarray(
            np.asarray(W, np.float32).reshape(-1, order="F"))
        h = C.c_void_p()
        check(_lib.lib().mm_sense_create(_lib.context(), S._plan.h, R._plan.h, ptr(self._W), C.byref(h)))
        self._plan = _Plan(h, keep=(S._plan, R._plan))
        self.nslice = getattr(S, "nslice", 1)
        self.image_shape = S.image_shape
        ni, nd = self._shapes()
        if self.nslice == 1:
            self.data_shape = tuple(S.data_shape) + (R.C,)
        else:
            self.data_shape = (nd,)


class cartOp(LinearOp):
    """A Cartesian k-space mask used as mrSampFunc's ``samplingOp``
    (mrSampFunc.m:108-114, :137-143) as a device operator: ``cartOp(mask) * x``
    is ``mask .* ifftnc(x, nDim)``, its adjoint ``fftnc(mask .* y, nDim)``
    (nDim = ndims(mask), 1 for a column).  Real masks of 1 or 2 dimensions."""

    def __init__(self, mask):
        m = np.asarray(mask)
        nd = m.ndim
        if nd == 2 and m.shape[1] == 1:
            nd = 1
        if nd not in (1, 2):
            raise NotImplementedError("Cartesian masks of 1 or 2 dimensions only")
        if np.iscomplexobj(m) and np.any(np.imag(m) != 0):
            raise NotImplementedError("complex Cartesian masks are not supported")
        imN = tuple(int(n) for n in m.shape[:nd])
        self._mask = np.asfortranarray(np.real(m).astype(np.float32)).reshape(-1, order="F")
        n = np.array(imN + ((1,) if nd == 1 else ()), dtype=np.int64)
        h = C.c_void_p()
        check(_lib.lib().mm_cart_plan_create(_lib.context(), nd, n.ctypes.data_as(C.POINTER(C.c_int64)),
                                             self._mask.ctypes.data, C.byref(h)))
        self._plan = _Plan(h)
        self.image_shape = imN
        self.data_shape = imN
        self.nslice = 1


class sumDimsOp(LinearOp):
    """``opt.sumDims`` of mrSampFunc (mrSampFunc.m:100-105): ``*`` sums an
    array of shape ``dims`` over the (1-based) dimensions ``sumDims``,
    keeping them as singletons; the adjoint broadcasts back (MATLAB's
    implicit expansion in the 'transp' branch)."""

    def __init__(self, dims, sumDims):
        dims = tuple(int(d) for d in dims)
        flags = np.array([1 if (d + 1) in set(int(k) for k in sumDims) else 0 for d in range(len(dims))],
                         dtype=np.int32)
        if any(int(k) > len(dims) for k in sumDims):
            dims = dims + (1,) * (max(int(k) for k in sumDims) - len(dims))
            flags = np.array([1 if (d + 1) in set(int(k) for k in sumDims) else 0 for d in range(len(dims))],
                             dtype=np.int32)
        dd = np.array(dims, dtype=np.int64)
        h = C.c_void_p()
        check(_lib.lib().mm_sumdims_plan_create(_lib.context(), len(dims), dd.ctypes.data_as(C.POINTER(C.c_int64)),
                                                flags.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h)))
        self._plan = _Plan(h)
        self.image_shape = dims
        self.data_shape = tuple(1 if f else n for n, f in zip(dims, flags))
        self.nslice = 1


_sense_cache = {}
_aux_cache = {}


def _sense_for(S, R, W=None):
    key = (id(S._plan), id(R._plan), None if W is None else id(W))
    e = _sense_cache.get(key)
    if e is None or e[1]() is None or e[2]() is None:
        op = senseOp(S, R, W)
        _sense_cache[key] = (op, weakref.ref(S._plan), weakref.ref(R._plan))
        return op
    return e[0]


def _cart_for(mask):
    m = np.asarray(mask)
    key = ("cart", m.shape, m.dtype.str, m.tobytes())
    op = _aux_cache.get(key)
    if op is None:
        op = _aux_cache[key] = cartOp(m)
    return op


def _sumdims_for(dims, sumDims):
    key = ("sum", tuple(dims), tuple(int(k) for k in np.atleast_1d(sumDims)))
    op = _aux_cache.get(key)
    if op is None:
        op = _aux_cache[key] = sumDimsOp(dims, np.atleast_1d(sumDims))
    return op


def _np_to_dev(a):
    """A NumPy array (column-major meaning) as a complex64 CUDA tensor with
    the reversed shape the device operators use."""
    a = np.asarray(a)
    return torch.as_tensor(np.ascontiguousarray(np.transpose(a)).astype(np.complex64), device="cuda")


def _dev_to_np(t):
    return np.transpose(t.cpu().numpy())


def mrSampFunc(x, transp, samplingOp, R=None, opt=None):
    """``mrSampFunc(x, transp, samplingOp, R, opt)`` (mrSampFunc.m:1-161) on
    the device.  ``samplingOp`` is a device operator (nufftOp, sampHighOrder)
    or a Cartesian mask (run as :class:`cartOp`); options ``imNFull`` /
    ``daNFull`` (vectorisation), ``W``, ``tikReg``, ``hardReal`` and
    ``sumDims`` (a :class:`sumDimsOp` between R and the sampling, with the
    vectors kept on the GPU between the three operators)."""
    opt = dict(opt or {})
    S = samplingOp if isinstance(samplingOp, LinearOp) else _cart_for(samplingOp)
    imNFull, daNFull = opt.get("imNFull"), opt.get("daNFull")
    tikReg, W = opt.get("tikReg"), opt.get("W")
    sumDims = opt.get("sumDims")
    if sumDims is not None and np.size(sumDims) == 0:
        sumDims = None
    hardReal = opt.get("hardReal", 0) or 0
    if tikReg is not None and (imNFull is None or daNFull is None):
        raise ValueError("Tikhonov regularization requires opt.imNFull and opt.daNFull")
    host = not _is_torch(x)
    if sumDims is not None:
        # R, the sum and S as three device operators on device tensors
        full = tuple(R.data_shape) if R is not None else tuple(imNFull if imNFull is not None else np.shape(x))
        Sm = _sumdims_for(full, sumDims)

        def fwd(v):
            v = R.mtimes(v) if R is not None else v
            return S.mtimes(Sm.mtimes(v))

        def adj(v):
            v = Sm.H.mtimes(S.H.mtimes(v))
            return R.H.mtimes(v) if R is not None else v
    else:
        E = _sense_for(S, R) if R is not None else S
        fwd, adj = E.mtimes, E.H.mtimes
    if transp == "notransp":
        xv = np.asarray(x) if host else x
        if hardReal:
            xv = np.real(xv) + 0j if host else torch.real(xv).to(torch.complex64)
        if tikReg is not None:
            t = np.atleast_1d(tikReg)
            x_tik = t[0] * xv.real + 1j * t[1] * xv.imag if t.size == 2 else t[0] * xv
        if imNFull is not None and host:
            xv = np.reshape(xv, tuple(imNFull), order="F")
        if sumDims is not None and host:
            y = _dev_to_np(fwd(_np_to_dev(xv)))
        else:
            y = fwd(xv)
        if daNFull is not None:
            y = y.reshape(-1, order="F") if host else y.reshape(-1)
        if tikReg is not None:
            y = np.concatenate([y.reshape(-1, order="F"), x_tik.reshape(-1, order="F")]) if host else \
                torch.cat([y.reshape(-1), x_tik.reshape(-1)])
        return y
    if transp == "transp":
        xv = np.asarray(x) if host else x
        if tikReg is not None:
            nda = int(np.prod(daNFull))
            x_tik = xv[nda:]
            xv = xv[:nda]
        if W is not None:
            xv = xv * (np.asarray(W) if host else torch.as_tensor(np.asarray(W), device=xv.device))
        if daNFull is not None and host:
            xv = np.reshape(xv, tuple(daNFull), order="F")
        if sumDims is not None and host:
            out = _dev_to_np(adj(_np_to_dev(xv)))
        else:
            out = adj(xv)
        if imNFull is not None:
            out = out.reshape(-1, order="F") if host else out.reshape(-1)
        if tikReg is not None:
            t = np.atleast_1d(tikReg)
            out = out + (t[0] * x_tik.real + 1j * t[1] * x_tik.imag if t.size == 2 else t[0] * x_tik)
        if hardReal:
            out = out.real + 0j if host else torch.real(out).to(torch.complex64)
        return out
    raise ValueError("transp must be 'notransp' or 'transp'")


def mrSampFuncMat(y, Crcvr, b0, sampTimes, phs_spha, phs_conc, phs_grid, imMask=None, on_device=False):
    """``[AtA, Aty] = mrSampFuncMat(y, Crcvr, b0, sampTimes, phs_spha, phs_conc,
    phs_grid, imMask)`` (operators/mrSampFuncMat.m:1-47) on the device: the
    dense ``[Vm x Vm]`` normal matrix over the masked voxels (the
    single-precision dense encoding matrix built by a HIP kernel, its Gram
    matrix on rocBLAS, the coil maps folded in by a second kernel) and
    ``Aty = R'*(S'*y)`` over the mask.  With ``on_device`` AtA stays in HBM
    as a CUDA tensor (row-major ``(Vm, Vm)`` = the column-major matrix
    transposed); otherwise NumPy arrays."""
    S = sampHighOrder(b0, sampTimes, phs_spha, phs_conc, phs_grid)
    R = rcvrOp(Crcvr)
    V = int(np.prod(S.image_shape))
    mask = np.ones(V, np.uint8) if imMask is None else \
        np.ascontiguousarray(np.asarray(imMask).reshape(-1, order="F").astype(np.uint8))
    Aty = np.asarray(senseOp(S, R).H * np.asarray(y)).reshape(-1, order="F")[mask.astype(bool)]
    L = _lib.lib()
    vm = C.c_int64()
    check(L.mm_hio_dense_normal(S._plan.h, R.C, R._maps.ctypes.data, mask.ctypes.data, None, C.byref(vm), 0))
    n = vm.value
    if on_device:
        AtA = torch.empty((n, n), dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        check(L.mm_hio_dense_normal(S._plan.h, R.C, R._maps.ctypes.data, mask.ctypes.data, AtA.data_ptr(),
                                    C.byref(vm), 1))
        return AtA, Aty
    AtA = np.empty((n, n), dtype=np.complex64, order="F")
    check(L.mm_hio_dense_normal(S._plan.h, R.C, R._maps.ctypes.data, mask.ctypes.data, AtA.ctypes.data,
                                C.byref(vm), 0))
    return AtA, Aty


class mrSampHandle:
    """The handle ``@(x, transp) mrSampFunc(x, transp, S, R, opt)`` the demos
    pass to cgne (demo_regSENSE.m:89-99).  cgne recognises it and runs the
    whole solve on the device."""

    def __init__(self, S, R, opt=None):
        self.S, self.R, self.opt = S, R, dict(opt or {})

    def __call__(self, x, transp):
        return mrSampFunc(x, transp, self.S, self.R, self.opt)

    def sense(self, W=None):
        S = self.S if isinstance(self.S, LinearOp) else _cart_for(self.S)
        return _sense_for(S, self.R, W)
