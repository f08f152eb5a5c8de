"""Trajectory data and delay estimation (the reference's ``findDel/``).

* :func:`load_traj_mat`  -- reads a MATLAB v5 trajectory file such as
  ``data_traj_spiral_R4.mat`` (``phs_spha``, ``phs_conc``, ``datatime``,
  ``tdwelltraj``) with ``scipy.io.loadmat`` (no code in the file runs).
* :func:`interpTrajTime` -- ``findDel/interpTrajTime.m:1-27`` on the device
  (``mm_interp_traj_time``: FP64 makima resampling kernel).
* :func:`findDelAuto`    -- ``findDel/findDelAuto.m:1-143``: the coarse
  search and the Gauss-Newton delay update over device ``sampHighOrder``
  operators, the device ``cgne`` and ``setPhiDiv``; only the scalar update
  ``bhathat(:) \\ bhat(:)`` is formed on the host from two FP64 device
  inner products.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .operators import mrSampFunc, rcvrOp, sampHighOrder, senseOp
from .solvers import cgne

try:
    import torch
except Exception:  # pragma: no cover
    torch = None


def load_traj_mat(path):
    """The arrays of a MATLAB v5 trajectory file, as float64 NumPy arrays:
    ``phs_spha`` [L x Nh], ``phs_conc`` [L x Nc] (or None), ``datatime``
    [K] and ``tdwelltraj`` (a float)."""
    import scipy.io as sio
    d = sio.loadmat(path)
    get = lambda k: np.asarray(d[k], dtype=np.float64) if k in d and np.size(d[k]) else None
    return {"phs_spha": get("phs_spha"), "phs_conc": get("phs_conc"),
            "datatime": get("datatime").reshape(-1), "tdwelltraj": float(np.asarray(d["tdwelltraj"]).reshape(-1)[0])}


def _interp(a, tdwell_in, delTime, tq):
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    L = a.shape[0]
    ncol = int(np.prod(a.shape[1:])) if a.ndim > 1 else 1
    out = np.empty((len(tq), ncol), dtype=np.float64, order="F")
    check(_lib.lib().mm_interp_traj_time(_lib.context(), a.ctypes.data, L, ncol, float(tdwell_in), float(delTime),
                                         tq.ctypes.data, len(tq), out.ctypes.data, 0))
    return out.reshape((len(tq),) + a.shape[1:], order="F") if a.ndim > 1 else out[:, 0]


def interpTrajTime(phs_spha, phs_conc, tdwell_in, delTime, datatime):
    """``[phs_spha, phs_conc] = interpTrajTime(phs_spha, phs_conc, tdwell_in,
    delTime, datatime)``: each array resampled along its first dimension from
    ``-delTime + tdwell_in*(0:L-1)`` to ``datatime`` (makima, NaN outside)."""
    empty = lambda a: a is None or np.size(a) == 0
    if empty(phs_spha) and empty(phs_conc):
        raise ValueError("at least one of phs_spha or phs_conc must be non-empty!")
    tq = np.ascontiguousarray(np.asarray(datatime, dtype=np.float64).reshape(-1))
    a = None if empty(phs_spha) else _interp(phs_spha, tdwell_in, delTime, tq)
    b = None if empty(phs_conc) else _interp(phs_conc, tdwell_in, delTime, tq)
    return a, b


def _deriv(a, tdwell):
    """``conv2(a(:,:), [1/8 1/4 0 -1/4 -1/8]', 'same') / tdwell``
    (findDelAuto.m:73-75): a smoothed time derivative along dimension 1."""
    if a is None:
        return None
    a = np.asarray(a, dtype=np.float64)
    L = a.shape[0]
    # out[r] = (a[r+2] + 2 a[r+1] - 2 a[r-1] - a[r-2]) / 8, zero outside
    pad = np.zeros((L + 4,) + a.shape[1:])
    pad[2:L + 2] = a
    out = (pad[4:] + 2 * pad[3:L + 3] - 2 * pad[1:L + 1] - pad[:L]) / 8.0
    return out / tdwell


def findDelAuto(delSk0, data_in, tdwell, phs_spha, phs_conc, datatime, phs_sph_grid, b0map, Crcvr, maxNit_cgne,
                delJumpFact=3, numCoarseSearch=2, verbose=False):
    """``[delSk, delSk_perIt] = findDelAuto(delSk0, data_in, tdwell, phs_spha,
    phs_conc, datatime, phs_sph_grid, b0map, Crcvr, maxNit_cgne, delJumpFact,
    numCoarseSearch)`` (findDel/findDelAuto.m:1-143).  Times in one unit
    (tdwell, datatime, b0map in rad/unit); ``data_in`` [K x C] (a host
    array, or a GPU tensor in this package's torch layout, [C, K]: then it is
    used in place); ``Crcvr`` has the image size plus a coil dimension.
    Every operator application and solve runs on the GPU."""
    R = rcvrOp(Crcvr)
    t = np.asarray(datatime, dtype=np.float64).reshape(-1)
    dts = float(t[1] - t[0])
    dspha, dconc = _deriv(phs_spha, tdwell), _deriv(phs_conc, tdwell)
    if torch is not None and torch.is_tensor(data_in):
        if not data_in.is_cuda:
            raise ValueError("torch tensors must be on the GPU")
        data = data_in.to(torch.complex64).contiguous()
    else:
        data = torch.as_tensor(np.ascontiguousarray(np.transpose(np.asarray(data_in))).astype(np.complex64),
                               device="cuda")

    def build(dl):
        a, b = interpTrajTime(phs_spha, phs_conc, tdwell, dl, t)
        return sampHighOrder(b0map, t, a.T, None if b is None else b.T, phs_sph_grid)

    delSk = float(delSk0)
    if numCoarseSearch > 0:                                                        # :77-101
        resmin, delmin = np.inf, delSk
        for ct in range(-numCoarseSearch, numCoarseSearch + 1):
            dl = delSk0 + ct * dts
            E = senseOp(build(dl), R)
            _, resvec, *_ = cgne(E, data, None, maxNit_cgne, {"stopOnResInc": 0})
            if resvec[-1] < resmin:
                delmin, resmin = dl, resvec[-1]
        delSk = delmin
    per_it = [delSk]
    ntry, delChange, delChangePrev = 0, np.inf, 0.0
    nit, resvec = maxNit_cgne, None
    while abs(delChange) > 0.005:                                                  # :106-141
        S = build(delSk)
        E = senseOp(S, R)
        opt = {"stopOnResInc": 1}
        if ntry > 0:
            opt = {"stopOnResInc": 0}
            nit = len(resvec) - 1
        x, resvec, *_ = cgne(E, data, None, nit, opt)
        bhat = data - E * x
        a, b = interpTrajTime(dspha, dconc, tdwell, delSk, t)
        S.setPhiDiv(a.T, None if b is None else b.T)
        bhh = senseOp(S, R) * x
        hv, bv = bhh.reshape(-1).to(torch.complex128), bhat.reshape(-1).to(torch.complex128)
        delChange = delJumpFact * float(torch.real(torch.vdot(hv, bv) / torch.vdot(hv, hv)).item())
        delSk = delSk + delChange
        if verbose:
            print(f" {delSk:.2f},", end="", flush=True)
        if ntry > 0 and np.sign(delChangePrev) != np.sign(delChange) and delJumpFact > 1:
            delJumpFact = delJumpFact / 2
        delChangePrev = delChange
        ntry += 1
        per_it.append(delSk)
        if ntry > 50:
            break
    return delSk, np.array(per_it)
