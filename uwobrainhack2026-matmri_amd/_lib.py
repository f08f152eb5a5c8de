"""ctypes binding of libmatmri_hip.so (the C ABI in include/matmri_hip.h).

The library is loaded from this package directory (or from the path in the
environment variable MATMRI_HIP_LIB, for a library built elsewhere).  There is
no fallback: if it is missing or no gfx950 device is visible, the operators
raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MATMRI_HIP_LIB") or os.path.join(HERE, "libmatmri_hip.so")

MM_OK, MM_ERR_ARG, MM_ERR_OOM, MM_ERR_DEVICE, MM_ERR_HIP, MM_ERR_UNSUPPORTED = range(6)

# every symbol include/matmri_hip.h declares (checked by the CPU test suite)
EXPORTS = [
    "mm_last_error", "mm_version", "mm_device_count", "mm_ctx_create", "mm_ctx_destroy",
    "mm_ctx_sync", "mm_nufft_plan_create", "mm_rcvr_plan_create", "mm_rcvr_plan_create_multi",
    "mm_hio_plan_create", "mm_hio_set_phi_div", "mm_nufft_set_toeplitz", "mm_sense_create", "mm_op_shape", "mm_op_nslice",
    "mm_op_apply", "mm_op_apply_normal", "mm_cg_opts_default", "mm_cgne_solve", "mm_cg_set_graph",
    "mm_op_profile", "mm_op_set_option", "mm_op_get_option", "mm_op_autotune", "mm_ctx_stream", "mm_op_destroy", "mm_cg_last_fallback",
    "mm_op_profile_stage",
    "mm_wsnet_create", "mm_wsnet_num_params", "mm_wsnet_set_params", "mm_wsnet_get_params",
    "mm_wsnet_forward", "mm_wsnet_loss_grad", "mm_wsnet_train_step", "mm_wsnet_destroy",
    "mm_rcvr_normalize_ssq",
    "mm_cart_plan_create",
    "mm_sumdims_plan_create",
    "mm_cgne_solve_ex",
    "mm_interp_traj_time",
    "mm_hio_dense_normal",
    "mm_hio_interp_create",
    "mm_rcvr_normalize_ssq_noalloc",
]


class MMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[matmri_hip error {code}] {msg}")
        self.code = code


class OutOfMemory(MMError, MemoryError):
    pass


class CgOpts(C.Structure):
    _fields_ = [
        ("resThresh", C.c_double), ("resThreshRel", C.c_double), ("resChangeThresh", C.c_double),
        ("xChangeThresh", C.c_double), ("expResN", C.c_int), ("stopOnResInc", C.c_int),
        ("resIncThresh", C.c_double), ("stopOnXdifInc", C.c_int), ("stopThresh", C.c_double),
        ("nseExtraIt", C.c_int), ("x0Fallback", C.c_int),
    ]


_lib = None


def lib():
    """Load (once) and return the shared library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          "(or python uwobrainhack2026-matmri_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    P = C.POINTER
    L.mm_last_error.restype = C.c_char_p
    L.mm_version.restype = C.c_char_p
    L.mm_device_count.restype = i32
    L.mm_ctx_create.argtypes = [i32, vp, P(vp)]
    L.mm_ctx_destroy.argtypes = [vp]
    L.mm_ctx_sync.argtypes = [vp]
    L.mm_nufft_plan_create.argtypes = [vp, i32, i32, P(i64), P(i64), P(vp), P(vp), dbl, i32, P(vp)]
    L.mm_rcvr_plan_create.argtypes = [vp, i64, i64, vp, P(vp)]
    L.mm_rcvr_plan_create_multi.argtypes = [vp, i32, P(i64), i64, vp, P(vp)]
    L.mm_hio_plan_create.argtypes = [vp, i64, i64, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, P(vp)]
    L.mm_hio_set_phi_div.argtypes = [vp, vp, vp]
    L.mm_cart_plan_create.argtypes = [vp, i32, P(i64), vp, P(vp)]
    L.mm_sumdims_plan_create.argtypes = [vp, i32, P(i64), P(i32), P(vp)]
    L.mm_interp_traj_time.argtypes = [vp, vp, i64, i32, dbl, dbl, vp, i64, vp, i32]
    L.mm_hio_dense_normal.argtypes = [vp, i64, vp, vp, vp, P(i64), i32]
    L.mm_hio_interp_create.argtypes = [vp, vp, i64, i64, i32, vp, vp, vp, P(vp)]
    L.mm_nufft_set_toeplitz.argtypes = [vp, vp]
    L.mm_sense_create.argtypes = [vp, vp, vp, vp, P(vp)]
    L.mm_op_shape.argtypes = [vp, P(i64), P(i64)]
    L.mm_op_nslice.argtypes = [vp]
    L.mm_op_nslice.restype = i32
    L.mm_op_apply.argtypes = [vp, i32, vp, vp, i64, i32]
    L.mm_op_apply_normal.argtypes = [vp, vp, vp, i64, i32]
    L.mm_cg_opts_default.argtypes = [P(CgOpts)]
    L.mm_cg_opts_default.restype = None
    L.mm_cgne_solve.argtypes = [vp, vp, vp, vp, i32, P(CgOpts), vp, vp, vp, vp, vp, i32]
    L.mm_cgne_solve_ex.argtypes = [vp, vp, vp, vp, i32, P(CgOpts), i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]
    L.mm_cg_set_graph.argtypes = [vp, i32]
    L.mm_cg_last_fallback.argtypes = [vp, vp]
    L.mm_op_destroy.argtypes = [vp]
    L.mm_op_profile.argtypes = [vp, i32, vp, vp, i64, i32, P(dbl), C.c_char_p, i32]
    L.mm_op_set_option.argtypes = [vp, C.c_char_p, i32]
    L.mm_op_get_option.argtypes = [vp, C.c_char_p, C.POINTER(i32)]
    L.mm_op_autotune.argtypes = [vp, i32]
    L.mm_op_profile_stage.argtypes = [vp, i32, P(dbl), P(i32), C.c_char_p, i32]
    L.mm_ctx_stream.argtypes = [vp]
    L.mm_ctx_stream.restype = vp
    L.mm_wsnet_create.argtypes = [vp, i64, i64, P(vp)]
    L.mm_wsnet_num_params.argtypes = [vp, P(i64)]
    L.mm_wsnet_set_params.argtypes = [vp, vp, i32]
    L.mm_wsnet_get_params.argtypes = [vp, vp, i32]
    L.mm_wsnet_forward.argtypes = [vp, vp, vp, i64, i32]
    L.mm_wsnet_loss_grad.argtypes = [vp, vp, vp, i64, P(dbl), vp, i32]
    L.mm_wsnet_train_step.argtypes = [vp, vp, vp, i64, C.c_float, P(dbl), i32]
    L.mm_rcvr_normalize_ssq.argtypes = [vp, vp, vp, C.c_float, i32]
    L.mm_rcvr_normalize_ssq_noalloc.argtypes = [vp, vp, vp, C.c_float]
    L.mm_wsnet_destroy.argtypes = [vp]
    _lib = L
    return L


def check(rc):
    if rc != MM_OK:
        msg = lib().mm_last_error().decode(errors="replace")
        if rc == MM_ERR_OOM:
            raise OutOfMemory(rc, msg)
        if rc == MM_ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        if rc == MM_ERR_ARG:
            raise ValueError(msg)
        raise MMError(rc, msg)


def ptr(a):
    """Address of a numpy array or torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


_ctx = {}


def context(device=0, stream=None):
    """The process-wide context of `device` (one HIP stream).  `stream` (a
    hipStream_t as int, e.g. ``torch.cuda.Stream().cuda_stream``) binds the
    context to an existing stream; it must be given before the first use."""
    if device in _ctx:
        if stream is not None and stream != lib().mm_ctx_stream(_ctx[device]):
            raise RuntimeError("the context of this device already runs on another stream")
        return _ctx[device]
    L = lib()
    h = C.c_void_p()
    check(L.mm_ctx_create(device, None if stream is None else C.c_void_p(stream), C.byref(h)))
    _ctx[device] = h
    return h


def context_stream(device=0):
    """hipStream_t (int) the device context runs on."""
    return lib().mm_ctx_stream(context(device))


def device_count():
    return int(lib().mm_device_count())
