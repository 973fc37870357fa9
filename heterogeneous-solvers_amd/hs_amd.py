"""ctypes binding of libhs_amd.so (the C ABI in include/hs_abi.h).

This is the Python side used by the tests, the benchmark and the adjoint
check harness. It loads the in-tree library built by build.py and fails
loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libhs_amd.so")

HS_OK = 0
HS_ERR_HIP = -1000
HS_ERR_NOMEM = -1001
HS_ERR_UNSUPPORTED = -1002
HS_ERR_NOT_CONVERGED = -1003
HS_ERR_WORKSPACE = -1004

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double

# name -> argument types (all return int)
_SIGS = {
    "hs_handle_create": [C.POINTER(_P), _I],
    "hs_handle_destroy": [_P],
    "hs_sync": [_P],
    "hs_workspace_query": [_L, _I, _L, C.POINTER(C.c_size_t)],
    "hs_workspace_reserve": [_P, _L, _I, _L],
    "hs_workspace_size": [_P, C.POINTER(C.c_size_t)],
    "hs_device_count": [C.POINTER(_I)],
    "hs_malloc": [C.POINTER(_P), C.c_size_t],
    "hs_free": [_P],
    "hs_host_malloc": [C.POINTER(_P), C.c_size_t],
    "hs_host_free": [_P],
    "hs_device_memory": [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "hs_memcpy_htod": [_P, _P, _P, C.c_size_t],
    "hs_memcpy_dtoh": [_P, _P, _P, C.c_size_t],
    "hs_memcpy_dtod": [_P, _P, _P, C.c_size_t],
    "hs_memset_zero": [_P, _P, C.c_size_t],
    "hs_event_create": [C.POINTER(_P)],
    "hs_event_destroy": [_P],
    "hs_event_record": [_P, _P],
    "hs_event_elapsed_ms": [_P, _P, C.POINTER(C.c_float)],
    "hs_potrf_diag_f64": [_P, _P, _L, _I, _I, _I],
    "hs_get_info": [_P, C.POINTER(_L)],
    "hs_reset_info": [_P],
    "hs_trsm_panel_f64": [_P, _P, _I, _I, _I, _I, _I],
    "hs_syrk_diag_f64": [_P, _P, _I, _I, _I, _I, _I, _I],
    "hs_gemm_trailing_f64": [_P, _P, _I, _I, _I, _I, _I, _I],
    "hs_syrk_gemm_trailing_f64": [_P, _P, _I, _I, _I],
    "hs_symv_block_f64": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "hs_symv_packed_f64": [_P, _P, _P, _P, _L, _I],
    "hs_symv_rows_f64": [_P, _P, _P, _P, _L, _I, _I],
    "hs_symv_packed_f64_timed": [_P, _P, _P, _P, _L, _I, _I, C.POINTER(_D)],
    "hs_trsv_block_f64": [_P, _P, _P, _L, _I, _I, _I, _I],
    "hs_col_update_f64": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "hs_vec_scale_f64": [_P, _P, _D, _P, _I, _I, _I],
    "hs_vec_add_f64": [_P, _P, _P, _P, _I, _I, _I],
    "hs_vec_sub_f64": [_P, _P, _P, _P, _I, _I, _I],
    "hs_vec_scale_add_f64": [_P, _P, _D, _P, _P, _I, _I, _I],
    "hs_dot_f64": [_P, _P, _P, _I, _I, _I, _P],
    "hs_cholesky_f64": [_P, _P, _L, _I],
    "hs_cholesky_f64_timed": [_P, _P, _L, _I, C.POINTER(_D), C.POINTER(_L)],
    "hs_tri_solve_f64": [_P, _P, _P, _L, _I],
    "hs_tri_prepare_query": [_L, _I, C.POINTER(C.c_size_t)],
    "hs_tri_prepare_f64": [_P, _P, _L, _I, _P],
    "hs_tri_solve_prepared_f64": [_P, _P, _P, _P, _L, _I],
    "hs_cholesky_schedule_f64": [_L, _I, C.POINTER(_L), C.POINTER(_D)],
    "hs_lookahead_dry_run": [_I, _I, _I, _P, _L, C.POINTER(_L)],
    "hs_trail_map_dry_run": [_I, _I, _I, _I, _I, _I, _I, _I, _I, _P, _L, C.POINTER(_L)],
    "hs_cholesky_mixed_schedule": [_L, _I, _I, C.POINTER(_L), C.POINTER(_D)],
    "hs_read_sum_f64": [_P, _P, _L, _P],
    "hs_cg_f64": [_P, _P, _P, _P, _L, _I, _D, _L, C.POINTER(_L), C.POINTER(_D), C.POINTER(_D)],
    "hs_cg_f64_trace": [_P, _P, _P, _P, _L, _I, _D, _L, C.POINTER(_L), C.POINTER(_D), C.POINTER(_D), _P],
    "hs_cg_ref_order_f64_trace": [_P, _P, _P, _P, _L, _I, _D, _L, _I, _I, _I, C.POINTER(_L), C.POINTER(_D),
                                  C.POINTER(_D), _P],
    "hs_cholesky_f64_trace": [_P, _P, _L, _I, _P, _P, _P],
    "hs_cg_f64_noalloc": [_P, _P, _P, _P, _L, _I, _D, _L, C.POINTER(_L), C.POINTER(_D), C.POINTER(_D)],
    "hs_cg_noalloc_limit": [C.POINTER(C.c_size_t)],
    "hs_cg_ref_order_f64": [_P, _P, _P, _P, _L, _I, _D, _L, _I, _I, _I, C.POINTER(_L), C.POINTER(_D),
                            C.POINTER(_D)],
    "hs_symv_ref_order_f64": [_P, _P, _P, _P, _L, _I, _I],
    "hs_dot_ref_order_f64": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P],
    "hs_cholesky_ref_order_f64": [_P, _P, _L, _I, _I],
    "hs_tri_solve_ref_order_f64": [_P, _P, _P, _L, _I, _I],
    "hs_potrf_ref_order_f64": [_P, _P, _L, _I, _I, _I, _I],
    "hs_trsm_ref_order_f64": [_P, _P, _I, _I, _I, _I, _I, _I],
    "hs_syrk_ref_order_f64": [_P, _P, _I, _I, _I, _I, _I, _I, _I],
    "hs_gemm_ref_order_f64": [_P, _P, _I, _I, _I, _I, _I, _I, _I],
    "hs_rbf_assemble_f64": [_P, _P, _L, _I, _I, _D, _D, _D, _P],
    "hs_rbf_window_f64": [_P, _P, _L, _I, _P],
    "hs_to_f32_packed": [_P, _P, _P, _L, _I, _D],
    "hs_cholesky_mixed": [_P, _P, _P, _L, _I, _I, _D],
    "hs_syrk_gemm_trailing_f32": [_P, _P, _I, _I, _I],
    "hs_syrk_gemm_trailing_bf16x3": [_P, _P, _I, _I, _I],
    "hs_syrk_gemm_trailing_f16x3": [_P, _P, _I, _I, _I],
    "hs_cholesky_mixed_timed": [_P, _P, _P, _L, _I, _I, _D, C.POINTER(_D), C.POINTER(_L)],
    "hs_refine_solve": [_P, _P, _P, _P, _P, _L, _I, _D, _I, _I, _P],
    "hs_refine_solve_to": [_P, _P, _P, _P, _P, _L, _I, _D, _D, _I, _I, _P],
    "hs_trsv_packed_f64": [_P, _P, _P, _L, _I, _I],
    "hs_gp_nll_f64": [_P, _P, _L, _I, _I, _D, _D, _D, _P, _P, C.POINTER(_D)],
    "hs_gp_nll_grad_f64": [_P, _P, _L, _I, _I, _D, _D, _D, _P, _P, _P, C.POINTER(_D), C.POINTER(_D)],
    "hs_rbf_cross_f64": [_P, _P, _L, _P, _L, _I, _D, _D, _P],
    "hs_gemv_rows_f64": [_P, _P, _P, _P, _L, _L],
    "hs_gp_predict_f64": [_P, _P, _L, _P, _L, _I, _D, _D, _P, _P],
    "hs_window_features_f64": [_P, _P, _L, _P],
}

HS_MIXED_FP32 = 0
HS_MIXED_BF16X3 = 1
HS_MIXED_F16X3 = 2


class RefineReport(C.Structure):
    """hs_refine_report of include/hs_abi.h"""
    _fields_ = [("ir_iterations", C.c_int64), ("pcg_iterations", C.c_int64), ("used_pcg", C.c_int32),
                ("converged", C.c_int32), ("backward_error", C.c_double), ("relative_residual", C.c_double),
                ("knorm_inf", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

_lib = None


class HsError(RuntimeError):
    def __init__(self, rc, what):
        super().__init__("%s failed with status %d: %s" % (what, rc, last_error()))
        self.rc = rc


def load(path=LIB_PATH):
    """Load libhs_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libhs_amd.so not built at %s: run heterogeneous-solvers_amd/build.py" % path)
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.hs_last_error_string.argtypes = []
    lib.hs_last_error_string.restype = C.c_char_p
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS.keys()) + ["hs_last_error_string"]


def last_error():
    if _lib is None:
        return ""
    s = _lib.hs_last_error_string()
    return s.decode() if s else ""


def check(rc, what="hs call"):
    if rc != HS_OK:
        raise HsError(rc, what)
    return rc


def call(name, *args):
    """Call an hs_* entry and raise on a negative / runtime status.
    Positive returns (factorisation info) are handed back."""
    rc = getattr(load(), name)(*args)
    if rc < 0:
        raise HsError(rc, name)
    return rc


class DeviceBuffer:
    def __init__(self, dev, nbytes):
        self.dev = dev
        self.nbytes = int(nbytes)
        p = _P()
        call("hs_malloc", C.byref(p), max(1, self.nbytes))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            call("hs_free", _P(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def p(self):
        return _P(self.ptr)

    def offset(self, nbytes):
        return _P(self.ptr + int(nbytes))


class Device:
    """One handle (device + stream + workspace). `reserve` lists the systems
    (N, b) or (N, b, M) the handle will solve: the workspace is sized for
    them at creation, since no solver entry allocates."""

    def __init__(self, device=0, reserve=()):
        load()
        h = _P()
        call("hs_handle_create", C.byref(h), int(device))
        self.h = h
        for sizes in reserve:
            self.reserve(*sizes)

    def reserve(self, N, b, M=0):
        call("hs_workspace_reserve", self.h, int(N), int(b), int(M))

    def workspace_bytes(self):
        v = C.c_size_t()
        call("hs_workspace_size", self.h, C.byref(v))
        return v.value

    def close(self):
        if self.h:
            call("hs_handle_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        call("hs_memcpy_htod", self.h, buf.p, arr.ctypes.data_as(_P), arr.nbytes)
        return buf

    def write(self, buf, arr, offset_bytes=0):
        arr = np.ascontiguousarray(arr)
        call("hs_memcpy_htod", self.h, buf.offset(offset_bytes), arr.ctypes.data_as(_P), arr.nbytes)

    def download(self, buf, count, dtype=np.float64, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        call("hs_memcpy_dtoh", self.h, out.ctypes.data_as(_P), buf.offset(offset_bytes), out.nbytes)
        return out

    def zeros(self, count, dtype=np.float64):
        buf = DeviceBuffer(self, count * np.dtype(dtype).itemsize)
        call("hs_memset_zero", self.h, buf.p, buf.nbytes)
        return buf

    def sync(self):
        call("hs_sync", self.h)

    def info(self):
        v = _L()
        call("hs_get_info", self.h, C.byref(v))
        return v.value


class Timer:
    """HIP events on the handle's stream."""

    def __init__(self, dev):
        self.dev = dev
        a, b = _P(), _P()
        call("hs_event_create", C.byref(a))
        call("hs_event_create", C.byref(b))
        self.a, self.b = a, b

    def start(self):
        call("hs_event_record", self.dev.h, self.a)

    def stop(self):
        call("hs_event_record", self.dev.h, self.b)

    def ms(self):
        v = C.c_float()
        call("hs_event_elapsed_ms", self.a, self.b, C.byref(v))
        return v.value

    def close(self):
        call("hs_event_destroy", self.a)
        call("hs_event_destroy", self.b)


def packed_len(N, b):
    n = -(-N // b)
    return n * (n + 1) // 2 * b * b


def block_count(N, b):
    return -(-N // b)
