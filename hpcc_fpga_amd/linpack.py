"""Python binding of libhpl_hip.so (ctypes over the C ABI in include/hpl_hip.h).

Mirrors the reference's LINPACK execution interface
(LINPACK/src/host/execution_types/execution_pcie.hpp:53-728 and
LINPACK/src/host/linpack_data.cpp:166-229): the matrix is the row-major float32
array ``a`` whose transpose M = a.T is factorised without pivoting; after GEFA
``a`` holds the dgefa encoding (negated multipliers below the diagonal of M,
u above, -1/u_kk on the diagonal) and GESL solves M x = b.

There is no CPU fallback: importing works without a GPU, but every compute
call raises HplError when the HIP library or device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhpl_hip.so")

HPL_SUCCESS = 0
STATUS_NAMES = {
    0: "HPL_SUCCESS", 1: "HPL_ERR_INVALID_ARG", 2: "HPL_ERR_HIP", 3: "HPL_ERR_OUT_OF_MEMORY",
    4: "HPL_ERR_ZERO_PIVOT", 5: "HPL_ERR_BUDGET", 6: "HPL_ERR_STATE",
    7: "HPL_ERR_NOT_CONVERGED", 8: "HPL_ERR_UNSUPPORTED", 9: "HPL_ERR_NO_DEVICE",
}
PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}


class HplError(RuntimeError):
    """A non-zero hpl_status (the C++ host maps it to an exception too)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class _IRResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("backward_error", C.c_double),
                ("residual_inf", C.c_double), ("norm_a", C.c_double), ("norm_x", C.c_double),
                ("norm_b", C.c_double), ("gefa_s", C.c_double), ("total_s", C.c_double)]


class _BatchStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("flops", C.c_double), ("gflops", C.c_double),
                ("max_iterations", C.c_int), ("max_backward_error_per_n", C.c_double),
                ("failed", C.c_int), ("zero_pivots", C.c_int)]


class _ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("hbm_budget", C.c_uint64),
                ("prec", C.c_int), ("max_iter", C.c_int), ("tol_per_n", C.c_double),
                ("max_n", C.c_uint64), ("max_nrhs", C.c_int)]


class _Request(C.Structure):
    _fields_ = [("n", C.c_uint64), ("nrhs", C.c_int), ("a", C.c_void_p), ("lda", C.c_uint64),
                ("seed", C.c_uint64), ("factor", C.c_uint64), ("keep", C.c_int),
                ("b", C.c_void_p), ("ldb", C.c_uint64), ("x", C.c_void_p), ("ldx", C.c_uint64)]


class _Response(C.Structure):
    _fields_ = [("status", C.c_int), ("factor", C.c_uint64), ("reused", C.c_int),
                ("iterations", C.c_int), ("backward_error", C.c_double), ("queue_s", C.c_double),
                ("service_s", C.c_double), ("latency_s", C.c_double), ("block", C.c_uint32),
                ("npad", C.c_uint64)]


class _ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_int), ("in_flight", C.c_int), ("max_in_flight", C.c_int),
                ("resident_factors", C.c_uint64), ("factor_bytes_in_use", C.c_uint64),
                ("peak_factor_bytes", C.c_uint64), ("factor_budget_bytes", C.c_uint64),
                ("workspace_bytes", C.c_uint64)]


class _Timing(C.Structure):
    _fields_ = [("gefa_s", C.c_double), ("gesl_s", C.c_double), ("h2d_s", C.c_double),
                ("d2h_s", C.c_double)]


@dataclass
class Timing:
    gefa_s: float = 0.0
    gesl_s: float = 0.0
    h2d_s: float = 0.0
    d2h_s: float = 0.0


_lib = None


def load_library():
    """Load libhpl_hip.so; raise HplError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HplError(2, f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u64, u32, i32, dbl = C.c_uint64, C.c_uint32, C.c_int, C.c_double
    T = C.POINTER(_Timing)
    sigs = {
        "hpl_create": ([i32, C.POINTER(P)], i32),
        "hpl_destroy": ([P], None),
        "hpl_last_error": ([P], C.c_char_p),
        "hpl_status_string": ([i32], C.c_char_p),
        "hpl_device_name": ([P, C.c_char_p, C.c_size_t], i32),
        "hpl_version": ([], C.c_char_p),
        "hpl_gefa": ([P, P, u64, u64, u32, T], i32),
        "hpl_gesl": ([P, P, P, u64, u64, T], i32),
        "hpl_upload": ([P, P, u64, u64, P, T], i32),
        "hpl_generate": ([P, u64, i32, u64, dbl], i32),
        "hpl_keep_original": ([P, i32], i32),
        "hpl_gefa_dev": ([P, u32, i32, T], i32),
        "hpl_gesl_dev": ([P, T], i32),
        "hpl_restore": ([P], i32),
        "hpl_download": ([P, P, u64, P, T], i32),
        "hpl_residual": ([P, C.POINTER(C.c_double)], i32),
        "hpl_download_vectors": ([P, P, P, P], i32),
        "hpl_solve_ir": ([P, u32, i32, i32, dbl, C.POINTER(_IRResult)], i32),
        "hpl_download_solution64": ([P, P], i32),
        "hpl_upload_rhs64": ([P, P], i32),
        "hpl_solve_rhs64": ([P, i32, P, u64, P, u64, i32, dbl, P], i32),
        "hpl_solve_gmres": ([P, u32, i32, i32, i32, dbl, C.POINTER(_IRResult)], i32),
        "hpl_serve_config_default": ([C.POINTER(_ServeConfig)], None),
        "hpl_serve_create": ([C.POINTER(_ServeConfig), C.POINTER(P)], i32),
        "hpl_serve_destroy": ([P], None),
        "hpl_serve_last_error": ([P], C.c_char_p),
        "hpl_serve_submit": ([P, C.POINTER(_Request), C.POINTER(u64)], i32),
        "hpl_serve_wait": ([P, u64, C.POINTER(_Response)], i32),
        "hpl_serve_release": ([P, u64], i32),
        "hpl_serve_stats_get": ([P, C.POINTER(_ServeStats)], i32),
        "hpl_batch_create": ([P, P, i32, C.POINTER(P)], i32),
        "hpl_batch_destroy": ([P], None),
        "hpl_batch_last_error": ([P], C.c_char_p),
        "hpl_batch_generate": ([P, u64], i32),
        "hpl_batch_set_system": ([P, i32, P, u64, P], i32),
        "hpl_batch_autotune": ([P, i32, i32, dbl, P], i32),
        "hpl_batch_set_blocked": ([P, i32], i32),
        "hpl_batch_solve": ([P, i32, i32, dbl, C.POINTER(_BatchStats)], i32),
        "hpl_batch_get": ([P, i32, P, C.POINTER(i32), C.POINTER(dbl)], i32),
        "hpl_batch_get_system": ([P, i32, P, P], i32),
        "hpl_gefa_pvt": ([P, P, u64, u64, u32, P, T], i32),
        "hpl_gesl_pvt": ([P, P, P, P, u64, u64, T], i32),
        "hpl_set_pivoting": ([P, i32], i32),
        "hpl_get_pivots": ([P, P], i32),
        "hpl_gefa_f64": ([P, P, u64, u64, T], i32),
        "hpl_gesl_f64": ([P, P, P, u64, u64, T], i32),
        "hpl_upload_f64": ([P, P, u64, u64, P, T], i32),
        "hpl_generate_f64": ([P, u64, i32, u64, dbl], i32),
        "hpl_gefa_dev_f64": ([P, T], i32),
        "hpl_gesl_dev_f64": ([P, T], i32),
        "hpl_restore_f64": ([P], i32),
        "hpl_download_f64": ([P, P, u64, P, T], i32),
        "hpl_residual_f64": ([P, C.POINTER(C.c_double)], i32),
        "hpl_set_profiling": ([P, i32], i32),
        "hpl_kernel_stats": ([P, C.POINTER(u64), C.POINTER(dbl), C.POINTER(dbl)], i32),
        "hpl_kernel_busy": ([P, C.POINTER(dbl)], i32),
        "hpl_kernel_bytes": ([P, C.POINTER(dbl)], i32),
        "hpl_gesl_noalloc": ([P, u64, u64, P, P], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _f64(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a


def _f32(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float32 array")
    return a


class Context:
    """One HIP device context holding one resident linear system."""

    def __init__(self, device: int = 0):
        self._L = load_library()
        self._c = C.c_void_p()
        st = self._L.hpl_create(device, C.byref(self._c))
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_status_string(st).decode())
        self.n = 0
        self.n64 = 0

    # -- plumbing
    def _check(self, st: int):
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_last_error(self._c).decode())

    def close(self):
        if self._c:
            self._L.hpl_destroy(self._c)
            self._c = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        self._check(self._L.hpl_device_name(self._c, buf, 256))
        return buf.value.decode()

    # -- host-buffer path (drop-in for execution::pcie::calculate)
    def gefa(self, a: np.ndarray, block: int) -> Timing:
        """Factorise ``a`` (n x n row-major float32) in place."""
        a = _f32(a)
        t = _Timing()
        self._check(self._L.hpl_gefa(self._c, a.ctypes.data, a.shape[0], a.shape[1], block,
                                     C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def gesl(self, lu: np.ndarray, b: np.ndarray) -> Timing:
        """Solve M x = b with a GEFA factor; b is overwritten with x."""
        lu, b = _f32(lu), _f32(b)
        t = _Timing()
        self._check(self._L.hpl_gesl(self._c, lu.ctypes.data, b.ctypes.data, lu.shape[0],
                                     lu.shape[1], C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    # -- partial pivoting (gefa_ref / gesl_ref)
    def gefa_pvt(self, a: np.ndarray, block: int):
        """Factorise ``a`` in place with partial pivoting (gefa_ref's layout);
        returns (ipvt, Timing)."""
        a = _f32(a)
        ipvt = np.empty(a.shape[0], np.int32)
        t = _Timing()
        self._check(self._L.hpl_gefa_pvt(self._c, a.ctypes.data, a.shape[0], a.shape[1], block,
                                         ipvt.ctypes.data, C.byref(t)))
        return ipvt, Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def gesl_pvt(self, lu: np.ndarray, b: np.ndarray, ipvt: np.ndarray) -> Timing:
        """Solve M x = b with a gefa_pvt factor; b is overwritten with x."""
        lu, b = _f32(lu), _f32(b)
        ipvt = np.ascontiguousarray(ipvt, np.int32)
        t = _Timing()
        self._check(self._L.hpl_gesl_pvt(self._c, lu.ctypes.data, b.ctypes.data,
                                         ipvt.ctypes.data, lu.shape[0], lu.shape[1],
                                         C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def set_pivoting(self, enable: bool = True):
        """Device-resident factorisations with (True) or without partial pivoting."""
        self._check(self._L.hpl_set_pivoting(self._c, int(enable)))

    def pivots(self) -> np.ndarray:
        ipvt = np.empty(self.n, np.int32)
        self._check(self._L.hpl_get_pivots(self._c, ipvt.ctypes.data))
        return ipvt

    # -- FP64 (the reference's DATA_TYPE=double build)
    def gefa_f64(self, a: np.ndarray) -> Timing:
        """Factorise ``a`` (n x n row-major float64) in place."""
        a = _f64(a)
        t = _Timing()
        self._check(self._L.hpl_gefa_f64(self._c, a.ctypes.data, a.shape[0], a.shape[1],
                                         C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def gesl_f64(self, lu: np.ndarray, b: np.ndarray) -> Timing:
        lu, b = _f64(lu), _f64(b)
        t = _Timing()
        self._check(self._L.hpl_gesl_f64(self._c, lu.ctypes.data, b.ctypes.data, lu.shape[0],
                                         lu.shape[1], C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def upload_f64(self, a: np.ndarray, b: np.ndarray | None = None) -> Timing:
        a = _f64(a)
        t = _Timing()
        bp = _f64(b).ctypes.data if b is not None else None
        self._check(self._L.hpl_upload_f64(self._c, a.ctypes.data, a.shape[0], a.shape[1], bp,
                                           C.byref(t)))
        self.n64 = a.shape[0]
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def generate_f64(self, n: int, mode: int = 0, seed: int = 42, d: float = 1.0):
        self._check(self._L.hpl_generate_f64(self._c, n, mode, seed, d))
        self.n, self.n64 = n, n

    def gefa_dev_f64(self) -> Timing:
        t = _Timing()
        self._check(self._L.hpl_gefa_dev_f64(self._c, C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def gesl_dev_f64(self) -> Timing:
        t = _Timing()
        self._check(self._L.hpl_gesl_dev_f64(self._c, C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def restore_f64(self):
        self._check(self._L.hpl_restore_f64(self._c))

    def download_f64(self, matrix: bool = True, solution: bool = True):
        n = self.n64
        a = np.empty((n, n), np.float64) if matrix else None
        x = np.empty(n, np.float64) if solution else None
        t = _Timing()
        self._check(self._L.hpl_download_f64(self._c, a.ctypes.data if matrix else None, n,
                                             x.ctypes.data if solution else None, C.byref(t)))
        return a, x

    def residual_f64(self):
        out = (C.c_double * 4)()
        self._check(self._L.hpl_residual_f64(self._c, out))
        return tuple(out)

    # -- device-resident path
    def keep_original(self, keep: bool = True):
        self._check(self._L.hpl_keep_original(self._c, int(keep)))

    def upload(self, a: np.ndarray, b: np.ndarray | None = None) -> Timing:
        a = _f32(a)
        t = _Timing()
        bp = _f32(b).ctypes.data if b is not None else None
        self._check(self._L.hpl_upload(self._c, a.ctypes.data, a.shape[0], a.shape[1], bp,
                                       C.byref(t)))
        self.n = a.shape[0]
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def generate(self, n: int, mode: int = 0, seed: int = 42, d: float = 1.0):
        self._check(self._L.hpl_generate(self._c, n, mode, seed, d))
        self.n = n

    def gefa_dev(self, block: int, prec: str = "fp32") -> Timing:
        t = _Timing()
        self._check(self._L.hpl_gefa_dev(self._c, block, PRECISIONS[prec], C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def gesl_dev(self) -> Timing:
        t = _Timing()
        self._check(self._L.hpl_gesl_dev(self._c, C.byref(t)))
        return Timing(t.gefa_s, t.gesl_s, t.h2d_s, t.d2h_s)

    def restore(self):
        self._check(self._L.hpl_restore(self._c))

    def download(self, matrix: bool = True, solution: bool = True):
        n = self.n
        a = np.empty((n, n), np.float32) if matrix else None
        x = np.empty(n, np.float32) if solution else None
        t = _Timing()
        self._check(self._L.hpl_download(self._c, a.ctypes.data if matrix else None, n,
                                         x.ctypes.data if solution else None, C.byref(t)))
        return a, x

    def download_vectors(self):
        """(b float32, b float64, x_true float64) of the resident system."""
        n = self.n
        b = np.empty(n, np.float32)
        b64 = np.empty(n, np.float64)
        xt = np.empty(n, np.float64)
        self._check(self._L.hpl_download_vectors(self._c, b.ctypes.data, b64.ctypes.data,
                                                 xt.ctypes.data))
        return b, b64, xt

    def residual(self):
        """(||b - Mx||inf, ||M||inf, ||x||inf, ||b||inf) in FP64 on the device."""
        out = (C.c_double * 4)()
        self._check(self._L.hpl_residual(self._c, out))
        return tuple(out)

    def solve_ir(self, block: int, prec: str = "bf16", max_iter: int = 50, tol: float = None,
                 check: bool = True) -> dict:
        """Mixed-precision factorisation + FP64 classic refinement (see hpl_solve_ir)."""
        if tol is None:
            tol = 1e-14 * self.n
        res = _IRResult()
        st = self._L.hpl_solve_ir(self._c, block, PRECISIONS[prec], max_iter, tol, C.byref(res))
        if st != HPL_SUCCESS and (check or st != 7):
            self._check(st)
        return {f: getattr(res, f) for f, _ in _IRResult._fields_}

    def solve_gmres(self, block: int, prec: str = "bf16", restart: int = 20, max_iter: int = 10,
                    tol: float = None, check: bool = True) -> dict:
        """Mixed-precision factorisation + GMRES-based FP64 refinement (see
        hpl_solve_gmres); ``iterations`` counts Arnoldi steps."""
        if tol is None:
            tol = 1e-14 * self.n
        res = _IRResult()
        st = self._L.hpl_solve_gmres(self._c, block, PRECISIONS[prec], restart, max_iter, tol,
                                     C.byref(res))
        if st != HPL_SUCCESS and (check or st != 7):
            self._check(st)
        return {f: getattr(res, f) for f, _ in _IRResult._fields_}

    def download_solution64(self) -> np.ndarray:
        x = np.empty(self.n, np.float64)
        self._check(self._L.hpl_download_solution64(self._c, x.ctypes.data))
        return x

    def upload_rhs64(self, b: np.ndarray):
        b = np.ascontiguousarray(b, np.float64)
        self._check(self._L.hpl_upload_rhs64(self._c, b.ctypes.data))

    def solve_rhs64(self, b: np.ndarray, max_iter: int = 50, tol: float = None,
                    check: bool = True):
        """Solve and refine further right-hand sides (n or n x nrhs, FP64) with
        the resident factor; returns (x, [per-column result dicts])."""
        b2 = np.asarray(b, np.float64)
        one = b2.ndim == 1
        cols = np.ascontiguousarray((b2[:, None] if one else b2).T)  # nrhs x n, column-major
        nrhs = cols.shape[0]
        x = np.empty_like(cols)
        if tol is None:
            tol = 1e-14 * self.n
        res = (_IRResult * nrhs)()
        st = self._L.hpl_solve_rhs64(self._c, nrhs, cols.ctypes.data, self.n, x.ctypes.data,
                                     self.n, max_iter, tol, res)
        if st != HPL_SUCCESS and (check or st != 7):
            self._check(st)
        out = [{f: getattr(r, f) for f, _ in _IRResult._fields_} for r in res]
        x = x.T
        return (x[:, 0].copy() if one else np.ascontiguousarray(x)), out

    def set_profiling(self, enable: bool = True):
        self._check(self._L.hpl_set_profiling(self._c, int(enable)))

    def kernel_stats(self):
        """(launches, seconds, flops) of the trailing-update GEMM since profiling on."""
        n, s, f = C.c_uint64(), C.c_double(), C.c_double()
        self._check(self._L.hpl_kernel_stats(self._c, C.byref(n), C.byref(s), C.byref(f)))
        return int(n.value), float(s.value), float(f.value)

    def kernel_bytes(self) -> float:
        """Algorithmic HBM bytes of the profiled trailing-update launches."""
        b = C.c_double()
        self._check(self._L.hpl_kernel_bytes(self._c, C.byref(b)))
        return float(b.value)

    def kernel_busy(self) -> float:
        """Seconds during which at least one profiled trailing-update launch ran."""
        s = C.c_double()
        self._check(self._L.hpl_kernel_busy(self._c, C.byref(s)))
        return float(s.value)


class Batch:
    """Many independent systems of uneven order solved by one graph replay."""

    def __init__(self, ctx: Context, orders):
        self._L = ctx._L
        self.ctx = ctx  # keeps the device context alive
        self.n = np.ascontiguousarray(orders, np.int32)
        self._b = C.c_void_p()
        st = self._L.hpl_batch_create(ctx._c, self.n.ctypes.data, len(self.n), C.byref(self._b))
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_status_string(st).decode())

    def _check(self, st):
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_batch_last_error(self._b).decode())

    def close(self):
        if self._b:
            self._L.hpl_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate(self, seed: int = 7):
        self._check(self._L.hpl_batch_generate(self._b, seed))

    def set_system(self, s: int, a: np.ndarray, b: np.ndarray):
        a = _f32(np.ascontiguousarray(a, np.float32))
        b = np.ascontiguousarray(b, np.float64)
        self._check(self._L.hpl_batch_set_system(self._b, s, a.ctypes.data, a.shape[1],
                                                 b.ctypes.data))

    def autotune(self, prec: str = "fp16", max_iter: int = 20, tol_per_n: float = 1e-14):
        th = np.zeros(4, np.int32)
        self._check(self._L.hpl_batch_autotune(self._b, PRECISIONS[prec], max_iter, tol_per_n,
                                               th.ctypes.data))
        return th

    def set_blocked(self, blocked: bool):
        """Factorise every bucket blocked (one launch per 32-wide step over
        the bucket) or one workgroup per system, until the next autotune."""
        self._check(self._L.hpl_batch_set_blocked(self._b, int(bool(blocked))))

    def solve(self, prec: str = "fp16", max_iter: int = 20, tol_per_n: float = 1e-14) -> dict:
        st = _BatchStats()
        self._check(self._L.hpl_batch_solve(self._b, PRECISIONS[prec], max_iter, tol_per_n,
                                            C.byref(st)))
        return {f: getattr(st, f) for f, _ in _BatchStats._fields_}

    def get(self, s: int):
        x = np.empty(int(self.n[s]), np.float64)
        it, be = C.c_int(), C.c_double()
        self._check(self._L.hpl_batch_get(self._b, s, x.ctypes.data, C.byref(it), C.byref(be)))
        return x, it.value, be.value

    def get_system(self, s: int):
        npad = (int(self.n[s]) + 31) // 32 * 32
        a = np.empty((npad, npad), np.float32)
        b = np.empty(npad, np.float64)
        self._check(self._L.hpl_batch_get_system(self._b, s, a.ctypes.data, b.ctypes.data))
        return a, b


class Ticket:
    """A submitted request; holds its host buffers until it is waited for."""

    def __init__(self, server, ticket, n, nrhs, a, b, x):
        self.server, self.ticket, self.n, self.nrhs = server, ticket, n, nrhs
        self._keep = (a, b)
        self.x = x  # nrhs x n (column j of the solution is row j)
        self.done = False


class Server:
    """Request loop over one device (see hpl_serve_* in include/hpl_hip.h).

    submit() returns at once with a Ticket; wait(ticket) blocks until that
    request is solved and returns (x, response).  Both release the GIL inside
    the library, so one thread can submit while another waits."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 200 * 10 ** 9,
                 prec: str = "fp8", max_iter: int = 30, tol_per_n: float = 1e-14,
                 max_n: int = 8192, max_nrhs: int = 16):
        self._L = load_library()
        cfg = _ServeConfig()
        self._L.hpl_serve_config_default(C.byref(cfg))
        cfg.device, cfg.slots, cfg.hbm_budget = device, slots, int(hbm_budget)
        cfg.prec, cfg.max_iter, cfg.tol_per_n = PRECISIONS[prec], max_iter, tol_per_n
        cfg.max_n, cfg.max_nrhs = max_n, max_nrhs
        self._s = C.c_void_p()
        st = self._L.hpl_serve_create(C.byref(cfg), C.byref(self._s))
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_serve_last_error(None).decode())

    def _check(self, st):
        if st != HPL_SUCCESS:
            raise HplError(st, self._L.hpl_serve_last_error(self._s).decode())

    def submit(self, b: np.ndarray, a: np.ndarray | None = None, seed: int | None = None,
               factor: int = 0, keep: bool = False) -> Ticket:
        """b: n or n x nrhs FP64.  The system is `a` (row-major float32, M =
        a.T), or generated on the device from `seed`, or the resident
        `factor`."""
        b2 = np.asarray(b, np.float64)
        cols = np.ascontiguousarray((b2[:, None] if b2.ndim == 1 else b2).T)
        nrhs, n = cols.shape
        x = np.empty_like(cols)
        r = _Request()
        r.n, r.nrhs = n, nrhs
        if a is not None:
            a = _f32(np.ascontiguousarray(a, np.float32))
            r.a, r.lda = a.ctypes.data, a.shape[1]
        r.seed = 0 if seed is None else int(seed)
        r.factor, r.keep = int(factor), int(keep)
        r.b, r.ldb, r.x, r.ldx = cols.ctypes.data, n, x.ctypes.data, n
        t = C.c_uint64()
        self._check(self._L.hpl_serve_submit(self._s, C.byref(r), C.byref(t)))
        return Ticket(self, t.value, n, nrhs, a, cols, x)

    def wait(self, ticket: Ticket, check: bool = True):
        """(x as n or n x nrhs, response dict)."""
        resp = _Response()
        st = self._L.hpl_serve_wait(self._s, ticket.ticket, C.byref(resp))
        ticket.done = True
        out = {f: getattr(resp, f) for f, _ in _Response._fields_}
        if st != HPL_SUCCESS:
            out["error"] = self._L.hpl_serve_last_error(self._s).decode()
            if check:
                raise HplError(st, out["error"])
        x = ticket.x.T
        return (x[:, 0].copy() if ticket.nrhs == 1 else np.ascontiguousarray(x)), out

    def release(self, factor: int):
        self._check(self._L.hpl_serve_release(self._s, int(factor)))

    def stats(self) -> dict:
        s = _ServeStats()
        self._check(self._L.hpl_serve_stats_get(self._s, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ServeStats._fields_}

    def close(self):
        if self._s:
            self._L.hpl_serve_destroy(self._s)
            self._s = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gefa(a: np.ndarray, block: int, device: int = 0) -> np.ndarray:
    """Return the dgefa-encoded factor of ``a`` (a copy), computed on the GPU."""
    out = np.ascontiguousarray(a, np.float32).copy()
    with Context(device) as ctx:
        ctx.gefa(out, block)
    return out


def gesl(lu: np.ndarray, b: np.ndarray, device: int = 0) -> np.ndarray:
    """Return x with M x = b for a dgefa-encoded factor ``lu``, on the GPU."""
    x = np.ascontiguousarray(b, np.float32).copy()
    with Context(device) as ctx:
        ctx.gesl(np.ascontiguousarray(lu, np.float32), x)
    return x
