"""ctypes wrapper of the LINPACK oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module.  It loads oracle/liblinpack_oracle.so (built by oracle/Makefile) and,
where it was built, oracle/_ref/libhpl_ref_blas.so (the reference's own
LINPACK/src/host/blas.c and gmres.c compiled unchanged).

All matrices use the reference's storage: row-major float32 A with
M(i, j) = A[j, i] the matrix actually solved (M x = b).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblinpack_oracle.so")
_REF = os.path.join(_HERE, "_ref", "libhpl_ref_blas.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_lib = None
_ref = None


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference tree exists)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/LINPACK/src/host"):
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_generate_ref.argtypes = [_f32p, _f32p, _i64, _i64, C.c_uint32, C.c_int,
                                       C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_generate_sized.argtypes = [_f32p, C.c_void_p, C.c_void_p, C.c_void_p, _i64, _i64,
                                         C.c_uint64, C.c_int, C.c_double]
        L.orc_hash_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_hash_uniform.restype = C.c_float
        L.orc_generate_batch_system.argtypes = [_f32p, _f64p, _i64, _i64, C.c_uint64, C.c_uint64]
        L.orc_gefa_nopvt.argtypes = [_f32p, _i64, _i64]
        L.orc_gefa_nopvt_omp.argtypes = [_f32p, _i64, _i64, C.c_int]
        L.orc_gesl_nopvt.argtypes = [_f32p, _f32p, _i64, _i64]
        L.orc_gefa_nopvt_f64.argtypes = [_f64p, _i64, _i64]
        L.orc_generate_ref_f64.argtypes = [_f64p, _f64p, _i64, _i64, C.c_uint32, C.c_int]
        _i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_gefa_pvt.argtypes = [_f32p, _i64, _i64, _i32p]
        L.orc_gesl_pvt.argtypes = [_f32p, _f32p, _i32p, _i64, _i64]
        L.orc_gesl_nopvt_f64.argtypes = [_f64p, _f64p, _i64, _i64]
        for fn in (L.orc_top_update, L.orc_left_update):
            fn.argtypes = [C.c_void_p, _i64, C.c_void_p, _i64, _i64]
        L.orc_inner_update.argtypes = [C.c_void_p, _i64, C.c_void_p, _i64, C.c_void_p, _i64, _i64]
        L.orc_gefa_blocked.argtypes = [_f32p, _i64, _i64, _i64]
        L.orc_gefa_blocked_lp.argtypes = [_f32p, _i64, _i64, _i64, C.c_int]
        L.orc_round_lp.argtypes = [C.c_double, C.c_int]
        L.orc_round_lp.restype = C.c_double
        L.orc_fp8_scale.argtypes = [C.c_double]
        L.orc_fp8_scale.restype = C.c_double
        L.orc_residn_ref.argtypes = [_f32p, _i64, C.c_float, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.orc_norms.argtypes = [_f32p, _i64, _i64, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's blas.c/gmres.c built into oracle/_ref, or None."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        R = C.CDLL(_REF)
        R.dtrsm_ref.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, C.c_int, C.c_int,
                                C.c_double, _f64p, C.c_int, _f64p, C.c_int]
        R.dgemv_ref.argtypes = [C.c_char, C.c_int, C.c_int, C.c_double, _f64p, C.c_int, _f64p,
                                C.c_int, C.c_double, _f64p, C.c_int]
        R.dlange_ref.argtypes = [C.c_char, C.c_int, C.c_int, _f64p, C.c_int]
        R.dlange_ref.restype = C.c_double
        R.gmres_ref.argtypes = [C.c_int, _f64p, C.c_int, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                                C.c_int, C.c_double]
        _ref = R
    return _ref


# --------------------------------------------------------------------------- generators
def generate_ref(n: int, seed: int = 0, diag_dominant: bool = True):
    """Reference generator (linpack_benchmark.hpp:211-297), one rank.

    Returns (A, b, norma, normb)."""
    A = np.empty((n, n), np.float32)
    b = np.empty(n, np.float32)
    na, nb = C.c_float(), C.c_float()
    lib().orc_generate_ref(A, b, n, n, seed, int(diag_dominant), C.byref(na), C.byref(nb))
    return A, b, float(na.value), float(nb.value)


def generate_ref_f64(n: int, seed: int = 0, diag_dominant: bool = True):
    """Reference generator of the DATA_TYPE=double build. Returns (A, b)."""
    A = np.empty((n, n), np.float64)
    b = np.empty(n, np.float64)
    lib().orc_generate_ref_f64(A, b, n, n, seed, int(diag_dominant))
    return A, b


def generate_sized(n: int, seed: int = 42, mode: int = 0, d: float = 1.0):
    """Counter-based sized generator. Returns (A, b32, b64, xtrue)."""
    A = np.empty((n, n), np.float32)
    b = np.empty(n, np.float32)
    b64 = np.empty(n, np.float64)
    xt = np.empty(n, np.float64)
    lib().orc_generate_sized(A, b.ctypes.data, b64.ctypes.data, xt.ctypes.data, n, n, seed,
                             mode, d)
    return A, b, b64, xt


def generate_batch_system(n: int, seed: int, s: int):
    """System s of a batch (padded to a multiple of 32 with an identity).

    Returns (A_pad, b64_pad, npad)."""
    npad = (n + 31) // 32 * 32
    A = np.empty((npad, npad), np.float32)
    b = np.empty(npad, np.float64)
    lib().orc_generate_batch_system(A, b, n, npad, seed, s)
    return A, b, npad


def batch_orders(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Orders n_i log-uniform in [lo, hi] from the counter-based generator
    (stream 1 of the seed): n = round(exp(log lo + u * (log hi - log lo)))."""
    u = np.array([hash_uniform(seed, 1, i) for i in range(count)], np.float64)
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int32)


def hash_uniform(seed: int, stream: int, k: int) -> float:
    return float(lib().orc_hash_uniform(seed, stream, k))


# --------------------------------------------------------------------------- solvers
def gefa_nopvt(A: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(A, np.float32).copy()
    lib().orc_gefa_nopvt(a, a.shape[0], a.shape[1])
    return a


def gefa_nopvt_omp(A: np.ndarray, threads: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(A, np.float32).copy()
    lib().orc_gefa_nopvt_omp(a, a.shape[0], a.shape[1], threads)
    return a


def gesl_nopvt(LU: np.ndarray, b: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(b, np.float32).copy()
    lu = np.ascontiguousarray(LU, np.float32)
    lib().orc_gesl_nopvt(lu, x, lu.shape[0], lu.shape[1])
    return x


def gefa_pvt(A: np.ndarray):
    """gefa_ref (partial pivoting, LINPACK/src/host/linpack_data.cpp:72-119).

    Returns (F, ipvt): F in gefa_ref's layout (negated multipliers below M's
    diagonal in the row order of their own step, u_kk on the diagonal)."""
    a = np.ascontiguousarray(A, np.float32).copy()
    ipvt = np.empty(a.shape[0], np.int32)
    lib().orc_gefa_pvt(a, a.shape[0], a.shape[1], ipvt)
    return a, ipvt


def gesl_pvt(F: np.ndarray, b: np.ndarray, ipvt: np.ndarray) -> np.ndarray:
    """gesl_ref (linpack_data.cpp:121-156) with a factor of gefa_pvt."""
    x = np.ascontiguousarray(b, np.float32).copy()
    f = np.ascontiguousarray(F, np.float32)
    lib().orc_gesl_pvt(f, x, np.ascontiguousarray(ipvt, np.int32), f.shape[0], f.shape[1])
    return x


def pvt_to_lapack(F: np.ndarray, ipvt: np.ndarray) -> np.ndarray:
    """The multipliers of a gefa_pvt factor with every later row exchange
    applied (LAPACK getrf's convention: L of P M = L U).  Pure data movement:
    exchange j of step j moves M rows j and ipvt[j] in the columns k < j."""
    a = np.array(F, np.float32, copy=True)
    for j, p in enumerate(np.asarray(ipvt)):
        if p != j:
            a[:j, [j, p]] = a[:j, [p, j]]
    return a


def lapack_to_pvt(F: np.ndarray, ipvt: np.ndarray) -> np.ndarray:
    """Inverse of pvt_to_lapack (the exchanges undone in reverse order)."""
    a = np.array(F, np.float32, copy=True)
    piv = np.asarray(ipvt)
    for j in range(len(piv) - 1, -1, -1):
        p = piv[j]
        if p != j:
            a[:j, [j, p]] = a[:j, [p, j]]
    return a


def pivot_permutation(ipvt: np.ndarray) -> np.ndarray:
    """perm with (P M)(i, :) = M(perm[i], :) for the exchanges of ipvt applied
    in order."""
    perm = np.arange(len(ipvt))
    for j, p in enumerate(np.asarray(ipvt)):
        perm[[j, p]] = perm[[p, j]]
    return perm


def gefa_nopvt_f64(A: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(A, np.float64).copy()
    lib().orc_gefa_nopvt_f64(a, a.shape[0], a.shape[1])
    return a


def gesl_nopvt_f64(LU: np.ndarray, b: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(b, np.float64).copy()
    lu = np.ascontiguousarray(LU, np.float64)
    lib().orc_gesl_nopvt_f64(lu, x, lu.shape[0], lu.shape[1])
    return x


def top_update(X: np.ndarray, LU: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(X, np.float32).copy()
    lu = np.ascontiguousarray(LU, np.float32)
    lib().orc_top_update(x.ctypes.data, x.shape[1], lu.ctypes.data, lu.shape[1], x.shape[0])
    return x


def left_update(X: np.ndarray, LU: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(X, np.float32).copy()
    lu = np.ascontiguousarray(LU, np.float32)
    lib().orc_left_update(x.ctypes.data, x.shape[1], lu.ctypes.data, lu.shape[1], x.shape[0])
    return x


def inner_update(Cm: np.ndarray, L: np.ndarray, T: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(Cm, np.float32).copy()
    l_ = np.ascontiguousarray(L, np.float32)
    t = np.ascontiguousarray(T, np.float32)
    lib().orc_inner_update(c.ctypes.data, c.shape[1], l_.ctypes.data, l_.shape[1], t.ctypes.data,
                           t.shape[1], c.shape[0])
    return c


def gefa_blocked(A: np.ndarray, B: int) -> np.ndarray:
    a = np.ascontiguousarray(A, np.float32).copy()
    lib().orc_gefa_blocked(a, a.shape[0], a.shape[1], B)
    return a


LP_PRECISIONS = {"bf16": 1, "fp16": 2, "fp8": 3}


def gefa_blocked_lp(A: np.ndarray, B: int, prec: str) -> np.ndarray:
    """Blocked factorisation whose trailing update uses BF16/FP16 copies or
    FP8 E4M3 copies with per-(32 vectors x B) power-of-two scales."""
    a = np.ascontiguousarray(A, np.float32).copy()
    lib().orc_gefa_blocked_lp(a, a.shape[0], a.shape[1], B, LP_PRECISIONS[prec])
    return a


def round_lp(x: float, prec: str) -> float:
    return float(lib().orc_round_lp(float(x), LP_PRECISIONS[prec]))


def fp8_scale(amax: float) -> float:
    return float(lib().orc_fp8_scale(float(amax)))


# --------------------------------------------------------------------------- metrics
def residn_ref(x: np.ndarray, normb: float):
    """Reference metric (linpack_benchmark.hpp:400-417): (resid, residn)."""
    r, rn = C.c_double(), C.c_double()
    xx = np.ascontiguousarray(x, np.float32)
    lib().orc_residn_ref(xx, xx.shape[0], normb, C.byref(r), C.byref(rn))
    return float(r.value), float(rn.value)


def norms(A: np.ndarray, x: np.ndarray, b: np.ndarray):
    """(||b - Mx||inf, ||M||inf, ||x||inf, ||b||inf) in double, M = A^T."""
    out = np.zeros(4, np.float64)
    a = np.ascontiguousarray(A, np.float32)
    lib().orc_norms(a, a.shape[0], a.shape[1], np.ascontiguousarray(x, np.float64),
                    np.ascontiguousarray(b, np.float64), out)
    return tuple(float(v) for v in out)


def hpl_scaled_residual(A, x, b, eps=float(np.finfo(np.float32).eps)):
    """HPL check ||Mx-b||inf / (eps ||M||inf ||x||inf n)."""
    rn, an, xn, _ = norms(A, x, b)
    return rn / (eps * an * xn * A.shape[0])


def backward_error(A, x, b):
    """Normwise backward error ||b-Mx||inf / (||M||inf ||x||inf + ||b||inf)."""
    rn, an, xn, bn = norms(A, x, b)
    return rn / (an * xn + bn)


def decode_factor(LU: np.ndarray):
    """Decode the dgefa encoding of M = A^T into LAPACK form (L unit lower, U).

    Returns (L, U) as float64 n x n arrays in M's orientation."""
    M = np.asarray(LU, np.float64).T
    n = M.shape[0]
    L = -np.tril(M, -1) + np.eye(n)
    U = np.triu(M, 1) + np.diag(-1.0 / np.diag(M))
    return L, U
