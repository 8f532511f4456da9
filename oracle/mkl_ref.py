"""The reference's stage-2/3 CPU path through Intel MKL -- CPU BASELINE ONLY.

Only the ``cpu_baseline`` leg of bench.py and tests/ use this module; the
product (lightne_amd) never imports it.

The reference links MKL ILP64 with GNU OpenMP threads
(``-lmkl_intel_ilp64 -lmkl_core -lmkl_gnu_thread``, LightNE/Makefile:22) and
makes these calls, which this module makes in the same order through
``libmkl_rt`` (MKL's single dynamic library, ILP64 interface and GNU threading
selected at load time) with prototypes declared here from MKL's published API:

* randomized SVD (mklredsvd.cc:93-403): ``mkl_sparse_s_create_csr``,
  ``mkl_sparse_set_mm_hint`` / ``set_memory_hint`` / ``optimize`` (-analyze 1),
  ``mkl_sparse_s_mm`` (Y = A Omega, B = A Y), ``LAPACKE_sgeqrf`` +
  ``LAPACKE_sorgqr`` (util.cc:41-79), ``cblas_sgemm`` (Z = B P, C = Z^T B,
  U = Z U_C), ``LAPACKE_sgesvd`` (C);
* propagation (SpectralPropagation.h:45-229): ``mkl_sparse_s_mm`` with the
  explicit operator M, ``cblas_saxpy`` for the recurrence, ``LAPACKE_sgesvd``
  for get_embedding_dense.

The sketch is the shared counter-based Gaussian (oracle.gaussian), not VSL's
SFMT19937 stream, so the values differ from the reference's by design; the
work and the calls are the same.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MKL_DIRS = ["/opt/conda/lib", "/opt/intel/oneapi/mkl/latest/lib/intel64", "/opt/intel/mkl/lib/intel64"]

# mkl_spblas.h enumerations
_STATUS_SUCCESS = 0
_OP_NON_TRANSPOSE = 10
_TYPE_GENERAL = 20
_INDEX_BASE_ZERO = 0
_FILL_LOWER = 40
_DIAG_NON_UNIT = 50
_LAYOUT_ROW_MAJOR = 101
_MEMORY_NONE = 80
# LAPACKE / CBLAS
_LAPACK_ROW_MAJOR = 101
_CBLAS_ROW_MAJOR = 101
_CBLAS_NO_TRANS = 111
_CBLAS_TRANS = 112
# mkl_service.h
_INTERFACE_ILP64 = 1
_THREADING_GNU = 3

I64 = C.c_int64
F32 = C.c_float
P = C.c_void_p


class _Descr(C.Structure):
    _fields_ = [("type", C.c_int), ("mode", C.c_int), ("diag", C.c_int)]


_lib = None
_path = None


def find() -> str | None:
    for d in MKL_DIRS:
        p = os.path.join(d, "libmkl_rt.so.1")
        if os.path.exists(p):
            return p
        p = os.path.join(d, "libmkl_rt.so")
        if os.path.exists(p):
            return p
    return None


def load(threads: int = 0):
    """libmkl_rt with the ILP64 interface and GNU threads (the reference's link
    line); None when MKL is not installed."""
    global _lib, _path
    if _lib is None:
        path = find()
        if path is None:
            return None
        lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        lib.MKL_Set_Interface_Layer.restype = C.c_int
        lib.MKL_Set_Interface_Layer.argtypes = [C.c_int]
        lib.MKL_Set_Threading_Layer.restype = C.c_int
        lib.MKL_Set_Threading_Layer.argtypes = [C.c_int]
        if lib.MKL_Set_Interface_Layer(_INTERFACE_ILP64) != _INTERFACE_ILP64:
            raise RuntimeError("MKL: ILP64 interface not available")
        lib.MKL_Set_Threading_Layer(_THREADING_GNU)
        sig = {
            "MKL_Set_Num_Threads": (None, [C.c_int]),
            "MKL_Get_Max_Threads": (C.c_int, []),
            "MKL_Get_Version_String": (None, [C.c_char_p, C.c_int]),
            "mkl_sparse_s_create_csr": (C.c_int, [C.POINTER(P), C.c_int, I64, I64, P, P, P, P]),
            "mkl_sparse_set_mm_hint": (C.c_int, [P, C.c_int, _Descr, C.c_int, I64, I64]),
            "mkl_sparse_set_memory_hint": (C.c_int, [P, C.c_int]),
            "mkl_sparse_optimize": (C.c_int, [P]),
            "mkl_sparse_destroy": (C.c_int, [P]),
            "mkl_sparse_s_mm": (C.c_int, [C.c_int, F32, P, _Descr, C.c_int, P, I64, I64, F32, P, I64]),
            "LAPACKE_sgeqrf": (I64, [C.c_int, I64, I64, P, I64, P]),
            "LAPACKE_sorgqr": (I64, [C.c_int, I64, I64, I64, P, I64, P]),
            "LAPACKE_sgesvd": (I64, [C.c_int, C.c_char, C.c_char, I64, I64, P, I64, P, P, I64, P, I64, P]),
            "cblas_sgemm": (None, [C.c_int, C.c_int, C.c_int, I64, I64, I64, F32, P, I64, P, I64, F32, P, I64]),
            "cblas_saxpy": (None, [I64, F32, P, I64, P, I64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib, _path = lib, path
    if threads > 0:
        _lib.MKL_Set_Num_Threads(threads)
    return _lib


def version() -> str:
    lib = load()
    if lib is None:
        return "absent"
    buf = C.create_string_buffer(256)
    lib.MKL_Get_Version_String(buf, 256)
    return buf.value.decode(errors="replace").strip()


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _ok(status: int, what: str) -> None:
    if status != _STATUS_SUCCESS:
        raise RuntimeError(f"MKL {what} failed with status {status}")


class SparseCSR:
    """mkl_sparse_s_create_csr over int64 (ILP64) arrays, optionally with the
    -analyze hints (mklredsvd.cc:117-136)."""

    def __init__(self, rp, col, val, ncols: int | None = None, hint_k: int = 0, hint_calls: int = 0):
        lib = load()
        self.rp = np.ascontiguousarray(rp, np.int64)
        self.col = np.ascontiguousarray(col, np.int64)
        self.val = np.ascontiguousarray(val, np.float32)
        self.n = self.rp.shape[0] - 1
        self.h = P()
        _ok(lib.mkl_sparse_s_create_csr(C.byref(self.h), _INDEX_BASE_ZERO, self.n, ncols or self.n, _p(self.rp),
                                        self.rp.ctypes.data + 8, _p(self.col), _p(self.val)), "create_csr")
        self.descr = _Descr(_TYPE_GENERAL, _FILL_LOWER, _DIAG_NON_UNIT)
        if hint_calls > 0:
            _ok(lib.mkl_sparse_set_mm_hint(self.h, _OP_NON_TRANSPOSE, self.descr, _LAYOUT_ROW_MAJOR, hint_k,
                                           hint_calls), "set_mm_hint")
            _ok(lib.mkl_sparse_set_memory_hint(self.h, _MEMORY_NONE), "set_memory_hint")
            _ok(lib.mkl_sparse_optimize(self.h), "optimize")

    def mm(self, X: np.ndarray, alpha: float = 1.0, beta: float = 0.0, Y: np.ndarray | None = None,
           cols: int = 0) -> np.ndarray:
        """Y = alpha A X[:, :cols] + beta Y, row-major (mkl_sparse_s_mm); X's
        row stride is its full width."""
        k = cols or X.shape[1]
        if Y is None:
            Y = np.empty((self.n, k), np.float32)
        _ok(load().mkl_sparse_s_mm(_OP_NON_TRANSPOSE, alpha, self.h, self.descr, _LAYOUT_ROW_MAJOR, _p(X), k,
                                   X.shape[1], beta, _p(Y), Y.shape[1]), "sparse_s_mm")
        return Y

    def close(self):
        if self.h:
            load().mkl_sparse_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gram_schmidt(Y: np.ndarray) -> None:
    """util.cc:41-79: Householder QR (geqrf) + explicit Q (orgqr), row-major, in place."""
    lib = load()
    n, k = Y.shape
    tau = np.empty(k, np.float32)
    r = lib.LAPACKE_sgeqrf(_LAPACK_ROW_MAJOR, n, k, _p(Y), k, _p(tau))
    if r != 0:
        raise RuntimeError(f"LAPACKE_sgeqrf: {r}")
    r = lib.LAPACKE_sorgqr(_LAPACK_ROW_MAJOR, n, k, k, _p(Y), k, _p(tau))
    if r != 0:
        raise RuntimeError(f"LAPACKE_sorgqr: {r}")


def sgemm(A: np.ndarray, B: np.ndarray, trans_a: bool = False) -> np.ndarray:
    """C = op(A) B, row-major (cblas_sgemm)."""
    m = A.shape[1] if trans_a else A.shape[0]
    kk = A.shape[0] if trans_a else A.shape[1]
    n = B.shape[1]
    Cm = np.empty((m, n), np.float32)
    load().cblas_sgemm(_CBLAS_ROW_MAJOR, _CBLAS_TRANS if trans_a else _CBLAS_NO_TRANS, _CBLAS_NO_TRANS, m, n, kk, 1.0,
                       _p(A), A.shape[1], _p(B), n, 0.0, _p(Cm), n)
    return Cm


def saxpy(a: float, x: np.ndarray, y: np.ndarray) -> None:
    load().cblas_saxpy(x.size, a, _p(x), 1, _p(y), 1)


def gesvd_left(A: np.ndarray):
    """LAPACKE_sgesvd('S', 'N') of a row-major m x n matrix (destroyed): U (m x n), S."""
    lib = load()
    m, n = A.shape
    U = np.empty((m, min(m, n)), np.float32)
    S = np.empty(min(m, n), np.float32)
    superb = np.empty(max(min(m, n) - 1, 1), np.float32)
    vt = np.empty(1, np.float32)
    r = lib.LAPACKE_sgesvd(_LAPACK_ROW_MAJOR, b"S", b"N", m, n, _p(A), n, _p(S), _p(U), U.shape[1], _p(vt), n,
                           _p(superb))
    if r != 0:
        raise RuntimeError(f"LAPACKE_sgesvd: {r}")
    return U, S


def redsvd(A: SparseCSR, Omega: np.ndarray, P_small: np.ndarray):
    """mklredsvd.cc:229-400 with q = 1: Y = A Omega, Q = orth(Y), B = A Q,
    Z = orth(B P), C = Z^T B, C = U_C S V^T, U = Z U_C.  Returns (U, S)."""
    Y = A.mm(Omega)
    gram_schmidt(Y)
    B = A.mm(Y)
    del Y
    Z = sgemm(B, P_small)
    gram_schmidt(Z)
    Cm = sgemm(Z, B, trans_a=True)
    del B
    Uc, S = gesvd_left(Cm)
    return sgemm(Z, Uc), S


def prop_operator(offsets, edges, mu: float, rows_end: int | None = None) -> SparseCSR:
    """SpectralPropagation.h:51-92: rows of M = (1 - mu) I - D~^-1 A~ with the
    self loop first, values formed in double and stored as float.  With
    ``rows_end`` only the first rows (all n columns)."""
    off = np.asarray(offsets, np.int64)
    ncols = off.shape[0] - 1
    if rows_end is not None:
        off = off[:rows_end + 1]
        edges = edges[:int(off[-1])]
    n = off.shape[0] - 1
    deg = np.diff(off)
    rp = off + np.arange(n + 1, dtype=np.int64)
    col = np.empty(int(rp[-1]), np.int64)
    val = np.empty(int(rp[-1]), np.float32)
    d = (deg + 1).astype(np.float32).astype(np.float64)
    mu64 = float(np.float32(mu))
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    pos = np.arange(off[-1], dtype=np.int64) + rows + 1  # neighbour j of row i -> rp[i] + 1 + (j - off[i])
    col[rp[:-1]] = np.arange(n, dtype=np.int64)
    val[rp[:-1]] = (1.0 - mu64 - 1.0 / d).astype(np.float32)
    col[pos] = np.asarray(edges, np.int64)
    val[pos] = (-1.0 / d[rows]).astype(np.float32)
    return SparseCSR(rp, col, val, ncols=ncols)


def aggregate_operator(offsets, edges) -> SparseCSR:
    """A + I with unit values (SpectralPropagation.h:210-221's loop as a product)."""
    off = np.asarray(offsets, np.int64)
    n = off.shape[0] - 1
    deg = np.diff(off)
    rp = off + np.arange(n + 1, dtype=np.int64)
    col = np.empty(int(rp[-1]), np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    col[rp[:-1]] = np.arange(n, dtype=np.int64)
    col[np.arange(off[-1], dtype=np.int64) + rows + 1] = np.asarray(edges, np.int64)
    return SparseCSR(rp, col, np.ones(int(rp[-1]), np.float32))


def chebyshev(M: SparseCSR, a: np.ndarray, coeffs) -> np.ndarray:
    """SpectralPropagation.h:94-208: the recurrence with mkl_sparse_s_mm and
    cblas_saxpy exactly as written; returns y = a - conv (a is left intact)."""
    Lx0 = a.copy()
    tmp = M.mm(a)
    Lx1 = a.copy()
    M.mm(tmp, 0.5, -1.0, Lx1)
    conv = np.zeros_like(a)
    saxpy(float(np.float32(coeffs[0])), Lx0, conv)
    saxpy(float(np.float32(coeffs[1])), Lx1, conv)
    Lx2 = np.empty_like(a)
    for i in range(2, len(coeffs)):
        M.mm(Lx1, 1.0, 0.0, tmp)
        np.copyto(Lx2, Lx0)
        M.mm(tmp, 1.0, -1.0, Lx2)
        saxpy(-2.0, Lx1, Lx2)
        saxpy(float(np.float32(coeffs[i])), Lx2, conv)
        Lx0, Lx1, Lx2 = Lx1, Lx2, Lx0
    y = a.copy()
    saxpy(-1.0, conv, y)
    return y
