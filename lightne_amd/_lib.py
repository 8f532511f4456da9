"""ctypes binding of liblightne_amd.so (the C-ABI declared in include/lightne.h).

The shared library is built in-tree (lightne_amd/liblightne_amd.so) by
``__graft_entry__.build()``; nothing here compiles code.  Loading fails loudly
when the library is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblightne_amd.so")

OK = 0
ERR_BAD_ARG = 1
ERR_OOM = 2
ERR_TABLE_FULL = 3
ERR_HIP = 4
ERR_BAD_GRAPH = 5
ERR_NUMERIC = 6
ERR_IO = 7
ERR_NO_DEVICE = 8

FP32 = 0
BF16 = 1
FP64 = 2  # randomized SVD only (accuracy reference)
BF16_POWER = 3  # BF16, and the power iterations gather BF16 too
FP64_LEAN = 4  # randomized SVD only: FP64 in two panels (the FP32 form's schedule); output may be host memory

_ERR_NAMES = {
    OK: "OK", ERR_BAD_ARG: "BAD_ARG", ERR_OOM: "OOM", ERR_TABLE_FULL: "TABLE_FULL", ERR_HIP: "HIP",
    ERR_BAD_GRAPH: "BAD_GRAPH", ERR_NUMERIC: "NUMERIC", ERR_IO: "IO", ERR_NO_DEVICE: "NO_DEVICE",
}


class LightNEError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


ITH_PLAIN = 0  # walks take neighbour i in input order
ITH_BYTEPD = 1  # walks take the compressed-format lookup (neighbour i - 1 unless 64 | i)


class Graph(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("offsets", C.c_void_p), ("edges", C.c_void_p),
                ("ith_rule", C.c_int32), ("reserved", C.c_int32)]


class SparsifyParams(C.Structure):
    _fields_ = [
        ("walk_len", C.c_uint64), ("negative", C.c_uint64), ("sample", C.c_int32), ("upper", C.c_int32),
        ("sample_ratio", C.c_float), ("mem_ratio", C.c_float), ("tablesz", C.c_uint64), ("seed", C.c_uint64),
        ("max_table_bytes", C.c_uint64), ("table_layout", C.c_int32), ("csr_parts", C.c_int32),
        ("value_cache", C.c_int32), ("wide_offsets", C.c_int32),
    ]


class SparsifyStats(C.Structure):
    _fields_ = [
        ("logn", C.c_double), ("M", C.c_double), ("one_over_lambda", C.c_double), ("fractional_bits", C.c_double),
        ("exp_rv_rounded", C.c_uint64), ("table_size", C.c_uint64),
        ("table_capacity", C.c_uint64), ("attempts", C.c_uint64), ("samples", C.c_uint64),
        ("walk_steps", C.c_uint64), ("distinct_keys", C.c_uint64),
        ("M_exact", C.c_double), ("factor", C.c_double), ("kept_upper", C.c_uint64), ("nnz", C.c_uint64),
        ("t_sample", C.c_double), ("t_table", C.c_double), ("t_csr", C.c_double),
        ("table_layout", C.c_uint32), ("csr_parts", C.c_uint32), ("value_cache", C.c_uint32),
        ("sampler_runs", C.c_uint32), ("t_canon", C.c_double), ("row_order", C.c_uint32), ("walk_rows", C.c_uint32),
    ]


class RsvdParams(C.Structure):
    _fields_ = [
        ("rank", C.c_int64), ("power_iters", C.c_int32), ("precision", C.c_int32),
        ("random_project_only", C.c_int32), ("upper", C.c_int32), ("seed", C.c_uint64),
        ("sparse_project", C.c_int32), ("sparse_project_s", C.c_float),
    ]


class RsvdStats(C.Structure):
    _fields_ = [
        ("t_sketch", C.c_double), ("t_spmm", C.c_double), ("t_qr", C.c_double), ("t_gemm", C.c_double),
        ("t_small_svd", C.c_double), ("t_total", C.c_double), ("spmm_calls", C.c_int32), ("qr_calls", C.c_int32),
        ("t_spmm_sketch", C.c_double),
    ]


class PropParams(C.Structure):
    _fields_ = [("order", C.c_int64), ("mu", C.c_float), ("theta", C.c_float), ("precision", C.c_int32)]


class PropStats(C.Structure):
    _fields_ = [
        ("t_cheb", C.c_double), ("t_aggregate", C.c_double), ("t_svd", C.c_double), ("t_total", C.c_double),
        ("spmm_calls", C.c_int32),
    ]


class NetsmfStats(C.Structure):
    _fields_ = [
        ("t_sparsify", C.c_double), ("t_svd", C.c_double), ("t_total", C.c_double),
        ("sparsify", SparsifyStats), ("rsvd", RsvdStats),
    ]


_P = C.c_void_p
_I64 = C.c_int64
_PI64 = C.POINTER(C.c_int64)
_PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    "lightne_version": (C.c_char_p, []),
    "lightne_global_error": (C.c_char_p, []),
    "lightne_last_error": (C.c_char_p, [_P]),
    "lightne_ctx_create": (C.c_int, [C.c_int, _P, _PP]),
    "lightne_ctx_destroy": (C.c_int, [_P]),
    "lightne_ctx_synchronize": (C.c_int, [_P]),
    "lightne_ctx_release": (C.c_int, [_P]),
    "lightne_sample_scalars": (C.c_int, [_I64, _I64, C.POINTER(SparsifyParams), C.POINTER(SparsifyStats)]),
    "lightne_sparsify_counts": (C.c_int, [_P, C.POINTER(Graph), C.POINTER(SparsifyParams), _PP,
                                          C.POINTER(SparsifyStats)]),
    "lightne_table_info": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "lightne_table_copy": (C.c_int, [_P, _P, _P, _P, _P]),
    "lightne_table_free": (C.c_int, [_P]),
    "lightne_sparsify": (C.c_int, [_P, C.POINTER(Graph), C.POINTER(SparsifyParams), _PP,
                                   C.POINTER(SparsifyStats)]),
    "lightne_csr_create": (C.c_int, [_P, _I64, _I64, _P, _P, _P, _PP]),
    "lightne_csr_info": (C.c_int, [_P, _PI64, _PI64]),
    "lightne_csr_device_ptrs": (C.c_int, [_P, _PP, _PP, _PP]),
    "lightne_csr_copy": (C.c_int, [_P, _P, _P, _P, _P]),
    "lightne_csr_copy_prefix": (C.c_int, [_P, _P, _I64, _P, _P]),
    "lightne_csr_free": (C.c_int, [_P]),
    "lightne_rsvd": (C.c_int, [_P, _P, C.POINTER(RsvdParams), _P, _P, C.POINTER(RsvdStats)]),
    "lightne_u_sigma_root": (C.c_int, [_P, _P, _P, _I64, _I64, _I64, C.c_int32, _P]),
    "lightne_propagate": (C.c_int, [_P, C.POINTER(Graph), _P, _I64, C.POINTER(PropParams), _P, _P,
                                    C.POINTER(PropStats)]),
    "lightne_embed_dense": (C.c_int, [_P, _P, _I64, _I64, _I64, _P, _P]),
    "lightne_chebyshev_expansion": (C.c_int, [_P, C.POINTER(Graph), _P, _I64, C.POINTER(PropParams), _P,
                                              C.POINTER(PropStats)]),
    "lightne_propagate_backward": (C.c_int, [_P, C.POINTER(Graph), _P, _I64, C.POINTER(PropParams), C.c_int32,
                                             _P, _P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32]),
    "lightne_netsmf": (C.c_int, [_P, C.POINTER(Graph), C.POINTER(SparsifyParams), C.POINTER(RsvdParams), _I64,
                                 C.c_int32, _P, _P, C.POINTER(NetsmfStats)]),
    "lightne_ne_zhang": (C.c_int, [_P, C.POINTER(Graph), C.POINTER(RsvdParams), _I64, _P, _P,
                                   C.POINTER(RsvdStats)]),
    "lightne_ne_zhang_matrix": (C.c_int, [_P, C.POINTER(Graph), _PP]),
    "lightne_dense_gram": (C.c_int, [_P, _P, _P, _I64, _I64, _I64, C.c_int32, _P]),
    "lightne_dense_gram_noalloc": (C.c_int, [_P, _P, _P, _I64, _I64, _I64, C.c_int32, _P]),
    "lightne_dense_panel_gemm": (C.c_int, [_P, _P, _I64, _I64, _P, _I64, _P, C.c_int32]),
    "lightne_dense_cholqr": (C.c_int, [_P, _P, _I64, _I64, _P, C.POINTER(C.c_int32)]),
    "lightne_graph_read_adj": (C.c_int, [C.c_char_p, _PI64, _PI64, _PP, _PP]),
    "lightne_graph_write_adj": (C.c_int, [C.c_char_p, _I64, _P, _P]),
    "lightne_graph_read_bin": (C.c_int, [C.c_char_p, _PI64, _PI64, _PP, _PP]),
    "lightne_graph_write_bin": (C.c_int, [C.c_char_p, _I64, _P, _P]),
    "lightne_graph_decode_compressed": (C.c_int, [_P, _P, C.c_uint64, _I64, _I64, _P, _P]),
    "lightne_gen_chung_lu": (C.c_int, [_I64, C.c_double, C.c_double, C.c_double, C.c_uint64, _PI64, _PP, _PP]),
    "lightne_gen_rmat": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                   _PI64, _PI64, _PP, _PP]),
    "lightne_host_free": (None, [_P]),
    "lightne_bessel_i": (C.c_double, [C.c_int, C.c_double]),
}

_lib = None


def load() -> C.CDLL:
    """Load liblightne_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or make -C lightne_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def global_error() -> str:
    return load().lightne_global_error().decode()


def check(rc: int, ctx=None) -> None:
    if rc != OK:
        lib = load()
        msg = lib.lightne_last_error(ctx).decode() if ctx else ""
        if not msg:
            msg = lib.lightne_global_error().decode()
        raise LightNEError(rc, msg)
