"""lightne_amd -- MI355X-native LightNE embedding pipeline.

Stage 1 (NetSMF sparsifier), stage 2 (randomized SVD) and stage 3 (spectral
propagation) run as hand-written HIP kernels for gfx950 behind the C-ABI in
include/lightne.h; this package is the host-side mirror of the reference's
stage API (see ``lightne_amd.api``).
"""
from ._lib import BF16, BF16_POWER, FP32, FP64, FP64_LEAN, ITH_BYTEPD, ITH_PLAIN, LightNEError, LIB_PATH  # noqa: F401
from .api import (CSR, Context, DeviceGraph, HostGraph, MKLRedSVD, NE_Zhang_et_al, NetSMF, SparseTable,  # noqa: F401
                  bessel_i, chebyshev_expansion, chung_lu, compute_u_sigma_root, context, decode_compressed_graph,
                  dense_cholqr, dense_gram, dense_panel_gemm, generate_trunc_log_matrix, get_embedding_dense,
                  ne_zhang_matrix, propagate, propagate_backward, read_compressed_graph, read_graph_bin,
                  read_unweighted_symmetric_graph, rmat, sample_scalars, sparsify_counts, write_adjacency_graph,
                  write_compressed_graph, write_graph_bin)

__all__ = [
    "BF16", "BF16_POWER", "FP32", "FP64", "FP64_LEAN", "ITH_BYTEPD", "ITH_PLAIN", "LightNEError", "CSR", "Context", "DeviceGraph", "HostGraph",
    "MKLRedSVD", "NE_Zhang_et_al", "NetSMF", "SparseTable", "bessel_i", "chebyshev_expansion", "chung_lu",
    "compute_u_sigma_root", "context", "dense_cholqr", "dense_gram", "dense_panel_gemm",
    "generate_trunc_log_matrix", "get_embedding_dense", "ne_zhang_matrix", "propagate", "propagate_backward",
    "read_unweighted_symmetric_graph", "rmat", "sample_scalars", "sparsify_counts", "write_adjacency_graph",
    "decode_compressed_graph", "read_compressed_graph", "read_graph_bin", "write_compressed_graph", "write_graph_bin",
]
