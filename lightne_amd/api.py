"""Host-side mirror of LightNE's stage API over the liblightne_amd C-ABI.

Names and argument meaning follow the reference:

* ``generate_trunc_log_matrix``  -- path_embed::generate_trunc_log_matrix_v2
  (LightNE/PathEmbed.h:81-84)
* ``MKLRedSVD``                  -- mkl_redsvd::MKLRedSVD<FP> (LightNE/mklredsvd/mklredsvd.h:8-26)
* ``compute_u_sigma_root``       -- path_embed::compute_u_sigma_root (LightNE/PathEmbed.h:371-408)
* ``chebyshev_expansion``        -- spectral_propagation::chebyshev_expansion
  (LightNE/SpectralPropagation.h:45-229)
* ``get_embedding_dense``        -- spectral_propagation::get_embedding_dense
  (LightNE/SpectralPropagation.h:16-43)
* ``NetSMF``                     -- path_embed::NetSMF (LightNE/PathEmbed.h:494-563)
* ``read_unweighted_symmetric_graph`` -- gbbs_io::read_unweighted_symmetric_graph
  (ligra/graph_io.h:233-252)

Device buffers are torch tensors on ``cuda``; torch only provides memory and
the stream, every computation runs in the HIP library.  Errors raise
``LightNEError`` (the reference calls exit(0) instead, par_table.h:104-107).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _lib
from ._lib import (BF16, BF16_POWER, FP32, FP64, FP64_LEAN, Graph, LightNEError, NetsmfStats, PropParams,
                   PropStats, RsvdParams, RsvdStats, SparsifyParams, SparsifyStats, check, load)


def _torch():
    import torch  # imported lazily: host-only helpers work without torch

    return torch


def _bad(msg: str):
    raise LightNEError(_lib.ERR_BAD_ARG, msg)


def _need(t, name: str, dtype, shape=None, device: Optional[int] = None):
    """A device tensor the library reads or writes through a raw pointer:
    it must be a contiguous torch tensor of ``dtype`` (and ``shape``) on the
    context's device, or the C side would read the wrong bytes."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        _bad(f"{name}: expected a torch tensor on the device, got {type(t).__name__}")
    if t.device.type != "cuda" or (device is not None and (t.device.index or 0) != device):
        _bad(f"{name}: tensor on {t.device}, expected cuda:{device if device is not None else 0}")
    if t.dtype != dtype:
        _bad(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        _bad(f"{name}: tensor is not contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        _bad(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


def _stats_dict(s) -> dict:
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = _stats_dict(v) if isinstance(v, C.Structure) else v
    return out


# ---------------------------------------------------------------------------
# host graphs
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class HostGraph:
    """Symmetric CSR on the host: offsets int64[n+1], edges uint32[m] (input order)."""

    offsets: np.ndarray
    edges: np.ndarray

    @property
    def n(self) -> int:
        return int(self.offsets.shape[0] - 1)

    @property
    def m(self) -> int:
        return int(self.edges.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)

    def to_device(self, device: str = "cuda", ith_rule: int = _lib.ITH_PLAIN) -> "DeviceGraph":
        torch = _torch()
        off = torch.from_numpy(self.offsets.astype(np.int64)).to(device)
        # torch has no uint32 kernels we need; int32 storage carries the same bits
        edg = torch.from_numpy(self.edges.astype(np.uint32).view(np.int32)).to(device)
        return DeviceGraph(off, edg, self.n, self.m, ith_rule)


@dataclasses.dataclass
class DeviceGraph:
    offsets: "object"  # torch.int64 [n+1]
    edges: "object"  # torch.int32 [m], bits of uint32
    n: int
    m: int
    # how walks map draw i to a neighbour: ITH_PLAIN (input order), or
    # ITH_BYTEPD for graphs read from the compressed format, whose lookup in
    # the reference returns neighbour i - 1 unless i is a multiple of 64
    # (ligra/encodings/byte_pd_amortized.h:653-688)
    ith_rule: int = _lib.ITH_PLAIN

    def desc(self) -> Graph:
        torch = _torch()
        _need(self.offsets, "graph offsets", torch.int64, (self.n + 1,))
        if self.m:
            _need(self.edges, "graph edges", torch.int32, (self.m,), self.offsets.device.index or 0)
        return Graph(self.n, self.m, self.offsets.data_ptr(), self.edges.data_ptr() if self.m else None,
                     self.ith_rule, 0)


def _take_host_graph(n: int, m: int, off_p: C.c_void_p, edg_p: C.c_void_p) -> HostGraph:
    lib = load()
    try:
        off = np.ctypeslib.as_array(C.cast(off_p, C.POINTER(C.c_int64)), shape=(n + 1,)).copy()
        edg = (np.ctypeslib.as_array(C.cast(edg_p, C.POINTER(C.c_uint32)), shape=(m,)).copy()
               if m else np.zeros(0, np.uint32))
    finally:
        lib.lightne_host_free(off_p)
        lib.lightne_host_free(edg_p)
    return HostGraph(off, edg)


def read_unweighted_symmetric_graph(path: str) -> HostGraph:
    """Read a Ligra AdjacencyGraph text file (ligra/graph_io.h:185-252)."""
    lib = load()
    n, m = C.c_int64(), C.c_int64()
    op, ep = C.c_void_p(), C.c_void_p()
    check(lib.lightne_graph_read_adj(path.encode(), C.byref(n), C.byref(m), C.byref(op), C.byref(ep)))
    return _take_host_graph(n.value, m.value, op, ep)


def write_adjacency_graph(path: str, g: HostGraph) -> None:
    lib = load()
    off = np.ascontiguousarray(g.offsets, dtype=np.int64)
    edg = np.ascontiguousarray(g.edges, dtype=np.uint32)
    check(lib.lightne_graph_write_adj(path.encode(), g.n, off.ctypes.data, edg.ctypes.data if g.m else None))


def read_graph_bin(path: str) -> HostGraph:
    """Read the binary CSR cache ("LNECSR01", int64 n, m, offsets, uint32 edges)."""
    lib = load()
    n, m = C.c_int64(), C.c_int64()
    op, ep = C.c_void_p(), C.c_void_p()
    check(lib.lightne_graph_read_bin(path.encode(), C.byref(n), C.byref(m), C.byref(op), C.byref(ep)))
    return _take_host_graph(n.value, m.value, op, ep)


def write_graph_bin(path: str, g: HostGraph) -> None:
    lib = load()
    off = np.ascontiguousarray(g.offsets, dtype=np.int64)
    edg = np.ascontiguousarray(g.edges, dtype=np.uint32)
    check(lib.lightne_graph_write_bin(path.encode(), g.n, off.ctypes.data, edg.ctypes.data if g.m else None))


def write_compressed_graph(path: str, g: HostGraph) -> None:
    """Write ``g`` in Ligra's compressed format (the ``-c`` input; encoder in
    lightne_amd.bytepd)."""
    from . import bytepd

    bytepd.write(path, g.offsets, g.edges)


def decode_compressed_graph(image, ith_rule: int = _lib.ITH_BYTEPD, ctx: Optional["Context"] = None) -> DeviceGraph:
    """Decode a compressed-format file image (uint8 torch tensor on the
    device, the whole file) into a DeviceGraph (gbbs_io::
    read_compressed_symmetric_graph, ligra/graph_io.h:344-377, with the block
    decoder of ligra/encodings/byte_pd_amortized.h).  ``ith_rule`` defaults to
    the reference's walk lookup on compressed graphs."""
    torch = _torch()
    ctx = ctx or context(image.device.index or 0)
    nbytes = int(image.numel())
    if nbytes < 24:
        raise LightNEError(_lib.ERR_BAD_GRAPH, "compressed graph: file shorter than its header")
    n, m = (int(x) for x in image[:16].cpu().numpy().view(np.int64))
    if n <= 0 or m < 0:
        raise LightNEError(_lib.ERR_BAD_GRAPH, f"compressed graph: bad header n = {n}, m = {m}")
    off = torch.empty(n + 1, dtype=torch.int64, device=image.device)
    edg = torch.empty(max(m, 1), dtype=torch.int32, device=image.device)
    ctx.check(ctx.lib.lightne_graph_decode_compressed(ctx.ptr, image.data_ptr(), nbytes, n, m, off.data_ptr(),
                                                      edg.data_ptr()))
    return DeviceGraph(off, edg[:m], n, m, ith_rule)


def read_compressed_graph(path: str, device: int = 0, ith_rule: int = _lib.ITH_BYTEPD,
                          ctx: Optional["Context"] = None) -> DeviceGraph:
    """Read a compressed-format file (``lightne -c``) and decode it on the device."""
    torch = _torch()
    img = torch.from_numpy(np.fromfile(path, dtype=np.uint8)).to(f"cuda:{device}")
    return decode_compressed_graph(img, ith_rule, ctx)


def chung_lu(n: int, beta: float, mean_deg: float, max_deg: float, seed: int) -> HostGraph:
    lib = load()
    m = C.c_int64()
    op, ep = C.c_void_p(), C.c_void_p()
    check(lib.lightne_gen_chung_lu(n, beta, mean_deg, max_deg, seed, C.byref(m), C.byref(op), C.byref(ep)))
    return _take_host_graph(n, m.value, op, ep)


def rmat(scale: int, edge_factor: float, a: float, b: float, c: float, seed: int) -> HostGraph:
    lib = load()
    n, m = C.c_int64(), C.c_int64()
    op, ep = C.c_void_p(), C.c_void_p()
    check(lib.lightne_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(n), C.byref(m), C.byref(op), C.byref(ep)))
    return _take_host_graph(n.value, m.value, op, ep)


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
class _StreamOrdered:
    """The library's entry points, each called with the context's stream
    ordered against torch's current stream on the device at the time of the
    call: the context stream first waits for the work queued on the current
    stream (which produced the inputs), and the current stream then waits for
    the call's work (which produced the outputs).  Work on another torch
    stream therefore never races with the library."""

    def __init__(self, lib, ctx: "Context"):
        self._lib = lib
        self._ctx = ctx

    def __getattr__(self, name):
        fn = getattr(self._lib, name)
        ctx = self._ctx

        def call(*args):
            torch = _torch()
            cur = torch.cuda.current_stream(ctx.device)
            other = cur.cuda_stream != ctx.stream.cuda_stream
            if other:
                ctx.stream.wait_stream(cur)
            rc = fn(*args)
            if other:
                cur.wait_stream(ctx.stream)
            return rc

        return call


class Context:
    """One library context bound to a device and a stream on it (torch's
    current stream when the context is created); calls made while another
    torch stream is current are ordered against it (``_StreamOrdered``)."""

    def __init__(self, device: int = 0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise LightNEError(_lib.ERR_NO_DEVICE, "no GPU visible to torch")
        self.raw = load()
        self.lib = _StreamOrdered(self.raw, self)
        self.device = device
        with torch.cuda.device(device):
            self.stream = torch.cuda.current_stream(device)
        self.ptr = C.c_void_p()
        check(self.raw.lightne_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(self.ptr)))

    def check(self, rc: int) -> None:
        check(rc, self.ptr)

    def release(self) -> None:
        """Free the device memory the context keeps between calls (workspace,
        scratch); the next call allocates again."""
        if self.ptr:
            self.check(self.lib.lightne_ctx_release(self.ptr))

    def close(self) -> None:
        if self.ptr:
            self.raw.lightne_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX: dict = {}


def context(device: int = 0) -> Context:
    c = _CTX.get(device)
    if c is None:
        c = Context(device)
        _CTX[device] = c
    return c


def _dev(t) -> int:
    return t.data_ptr()


# ---------------------------------------------------------------------------
# stage 1
# ---------------------------------------------------------------------------
def sample_scalars(n: int, m: int, walk_len: int = 10, sample_ratio: float = 1.0, mem_ratio: float = 1.0,
                   table_size: int = 0) -> dict:
    """Host evaluation of PathEmbed.h:89,99,108-117 (logn, M, 1/lambda, ...)."""
    lib = load()
    p = SparsifyParams(walk_len=walk_len, negative=1, sample=0, upper=0, sample_ratio=sample_ratio,
                       mem_ratio=mem_ratio, tablesz=table_size, seed=0, max_table_bytes=0)
    st = SparsifyStats()
    check(lib.lightne_sample_scalars(n, m, C.byref(p), C.byref(st)))
    return _stats_dict(st)


def _sparsify_params(seed, walk_len, upper, sample, table_size, negative, sample_ratio, mem_ratio,
                     max_table_bytes, table_layout=0, csr_parts=0, value_cache=0, wide_offsets=0) -> SparsifyParams:
    """``table_layout``: 0 auto, 1 16-byte slots, 2 8-byte slots (n <= 2^26).
    ``csr_parts``: 0 auto, k >= 2 builds the CSR in at least k row ranges.
    ``value_cache`` (pipeline): 0 auto, 1 always, -1 never cache the count
    pass's CSR value of every distinct key for the emission (same CSR either
    way).
    ``wide_offsets``: 0 auto (the walks read each neighbour with its row,
    entries built per call, when the graph has fewer than 2^32 directed edges;
    else the graph's 64-bit offsets), 1 neighbours and the 64-bit offsets
    always, 2 neighbours and 32-bit row pointers where they fit (same samples
    every way)."""
    return SparsifyParams(walk_len=walk_len, negative=negative, sample=int(bool(sample)), upper=int(bool(upper)),
                          sample_ratio=sample_ratio, mem_ratio=mem_ratio, tablesz=table_size,
                          seed=seed & 0xFFFFFFFFFFFFFFFF, max_table_bytes=max_table_bytes,
                          table_layout=table_layout, csr_parts=csr_parts, value_cache=value_cache,
                          wide_offsets=int(wide_offsets))


@dataclasses.dataclass
class SparseTable:
    keys: np.ndarray  # uint64, sorted: (min(x,z) << 32) + max(x,z)
    counts: np.ndarray  # uint32 samples per key
    weights: np.ndarray  # float64 summed weight per key
    stats: dict


def sparsify_counts(graph: DeviceGraph, seed: int, walk_len: int = 10, sample: bool = False,
                    negative: int = 1, sample_ratio: float = 1.0, mem_ratio: float = 1.0, table_size: int = 0,
                    max_table_bytes: int = 0, table_layout: int = 0, wide_offsets: int = 0,
                    ctx: Optional[Context] = None) -> SparseTable:
    """Key -> (count, weight) multiset of the sampler (PathEmbed.h:128-188)."""
    ctx = ctx or context()
    p = _sparsify_params(seed, walk_len, False, sample, table_size, negative, sample_ratio, mem_ratio,
                         max_table_bytes, table_layout, wide_offsets=wide_offsets)
    st = SparsifyStats()
    tab = C.c_void_p()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_sparsify_counts(ctx.ptr, C.byref(g), C.byref(p), C.byref(tab), C.byref(st)))
    try:
        nk = C.c_uint64()
        check(ctx.lib.lightne_table_info(tab, C.byref(nk)))
        keys = np.empty(nk.value, np.uint64)
        cnt = np.empty(nk.value, np.uint32)
        wgh = np.empty(nk.value, np.float64)
        ctx.check(ctx.lib.lightne_table_copy(ctx.ptr, tab, keys.ctypes.data, cnt.ctypes.data, wgh.ctypes.data))
    finally:
        ctx.lib.lightne_table_free(tab)
    return SparseTable(keys, cnt, wgh, _stats_dict(st))


class CSR:
    """Device CSR owned by the library (int64 row_ptr, int32 col, float val)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle
        n, nnz = C.c_int64(), C.c_int64()
        check(ctx.lib.lightne_csr_info(handle, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    @classmethod
    def from_arrays(cls, row_ptr, col, val, ctx: Optional[Context] = None) -> "CSR":
        """Import a CSR (numpy host arrays or torch device tensors)."""
        ctx = ctx or context()
        keep = []

        torch = _torch()
        tdt = {np.int64: torch.int64, np.int32: torch.int32, np.float32: torch.float32}

        def ptr(a, dt, name, length):
            if isinstance(a, np.ndarray):
                a = np.ascontiguousarray(a, dtype=dt)
                if a.shape != (length,):
                    _bad(f"CSR.from_arrays: {name} has shape {a.shape}, expected ({length},)")
                keep.append(a)
                return a.ctypes.data
            return _need(a, f"CSR.from_arrays: {name}", tdt[dt], (length,), ctx.device).data_ptr()

        n = int(len(row_ptr) - 1)
        nnz = int(len(col))
        if n <= 0:
            _bad("CSR.from_arrays: row_ptr needs n + 1 >= 2 entries")
        h = C.c_void_p()
        ctx.check(ctx.lib.lightne_csr_create(ctx.ptr, n, nnz, ptr(row_ptr, np.int64, "row_ptr", n + 1),
                                             ptr(col, np.int32, "col", nnz), ptr(val, np.float32, "val", nnz),
                                             C.byref(h)))
        return cls(ctx, h)

    def to_numpy(self, nnz_fraction: Optional[float] = None):
        """Host copies (row_ptr, col, val).  With ``nnz_fraction`` only the
        entries of the first rows holding that fraction of the nonzeros are
        copied (row_ptr stays whole)."""
        rp = np.empty(self.n + 1, np.int64)
        if nnz_fraction is None:
            col = np.empty(self.nnz, np.int32)
            val = np.empty(self.nnz, np.float32)
            self.ctx.check(self.ctx.lib.lightne_csr_copy(self.ctx.ptr, self.handle, rp.ctypes.data,
                                                         col.ctypes.data, val.ctypes.data))
            return rp, col, val
        self.ctx.check(self.ctx.lib.lightne_csr_copy(self.ctx.ptr, self.handle, rp.ctypes.data, None, None))
        r = int(np.searchsorted(rp, nnz_fraction * self.nnz))
        nz = int(rp[min(r, self.n)])
        col = np.empty(nz, np.int32)
        val = np.empty(nz, np.float32)
        self.ctx.check(self.ctx.lib.lightne_csr_copy_prefix(self.ctx.ptr, self.handle, nz, col.ctypes.data,
                                                            val.ctypes.data))
        return rp, col, val

    def free(self) -> None:
        if self.handle:
            self.ctx.lib.lightne_csr_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate_trunc_log_matrix(graph: DeviceGraph, seed: int, walks_per_edge: int = 100, walk_len: int = 10,
                              upper: bool = False, sample: bool = False, table_size: int = 0, negative: int = 1,
                              sample_ratio: float = 1.0, mem_ratio: float = 1.0, max_table_bytes: int = 0,
                              table_layout: int = 0, csr_parts: int = 0, wide_offsets: int = 0,
                              ctx: Optional[Context] = None):
    """Stage 1 (PathEmbed.h:81-316): returns (CSR, stats).  ``walks_per_edge`` is
    accepted and unused, as in the reference (PathEmbed.h:107-108)."""
    del walks_per_edge
    ctx = ctx or context()
    p = _sparsify_params(seed, walk_len, upper, sample, table_size, negative, sample_ratio, mem_ratio,
                         max_table_bytes, table_layout, csr_parts, wide_offsets=wide_offsets)
    st = SparsifyStats()
    h = C.c_void_p()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_sparsify(ctx.ptr, C.byref(g), C.byref(p), C.byref(h), C.byref(st)))
    return CSR(ctx, h), _stats_dict(st)


# ---------------------------------------------------------------------------
# stage 2
# ---------------------------------------------------------------------------
class MKLRedSVD:
    """Randomized SVD of a symmetric sparse matrix (mklredsvd.h:8-26 / .cc:93-403).

    After ``run()``: ``matU`` (torch n x rank, device) and ``S`` (numpy rank),
    ``S`` is None with ``random_project_only`` as in the reference."""

    def __init__(self, csr: CSR, rank: int, power_iters: int = 1, precision: int = FP32,
                 random_project_only: bool = False, upper: bool = False, seed: int = 777,
                 sparse_project: bool = False, sparse_project_s: float = 0.0, ctx: Optional[Context] = None):
        self.csr = csr
        self.ctx = ctx or csr.ctx
        self.params = RsvdParams(rank=rank, power_iters=power_iters, precision=precision,
                                 random_project_only=int(bool(random_project_only)), upper=int(bool(upper)),
                                 seed=seed, sparse_project=int(bool(sparse_project)),
                                 sparse_project_s=sparse_project_s)
        self.matU = None
        self.S = None
        self.stats = None

    def run(self, out=None):
        """``out`` (n x rank float32) receives U: a device tensor, or with the
        FP64_LEAN form also a host numpy array."""
        torch = _torch()
        n, k = self.csr.n, self.params.rank
        if out is None:
            U = torch.empty((n, k), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        else:
            U = out
        S = np.zeros(k, np.float32)
        st = RsvdStats()
        uptr = _out_ptr(U, (n, k), "MKLRedSVD.run: out", self.ctx, self.params.precision == FP64_LEAN)
        self.ctx.check(self.ctx.lib.lightne_rsvd(self.ctx.ptr, self.csr.handle, C.byref(self.params), uptr,
                                                 S.ctypes.data, C.byref(st)))
        self.matU = U
        self.S = None if self.params.random_project_only else S
        self.stats = _stats_dict(st)
        return self


def _out_ptr(out, shape, name: str, ctx: Context, host_ok: bool) -> int:
    """Pointer of an output buffer: a device tensor, or (where the entry point
    accepts host memory) a writable C-contiguous float32 numpy array."""
    torch = _torch()
    if isinstance(out, np.ndarray):
        if not host_ok:
            _bad(f"{name}: must be a device tensor")
        if out.dtype != np.float32 or not out.flags.c_contiguous or not out.flags.writeable or out.shape != shape:
            _bad(f"{name}: host output must be a writable C-contiguous float32 array of shape {shape}")
        return out.ctypes.data
    return _need(out, name, torch.float32, shape, ctx.device).data_ptr()


def compute_u_sigma_root(U, S, n: int, rank: int, dim: int, normalize: bool, ctx: Optional[Context] = None):
    """emb[:, j] = sqrt(S_j) U[:, j], optional row normalisation (PathEmbed.h:371-408)."""
    torch = _torch()
    ctx = ctx or context()
    _need(U, "compute_u_sigma_root: U", torch.float32, (n, rank), ctx.device)
    emb = torch.empty((n, dim), dtype=torch.float32, device=U.device)
    s = None if S is None else np.ascontiguousarray(S, dtype=np.float32)
    ctx.check(ctx.lib.lightne_u_sigma_root(ctx.ptr, _dev(U), s.ctypes.data if s is not None else None, n, rank, dim,
                                           int(bool(normalize)), _dev(emb)))
    return emb


def NetSMF(graph: DeviceGraph, walks_per_edge: int = 100, walk_len: int = 10, upper: bool = False,
           sample: bool = False, rank: int = 256, dim: int = 128, analyze: bool = False, table_size: int = 0,
           negative: int = 1, normalize: bool = False, sample_ratio: float = 1.0, mem_ratio: float = 1.0,
           random_project_only: bool = False, seed: int = 0, svd_seed: int = 777, power_iters: int = 1,
           precision: int = FP32, max_table_bytes: int = 0, sparse_project: bool = False,
           sparse_project_s: float = 100.0, table_layout: int = 0, csr_parts: int = 0,
           value_cache: int = 0, out=None, ctx: Optional[Context] = None):
    """Stages 1 + 2 (PathEmbed.h:494-563): returns (emb n x dim, S, stats).
    ``sparse_project_s`` defaults to 100 as LightNE.cc:50 does.  ``out``
    (n x dim float32) receives emb: a device tensor, or with the FP64_LEAN
    form also a host numpy array."""
    del walks_per_edge, analyze
    torch = _torch()
    ctx = ctx or context()
    sp = _sparsify_params(seed, walk_len, upper, sample, table_size, negative, sample_ratio, mem_ratio,
                          max_table_bytes, table_layout, csr_parts, value_cache)
    rp = RsvdParams(rank=rank, power_iters=power_iters, precision=precision,
                    random_project_only=int(bool(random_project_only)), upper=int(bool(upper)), seed=svd_seed,
                    sparse_project=int(bool(sparse_project)), sparse_project_s=sparse_project_s)
    emb = out if out is not None else torch.empty((graph.n, dim), dtype=torch.float32, device=f"cuda:{ctx.device}")
    eptr = _out_ptr(emb, (graph.n, dim), "NetSMF: out", ctx, precision == FP64_LEAN)
    S = np.zeros(rank, np.float32)
    st = NetsmfStats()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_netsmf(ctx.ptr, C.byref(g), C.byref(sp), C.byref(rp), dim, int(bool(normalize)),
                                     eptr, S.ctypes.data, C.byref(st)))
    return emb, S, _stats_dict(st)


def ne_zhang_matrix(graph: DeviceGraph, ctx: Optional[Context] = None) -> "CSR":
    """The ne_zhang_et_al log matrix (PathEmbed.h:412-444) on the graph's pattern."""
    ctx = ctx or context()
    h = C.c_void_p()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_ne_zhang_matrix(ctx.ptr, C.byref(g), C.byref(h)))
    return CSR(ctx, h)


def NE_Zhang_et_al(graph: DeviceGraph, rank: int = 256, dim: int = 128, svd_seed: int = 777, power_iters: int = 1,
                   precision: int = FP32, ctx: Optional[Context] = None):
    """path_embed::NE_Zhang_et_al (PathEmbed.h:410-463): returns (emb n x dim,
    S, stats); rows of emb are L2-normalised as in the reference."""
    torch = _torch()
    ctx = ctx or context()
    rp = RsvdParams(rank=rank, power_iters=power_iters, precision=precision, random_project_only=0, upper=0,
                    seed=svd_seed)
    emb = torch.empty((graph.n, dim), dtype=torch.float32, device=f"cuda:{ctx.device}")
    S = np.zeros(rank, np.float32)
    st = RsvdStats()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_ne_zhang(ctx.ptr, C.byref(g), C.byref(rp), dim, _dev(emb), S.ctypes.data,
                                       C.byref(st)))
    return emb, S, _stats_dict(st)


# ---------------------------------------------------------------------------
# stage 3
# ---------------------------------------------------------------------------
def propagate(a, graph: DeviceGraph, order: int = 10, mu: float = 0.2, theta: float = 0.5,
              precision: int = FP32, aggregate: bool = True, ctx: Optional[Context] = None):
    """y = a - sum_i c_i T_i a and z = (A + I) y (SpectralPropagation.h:94-221).
    ``a`` (n x dim float32 on the device) is not modified.  Returns (y, z, stats)."""
    torch = _torch()
    ctx = ctx or context()
    if getattr(a, "ndim", 0) != 2:
        _bad("propagate: a must be an n x dim tensor")
    _need(a, "propagate: a", torch.float32, (graph.n, a.shape[1]), ctx.device)
    n, dim = a.shape
    y = torch.empty_like(a)
    z = torch.empty_like(a) if aggregate else None
    p = PropParams(order=order, mu=mu, theta=theta, precision=precision)
    st = PropStats()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_propagate(ctx.ptr, C.byref(g), _dev(a), dim, C.byref(p), _dev(y),
                                        _dev(z) if z is not None else None, C.byref(st)))
    return y, z, _stats_dict(st)


def get_embedding_dense(mat, n: int, rank: int, dim: int, ctx: Optional[Context] = None):
    """Thin SVD of mat (n x rank), emb = rownorm(U[:, :dim] sqrt(S)) (SpectralPropagation.h:16-43).
    Returns (emb, S)."""
    torch = _torch()
    ctx = ctx or context()
    _need(mat, "get_embedding_dense: mat", torch.float32, (n, rank), ctx.device)
    emb = torch.empty((n, dim), dtype=torch.float32, device=mat.device)
    S = np.zeros(rank, np.float32)
    ctx.check(ctx.lib.lightne_embed_dense(ctx.ptr, _dev(mat), n, rank, dim, _dev(emb), S.ctypes.data))
    return emb, S


def chebyshev_expansion(a, graph: DeviceGraph, rank: int, dim: int, order: int = 10, mu: float = 0.2,
                        s: float = 0.5, precision: int = FP32, out=None, ctx: Optional[Context] = None):
    """Stage 3 (SpectralPropagation.h:45-229).  Returns (emb, stats).  ``out``
    (n x dim FP32) receives the result and may be ``a`` itself: the input is
    read in full before the output is written."""
    torch = _torch()
    if rank != dim:
        raise LightNEError(_lib.ERR_BAD_ARG, "chebyshev_expansion: the runner passes rank == dim (LightNE.cc:114)")
    ctx = ctx or context()
    n = graph.n
    _need(a, "chebyshev_expansion: a", torch.float32, (n, dim), ctx.device)
    emb = out if out is not None else torch.empty((n, dim), dtype=torch.float32, device=a.device)
    _need(emb, "chebyshev_expansion: out", torch.float32, (n, dim), ctx.device)
    p = PropParams(order=order, mu=mu, theta=s, precision=precision)
    st = PropStats()
    g = graph.desc()
    ctx.check(ctx.lib.lightne_chebyshev_expansion(ctx.ptr, C.byref(g), _dev(a), dim, C.byref(p), _dev(emb),
                                                  C.byref(st)))
    return emb, _stats_dict(st)


# ---------------------------------------------------------------------------
# dense panel building blocks
# ---------------------------------------------------------------------------
def dense_gram(X, Y=None, ctx: Optional[Context] = None) -> np.ndarray:
    """X^T Y in FP64 (device FP32 panels)."""
    torch = _torch()
    ctx = ctx or context()
    sym = Y is None
    Y = X if sym else Y
    if getattr(X, "ndim", 0) != 2 or getattr(Y, "ndim", 0) != 2:
        _bad("dense_gram: X and Y must be 2-D tensors")
    n, k1 = X.shape
    k2 = Y.shape[1]
    _need(X, "dense_gram: X", torch.float32, (n, k1), ctx.device)
    _need(Y, "dense_gram: Y", torch.float32, (n, k2), ctx.device)
    G = np.zeros((k1, k2), np.float64)
    ctx.check(ctx.lib.lightne_dense_gram(ctx.ptr, _dev(X), _dev(Y), n, k1, k2, int(sym), G.ctypes.data))
    return G


def dense_panel_gemm(X, W: np.ndarray, normalize_rows: bool = False, out=None, ctx: Optional[Context] = None):
    """X (n x k1, device FP32) times W (k1 x k2 host) -> device FP32 n x k2."""
    torch = _torch()
    ctx = ctx or context()
    if getattr(X, "ndim", 0) != 2:
        _bad("dense_panel_gemm: X must be a 2-D tensor")
    n, k1 = X.shape
    _need(X, "dense_panel_gemm: X", torch.float32, (n, k1), ctx.device)
    W = np.ascontiguousarray(W, np.float64)
    if W.ndim != 2 or W.shape[0] != k1:
        _bad(f"dense_panel_gemm: W must be {k1} x k2, got {W.shape}")
    k2 = W.shape[1]
    if out is None:
        out = torch.empty((n, k2), dtype=torch.float32, device=X.device)
    _need(out, "dense_panel_gemm: out", torch.float32, (n, k2), ctx.device)
    ctx.check(ctx.lib.lightne_dense_panel_gemm(ctx.ptr, _dev(X), n, k1, W.ctypes.data, k2, _dev(out),
                                               int(bool(normalize_rows))))
    return out


def dense_cholqr(Y, ctx: Optional[Context] = None):
    """CholeskyQR2 in place: returns (Q (= Y), R, passes)."""
    torch = _torch()
    ctx = ctx or context()
    if getattr(Y, "ndim", 0) != 2:
        _bad("dense_cholqr: Y must be a 2-D tensor")
    n, k = Y.shape
    _need(Y, "dense_cholqr: Y", torch.float32, (n, k), ctx.device)
    R = np.zeros((k, k), np.float64)
    passes = C.c_int32()
    ctx.check(ctx.lib.lightne_dense_cholqr(ctx.ptr, _dev(Y), n, k, R.ctypes.data, C.byref(passes)))
    return Y, R, passes.value


def bessel_i(k: int, x: float) -> float:
    """Modified Bessel I_k(x) as used for the Chebyshev weights."""
    return load().lightne_bessel_i(k, x)


# ---------------------------------------------------------------------------
# differentiation form
# ---------------------------------------------------------------------------
def propagate_backward(a, grad_out, graph: DeviceGraph, order: int = 10, mu: float = 0.2, theta: float = 0.5,
                       aggregate: bool = False, checkpoint_every: int = 3, ctx: Optional[Context] = None):
    """dL/da, dL/dmu, dL/dtheta of y (or z = (A+I) y when ``aggregate``) given
    dL/dy (dL/dz) -- the hand-written adjoint of the Chebyshev recurrence."""
    torch = _torch()
    ctx = ctx or context()
    if getattr(a, "ndim", 0) != 2:
        _bad("propagate_backward: a must be an n x dim tensor")
    _need(a, "propagate_backward: a", torch.float32, (graph.n, a.shape[1]), ctx.device)
    g_out = _need(grad_out, "propagate_backward: grad_out", torch.float32, tuple(a.shape), ctx.device)
    n, dim = a.shape
    grad_a = torch.empty_like(a)
    gm, gt = C.c_double(), C.c_double()
    p = PropParams(order=order, mu=mu, theta=theta, precision=FP32)
    g = graph.desc()
    ctx.check(ctx.lib.lightne_propagate_backward(ctx.ptr, C.byref(g), _dev(a), dim, C.byref(p), int(bool(aggregate)),
                                                 _dev(g_out), _dev(grad_a), C.byref(gm), C.byref(gt),
                                                 checkpoint_every))
    return grad_a, gm.value, gt.value
