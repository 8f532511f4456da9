"""Writer of Ligra's compressed graph format (the ``-c`` input), host side.

The reference reads these files (ligra/graph_io.h:344-377) but ships no
writer; this one produces the byte-pd-amortized encoding that its reader and
decoder expect (ligra/encodings/byte_pd_amortized.h:124-160,1262-1318), so
any symmetric CSR can be fed to ``lightne -c`` or to
``read_compressed_graph``.  Vectorised numpy, no per-edge Python loop.

File image (little endian):
    int64  n, m, total_space
    uint64 offsets[n + 1]   byte offset of each vertex record in the edge section
    uint32 degrees[n]
    uint8  edges[total_space]
Record of a vertex v of degree d > 0: uint32 d, uint32 block_off[nb - 1]
(relative to the record), then nb = ceil(d / 64) blocks; block b holds uint32
64 b, the first neighbour as a signed difference from v (first byte: 6
magnitude bits, 0x40 negative, 0x80 more), then each further neighbour as the
unsigned LEB128 difference (mod 2^32) from the previous one.  Vertices of
degree 0 have no record.
"""
from __future__ import annotations

import numpy as np

BLOCK = 64  # PARALLEL_DEGREE (ligra/macros.h:85)


def _groups7(x: np.ndarray) -> np.ndarray:
    """Number of 7-bit groups the LEB128 loop writes for x (0 for x == 0)."""
    g = np.zeros(x.shape, np.int64)
    t = x.copy()
    while True:
        nz = t > 0
        if not nz.any():
            return g
        g += nz
        t >>= np.uint64(7)


def _put32(buf: np.ndarray, pos: np.ndarray, val: np.ndarray) -> None:
    val = val.astype(np.uint64)
    for t in range(4):
        buf[pos + t] = ((val >> np.uint64(8 * t)) & np.uint64(0xFF)).astype(np.uint8)


def encode(offsets: np.ndarray, edges: np.ndarray) -> np.ndarray:
    """The whole compressed file for the symmetric CSR (offsets int64[n+1],
    edges uint32[m]), neighbours in the given order, as a uint8 array."""
    off = np.ascontiguousarray(offsets, np.int64)
    adj = np.ascontiguousarray(edges, np.uint32).astype(np.int64)
    n = off.shape[0] - 1
    m = adj.shape[0]
    if n <= 0 or off[0] != 0 or off[-1] != m or np.any(np.diff(off) < 0):
        raise ValueError("offsets must be a non-decreasing int64[n+1] from 0 to m")
    if m and (adj.max() >= n):
        raise ValueError("neighbour id >= n")
    deg = np.diff(off)
    row = np.repeat(np.arange(n, dtype=np.int64), deg)
    pos_in_row = np.arange(m, dtype=np.int64) - off[row]
    blk = pos_in_row // BLOCK
    first = (pos_in_row % BLOCK) == 0
    # first edge of a block: |target - source| and its sign (:124-160)
    diff = adj - row
    mag = np.abs(diff).astype(np.uint64)
    neg = diff < 0
    # other edges: difference from the previous neighbour, mod 2^32 (:1301)
    prev = np.zeros(m, np.int64)
    if m > 1:
        prev[1:] = adj[:-1]
    delta = ((adj - prev) & 0xFFFFFFFF).astype(np.uint64)
    if np.any(~first & (delta == 0)):
        raise ValueError("a repeated neighbour inside a block has no encoding (compressEdge writes 0 bytes for 0)")
    high = mag >> np.uint64(6)
    nbytes = np.where(first, 1 + _groups7(np.where(first, high, 0)), _groups7(np.where(first, 0, delta)))

    nb = (deg + BLOCK - 1) // BLOCK
    head = 4 * nb  # uint32 degree + nb - 1 block offsets
    per_row = np.bincount(row, weights=nbytes, minlength=n).astype(np.int64) if m else np.zeros(n, np.int64)
    rec = np.where(deg > 0, head + 4 * nb + per_row, 0)
    voff = np.zeros(n + 1, np.int64)
    np.cumsum(rec, out=voff[1:])
    total = int(voff[-1])

    # byte position of every edge: record start + table + (blk + 1) block
    # indices + the bytes of the earlier edges of the row
    csum = np.zeros(m + 1, np.int64)
    np.cumsum(nbytes, out=csum[1:])
    before = csum[:-1] - csum[off[row]]
    epos = voff[row] + head[row] + 4 * (blk + 1) + before

    e = np.zeros(total, np.uint8)
    has = deg > 0
    _put32(e, voff[:-1][has], deg[has])
    fi = np.nonzero(first)[0]
    bpos = epos[fi] - 4
    _put32(e, bpos, BLOCK * blk[fi])
    later = fi[blk[fi] > 0]
    _put32(e, voff[row[later]] + 4 * blk[later], epos[later] - 4 - voff[row[later]])
    for k in range(6):
        sel = np.nonzero(nbytes > k)[0]
        if sel.size == 0:
            break
        f = first[sel]
        last = nbytes[sel] == k + 1
        if k == 0:
            low = np.where(f, (mag[sel] & np.uint64(0x3F)) | (neg[sel].astype(np.uint64) << np.uint64(6)),
                           delta[sel] & np.uint64(0x7F))
        else:
            src = np.where(f, high[sel] >> np.uint64(7 * (k - 1)), delta[sel] >> np.uint64(7 * k))
            low = src & np.uint64(0x7F)
        byte = low | np.where(last, np.uint64(0), np.uint64(0x80))
        e[epos[sel] + k] = byte.astype(np.uint8)

    hdr = np.array([n, m, total], np.int64).view(np.uint8)
    return np.concatenate([hdr, voff.astype(np.uint64).view(np.uint8), deg.astype(np.uint32).view(np.uint8), e])


def write(path: str, offsets: np.ndarray, edges: np.ndarray) -> None:
    encode(offsets, edges).tofile(path)


def header(image: np.ndarray) -> tuple[int, int]:
    """(n, m) from a compressed file image."""
    if image.shape[0] < 24:
        raise ValueError("compressed graph: file shorter than its header")
    n, m = np.frombuffer(image[:16].tobytes(), np.int64)
    return int(n), int(m)
