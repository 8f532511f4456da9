// Device decoder for Ligra's compressed graph format (the -c input of the
// reference runner): the byte-pd-amortized encoding of
// ligra/encodings/byte_pd_amortized.h in the file layout read by
// gbbs_io::read_compressed_symmetric_graph (ligra/graph_io.h:344-377).
//
// File image (little endian, graph_io.h:350-360):
//   int64  n, m, total_space
//   uint64 offsets[n + 1]   byte offset of each vertex's record in `edges`
//   uint32 degrees[n]
//   uint8  edges[]
// Record of a vertex v of degree d > 0 (byte_pd_amortized.h:1262-1318):
//   uint32 d (the "virtual degree"), uint32 block_off[nb - 1] relative to the
//   record, nb = 1 + (d - 1) / 64; then nb blocks of at most 64 edges.  Block
//   b (block 0 right after the table, block b > 0 at block_off[b - 1]) holds
//   uint32 64 b (its first edge index), the first neighbour as a difference
//   from v (6 magnitude bits, 0x40 = negative, 0x80 = more bytes of 7 bits,
//   :86-100), then each further neighbour as the unsigned LEB128 difference
//   from the previous one (:105-119).
//
// One work item decodes one 64-edge block; it finds its vertex by a binary
// search over the exclusive scan of per-vertex block counts, so a vertex of
// degree 10^6 is spread over 15625 items instead of serialising one thread.
// The output is the symmetric CSR with neighbours in decode order (the order
// map_edges and decode visit them).  Header words sit at arbitrary byte
// offsets and are assembled from bytes; every read is checked against the
// image size and every decoded id against n, so a malformed image sets an
// error bit instead of reading out of range.
#include "lightne_internal.h"

namespace {

constexpr int BP_BLOCK = 64;  // PARALLEL_DEGREE (ligra/macros.h:85)

enum : uint32_t {
  BP_TRUNCATED = 1u,    // a record or block reaches past the image
  BP_VDEGREE = 2u,      // virtual degree differs from the degree array
  BP_BLOCKHDR = 4u,     // a block's first-edge index is not 64 b
  BP_NEIGHBOUR = 8u,    // a decoded neighbour id is >= n
  BP_VARINT = 16u,      // a varint longer than 5 bytes
};

__device__ __forceinline__ bool rd_u32(const uint8_t* __restrict__ e, uint64_t nbytes, uint64_t pos, uint32_t& v) {
  if (pos >= nbytes || nbytes - pos < 4) return false;
  v = (uint32_t)e[pos] | ((uint32_t)e[pos + 1] << 8) | ((uint32_t)e[pos + 2] << 16) | ((uint32_t)e[pos + 3] << 24);
  return true;
}

// per vertex: number of 64-edge blocks, and the record header checks
__global__ __launch_bounds__(256) void k_bp_count(const uint64_t* __restrict__ voff, const uint32_t* __restrict__ deg,
                                                  const uint8_t* __restrict__ e, uint64_t ebytes, int64_t n,
                                                  uint32_t* __restrict__ nblk, uint32_t* __restrict__ err) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t d = deg[v];
    uint32_t nb = d ? 1u + (d - 1u) / BP_BLOCK : 0u;
    uint32_t bad = 0;
    if (d) {
      const uint64_t base = voff[v];
      uint32_t vd = 0;
      if (!rd_u32(e, ebytes, base, vd)) {
        bad |= BP_TRUNCATED;
      } else if (vd != d) {
        bad |= BP_VDEGREE;
      } else if (base + 4ull * (nb + 1) > ebytes) {  // block table + block 0's index
        bad |= BP_TRUNCATED;
      }
    }
    if (bad) {
      atomicOr(err, bad);
      nb = 0;
    }
    nblk[v] = nb;
  }
}

template <bool FIRST>
__device__ __forceinline__ bool eat(const uint8_t* __restrict__ e, uint64_t ebytes, uint64_t& pos, uint32_t& out,
                                    bool& neg, uint32_t& bad) {
  // first edge: 6 bits + sign in the first byte (byte_pd_amortized.h:86-100);
  // later edges: 7 bits per byte (:105-119); 0x80 = another byte follows
  if (pos >= ebytes) {
    bad |= BP_TRUNCATED;
    return false;
  }
  uint32_t b = e[pos++];
  uint32_t val;
  int shift;
  if (FIRST) {
    val = b & 0x3fu;
    neg = (b & 0x40u) != 0;
    shift = 6;
  } else {
    val = b & 0x7fu;
    shift = 7;
  }
  int nbytes = 1;
  while (b & 0x80u) {
    if (nbytes == 5) {
      bad |= BP_VARINT;
      return false;
    }
    if (pos >= ebytes) {
      bad |= BP_TRUNCATED;
      return false;
    }
    b = e[pos++];
    val += (b & 0x7fu) << shift;  // bits past 32 drop, as in the 32-bit reference decode
    shift += 7;
    ++nbytes;
  }
  out = val;
  return true;
}

// one item per 64-edge block
__global__ __launch_bounds__(256) void k_bp_decode(const uint64_t* __restrict__ voff,
                                                   const uint32_t* __restrict__ deg, const uint8_t* __restrict__ e,
                                                   uint64_t ebytes, int64_t n, const uint64_t* __restrict__ bscan,
                                                   const uint64_t* __restrict__ csr_off,
                                                   uint32_t* __restrict__ adj, uint32_t* __restrict__ err) {
  const uint64_t total = bscan[n];
  for (uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += (uint64_t)gridDim.x * blockDim.x) {
    // largest v with bscan[v] <= item: bscan[v + 1] > item, so v has >= 1 block
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo + 1) >> 1);
      if (bscan[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const int64_t v = lo;
    const uint32_t b = (uint32_t)(item - bscan[v]);
    const uint32_t d = deg[v];
    const uint32_t nb = 1u + (d - 1u) / BP_BLOCK;
    const uint64_t base = voff[v];
    uint32_t bad = 0;
    uint64_t pos = base + 4ull * nb;  // block 0 follows the virtual degree and the nb - 1 block offsets
    if (b > 0) {
      uint32_t rel = 0;
      rd_u32(e, ebytes, base + 4ull * b, rel);  // in range: checked by k_bp_count
      pos = base + rel;
    }
    uint32_t start = 0;
    if (!rd_u32(e, ebytes, pos, start)) bad |= BP_TRUNCATED;
    else if (start != b * (uint32_t)BP_BLOCK) bad |= BP_BLOCKHDR;
    pos += 4;
    if (!bad) {
      const uint32_t cnt = min((uint32_t)BP_BLOCK, d - start);
      uint32_t* out = adj + csr_off[v] + start;
      uint32_t mag = 0;
      bool neg = false;
      if (eat<true>(e, ebytes, pos, mag, neg, bad)) {
        uint32_t ngh = neg ? (uint32_t)v - mag : (uint32_t)v + mag;
        if (ngh >= (uint64_t)n) bad |= BP_NEIGHBOUR;
        else out[0] = ngh;
        for (uint32_t j = 1; j < cnt && !bad; ++j) {
          uint32_t delta = 0;
          bool unused = false;
          if (!eat<false>(e, ebytes, pos, delta, unused, bad)) break;
          ngh += delta;
          if (ngh >= (uint64_t)n) bad |= BP_NEIGHBOUR;
          else out[j] = ngh;
        }
      }
    }
    if (bad) atomicOr(err, bad);
  }
}

const char* bp_reason(uint32_t err) {
  if (err & BP_TRUNCATED) return "a vertex record reaches past the end of the file";
  if (err & BP_VDEGREE) return "a record's degree differs from the degree array";
  if (err & BP_BLOCKHDR) return "a block does not start at a multiple of 64 edges";
  if (err & BP_NEIGHBOUR) return "a neighbour id is >= n";
  return "a varint longer than 5 bytes";
}

}  // namespace

extern "C" {

int lightne_graph_decode_compressed(lightne_ctx* ctx, const void* image, uint64_t image_bytes, int64_t n,
                                    int64_t m, int64_t* offsets_out, uint32_t* edges_out) {
  if (!ctx || !image || !offsets_out || (m > 0 && !edges_out))
    return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_graph_decode_compressed: NULL argument");
  if ((uintptr_t)image % 8 != 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_graph_decode_compressed: image must be 8-byte aligned");
  if (n < 0 || n >= ((int64_t)1 << 32) || m < 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_graph_decode_compressed: n must be in [0, 2^32), m >= 0");
  const uint64_t head = 24 + 8 * (uint64_t)(n + 1) + 4 * (uint64_t)n;
  if (image_bytes < head)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_GRAPH, "compressed graph: file of %llu bytes is shorter than its header",
                    (unsigned long long)image_bytes);
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const auto* img = (const uint8_t*)image;
  int64_t hdr[2] = {0, 0};
  LN_HIP(ctx, hipMemcpyAsync(hdr, img, sizeof(hdr), hipMemcpyDeviceToHost, st));
  LN_HIP(ctx, hipStreamSynchronize(st));
  if (hdr[0] != n || hdr[1] != m)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "compressed graph: header says n = %lld, m = %lld; caller passed %lld, %lld",
                    (long long)hdr[0], (long long)hdr[1], (long long)n, (long long)m);
  const auto* voff = (const uint64_t*)(img + 24);
  const auto* deg = (const uint32_t*)(img + 24 + 8 * (n + 1));
  const uint8_t* e = img + head;
  const uint64_t ebytes = image_bytes - head;
  auto* nblk = (uint32_t*)ctx_buf(ctx, "bp_nblk", (size_t)std::max<int64_t>(n, 1) * 4);
  auto* bscan = (uint64_t*)ctx_buf(ctx, "bp_bscan", (size_t)(n + 1) * 8);
  auto* err = (uint32_t*)ctx_buf(ctx, "bp_err", 4);
  if (!nblk || !bscan || !err) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemsetAsync(err, 0, 4, st));
  if (n > 0) {
    hipLaunchKernelGGL(k_bp_count, dim3(grid_blocks(n, 256)), dim3(256), 0, st, voff, deg, e, ebytes, n, nblk, err);
    LN_LAUNCH(ctx);
  }
  // CSR offsets = exclusive scan of the degrees; the sum must be m
  LN_CHECK(scan_u32_to_u64(ctx, deg, (uint64_t*)offsets_out, n));
  LN_CHECK(scan_u32_to_u64(ctx, nblk, bscan, n));
  uint64_t tot[2] = {0, 0};
  uint32_t herr = 0;
  LN_HIP(ctx, hipMemcpyAsync(&tot[0], offsets_out + n, 8, hipMemcpyDeviceToHost, st));
  LN_HIP(ctx, hipMemcpyAsync(&tot[1], bscan + n, 8, hipMemcpyDeviceToHost, st));
  LN_HIP(ctx, hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
  LN_HIP(ctx, hipStreamSynchronize(st));
  if (herr) return ctx_fail(ctx, LIGHTNE_ERR_BAD_GRAPH, "compressed graph: %s", bp_reason(herr));
  if (tot[0] != (uint64_t)m)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_GRAPH, "compressed graph: degrees sum to %llu, header m = %lld",
                    (unsigned long long)tot[0], (long long)m);
  if (tot[1] > 0) {
    hipLaunchKernelGGL(k_bp_decode, dim3(grid_blocks((int64_t)tot[1], 256)), dim3(256), 0, st, voff, deg, e, ebytes,
                       n, (const uint64_t*)bscan, (const uint64_t*)offsets_out, edges_out, err);
    LN_LAUNCH(ctx);
    LN_HIP(ctx, hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
    LN_HIP(ctx, hipStreamSynchronize(st));
    if (herr) return ctx_fail(ctx, LIGHTNE_ERR_BAD_GRAPH, "compressed graph: %s", bp_reason(herr));
  }
  return LIGHTNE_OK;
}

}  // extern "C"
