// Device primitives: deterministic exclusive scan and a stable LSD radix sort
// of (u64 key, 32-bit payload) pairs.  Written for 64-wide wavefronts: peer
// detection inside a wave uses 8 ballots over the digit bits, cross-wave
// ranking goes through LDS.
#include "lightne_internal.h"

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// exclusive scan of one value per thread across a 256-thread block
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* s_wave, uint64_t& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t inc = wave_incl_scan(v);
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  uint64_t off = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    uint64_t t = s_wave[w];
    if (w < wid) off += t;
    total += t;
  }
  __syncthreads();
  return off + inc - v;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const T* __restrict__ in, int64_t n,
                                                              uint64_t* __restrict__ sums) {
  __shared__ uint64_t s_wave[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint64_t acc = 0;
#pragma unroll
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    int64_t i = base + (int64_t)it * SCAN_THREADS + threadIdx.x;
    if (i < n) acc += (uint64_t)in[i];
  }
  uint64_t total;
  (void)block_excl_scan(acc, s_wave, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const T* __restrict__ in, int64_t n,
                                                             uint64_t* __restrict__ out,
                                                             const uint64_t* __restrict__ block_off) {
  __shared__ uint64_t s_data[SCAN_TILE];
  __shared__ uint64_t s_wave[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    int idx = it * SCAN_THREADS + threadIdx.x;
    int64_t i = base + idx;
    s_data[idx] = (i < n) ? (uint64_t)in[i] : 0;
  }
  __syncthreads();
  uint64_t loc[SCAN_ITEMS];
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    loc[j] = sum;
    sum += s_data[threadIdx.x * SCAN_ITEMS + j];
  }
  uint64_t total;
  uint64_t off = block_excl_scan(sum, s_wave, total) + (block_off ? block_off[blockIdx.x] : 0);
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s_data[threadIdx.x * SCAN_ITEMS + j] = loc[j] + off;
  __syncthreads();
#pragma unroll
  for (int it = 0; it < SCAN_ITEMS; ++it) {
    int idx = it * SCAN_THREADS + threadIdx.x;
    int64_t i = base + idx;
    if (i < n) out[i] = s_data[idx];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    uint64_t bo = block_off ? block_off[blockIdx.x] : 0;
    out[n] = bo + total;
  }
}

template <typename T>
int scan_impl(lightne_ctx* ctx, const T* in, uint64_t* out, int64_t n, int depth) {
  if (n <= 0) {
    LN_HIP(ctx, hipMemsetAsync(out, 0, sizeof(uint64_t), ctx->stream));
    return LIGHTNE_OK;
  }
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, in, n, out,
                       (const uint64_t*)nullptr);
    LN_LAUNCH(ctx);
    return LIGHTNE_OK;
  }
  char name_s[32], name_o[32];
  snprintf(name_s, sizeof(name_s), "scan_sums%d", depth);
  snprintf(name_o, sizeof(name_o), "scan_offs%d", depth);
  auto* sums = (uint64_t*)ctx_buf(ctx, name_s, nb * sizeof(uint64_t));
  auto* offs = (uint64_t*)ctx_buf(ctx, name_o, (nb + 1) * sizeof(uint64_t));
  if (!sums || !offs) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, n, sums);
  LN_LAUNCH(ctx);
  LN_CHECK(scan_impl<uint64_t>(ctx, sums, offs, nb, depth + 1));
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, ctx->stream, in, n, out,
                     (const uint64_t*)offs);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
// 8192-element tiles (512 threads x 16): digit runs of ~32 entries per tile
// keep the scattered writes long
constexpr int RS_THREADS = 512;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ uint32_t s_h[256];
  if (threadIdx.x < 256) s_h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int it = 0; it < RS_ITEMS; ++it) {
    int64_t i = base + (int64_t)it * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 256) hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

// Stable scatter of one 8-bit digit.  Each wave ranks its own contiguous
// 1024-element part of the tile (peers by 8 ballots per item, a running
// per-digit count in LDS that only this wave touches, so no block barrier per
// item); one block pass turns the waves' counts into LDS offsets; the tile is
// staged in digit order and each digit run is written to its global offset
// with consecutive lanes on consecutive addresses.
template <typename V>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                           const V* __restrict__ vin, uint64_t* __restrict__ kout,
                                                           V* __restrict__ vout, int64_t n, int shift,
                                                           const uint64_t* __restrict__ offs,
                                                           const uint32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ uint64_t s_keys[RS_TILE];
  __shared__ V s_vals[RS_TILE];
  __shared__ uint64_t s_base[256];
  __shared__ uint32_t s_dstart[256];
  __shared__ uint32_t s_wcnt[RS_WAVES][256];
  __shared__ uint32_t s_scan[RS_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // tile-local start of every digit (threads 0..255, one digit each):
  // exclusive scan of this tile's histogram
  {
    const uint32_t c = tid < 256 ? hist[(int64_t)tid * ntiles + blockIdx.x] : 0u;
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) s_scan[wid] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wid; ++w) off += s_scan[w];
    if (tid < 256) {
      s_dstart[tid] = off + incl - c;
      s_base[tid] = offs[(int64_t)tid * ntiles + blockIdx.x];
    }
  }
  for (int i = tid; i < RS_WAVES * 256; i += RS_THREADS) (&s_wcnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const int nvalid = (int)min<int64_t>(RS_TILE, n - base);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t kr[RS_ITEMS];
  V vr[RS_ITEMS];
  uint32_t pr[RS_ITEMS];  // digit | (position among the wave's elements of that digit) << 8
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int li = wid * (RS_ITEMS * 64) + it * 64 + lane;
    const bool valid = li < nvalid;
    const uint64_t k = valid ? kin[base + li] : 0;
    const V v = valid ? vin[base + li] : V(0);
    const uint32_t d = (uint32_t)((k >> shift) & 255u);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    // every lane reads the wave's running count before the leader adds to it
    // (one wave's LDS operations complete in order)
    const uint32_t before = s_wcnt[wid][d];
    if (valid && rank == 0) s_wcnt[wid][d] = before + (uint32_t)__popcll(peers);
    kr[it] = k;
    vr[it] = v;
    pr[it] = valid ? (d | ((before + rank) << 8)) : ~0u;
  }
  __syncthreads();
  // per digit: LDS start of each wave's run = digit start + earlier waves' counts
  if (tid < 256) {
    uint32_t run = s_dstart[tid];
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const uint32_t c = s_wcnt[w][tid];
      s_wcnt[w][tid] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    if (pr[it] != ~0u) {
      const uint32_t lp = s_wcnt[wid][pr[it] & 255u] + (pr[it] >> 8);
      s_keys[lp] = kr[it];
      s_vals[lp] = vr[it];
    }
  }
  __syncthreads();
  for (int li = tid; li < nvalid; li += RS_THREADS) {
    const uint64_t k = s_keys[li];
    const uint32_t d = (uint32_t)((k >> shift) & 255u);
    const uint64_t pos = s_base[d] + (uint64_t)(li - (int)s_dstart[d]);
    kout[pos] = k;
    vout[pos] = s_vals[li];
  }
}

template <typename V>
int radix_sort_impl(lightne_ctx* ctx, uint64_t* keys, V* vals, int64_t n, int bits, uint64_t* k2, V* v2, int lo,
                    uint64_t** k_res, V** v_res) {
  if (k_res) *k_res = keys;
  if (v_res) *v_res = vals;
  if (n <= 1 || bits <= 0) return LIGHTNE_OK;
  const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  if (!k2 || !v2) {
    k2 = (uint64_t*)ctx_buf(ctx, "rs_keys", n * sizeof(uint64_t));
    v2 = (V*)ctx_buf(ctx, "rs_vals", n * sizeof(V));
  }
  auto* hist = (uint32_t*)ctx_buf(ctx, "rs_hist", 256 * ntiles * sizeof(uint32_t));
  auto* offs = (uint64_t*)ctx_buf(ctx, "rs_offs", (256 * ntiles + 1) * sizeof(uint64_t));
  if (!k2 || !v2 || !hist || !offs) return LIGHTNE_ERR_OOM;
  uint64_t *ka = keys, *kb = k2;
  V *va = vals, *vb = v2;
  for (int shift = lo; shift < lo + bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, ctx->stream, ka, n, shift, hist,
                       ntiles);
    LN_LAUNCH(ctx);
    LN_CHECK(scan_u32_to_u64(ctx, hist, offs, 256 * ntiles));
    hipLaunchKernelGGL(k_rs_scatter<V>, dim3((unsigned)ntiles), dim3(RS_THREADS), 0, ctx->stream, ka, va, kb, vb, n,
                       shift, (const uint64_t*)offs, (const uint32_t*)hist, ntiles);
    LN_LAUNCH(ctx);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (k_res) {  // the caller reads the result where the last pass left it
    *k_res = ka;
    *v_res = va;
  } else if (ka != keys) {
    LN_HIP(ctx, hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
    LN_HIP(ctx, hipMemcpyAsync(vals, va, n * sizeof(V), hipMemcpyDeviceToDevice, ctx->stream));
  }
  return LIGHTNE_OK;
}

}  // namespace

int scan_u32_to_u64(lightne_ctx* ctx, const uint32_t* in, uint64_t* out, int64_t n) {
  return scan_impl<uint32_t>(ctx, in, out, n, 0);
}
int radix_sort_pairs(lightne_ctx* ctx, uint64_t* keys, float* vals, int64_t n, int bits, uint64_t* k_tmp,
                     float* v_tmp, int lo_bit, uint64_t** k_res, float** v_res) {
  return radix_sort_impl<float>(ctx, keys, vals, n, bits, k_tmp, v_tmp, lo_bit, k_res, v_res);
}
int radix_sort_pairs_u32(lightne_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bits) {
  return radix_sort_impl<uint32_t>(ctx, keys, vals, n, bits, nullptr, nullptr, 0, nullptr, nullptr);
}
