// Device primitives: deterministic exclusive scan and a stable LSD radix sort
// of (u64 key, 32-bit payload) pairs.  Written for 64-wide wavefronts: peer
// detection inside a wave uses 8 ballots over the digit bits, cross-wave
// ranking goes through LDS, cross-tile offsets through decoupled look-back.
#include "lightne_internal.h"

#include <algorithm>
#include <type_traits>

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
// radix sort: one histogram read for all digits, then one kernel per 8-bit
// digit that ranks a tile, finds its global offsets by decoupled look-back
// over the preceding tiles and scatters.
// ---------------------------------------------------------------------------
// 8192-element tiles (512 threads x 16): digit runs of ~32 entries per tile
// keep the scattered writes long
constexpr int RS_THREADS = 512;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_MAX_PASSES = 8;
constexpr int RS_MAX_DIGITS = 512;  // 9-bit digits at most
constexpr int RS_HIST_BLOCKS = 2048;

// Look-back word of (tile, digit): kind in bits 62-63 (0 none yet, 1 the
// tile's own count, 2 the inclusive prefix over tiles 0..tile), the pass tag
// (pass + 1) in bits 56-61, the value in bits 0-55.  Count and kind travel
// in one 64-bit store, so a reader never sees one without the other.
constexpr uint64_t LB_VALUE = (1ull << 56) - 1;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 22;
constexpr int LB_WINDOW = 8;

__device__ __forceinline__ uint64_t lb_word(uint64_t kind, uint32_t tag, uint64_t v) {
  return (kind << 62) | ((uint64_t)tag << 56) | v;
}

// the passes: digit p is bits [shift[p], shift[p] + width[p]) of the key
struct RsPlan {
  int npass;
  int shift[RS_MAX_PASSES];
  int width[RS_MAX_PASSES];
};

// per-digit totals of every pass in one read of the keys (pass p's counters
// at ghist[p * RS_MAX_DIGITS ..]).  Each wave counts into its own LDS copy
// (the high digits of row-major keys are skewed: shared counters would
// serialise on the hub rows' digit), four keys per thread in flight.
constexpr int HA_THREADS = 256;
constexpr int HA_WAVES = HA_THREADS / 64;
constexpr int HA_ITEMS = 4;

__global__ __launch_bounds__(HA_THREADS) void k_rs_hist_all(const uint64_t* __restrict__ keys, int64_t n, RsPlan pl,
                                                            unsigned long long* __restrict__ ghist) {
  extern __shared__ uint32_t s_hw[];  // [HA_WAVES][npass * RS_MAX_DIGITS]
  const int tid = threadIdx.x, wid = tid >> 6;
  const int W = pl.npass * RS_MAX_DIGITS;
  for (int i = tid; i < HA_WAVES * W; i += HA_THREADS) s_hw[i] = 0;
  __syncthreads();
  uint32_t* my = s_hw + wid * W;
  const int64_t step = (int64_t)gridDim.x * HA_THREADS * HA_ITEMS;
  for (int64_t b = (int64_t)blockIdx.x * HA_THREADS * HA_ITEMS + tid; b < n; b += step) {
    uint64_t k[HA_ITEMS];
#pragma unroll
    for (int j = 0; j < HA_ITEMS; ++j) {
      const int64_t i = b + (int64_t)j * HA_THREADS;
      k[j] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < HA_ITEMS; ++j) {
      if (b + (int64_t)j * HA_THREADS >= n) continue;
      for (int p = 0; p < pl.npass; ++p)
        atomicAdd(&my[p * RS_MAX_DIGITS + (uint32_t)((k[j] >> pl.shift[p]) & ((1u << pl.width[p]) - 1u))], 1u);
    }
  }
  __syncthreads();
  for (int i = tid; i < W; i += HA_THREADS) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < HA_WAVES; ++w) t += s_hw[w * W + i];
    if (t) atomicAdd(&ghist[i], (unsigned long long)t);
  }
}

// block p: exclusive scan of pass p's digit totals (512 slots; the unused
// ones of an 8-bit pass hold zeros)
__global__ __launch_bounds__(RS_MAX_DIGITS) void k_rs_digit_starts(const unsigned long long* __restrict__ ghist,
                                                                   uint64_t* __restrict__ dstart) {
  __shared__ uint64_t s_w[RS_MAX_DIGITS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint64_t c = ghist[blockIdx.x * RS_MAX_DIGITS + tid];
  const uint64_t incl = wave_incl_scan(c);
  if (lane == 63) s_w[wid] = incl;
  __syncthreads();
  uint64_t off = 0;
  for (int w = 0; w < wid; ++w) off += s_w[w];
  dstart[blockIdx.x * RS_MAX_DIGITS + tid] = off + incl - c;
}

// One stable pass over one RB-bit digit (RB = 8 or 9).  The block takes the
// next tile in launch order (a counter, so every earlier tile belongs to a
// block that is already running), ranks it (each wave ranks its own
// contiguous 1024 elements: peers by RB ballots per item and a per-wave
// running count per digit in LDS, so no block barrier per item), publishes
// the tile's digit counts, and reads back over the earlier tiles' words until
// an inclusive prefix, adding their counts.  The tile's keys, then its
// values, are staged in digit order in one LDS buffer (at most 78 KB per
// block, two blocks per CU) and each digit run is written at its global
// offset with consecutive lanes on consecutive addresses.  The look-back wait
// is bounded: past the limit it raises *err and writes nothing, so the kernel
// always ends.
template <typename V, int RB>
__global__ __launch_bounds__(RS_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_rs_onesweep(const uint64_t* __restrict__ kin,
                                                            const V* __restrict__ vin, uint64_t* __restrict__ kout,
                                                            V* __restrict__ vout, int64_t n, int shift,
                                                            const uint64_t* __restrict__ dstart,
                                                            uint64_t* __restrict__ status, uint32_t* __restrict__ tile_ctr,
                                                            uint32_t tag, uint32_t* __restrict__ err) {
  static_assert(sizeof(V) <= sizeof(uint64_t), "values are staged in the key buffer");
  constexpr int RD = 1 << RB;
  constexpr uint32_t DM = RD - 1;
  static_assert(RD <= RS_THREADS, "one thread per digit");
  __shared__ uint64_t s_stage[RS_TILE];
  __shared__ uint64_t s_base[RD];
  __shared__ uint32_t s_dstart[RD];
  // a wave's count of one digit (at most 1024), later a tile offset (< 8192):
  // 16 bits for 9-bit digits keep the block within half the LDS
  using WC = typename std::conditional<RB == 9, uint16_t, uint32_t>::type;
  __shared__ WC s_wcnt[RS_WAVES][RD];
  __shared__ uint32_t s_scan[RS_WAVES];
  __shared__ uint32_t s_tile, s_fail;
  V* s_vstage = reinterpret_cast<V*>(s_stage);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (tid == 0) {
    s_tile = atomicAdd(tile_ctr, 1u);
    s_fail = 0;
  }
  for (int i = tid; i < RS_WAVES * RD; i += RS_THREADS) (&s_wcnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * RS_TILE;
  const int nvalid = (int)min<int64_t>(RS_TILE, n - base);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t kr[RS_ITEMS];
  V vr[RS_ITEMS];
  uint32_t pr[RS_ITEMS];  // digit | (position among the wave's elements of that digit) << RB
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int li = wid * (RS_ITEMS * 64) + it * 64 + lane;
    const bool valid = li < nvalid;
    const uint64_t k = valid ? kin[base + li] : 0;
    const V v = valid ? vin[base + li] : V(0);
    const uint32_t d = (uint32_t)((k >> shift) & DM);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    // every lane reads the wave's running count before the leader adds to it
    // (one wave's LDS operations complete in order)
    const uint32_t before = s_wcnt[wid][d];
    if (valid && rank == 0) s_wcnt[wid][d] = (WC)(before + (uint32_t)__popcll(peers));
    kr[it] = k;
    vr[it] = v;
    pr[it] = valid ? (d | ((before + rank) << RB)) : ~0u;
  }
  __syncthreads();
  // thread d < RD owns digit d: the tile's count, published at once
  uint32_t c = 0;
  if (tid < RD) {
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) c += s_wcnt[w][tid];
    __hip_atomic_store(status + tile * RD + tid, lb_word(tile == 0 ? 2 : 1, tag, c), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // tile-local start of every digit: exclusive scan of the counts
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  if (lane == 63) s_scan[wid] = incl;
  __syncthreads();
  if (tid < RD) {
    uint32_t off = 0;
    for (int w = 0; w < wid; ++w) off += s_scan[w];
    const uint32_t ds = off + incl - c;
    s_dstart[tid] = ds;
    // LDS start of each wave's run of digit tid
    uint32_t run = ds;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
      const uint32_t cw = s_wcnt[w][tid];
      s_wcnt[w][tid] = (WC)run;
      run += cw;
    }
    // global offset: counts of this digit in all earlier tiles, read
    // LB_WINDOW tiles at a time (independent loads), summed in tile order up
    // to the first inclusive word
    uint64_t excl = 0;
    if (tile > 0) {
      int64_t t = tile - 1;
      uint32_t spins = 0;
      while (true) {
        uint64_t w[LB_WINDOW];
#pragma unroll
        for (int j = 0; j < LB_WINDOW; ++j)
          w[j] = t - j >= 0 ? __hip_atomic_load(status + (t - j) * RD + tid, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : 0ull;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int j = 0; j < LB_WINDOW; ++j) {
          if (used != j || done) continue;
          const uint32_t kind = (uint32_t)(w[j] >> 62);
          if (kind == 0 || (uint32_t)((w[j] >> 56) & 63u) != tag) continue;  // not yet published
          excl += w[j] & LB_VALUE;
          used = j + 1;
          done = kind == 2;
        }
        if (done) break;
        t -= used;
        if (t < 0) {  // tile 0 always publishes an inclusive word
          atomicOr(err, 2u);
          s_fail = 1;
          break;
        }
        if (used == 0) {
          if (++spins > LB_SPIN_LIMIT) {
            atomicOr(err, 1u);
            s_fail = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        } else {
          spins = 0;
        }
      }
      __hip_atomic_store(status + tile * RD + tid, lb_word(2, tag, excl + c), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    s_base[tid] = dstart[tid] + excl;
  }
  __syncthreads();
  if (s_fail) return;  // offsets unknown: write nothing (the host reports *err)
  // LDS slot of each item (reused for the values)
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    if (pr[it] != ~0u) {
      pr[it] = (uint32_t)s_wcnt[wid][pr[it] & DM] + (pr[it] >> RB);
      s_stage[pr[it]] = kr[it];
    }
  }
  __syncthreads();
  uint32_t dg[RS_ITEMS / 2] = {};  // digits of the staged elements this thread writes out, two per register
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int li = it * RS_THREADS + tid;
    if (li < nvalid) {
      const uint64_t k = s_stage[li];
      const uint32_t d = (uint32_t)((k >> shift) & DM);
      dg[it >> 1] |= d << (16 * (it & 1));
      kout[s_base[d] + (uint64_t)(li - (int)s_dstart[d])] = k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it)
    if (pr[it] != ~0u) s_vstage[pr[it]] = vr[it];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; ++it) {
    const int li = it * RS_THREADS + tid;
    if (li < nvalid) {
      const uint32_t d = (dg[it >> 1] >> (16 * (it & 1))) & 0xffffu;
      vout[s_base[d] + (uint64_t)(li - (int)s_dstart[d])] = s_vstage[li];
    }
  }
}

__global__ void k_rs_clear(uint64_t* __restrict__ p, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Digit widths: 8-bit passes, or 9-bit ones where that saves a pass (a key of
// at most 45 bits: 42 = 9 + 9 + 8 + 8 + 8); lowest bits first
RsPlan rs_plan(int lo, int bits) {
  RsPlan pl{};
  const int n8 = (bits + 7) / 8, n9 = (bits + 8) / 9;
  pl.npass = std::min(n8, n9);
  const int wide = pl.npass < n8 ? std::max(0, bits - 8 * pl.npass) : 0;  // passes that take 9 bits
  int sh = lo;
  for (int p = 0; p < pl.npass; ++p) {
    pl.shift[p] = sh;
    pl.width[p] = p < wide ? 9 : 8;
    sh += pl.width[p];
  }
  return pl;
}

template <typename V>
int radix_sort_impl(lightne_ctx* ctx, uint64_t* keys, V* vals, int64_t n, int bits, uint64_t* k2, V* v2, int lo,
                    uint64_t** k_res, V** v_res) {
  if (k_res) *k_res = keys;
  if (v_res) *v_res = vals;
  if (n <= 1 || bits <= 0) return LIGHTNE_OK;
  const RsPlan pl = rs_plan(lo, bits);
  if (pl.npass > RS_MAX_PASSES || pl.shift[pl.npass - 1] + pl.width[pl.npass - 1] > 64)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "radix sort of bits %d..%d (at most %d passes of a 64-bit key)", lo,
                    lo + bits, RS_MAX_PASSES);
  const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
  if (!k2 || !v2) {
    k2 = (uint64_t*)ctx_buf(ctx, "rs_keys", n * sizeof(uint64_t));
    v2 = (V*)ctx_buf(ctx, "rs_vals", n * sizeof(V));
  }
  // one zero-filled block: look-back words (ntiles x the widest pass's
  // digits), digit totals and starts (passes x 512 each), tile counters,
  // error flag
  int maxw = 8;
  for (int p = 0; p < pl.npass; ++p) maxw = std::max(maxw, pl.width[p]);
  const int64_t nstat = ntiles * ((int64_t)1 << maxw);
  const int64_t nwords = nstat + 2 * RS_MAX_PASSES * RS_MAX_DIGITS + RS_MAX_PASSES + 1;
  auto* state = (uint64_t*)ctx_buf(ctx, "rs_state", nwords * sizeof(uint64_t));
  if (!k2 || !v2 || !state) return LIGHTNE_ERR_OOM;
  uint64_t* status = state;
  auto* ghist = (unsigned long long*)(state + nstat);
  uint64_t* dstart = state + nstat + RS_MAX_PASSES * RS_MAX_DIGITS;
  auto* tile_ctr = (uint32_t*)(state + nstat + 2 * RS_MAX_PASSES * RS_MAX_DIGITS);
  auto* err = (uint32_t*)(state + nwords - 1);
  const hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_rs_clear, dim3((unsigned)std::min<int64_t>(4096, (nwords + 255) / 256)), dim3(256), 0, s, state,
                     nwords);
  LN_LAUNCH(ctx);
  hipLaunchKernelGGL(k_rs_hist_all,
                     dim3((unsigned)std::min<int64_t>(RS_HIST_BLOCKS, (n + HA_THREADS * HA_ITEMS - 1) /
                                                                          (HA_THREADS * HA_ITEMS))),
                     dim3(HA_THREADS), (size_t)HA_WAVES * pl.npass * RS_MAX_DIGITS * sizeof(uint32_t), s, keys, n, pl,
                     ghist);
  LN_LAUNCH(ctx);
  hipLaunchKernelGGL(k_rs_digit_starts, dim3(pl.npass), dim3(RS_MAX_DIGITS), 0, s, (const unsigned long long*)ghist,
                     dstart);
  LN_LAUNCH(ctx);
  uint64_t *ka = keys, *kb = k2;
  V *va = vals, *vb = v2;
  for (int p = 0; p < pl.npass; ++p) {
    const uint64_t* ds = dstart + p * RS_MAX_DIGITS;
    if (pl.width[p] == 9)
      hipLaunchKernelGGL((k_rs_onesweep<V, 9>), dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka, va, kb, vb, n,
                         pl.shift[p], ds, status, tile_ctr + p, (uint32_t)(p + 1), err);
    else
      hipLaunchKernelGGL((k_rs_onesweep<V, 8>), dim3((unsigned)ntiles), dim3(RS_THREADS), 0, s, ka, va, kb, vb, n,
                         pl.shift[p], ds, status, tile_ctr + p, (uint32_t)(p + 1), err);
    LN_LAUNCH(ctx);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  uint32_t herr = 0;
  LN_HIP(ctx, hipMemcpyAsync(&herr, err, sizeof(herr), hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  if (herr)
    return ctx_fail(ctx, LIGHTNE_ERR_HIP, "radix sort: tile look-back failed (code %u, %lld tiles)", herr,
                    (long long)ntiles);
  if (k_res) {  // the caller reads the result where the last pass left it
    *k_res = ka;
    *v_res = va;
  } else if (ka != keys) {
    LN_HIP(ctx, hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    LN_HIP(ctx, hipMemcpyAsync(vals, va, n * sizeof(V), hipMemcpyDeviceToDevice, s));
  }
  return LIGHTNE_OK;
}

}  // namespace

int scan_u32_to_u64(lightne_ctx* ctx, const uint32_t* in, uint64_t* out, int64_t n) {
  return scan_impl<uint32_t>(ctx, in, out, n, 0);
}
int radix_sort_passes(int bits) { return bits > 0 ? rs_plan(0, bits).npass : 0; }
int radix_sort_pairs(lightne_ctx* ctx, uint64_t* keys, float* vals, int64_t n, int bits, uint64_t* k_tmp,
                     float* v_tmp, int lo_bit, uint64_t** k_res, float** v_res) {
  return radix_sort_impl<float>(ctx, keys, vals, n, bits, k_tmp, v_tmp, lo_bit, k_res, v_res);
}
int radix_sort_pairs_u32(lightne_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bits) {
  return radix_sort_impl<uint32_t>(ctx, keys, vals, n, bits, nullptr, nullptr, 0, nullptr, nullptr);
}
