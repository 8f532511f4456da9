// CSR x dense-panel products for stages 2 and 3.
//
// Replaces mkl_sparse_?_mm (LightNE/mklredsvd/mklredsvd.cc:235,271 and
// LightNE/SpectralPropagation.h:101,116,153,168) and the neighbour loop of
// SpectralPropagation.h:210-221.
//
// Layout: X is n x k row-major (k % 16 == 0), FP32 or BF16.  A lane owns 16
// columns of one row, so a group of G = k/16 lanes reads a whole X row (576 B
// at k=144 FP32, 288 B BF16); every load instruction of a group covers G
// consecutive 16-byte pieces.  A 64-wide wave holds NG = 64/G groups.
//
// Rows are split by length (power-law graphs), in one launch:
//   * short rows (nnz <= SHORT): one group per row, no cross-lane reduction;
//     a wave's NG rows come from the plan's list of short rows grouped by
//     length, so they take about the same number of steps;
//   * long rows: chunks of CHUNK nnz, one wave per chunk, the NG groups of the
//     wave stride over the chunk and are reduced by in-wave shuffles; rows with
//     several chunks write partial sums that k_spmm_fixup adds in chunk order
//     (deterministic, no float atomics).
// Epilogues fuse the Chebyshev recurrence of SpectralPropagation.h:94-208 so
// each step is exactly two gathers over the graph.
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int SHORT_NNZ = 48;
constexpr int CHUNK_NNZ = 512;
// waves per launch: one grid dimension holds fewer than 2^32 work-items, so
// larger products run as several launches over consecutive wave ranges
constexpr int64_t MAX_WAVES = ((int64_t)1 << 26) - 64;

// Lane column layout.  A group of G = k/16 lanes covers one k-wide row; lane
// `sub` owns 16 columns as chunks of CW consecutive columns (CW = 4 floats or
// 8 bf16, i.e. 16 bytes), chunk q at column CW*sub + CW*G*q, so each load
// instruction of a group reads G consecutive 16-byte pieces.
template <typename XT>
struct Lay;
template <>
struct Lay<float> {
  static constexpr int CW = 4;
};
template <>
struct Lay<__hip_bfloat16> {
  static constexpr int CW = 8;
};

// column of accumulator quad i4 (accumulators 4*i4 .. 4*i4+3)
template <typename XT>
__device__ __forceinline__ int quad_col(int i4, int sub, int G) {
  constexpr int CW = Lay<XT>::CW;
  return CW * sub + CW * G * ((4 * i4) / CW) + ((4 * i4) % CW);
}

template <typename XT>
__device__ __forceinline__ void load16(const XT* __restrict__ row, int sub, int G, float* v);

template <>
__device__ __forceinline__ void load16<float>(const float* __restrict__ row, int sub, int G, float* v) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 t = *reinterpret_cast<const float4*>(row + 4 * sub + 4 * G * q);
    v[4 * q + 0] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}

template <>
__device__ __forceinline__ void load16<__hip_bfloat16>(const __hip_bfloat16* __restrict__ row, int sub, int G,
                                                       float* v) {
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const uint4 t = *reinterpret_cast<const uint4*>(row + 8 * sub + 8 * G * q);
    const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[8 * q + 2 * j + 0] = __uint_as_float(w[j] << 16);
      v[8 * q + 2 * j + 1] = __uint_as_float(w[j] & 0xffff0000u);
    }
  }
}

// Rows are walked GATHER_U nonzeros at a time so that each lane has
// GATHER_U independent row loads in flight; products are added one nonzero at
// a time in index order whatever GATHER_U is, so the FP32 and BF16 panels
// give bit-identical sums for BF16-representable inputs.
constexpr int GATHER_U = 2;

template <typename XT, bool IMPL>
__device__ __forceinline__ void gather_row(const SpmmArgs& a, int64_t j0, int64_t j1, int64_t stride, int sub,
                                           float* acc) {
  const XT* __restrict__ X = reinterpret_cast<const XT*>(a.X);
  const int G = a.G;
  int64_t j = j0;
  for (; j + (GATHER_U - 1) * stride < j1; j += GATHER_U * stride) {
    int64_t c[GATHER_U];
    float w[GATHER_U];
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) {
      if (IMPL) {
        c[u] = a.gcol[j + u * stride];
        w[u] = 1.f;
      } else {
        c[u] = a.col[j + u * stride];
        w[u] = a.val[j + u * stride];
      }
    }
    float x[GATHER_U][16];
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) load16<XT>(X + c[u] * a.ldx, sub, G, x[u]);
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[i] += w[u] * x[u][i];
  }
  for (; j < j1; j += stride) {
    int64_t c0;
    float w0 = 1.f;
    if (IMPL) {
      c0 = a.gcol[j];
    } else {
      c0 = a.col[j];
      w0 = a.val[j];
    }
    float x0[16];
    load16<XT>(X + c0 * a.ldx, sub, G, x0);
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] += w0 * x0[i];
  }
}

// One launch for both row classes.  Waves [0, W_short) take NG consecutive
// entries of the short-row list each, one group per row; the
// remaining waves take one chunk of a long row each: the NG groups stride over
// the chunk and are summed into group 0 by in-wave shuffles (group order, so
// the result is deterministic), which applies the epilogue or writes the
// chunk's partial sum.
template <typename XT, int EPI, bool IMPL>
__global__ __launch_bounds__(256) void k_spmm(SpmmArgs a, int64_t w_short, int64_t wave0) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, NG = 64 / G;
  const int g = lane / G, sub = lane - g * G;
  const int64_t wave = wave0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  float acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.f;
  if (wave < w_short) {
    if (g >= NG) return;
    const int64_t i = wave * NG + g;
    if (i >= a.nshort) return;
    const int64_t r = a.short_rows[i];
    const int64_t j0 = a.rp[r], j1 = a.rp[r + 1];
    gather_row<XT, IMPL>(a, j0, j1, 1, sub, acc);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
      epilogue<XT, EPI, IMPL>(a, r, quad_col<XT>(q, sub, G), v);
    }
    return;
  }
  const int64_t c = wave - w_short;
  if (c >= a.nchunks) return;  // uniform across the wave
  if (g < NG) gather_row<XT, IMPL>(a, a.chunk_start[2 * c] + g, a.chunk_start[2 * c + 1], NG, sub, acc);
  for (int q = 1; q < NG; ++q) {
    const int src = (lane + q * G) & 63;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float t = __shfl(acc[i], src, 64);
      if (lane < G) acc[i] += t;
    }
  }
  if (lane >= G) return;
  const int64_t row = a.chunk_row[c];
  const int slot = a.chunk_slot[c];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
    const int col = quad_col<XT>(q, sub, G);
    if (slot < 0)
      epilogue<XT, EPI, IMPL>(a, row, col, v);
    else
      *reinterpret_cast<float4*>(a.partial + (int64_t)slot * a.k + col) = v;
  }
}

// rows with several chunks: add partials in chunk order, then the epilogue
template <typename XT, int EPI, bool IMPL>
__global__ __launch_bounds__(256) void k_spmm_fixup(SpmmArgs a, int64_t wave0) {
  const int lane = threadIdx.x & 63;
  const int64_t w = wave0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= a.nmulti) return;
  const int ncol4 = (int)(a.k >> 2);
  const int64_t row = a.multi_row[w];
  const int64_t s0 = a.multi_first[w], s1 = a.multi_first[w + 1];
  for (int q = lane; q < ncol4; q += 64) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t sl = s0; sl < s1; ++sl) {
      const float4 t = *reinterpret_cast<const float4*>(a.partial + sl * a.k + 4 * q);
      s.x += t.x;
      s.y += t.y;
      s.z += t.z;
      s.w += t.w;
    }
    epilogue<XT, EPI, IMPL>(a, row, 4 * q, s);
  }
}

template <typename XT, int EPI, bool IMPL>
int launch(lightne_ctx* ctx, const SpmmPlan& P, SpmmArgs a) {
  a.short_nnz = P.short_nnz;
  a.nshort = P.nshort;
  a.short_rows = P.short_rows;
  a.nchunks = P.nchunks;
  a.chunk_start = P.chunk_start;
  a.chunk_row = P.chunk_row;
  a.chunk_slot = P.chunk_slot;
  a.nmulti = P.nmulti;
  a.multi_row = P.multi_row;
  a.multi_first = P.multi_first;
  a.G = (int)(a.k / 16);
  const int NG = 64 / a.G;
  if (P.nparts > 0 && !a.partial) {  // the caller may bring them (nparts x k floats)
    a.partial = (float*)ctx_buf(ctx, "spmm_partial", (size_t)P.nparts * a.k * sizeof(float));
    if (!a.partial) return LIGHTNE_ERR_OOM;
  }
  const int64_t waves_short = (a.nshort + NG - 1) / NG;
  const int64_t waves = waves_short + P.nchunks;
  for (int64_t w0 = 0; w0 < waves; w0 += MAX_WAVES) {
    hipLaunchKernelGGL((k_spmm<XT, EPI, IMPL>), dim3((unsigned)((std::min(MAX_WAVES, waves - w0) + 3) / 4)),
                       dim3(256), 0, ctx->stream, a, waves_short, w0);
    LN_LAUNCH(ctx);
  }
  for (int64_t w0 = 0; w0 < P.nmulti; w0 += MAX_WAVES) {
    hipLaunchKernelGGL((k_spmm_fixup<XT, EPI, IMPL>), dim3((unsigned)((std::min(MAX_WAVES, P.nmulti - w0) + 3) / 4)),
                       dim3(256), 0, ctx->stream, a, w0);
    LN_LAUNCH(ctx);
  }
  return LIGHTNE_OK;
}

// ---------------------------------------------------------------------------
// FP64 form (explicit CSR, Y = A X with X, Y double): the same work split and
// deterministic reduction; a lane owns 16 columns as 8 chunks of 2 doubles,
// chunk q at column 2 sub + 2 G q.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gather_row64(const SpmmArgs& a, int64_t j0, int64_t j1, int64_t stride, int sub,
                                             double* acc) {
  const double* __restrict__ X = reinterpret_cast<const double*>(a.X);
  const int G = a.G;
  for (int64_t j = j0; j < j1; j += stride) {
    const double w = (double)a.val[j];
    const double* xr = X + (int64_t)a.col[j] * a.k + 2 * sub;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double2 t = *reinterpret_cast<const double2*>(xr + 2 * G * q);
      acc[2 * q] += w * t.x;
      acc[2 * q + 1] += w * t.y;
    }
  }
}

__device__ __forceinline__ void store_row64(double* __restrict__ out, int64_t k, int G, int sub, const double* acc) {
#pragma unroll
  for (int q = 0; q < 8; ++q)
    *reinterpret_cast<double2*>(out + 2 * sub + 2 * G * q) = make_double2(acc[2 * q], acc[2 * q + 1]);
}

__global__ __launch_bounds__(256) void k_spmm64(SpmmArgs a, int64_t w_short, int64_t wave0) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, NG = 64 / G;
  const int g = lane / G, sub = lane - g * G;
  const int64_t wave = wave0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  double acc[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0;
  double* out = reinterpret_cast<double*>(a.out);
  auto* part = reinterpret_cast<double*>(a.partial);
  if (wave < w_short) {
    if (g >= NG) return;
    const int64_t i = wave * NG + g;
    if (i >= a.nshort) return;
    const int64_t r = a.short_rows[i];
    const int64_t j0 = a.rp[r], j1 = a.rp[r + 1];
    gather_row64(a, j0, j1, 1, sub, acc);
    store_row64(out + r * a.k, a.k, G, sub, acc);
    return;
  }
  const int64_t c = wave - w_short;
  if (c >= a.nchunks) return;  // uniform across the wave
  if (g < NG) gather_row64(a, a.chunk_start[2 * c] + g, a.chunk_start[2 * c + 1], NG, sub, acc);
  for (int q = 1; q < NG; ++q) {
    const int src = (lane + q * G) & 63;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double t = __shfl(acc[i], src, 64);
      if (lane < G) acc[i] += t;
    }
  }
  if (lane >= G) return;
  const int slot = a.chunk_slot[c];
  if (slot < 0)
    store_row64(out + (int64_t)a.chunk_row[c] * a.k, a.k, G, sub, acc);
  else
    store_row64(part + (int64_t)slot * a.k, a.k, G, sub, acc);
}

__global__ __launch_bounds__(256) void k_spmm64_fixup(SpmmArgs a, int64_t wave0) {
  const int lane = threadIdx.x & 63;
  const int64_t w = wave0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (w >= a.nmulti) return;
  const int64_t row = a.multi_row[w];
  const int64_t s0 = a.multi_first[w], s1 = a.multi_first[w + 1];
  const auto* part = reinterpret_cast<const double*>(a.partial);
  double* out = reinterpret_cast<double*>(a.out) + row * a.k;
  for (int c = lane; c < a.k; c += 64) {
    double s = 0.0;
    for (int64_t sl = s0; sl < s1; ++sl) s += part[sl * a.k + c];
    out[c] = s;
  }
}

}  // namespace

void spmm_plan_free(SpmmPlan& P) {
  if (P.mem) (void)hipFree(P.mem);
  P = SpmmPlan{};
}

namespace {

// per row: chunks (0 for short rows), multi-chunk flag, partial slots
__global__ void k_plan_count(const int64_t* __restrict__ rp, int64_t n, int short_nnz, int chunk_nnz,
                             uint32_t* __restrict__ nch, uint32_t* __restrict__ multi, uint32_t* __restrict__ parts) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t len = rp[r + 1] - rp[r];
  const uint32_t nc = len <= short_nnz ? 0u : (uint32_t)((len + chunk_nnz - 1) / chunk_nnz);
  nch[r] = nc;
  multi[r] = nc > 1 ? 1u : 0u;
  parts[r] = nc > 1 ? nc : 0u;
}

__global__ void k_plan_fill(const int64_t* __restrict__ rp, int64_t n, int chunk_nnz, const uint64_t* __restrict__ coff,
                            const uint64_t* __restrict__ moff, const uint64_t* __restrict__ poff,
                            int64_t* __restrict__ chunk_se, int32_t* __restrict__ chunk_row,
                            int32_t* __restrict__ chunk_slot, int32_t* __restrict__ multi_row,
                            int64_t* __restrict__ multi_first) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) multi_first[moff[n]] = (int64_t)poff[n];
  if (r >= n) return;
  const uint64_t c0 = coff[r], nc = coff[r + 1] - c0;
  if (nc == 0) return;
  const int64_t b = rp[r], e = rp[r + 1];
  for (uint64_t q = 0; q < nc; ++q) {
    const uint64_t c = c0 + q;
    chunk_se[2 * c] = b + (int64_t)q * chunk_nnz;
    chunk_se[2 * c + 1] = min(e, b + (int64_t)(q + 1) * chunk_nnz);
    chunk_row[c] = (int32_t)r;
    chunk_slot[c] = nc > 1 ? (int32_t)(poff[r] + q) : -1;
  }
  if (nc > 1) {
    multi_row[moff[r]] = (int32_t)r;
    multi_first[moff[r]] = (int64_t)poff[r];
  }
}

// short rows by length: per block of 256 rows its short rows of every length,
// length-major: cnt[len * gridDim.x + block]
constexpr int PL_BINS = SHORT_NNZ + 1;  // lengths 0 .. SHORT_NNZ

__global__ __launch_bounds__(256) void k_plan_short_count(const int64_t* __restrict__ rp, int64_t n,
                                                          uint32_t* __restrict__ cnt) {
  __shared__ uint32_t s_c[PL_BINS];
  for (int b = threadIdx.x; b < PL_BINS; b += 256) s_c[b] = 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) {
    const int64_t len = rp[r + 1] - rp[r];
    if (len <= SHORT_NNZ) atomicAdd(&s_c[len], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < PL_BINS; b += 256) cnt[(uint64_t)b * gridDim.x + blockIdx.x] = s_c[b];
}

// every short row to its place in the list: lengths in increasing order,
// blocks in order within a length, rows of a block in index order (wave
// ranks by ballot, waves in order)
__global__ __launch_bounds__(256) void k_plan_short_fill(const int64_t* __restrict__ rp, int64_t n,
                                                         const uint64_t* __restrict__ off, int32_t* __restrict__ rows) {
  __shared__ uint32_t s_w[256 / 64][PL_BINS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int bin = -1;
  if (r < n) {
    const int64_t len = rp[r + 1] - rp[r];
    if (len <= SHORT_NNZ) bin = (int)len;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t rank = 0;
  for (int b = 0; b < PL_BINS; ++b) {
    const uint64_t m = __ballot(bin == b);
    if (bin == b) rank = (uint32_t)__popcll(m & below);
    if (lane == 0) s_w[wid][b] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (bin >= 0) {
    uint32_t before = 0;
    for (int w = 0; w < wid; ++w) before += s_w[w][bin];
    rows[off[(uint64_t)bin * gridDim.x + blockIdx.x] + before + rank] = (int32_t)r;
  }
}

}  // namespace

// Work split for one CSR, built on the device: rows longer than SHORT_NNZ are
// cut into chunks of chunk_nnz (CHUNK_NNZ unless given); rows with several
// chunks get partial slots (nparts x k accumulators per product).
int spmm_plan_build(lightne_ctx* ctx, const int64_t* rp_dev, int64_t n, SpmmPlan& P, int chunk_nnz) {
  if (chunk_nnz <= 0) chunk_nnz = CHUNK_NNZ;
  if (chunk_nnz < SHORT_NNZ) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: chunk below the short-row length");
  spmm_plan_free(P);
  P.n = n;
  P.short_nnz = SHORT_NNZ;
  P.chunk_nnz = chunk_nnz;
  if (n >= ((int64_t)1 << 31)) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: more than 2^31 rows");
  hipStream_t s = ctx->stream;
  auto* nch = (uint32_t*)ctx_buf(ctx, "plan_nch", std::max<int64_t>(n, 1) * 4);
  auto* mul = (uint32_t*)ctx_buf(ctx, "plan_mul", std::max<int64_t>(n, 1) * 4);
  auto* par = (uint32_t*)ctx_buf(ctx, "plan_par", std::max<int64_t>(n, 1) * 4);
  auto* coff = (uint64_t*)ctx_buf(ctx, "plan_coff", (n + 1) * 8);
  auto* moff = (uint64_t*)ctx_buf(ctx, "plan_moff", (n + 1) * 8);
  auto* poff = (uint64_t*)ctx_buf(ctx, "plan_poff", (n + 1) * 8);
  if (!nch || !mul || !par || !coff || !moff || !poff) return LIGHTNE_ERR_OOM;
  const unsigned nb = (unsigned)((n + 255) / 256);
  auto* scnt = (uint32_t*)ctx_buf(ctx, "plan_scnt", (size_t)PL_BINS * nb * 4);
  auto* soff = (uint64_t*)ctx_buf(ctx, "plan_soff", ((size_t)PL_BINS * nb + 1) * 8);
  if (!scnt || !soff) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_plan_count, dim3(nb), dim3(256), 0, s, rp_dev, n, SHORT_NNZ, chunk_nnz, nch, mul, par);
  LN_LAUNCH(ctx);
  hipLaunchKernelGGL(k_plan_short_count, dim3(nb), dim3(256), 0, s, rp_dev, n, scnt);
  LN_LAUNCH(ctx);
  LN_CHECK(scan_u32_to_u64(ctx, nch, coff, n));
  LN_CHECK(scan_u32_to_u64(ctx, mul, moff, n));
  LN_CHECK(scan_u32_to_u64(ctx, par, poff, n));
  LN_CHECK(scan_u32_to_u64(ctx, scnt, soff, (int64_t)PL_BINS * nb));
  uint64_t tot[4];
  LN_HIP(ctx, hipMemcpyAsync(&tot[0], coff + n, 8, hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipMemcpyAsync(&tot[1], moff + n, 8, hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipMemcpyAsync(&tot[2], poff + n, 8, hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipMemcpyAsync(&tot[3], soff + (size_t)PL_BINS * nb, 8, hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  P.nchunks = (int64_t)tot[0];
  P.nmulti = (int64_t)tot[1];
  P.nparts = (int64_t)tot[2];
  P.nshort = (int64_t)tot[3];
  if (P.nparts >= ((int64_t)1 << 31)) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: too many partial chunks");
  // the plan's arrays in one allocation (freed with the plan)
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_cs = al((2 * P.nchunks + 2) * sizeof(int64_t)), b_cr = al(P.nchunks * sizeof(int32_t)),
               b_mr = al(P.nmulti * sizeof(int32_t)), b_mf = al((P.nmulti + 1) * sizeof(int64_t)),
               b_sr = al(P.nshort * sizeof(int32_t));
  if (ctx_malloc(ctx, &P.mem, b_cs + 2 * b_cr + b_mr + b_mf + b_sr) != hipSuccess)
    return ctx_fail(ctx, LIGHTNE_ERR_OOM, "spmm plan allocation failed");
  char* q = (char*)P.mem;
  P.chunk_start = (int64_t*)q;
  q += b_cs;
  P.chunk_row = (int32_t*)q;
  q += b_cr;
  P.chunk_slot = (int32_t*)q;
  q += b_cr;
  P.multi_row = (int32_t*)q;
  q += b_mr;
  P.multi_first = (int64_t*)q;
  q += b_mf;
  P.short_rows = (int32_t*)q;
  hipLaunchKernelGGL(k_plan_fill, dim3(nb), dim3(256), 0, s, rp_dev, n, chunk_nnz, (const uint64_t*)coff,
                     (const uint64_t*)moff, (const uint64_t*)poff, P.chunk_start, P.chunk_row, P.chunk_slot,
                     P.multi_row, P.multi_first);
  LN_LAUNCH(ctx);
  // the short rows grouped by length
  hipLaunchKernelGGL(k_plan_short_fill, dim3(nb), dim3(256), 0, s, rp_dev, n, (const uint64_t*)soff, P.short_rows);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

#define LN_DISPATCH(XT)                                                              \
  switch (epi) {                                                                     \
    case EPI_STORE: return impl ? launch<XT, EPI_STORE, true>(ctx, P, a)             \
                                : launch<XT, EPI_STORE, false>(ctx, P, a);           \
    case EPI_STORE_F32: return impl ? launch<XT, EPI_STORE_F32, true>(ctx, P, a)     \
                                    : launch<XT, EPI_STORE_F32, false>(ctx, P, a);   \
    case EPI_CHEB1: return launch<XT, EPI_CHEB1, true>(ctx, P, a);                   \
    case EPI_CHEB: return launch<XT, EPI_CHEB, true>(ctx, P, a);                     \
    case EPI_CHEB_LAST: return launch<XT, EPI_CHEB_LAST, true>(ctx, P, a);           \
    case EPI_ADJ: return launch<XT, EPI_ADJ, true>(ctx, P, a);                       \
    default: break;                                                                  \
  }

int spmm_run(lightne_ctx* ctx, const SpmmPlan& P, SpmmArgs a, int xt, int epi, bool impl) {
  if (a.k % 16 != 0 || a.k < 16 || a.k > 1024)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel width %lld must be a multiple of 16 in [16, 1024]",
                    (long long)a.k);
  if (a.ldx == 0) a.ldx = a.k;
  if (a.ldo == 0) a.ldo = a.k;
  if (a.ldx < a.k || a.ldo < a.k || a.ldx % 16 || a.ldo % 16)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: row strides must be multiples of 16 of at least the width");
  if ((a.ldx != a.k || a.ldo != a.k) && epi != EPI_STORE && epi != EPI_STORE_F32)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: fused epilogues need contiguous panels");
  if (impl && (epi == EPI_CHEB1 || epi == EPI_CHEB || epi == EPI_CHEB_LAST || epi == EPI_ADJ) &&
      a.impl_mode != IMPL_PROP)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "Chebyshev epilogue needs the propagation operator");
  if (!impl && epi != EPI_STORE && epi != EPI_STORE_F32)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "explicit SpMM supports store epilogues only");
  if (xt == PANEL_F32) {
    LN_DISPATCH(float)
  } else {
    LN_DISPATCH(__hip_bfloat16)
  }
  return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "unknown SpMM epilogue %d", epi);
}

int spmm_run64(lightne_ctx* ctx, const SpmmPlan& P, SpmmArgs a) {
  if (a.k % 16 != 0 || a.k < 16 || a.k > 1024)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel width %lld must be a multiple of 16 in [16, 1024]",
                    (long long)a.k);
  if ((a.ldx && a.ldx != a.k) || (a.ldo && a.ldo != a.k))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "spmm: the FP64 form needs contiguous panels");
  a.short_nnz = P.short_nnz;
  a.nshort = P.nshort;
  a.short_rows = P.short_rows;
  a.nchunks = P.nchunks;
  a.chunk_start = P.chunk_start;
  a.chunk_row = P.chunk_row;
  a.chunk_slot = P.chunk_slot;
  a.nmulti = P.nmulti;
  a.multi_row = P.multi_row;
  a.multi_first = P.multi_first;
  a.G = (int)(a.k / 16);
  const int NG = 64 / a.G;
  if (P.nparts > 0) {
    a.partial = (float*)ctx_buf(ctx, "spmm_partial64", (size_t)P.nparts * a.k * sizeof(double));
    if (!a.partial) return LIGHTNE_ERR_OOM;
  }
  const int64_t waves_short = (a.nshort + NG - 1) / NG;
  const int64_t waves = waves_short + P.nchunks;
  for (int64_t w0 = 0; w0 < waves; w0 += MAX_WAVES) {
    hipLaunchKernelGGL(k_spmm64, dim3((unsigned)((std::min(MAX_WAVES, waves - w0) + 3) / 4)), dim3(256), 0,
                       ctx->stream, a, waves_short, w0);
    LN_LAUNCH(ctx);
  }
  for (int64_t w0 = 0; w0 < P.nmulti; w0 += MAX_WAVES) {
    hipLaunchKernelGGL(k_spmm64_fixup, dim3((unsigned)((std::min(MAX_WAVES, P.nmulti - w0) + 3) / 4)), dim3(256),
                       0, ctx->stream, a, w0);
    LN_LAUNCH(ctx);
  }
  return LIGHTNE_OK;
}
