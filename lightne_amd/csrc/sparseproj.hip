// Very sparse random projection sketch (-sparse_project,
// LightNE/mklredsvd/mklredsvd.cc:148-226; Li, Hastie, Church, KDD'06).
//
// Omega (n x k) has, in every row, each column present with probability
// 1/s and value +-sqrt(s) (s = -sparse_project_s when >= 1, else sqrt(n),
// :150-154).  The reference draws the per-row counts from VSL's binomial
// generator and the columns by a Fisher-Yates shuffle seeded from time(0)
// (:156-189): the same distribution as an independent Bernoulli(1/s) per
// entry, which is what is drawn here, from Philox4x32-10 counters
// (row, column) so that the host oracle draws the same matrix.
//
// Y = A Omega is formed without a dense Omega: a row of Omega is non-empty
// with probability about k/s (3.5% at n = 2^24, k = 144), so the kernel walks
// the nonzeros of A's row i, skips the empty rows of Omega through its row
// pointers, and adds the few contributions in index order (deterministic),
// accumulating in FP64.  Cost: one read of A's CSR and one write of Y,
// instead of an n x k panel gather per nonzero.
#include "lightne_internal.h"

#include <cmath>

namespace {

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

constexpr uint32_t SP_STREAM = 2u;  // Philox stream of the sparse sketch

// entry (i, c): present iff u0 < thresh (u0 * 2^-32 < 1/s); sign from u1's low bit
__device__ __forceinline__ int sp_entry(uint64_t seed, int64_t i, int c, uint32_t thresh) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)c, SP_STREAM};
  philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  if (ctr[0] >= thresh) return 0;
  return (ctr[1] & 1u) ? -1 : 1;
}

__global__ void k_sp_count(uint64_t seed, int64_t n, int k, uint32_t thresh, uint32_t* __restrict__ cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t t = 0;
    for (int c = 0; c < k; ++c) t += sp_entry(seed, i, c, thresh) != 0;
    cnt[i] = t;
  }
}

// entries of row i as (column << 1 | negative)
__global__ void k_sp_fill(uint64_t seed, int64_t n, int k, uint32_t thresh, const uint64_t* __restrict__ rpo,
                          uint16_t* __restrict__ ent) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t p = rpo[i];
    for (int c = 0; c < k; ++c) {
      const int e = sp_entry(seed, i, c, thresh);
      if (e) ent[p++] = (uint16_t)((c << 1) | (e < 0 ? 1 : 0));
    }
  }
}

// one wave per row of A: lane l holds columns l, l + 64, ... of Y's row
template <typename OT>
__global__ __launch_bounds__(256) void k_sp_mm(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                               const float* __restrict__ val, int64_t n, int k,
                                               const uint64_t* __restrict__ rpo, const uint16_t* __restrict__ ent,
                                               double scale, OT* __restrict__ Y) {
  const int lane = threadIdx.x & 63;
  constexpr int MAXC = 4;  // k <= 256
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n;
       i += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    double acc[MAXC] = {0.0, 0.0, 0.0, 0.0};
    const int64_t j1 = rp[i + 1];
    for (int64_t b = rp[i]; b < j1; b += 64) {
      const int64_t e = b + lane;
      uint64_t o0 = 0, o1 = 0;
      int jc = 0;
      float w = 0.f;
      if (e < j1) {
        jc = col[e];
        w = val[e];
        o0 = rpo[jc];
        o1 = rpo[jc + 1];
      }
      // lanes whose Omega row is non-empty, taken in index order
      uint64_t mask = __ballot(o1 > o0);
      while (mask) {
        const int src = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        const uint64_t s0 = __shfl(o0, src, 64), s1 = __shfl(o1, src, 64);
        const double a = (double)__shfl(w, src, 64) * scale;
        for (uint64_t q = s0; q < s1; ++q) {
          const uint32_t t = ent[q];
          const int c = (int)(t >> 1);
          const double v = (t & 1u) ? -a : a;
          if ((c & 63) == lane) acc[c >> 6] += v;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MAXC; ++m) {
      const int c = lane + 64 * m;
      if (c < k) Y[i * k + c] = (OT)acc[m];
    }
  }
}

}  // namespace

// Y (n x k, FP32 or FP64 by `f64`) = A Omega for the very sparse Omega drawn
// from `seed`; s <= 0 selects sqrt(n)
int sparse_project_sketch(lightne_ctx* ctx, const int64_t* rp, const int32_t* col, const float* val, int64_t n,
                          int64_t k, uint64_t seed, double s_param, void* Y, bool f64) {
  if (k <= 0 || k > 256) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "sparse projection: rank must be in [1, 256]");
  // mklredsvd.cc:150-154: FP s = sqrt(n), replaced by sparse_project_s when >= 1
  const float s = s_param >= 1.0 ? (float)s_param : (float)std::sqrt((double)n);
  const double density = 1.0 / (double)s;
  const uint32_t thresh = density >= 1.0 ? 0xffffffffu : (uint32_t)(density * 4294967296.0);
  const double scale = (double)std::sqrt(s);  // +-sqrt(s) (:186)
  hipStream_t st = ctx->stream;
  auto* cnt = (uint32_t*)ctx_buf(ctx, "sp_cnt", (size_t)std::max<int64_t>(n, 1) * 4);
  auto* rpo = (uint64_t*)ctx_buf(ctx, "sp_rpo", (size_t)(n + 1) * 8);
  if (!cnt || !rpo) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_sp_count, dim3(grid_blocks(n, 256)), dim3(256), 0, st, seed, n, (int)k, thresh, cnt);
  LN_LAUNCH(ctx);
  LN_CHECK(scan_u32_to_u64(ctx, cnt, rpo, n));
  uint64_t tot = 0;
  LN_HIP(ctx, hipMemcpyAsync(&tot, rpo + n, 8, hipMemcpyDeviceToHost, st));
  LN_HIP(ctx, hipStreamSynchronize(st));
  auto* ent = (uint16_t*)ctx_buf(ctx, "sp_ent", (size_t)std::max<uint64_t>(tot, 1) * 2);
  if (!ent) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_sp_fill, dim3(grid_blocks(n, 256)), dim3(256), 0, st, seed, n, (int)k, thresh,
                     (const uint64_t*)rpo, ent);
  LN_LAUNCH(ctx);
  if (f64)
    hipLaunchKernelGGL(k_sp_mm<double>, dim3(grid_blocks(n, 4)), dim3(256), 0, st, rp, col, val, n, (int)k,
                       (const uint64_t*)rpo, (const uint16_t*)ent, scale, (double*)Y);
  else
    hipLaunchKernelGGL(k_sp_mm<float>, dim3(grid_blocks(n, 4)), dim3(256), 0, st, rp, col, val, n, (int)k,
                       (const uint64_t*)rpo, (const uint16_t*)ent, scale, (float*)Y);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}
