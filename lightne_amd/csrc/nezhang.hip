// The ne_zhang_et_al embedding method (LightNE/PathEmbed.h:410-463): a
// deterministic log matrix on the graph's own sparsity pattern, then the
// stage-2 randomized SVD and U sqrt(Sigma) with row normalisation.
//
//   neg_j   = (sum_{i in N(j)} 1 / d_i)^0.75                  (:418-427)
//   S       = sum_j neg_j                                     (:429)
//   A_uv    = (float)(log(1 / d_u) - log(neg_v / S))          (:431-443)
//
// The matrix keeps the graph's row pointers and neighbour order (:432-444
// writes the entries of row u in adjacency order), so the CSR aliases the
// graph's offsets and neighbour array and only the values are new.
#include "lightne_internal.h"
#include "spmm_epi.h"

#include <cmath>
#include <vector>

namespace {

constexpr int NZ_BLOCKS = 1024;

// one wave per vertex: lanes stride over the neighbours, then a fixed-order
// shuffle tree (deterministic)
__global__ __launch_bounds__(256) void k_nz_neg(const int64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                int64_t n, double* __restrict__ neg) {
  const int lane = threadIdx.x & 63;
  for (int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n;
       j += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    double s = 0.0;
    for (int64_t e = off[j] + lane; e < off[j + 1]; e += 64) {
      const uint32_t i = adj[e];
      s += 1.0 / (double)(off[i + 1] - off[i]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) neg[j] = pow(s, 0.75);
  }
}

// per-block partial sums of neg (summed on the host in block order)
__global__ __launch_bounds__(256) void k_nz_sum(const double* __restrict__ neg, int64_t n, double* __restrict__ part) {
  __shared__ double s_w[4];
  double s = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    s += neg[j];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// one wave per row: values in adjacency order
__global__ __launch_bounds__(256) void k_nz_val(const int64_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                int64_t n, const double* __restrict__ neg, double neg_sum,
                                                float* __restrict__ val) {
  const int lane = threadIdx.x & 63;
  for (int64_t u = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < n;
       u += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const double lnom = log(1.0 / (double)(off[u + 1] - off[u]));
    for (int64_t e = off[u] + lane; e < off[u + 1]; e += 64) val[e] = (float)(lnom - log(neg[adj[e]] / neg_sum));
  }
}

}  // namespace

// Builds the matrix into `val` (m floats, device); rows / columns alias the
// graph.
int ne_zhang_values(lightne_ctx* ctx, const lightne_graph* g, float* val) {
  const int64_t n = g->n;
  auto* neg = (double*)ctx_buf(ctx, "nz_neg", (size_t)n * sizeof(double));
  auto* part = (double*)ctx_buf(ctx, "nz_part", NZ_BLOCKS * sizeof(double));
  if (!neg || !part) return LIGHTNE_ERR_OOM;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_nz_neg, dim3(grid_blocks(n, 4)), dim3(256), 0, s, g->offsets, g->edges, n, neg);
  LN_LAUNCH(ctx);
  hipLaunchKernelGGL(k_nz_sum, dim3(NZ_BLOCKS), dim3(256), 0, s, (const double*)neg, n, part);
  LN_LAUNCH(ctx);
  std::vector<double> hp(NZ_BLOCKS);
  LN_HIP(ctx, hipMemcpyAsync(hp.data(), part, NZ_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  double neg_sum = 0.0;
  for (double v : hp) neg_sum += v;
  if (g->m > 0) {
    hipLaunchKernelGGL(k_nz_val, dim3(grid_blocks(n, 4)), dim3(256), 0, s, g->offsets, g->edges, n,
                       (const double*)neg, neg_sum, val);
    LN_LAUNCH(ctx);
  }
  return LIGHTNE_OK;
}

extern "C" {

int lightne_ne_zhang_matrix(lightne_ctx* ctx, const lightne_graph* g, lightne_csr** out) {
  if (!ctx || !g || !out) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_ne_zhang_matrix: NULL argument");
  *out = nullptr;
  if (g->n <= 0 || !g->offsets || (g->m > 0 && !g->edges) || g->n >= ((int64_t)1 << 31))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "lightne_ne_zhang_matrix: bad graph (n must be < 2^31)");
  CallScope scope(ctx);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  DevMem rp, col, val;
  LN_CHECK(dev_alloc(ctx, rp, (g->n + 1) * 8, "csr row_ptr"));
  LN_CHECK(dev_alloc(ctx, col, std::max<int64_t>(g->m, 1) * 4, "csr col"));
  LN_CHECK(dev_alloc(ctx, val, std::max<int64_t>(g->m, 1) * 4, "csr val"));
  LN_CHECK(ne_zhang_values(ctx, g, (float*)val.p));
  LN_HIP(ctx, hipMemcpyAsync(rp.p, g->offsets, (g->n + 1) * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (g->m > 0)
    LN_HIP(ctx, hipMemcpyAsync(col.p, g->edges, g->m * 4, hipMemcpyDeviceToDevice, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto* c = new lightne_csr();
  c->n = g->n;
  c->nnz = g->m;
  c->row_ptr = (int64_t*)rp.release();
  c->col = (int32_t*)col.release();
  c->val = (float*)val.release();
  *out = c;
  return LIGHTNE_OK;
}

}  // extern "C"
