// Dense panel kernels for the randomized SVD and the tall SVD of stage 3.
//
//  * k_gaussian      : counter-based Gaussian sketch (Philox4x32-10 +
//                      Box-Muller in double), replacing MKL VSL SFMT19937 +
//                      Box-Muller2 (LightNE/mklredsvd/mklredsvd.cc:141,229-231,
//                      mklredsvd/util.cc:21-38), whose stream cannot be
//                      reproduced.  Element e of stream t is fully determined by
//                      (seed, t, e), so CPU and GPU draw the same matrix; values
//                      are rounded to BF16 precision so both panel types hold
//                      the same sketch exactly.
//  * k_gram_f64      : G = X^T Y in FP64 on v_mfma_f64_16x16x4f64 (the Gram of
//                      CholeskyQR2 and C = Z^T B of mklredsvd.cc:327-345).
//  * k_panel_small   : Out = X W for a tall panel X and a small W on
//                      v_mfma_f32_16x16x4f32 (Z = B P, U = Z U_C,
//                      mklredsvd.cc:294-312,379-397, and Q = Y R^-1), with an
//                      optional row L2-normalisation epilogue
//                      (compute_u_sigma_root, PathEmbed.h:390-406).
#include "lightne_internal.h"

#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {

__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
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

// round to the nearest BF16-representable float (ties to even): the sketch is
// drawn at BF16 precision so that the FP32 and BF16 forms multiply by the same
// matrix
__host__ __device__ __forceinline__ float round_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

__global__ void k_gaussian(uint64_t seed, uint32_t stream, int64_t count, float* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * p < count;
       p += (int64_t)gridDim.x * blockDim.x) {
  uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586476925286766559 * u2;
  out[2 * p] = round_bf16((float)(r * cos(t)));
  if (2 * p + 1 < count) out[2 * p + 1] = round_bf16((float)(r * sin(t)));
  }
}

// the same values stored as BF16: every value is BF16-representable, so the
// 16-bit panel holds the sketch exactly at half the bytes
__global__ void k_gaussian_bf16(uint64_t seed, uint32_t stream, int64_t count, __hip_bfloat16* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * p < count;
       p += (int64_t)gridDim.x * blockDim.x) {
  uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586476925286766559 * u2;
  const unsigned short b0 = (unsigned short)(__float_as_uint(round_bf16((float)(r * cos(t)))) >> 16);
  out[2 * p] = *reinterpret_cast<const __hip_bfloat16*>(&b0);
  if (2 * p + 1 < count) {
    const unsigned short b1 = (unsigned short)(__float_as_uint(round_bf16((float)(r * sin(t)))) >> 16);
    out[2 * p + 1] = *reinterpret_cast<const __hip_bfloat16*>(&b1);
  }
  }
}

// the same n x k sketch written into a panel of row stride ld: element e of
// the stream goes to row e / k, column e % k
__global__ void k_gaussian_ld(uint64_t seed, uint32_t stream, int64_t n, int64_t k, int64_t ld,
                              float* __restrict__ out) {
  const int64_t count = n * k;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; 2 * p < count;
       p += (int64_t)gridDim.x * blockDim.x) {
  uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
  const double u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586476925286766559 * u2;
  const int64_t e0 = 2 * p, e1 = 2 * p + 1;
  out[(e0 / k) * ld + e0 % k] = round_bf16((float)(r * cos(t)));
  if (e1 < count) out[(e1 / k) * ld + e1 % k] = round_bf16((float)(r * sin(t)));
  }
}

// ---------------------------------------------------------------------------
// Gram: rows are staged through LDS in 32-KB slabs; every wave owns a
// fixed subset of the 16x16 output tiles and walks the slab 4 rows at a time.
// v_mfma_f64_16x16x4f64 operand layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register v of lane l holds D[(l/16) + 4v][l%16]
// (rows interleaved, unlike the f32 16x16x4 form whose register v holds row
// 4(l/16) + v).
// ---------------------------------------------------------------------------
constexpr int GR_THREADS = 512;
constexpr int GR_WAVES = GR_THREADS / 64;
constexpr int GR_MAX_PAIRS_PER_WAVE = 12;
constexpr int GR_MAX_V = 9;  // 16-B vectors per thread for one slab: 2 x 32 KB / 16 B / 512 = 8, +1

// rows per slab: 32 rows of an FP32 panel, 16 of an FP64 one (the same global
// bytes per load round).  Both are staged in LDS as FP64, so an FP32 slab takes
// twice the LDS of an FP64 one: 32 x (k1 + k2) x 8 bytes, 128 KiB for two
// different 256-wide panels, checked against the device's limit per block.
template <typename TI>
struct GramRows {
  static constexpr int R = 128 / sizeof(TI);
};

// One block owns a contiguous row range; the slab of X (and Y) is staged in
// LDS while the next slab is already in flight in registers.  TI = float for
// the FP32 panels, double for the FP64 form.  Every wave owns exactly PPW
// output tiles (the list is padded with dummy tiles whose results are not
// written), so a 4-row step issues all its LDS reads at once and then its
// MFMAs back to back, with no per-tile branch between them.
template <typename TI, int PPW>
__global__ __launch_bounds__(GR_THREADS) void k_gram_f64(const TI* __restrict__ X, const TI* __restrict__ Y,
                                                         int64_t n, int k1, int k2, int npairs,
                                                         const int2* __restrict__ pairs, int64_t rows_per_block,
                                                         double* __restrict__ part) {
  constexpr int ROWS = GramRows<TI>::R;
  constexpr int E = 16 / sizeof(TI);  // elements per 16-B vector
  // the slab is staged in LDS as FP64 (FP32 panels converted once per element
  // on the way in, not once per use)
  extern __shared__ double2 s_raw2[];
  double* s_mem = reinterpret_cast<double*>(s_raw2);
  const bool same = X == Y;
  double* sX = s_mem;
  double* sY = same ? s_mem : s_mem + ROWS * k1;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int2 pr[PPW];
  int pa[PPW], pb[PPW];
#pragma unroll
  for (int q = 0; q < PPW; ++q) {
    pr[q] = pairs[wid + q * GR_WAVES];  // padded to PPW * GR_WAVES entries
    pa[q] = pr[q].x * 16;
    pb[q] = pr[q].y * 16;
  }
  doublex4 acc[PPW];
#pragma unroll
  for (int q = 0; q < PPW; ++q) acc[q] = doublex4{0.0, 0.0, 0.0, 0.0};
  const int64_t r_begin = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r_end = min(n, r_begin + rows_per_block);
  const int nxv = ROWS * k1 / E, nyv = same ? 0 : ROWS * k2 / E;
  const int rowvx = k1 / E, rowvy = k2 / E;
  float4 buf[GR_MAX_V];
  auto fetch = [&](int64_t r0) {
    const int nrow = (int)min<int64_t>(ROWS, r_end - r0);
    const float4* gx = reinterpret_cast<const float4*>(X + r0 * k1);
    const float4* gy = reinterpret_cast<const float4*>(Y + r0 * k2);
#pragma unroll
    for (int t = 0; t < GR_MAX_V; ++t) {
      const int f = threadIdx.x + t * GR_THREADS;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (f < nxv) {
        if (f / rowvx < nrow) v = gx[f];
      } else if (f < nxv + nyv) {
        const int g = f - nxv;
        if (g / rowvy < nrow) v = gy[g];
      }
      buf[t] = v;
    }
  };
  if (r_begin < r_end) fetch(r_begin);
  for (int64_t r0 = r_begin; r0 < r_end; r0 += ROWS) {
    __syncthreads();
#pragma unroll
    for (int t = 0; t < GR_MAX_V; ++t) {
      const int f = threadIdx.x + t * GR_THREADS;
      if (f < nxv + nyv) {
        if constexpr (sizeof(TI) == 4) {
          s_raw2[2 * f] = make_double2((double)buf[t].x, (double)buf[t].y);
          s_raw2[2 * f + 1] = make_double2((double)buf[t].z, (double)buf[t].w);
        } else {
          const double2 d = *reinterpret_cast<const double2*>(&buf[t]);
          s_raw2[f] = d;
        }
      }
    }
    __syncthreads();
    if (r0 + ROWS < r_end) fetch(r0 + ROWS);
#pragma unroll 2
    for (int s = 0; s < ROWS; s += 4) {
      const double* xr = sX + (s + lk) * k1 + li;
      const double* yr = sY + (s + lk) * k2 + li;
      double av[PPW], bv[PPW];
#pragma unroll
      for (int q = 0; q < PPW; ++q) {
        av[q] = xr[pa[q]];
        bv[q] = yr[pb[q]];
      }
#pragma unroll
      for (int q = 0; q < PPW; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv[q], acc[q], 0, 0, 0);
    }
  }
  double* out = part + (int64_t)blockIdx.x * k1 * k2;
#pragma unroll
  for (int q = 0; q < PPW; ++q) {
    if (wid + q * GR_WAVES < npairs) {
#pragma unroll
      for (int v = 0; v < 4; ++v) out[(int64_t)(pr[q].x * 16 + lk + 4 * v) * k2 + pr[q].y * 16 + li] = acc[q][v];
    }
  }
}

__global__ void k_sum_parts(const double* __restrict__ part, int nparts, int64_t len, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = 0.0;
  for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * len + i];
  out[i] = s;
}

// ---------------------------------------------------------------------------
// Out(n x k2) = X(n x k1) W(k1 x k2).  One wave per 16-row slab, all k2
// columns.  The contraction index is permuted so that lane l reads one float4
// of its row per 4 MFMA steps: step (t, q) uses k = 16 t + 4 (l/16) + q for A
// and the same k for B.
// ---------------------------------------------------------------------------
constexpr int PS_THREADS = 512;
constexpr int PS_MAX_T2 = 16;  // k2 <= 256

// one 16-column step t of X's slab row (lane: row li, columns 16t + 4lk..+3)
// times rows 16t.. of the LDS copy of W; TRI (W upper triangular, k1 = k2):
// output tile j only takes steps t <= j, the others multiply zeros
template <int T2, bool TRI>
__device__ __forceinline__ void ps_step(const float4 xa, int t, int lk, int li, const float* sW, int ldw,
                                        floatx4* acc) {
  const float xs[4] = {xa.x, xa.y, xa.z, xa.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float* wr = sW + (16 * t + 4 * lk + q) * ldw + li;
#pragma unroll
    for (int j = 0; j < T2; ++j)
      if (!TRI || j >= t) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xs[q], wr[16 * j], acc[j], 0, 0, 0);
  }
}

template <int T2, bool TRI>
__global__ __launch_bounds__(PS_THREADS) void k_panel_small(const float* X, int64_t ldx, int64_t n,
                                                            int k1, const float* __restrict__ W, int ldW,
                                                            float* Out, int64_t ldo, int normalize) {
  extern __shared__ float sW[];  // k1 x (k2 + 4): columns [0, k2) of this block of W
  const int k2 = T2 * 16;
  const int ldw = k2 + 4;
  for (int e = threadIdx.x; e < k1 * k2; e += PS_THREADS) {
    const int r = e / k2;
    sW[r * ldw + (e - r * k2)] = W[(int64_t)r * ldW + (e - r * k2)];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t nslab = (n + 15) / 16;
  const int64_t wave0 = ((int64_t)blockIdx.x * PS_THREADS + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * PS_THREADS) >> 6;
  for (int64_t slab = wave0; slab < nslab; slab += nwaves) {
    const int64_t row = slab * 16 + li;
    const bool rv = row < n;
    floatx4 acc[T2];
#pragma unroll
    for (int j = 0; j < T2; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    if constexpr (TRI) {
      // k1 == 16 * T2: the slab row's k1 columns are loaded up front (all
      // loads in flight together), then multiplied; the steps unroll, so the
      // zero tiles skipped are known at compile time
      constexpr int MAXT = 16;
      const int nt = k1 / 16;
      float4 xa[MAXT];
#pragma unroll
      for (int t = 0; t < MAXT; ++t)
        if (t < nt)
          xa[t] = rv ? *reinterpret_cast<const float4*>(X + row * ldx + 16 * t + 4 * lk) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int t = 0; t < MAXT; ++t)
        if (t < nt) ps_step<T2, true>(xa[t], t, lk, li, sW, ldw, acc);
    } else {
      for (int t = 0; t < k1 / 16; ++t) {
        const float4 xa =
            rv ? *reinterpret_cast<const float4*>(X + row * ldx + 16 * t + 4 * lk) : make_float4(0.f, 0.f, 0.f, 0.f);
        ps_step<T2, false>(xa, t, lk, li, sW, ldw, acc);
      }
    }
    // acc[j][v] = Out[slab*16 + 4*lk + v][16 j + li]
    if (normalize) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float ss = 0.f;
#pragma unroll
        for (int j = 0; j < T2; ++j) ss += acc[j][v] * acc[j][v];
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) ss += __shfl_xor(ss, d, 64);
        const float nrm = sqrtf(ss);
        if (nrm > 1e-5f) {
          const float inv = 1.0f / nrm;
#pragma unroll
          for (int j = 0; j < T2; ++j) acc[j][v] *= inv;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t orow = slab * 16 + 4 * lk + v;
      if (orow < n) {
#pragma unroll
        for (int j = 0; j < T2; ++j) Out[orow * ldo + 16 * j + li] = acc[j][v];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 form: Out(n x k2, double) = X(n x k1, double) W(k1 x k2), W's column
// block [c0, c0 + 16 nt) in LDS, one wave per 16-row slab on
// v_mfma_f64_16x16x4f64 (lane l: A = X[row0 + l%16][kk + l/16],
// B = W[kk + l/16][c + l%16]; register v holds Out[row0 + l/16 + 4v][c + l%16]).
// ---------------------------------------------------------------------------
constexpr int P64_MAX_T = 4;

__global__ __launch_bounds__(256) void k_panel64(const double* __restrict__ X, int64_t n, int k1,
                                                 const double* __restrict__ W, int k2, int c0, int nt,
                                                 double* __restrict__ Out) {
  extern __shared__ double sW64[];  // k1 x (16 nt + 1)
  const int ldw = 16 * nt + 1;
  for (int e = threadIdx.x; e < k1 * 16 * nt; e += blockDim.x) {
    const int r = e / (16 * nt), c = e - r * (16 * nt);
    sW64[r * ldw + c] = W[(int64_t)r * k2 + c0 + c];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t nslab = (n + 15) / 16;
  const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t slab = wave0; slab < nslab; slab += nwaves) {
    const int64_t row = slab * 16 + li;
    const bool rv = row < n;
    doublex4 acc[P64_MAX_T];
#pragma unroll
    for (int j = 0; j < P64_MAX_T; ++j) acc[j] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int kk = 0; kk < k1; kk += 4) {
      const double a = rv ? X[row * k1 + kk + lk] : 0.0;
      const double* wr = sW64 + (kk + lk) * ldw + li;
#pragma unroll
      for (int j = 0; j < P64_MAX_T; ++j)
        if (j < nt) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wr[16 * j], acc[j], 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t orow = slab * 16 + lk + 4 * v;
      if (orow < n) {
#pragma unroll
        for (int j = 0; j < P64_MAX_T; ++j)
          if (j < nt) Out[orow * k2 + c0 + 16 * j + li] = acc[j][v];
      }
    }
  }
}

__global__ void k_f32_to_f64(const float* __restrict__ in, double* __restrict__ out, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

__global__ void k_f32_to_bf16(const float* __restrict__ in, __hip_bfloat16* __restrict__ out, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
  uint32_t u = __float_as_uint(in[i]);
  u += 0x7fffu + ((u >> 16) & 1u);
  const unsigned short b = (unsigned short)(u >> 16);
  out[i] = *reinterpret_cast<const __hip_bfloat16*>(&b);
  }
}

// emb[i, j] = sqrt(S_j) U[i, j] (j < dim, U row stride rank), optional row
// normalisation; one wave per row.
__global__ void k_u_sigma_root(const float* __restrict__ U, const float* __restrict__ sq, int64_t n, int64_t rank,
                               int64_t dim, int normalize, float* __restrict__ emb) {
  const int lane = threadIdx.x & 63;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < n;
       row += ((int64_t)gridDim.x * blockDim.x) >> 6) {
  float ss = 0.f;
  for (int64_t j = lane; j < dim; j += 64) {
    const float v = sq[j] * U[row * rank + j];
    ss += v * v;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) ss += __shfl_xor(ss, d, 64);
  const float nrm = sqrtf(ss);
  const float inv = (normalize && nrm > 1e-5f) ? 1.0f / nrm : 1.0f;
  for (int64_t j = lane; j < dim; j += 64) emb[row * dim + j] = sq[j] * U[row * rank + j] * inv;
  }
}

}  // namespace

void gaussian_fill_host(uint64_t seed, uint32_t stream, int64_t count, float* out) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t p = 0; 2 * p < count; ++p) {
    uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), stream, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;
    const double u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double t = two_pi * u2;
    out[2 * p] = round_bf16((float)(r * std::cos(t)));
    if (2 * p + 1 < count) out[2 * p + 1] = round_bf16((float)(r * std::sin(t)));
  }
}

int gaussian_fill_ld(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t n, int64_t k, int64_t ld,
                     float* out) {
  if (n <= 0 || k <= 0) return LIGHTNE_OK;
  const int64_t pairs = (n * k + 1) / 2;
  hipLaunchKernelGGL(k_gaussian_ld, dim3(grid_blocks(pairs, 256)), dim3(256), 0, ctx->stream, seed, stream, n, k,
                     ld, out);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

int gaussian_fill(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t count, float* out) {
  if (count <= 0) return LIGHTNE_OK;
  const int64_t pairs = (count + 1) / 2;
  hipLaunchKernelGGL(k_gaussian, dim3(grid_blocks(pairs, 256)), dim3(256), 0, ctx->stream, seed, stream,
                     count, out);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

int gaussian_fill_bf16(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t count, __hip_bfloat16* out) {
  if (count <= 0) return LIGHTNE_OK;
  const int64_t pairs = (count + 1) / 2;
  hipLaunchKernelGGL(k_gaussian_bf16, dim3(grid_blocks(pairs, 256)), dim3(256), 0, ctx->stream, seed, stream,
                     count, out);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free Gram (lightne_dense_gram_noalloc): the pair list, the
// per-block partial sums and the reduced sum live in this file-scope device
// array instead of context buffers.  Each device holds its own copy from the
// first use until the process ends; its device address is taken once per
// device at context set-up (dense_scratch_setup) and only that address is
// used from the host.  Every call on a device shares the array: calls that
// overlap on one device race.  Nothing in it needs clearing: every entry a
// call reads was written earlier in the same call.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_BYTES = (size_t)48 << 20;
constexpr size_t SCR_PAIRS_BYTES = 4096 * sizeof(int2);
constexpr size_t SCR_SUM_BYTES = 256 * 256 * sizeof(double);
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int SCR_MAX_DEVICES = 64;
std::mutex g_scr_mu;
void* g_scr_ptr[SCR_MAX_DEVICES] = {};
}  // namespace

int dense_scratch_setup(lightne_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= SCR_MAX_DEVICES)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "scratch: device %d beyond %d", ctx->device, SCR_MAX_DEVICES);
  std::lock_guard<std::mutex> g(g_scr_mu);
  if (!g_scr_ptr[ctx->device]) {
    void* p = nullptr;
    LN_HIP(ctx, hipSetDevice(ctx->device));
    LN_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return ctx_fail(ctx, LIGHTNE_ERR_HIP, "scratch: no device address for the scratch array");
    g_scr_ptr[ctx->device] = p;
  }
  ctx->scratch = g_scr_ptr[ctx->device];
  return LIGHTNE_OK;
}

namespace {
// device memory a Gram call uses in place of context buffers
struct GramScratch {
  int2* pairs;
  double* sum;
  double* part;
  int max_blocks;
};

template <typename TI>
int gram_impl(lightne_ctx* ctx, const TI* X, const TI* Y, int64_t n, int64_t k1, int64_t k2, int symmetric,
              double* G_host, const GramScratch* scr = nullptr) {
  constexpr int ROWS = GramRows<TI>::R;
  if (k1 % 16 || k2 % 16 || k1 <= 0 || k2 <= 0 || k1 > 256 || k2 > 256)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "gram: widths must be multiples of 16 in [16, 256]");
  if (symmetric && (X != Y || k1 != k2)) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "gram: symmetric needs X == Y");
  const int T1 = (int)(k1 / 16), T2 = (int)(k2 / 16);
  std::vector<int2> pairs;
  for (int a = 0; a < T1; ++a)
    for (int b = 0; b < T2; ++b)
      if (!symmetric || a <= b) pairs.push_back(make_int2(a, b));
  const int npairs_all = (int)pairs.size();
  const int per_launch = GR_MAX_PAIRS_PER_WAVE * GR_WAVES;
  // each launch's list padded to a whole number of tiles per wave
  std::vector<int2> padded;
  std::vector<int> launch_np, launch_ppw, launch_off;
  for (int p0 = 0; p0 < npairs_all; p0 += per_launch) {
    const int np = std::min(per_launch, npairs_all - p0);
    const int ppw = (np + GR_WAVES - 1) / GR_WAVES;
    launch_np.push_back(np);
    launch_ppw.push_back(ppw);
    launch_off.push_back((int)padded.size());
    for (int i = 0; i < ppw * GR_WAVES; ++i) padded.push_back(i < np ? pairs[p0 + i] : make_int2(0, 0));
  }
  int nblk = std::max(1, ctx->num_cus);
  if (scr) nblk = std::max(1, std::min(nblk, scr->max_blocks));
  int64_t rpb = (n + nblk - 1) / nblk;
  rpb = ((rpb + ROWS - 1) / ROWS) * ROWS;
  if (rpb == 0) rpb = ROWS;
  nblk = (int)std::max<int64_t>(1, (n + rpb - 1) / rpb);
  const int64_t len = k1 * k2;
  int2* dpairs;
  double *part, *gsum;
  if (scr) {
    if (padded.size() * sizeof(int2) > SCR_PAIRS_BYTES)
      return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "gram: %zu tile pairs exceed the scratch", padded.size());
    dpairs = scr->pairs;
    part = scr->part;
    gsum = scr->sum;
  } else {
    dpairs = (int2*)ctx_buf(ctx, "gram_pairs", std::max<size_t>(padded.size(), 1) * sizeof(int2));
    part = (double*)ctx_buf(ctx, "gram_part", (size_t)nblk * len * sizeof(double));
    gsum = (double*)ctx_buf(ctx, "gram_sum", len * sizeof(double));
    if (!dpairs || !part || !gsum) return LIGHTNE_ERR_OOM;
  }
  if (!padded.empty())
    LN_HIP(ctx, hipMemcpyAsync(dpairs, padded.data(), padded.size() * sizeof(int2), hipMemcpyHostToDevice,
                               ctx->stream));
  const size_t lds = (size_t)ROWS * (k1 + (X == Y ? 0 : k2)) * sizeof(double);  // staged as FP64
  if (lds > ctx->lds_per_block)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "gram: a %zu-byte LDS slab exceeds the device's %zu bytes per block",
                    lds, ctx->lds_per_block);
  if (n > 0) {
    // each launch owns a disjoint set of output tiles
    for (size_t L = 0; L < launch_np.size(); ++L) {
      const int2* dp = dpairs + launch_off[L];
      const int np = launch_np[L];
      switch (launch_ppw[L]) {
#define GRAM_CASE(P)                                                                                          \
  case P:                                                                                                     \
    hipLaunchKernelGGL((k_gram_f64<TI, P>), dim3(nblk), dim3(GR_THREADS), lds, ctx->stream, X, Y, n, (int)k1, \
                       (int)k2, np, dp, rpb, part);                                                          \
    break;
        GRAM_CASE(1) GRAM_CASE(2) GRAM_CASE(3) GRAM_CASE(4) GRAM_CASE(5) GRAM_CASE(6)
        GRAM_CASE(7) GRAM_CASE(8) GRAM_CASE(9) GRAM_CASE(10) GRAM_CASE(11) GRAM_CASE(12)
#undef GRAM_CASE
        default:
          return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "gram: %d tiles per wave", launch_ppw[L]);
      }
      LN_LAUNCH(ctx);
    }
  } else {
    // no rows: G = 0 (written on the host, no device fill)
    std::fill(G_host, G_host + len, 0.0);
    return LIGHTNE_OK;
  }
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, ctx->stream, part, nblk, len,
                     gsum);
  LN_LAUNCH(ctx);
  LN_HIP(ctx, hipMemcpyAsync(G_host, gsum, len * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (symmetric)
    for (int64_t i = 0; i < k1; ++i)
      for (int64_t j = 0; j < i; ++j) G_host[i * k2 + j] = G_host[j * k2 + i];
  return LIGHTNE_OK;
}
}  // namespace

int gram_f64(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2, int symmetric,
             double* G_host) {
  return gram_impl<float>(ctx, X, Y, n, k1, k2, symmetric, G_host);
}
int gram_f64(lightne_ctx* ctx, const double* X, const double* Y, int64_t n, int64_t k1, int64_t k2, int symmetric,
             double* G_host) {
  return gram_impl<double>(ctx, X, Y, n, k1, k2, symmetric, G_host);
}

int gram_f64_noalloc(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                     int symmetric, double* G_host) {
  if (!ctx->scratch) return ctx_fail(ctx, LIGHTNE_ERR_HIP, "gram: the scratch array was not set up");
  if (k1 % 16 || k2 % 16 || k1 <= 0 || k2 <= 0 || k1 > 256 || k2 > 256)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG,
                    "dense_gram_noalloc: widths must be multiples of 16 in [16, 256] (the %zu-byte scratch holds "
                    "the partial sums of at least %zu blocks at 256 x 256)", SCRATCH_BYTES,
                    (SCRATCH_BYTES - SCR_PAIRS_BYTES - SCR_SUM_BYTES) / SCR_SUM_BYTES);
  char* base = (char*)ctx->scratch;  // the device address of `scratch` on this device
  GramScratch scr;
  scr.pairs = (int2*)base;
  scr.sum = (double*)(base + SCR_PAIRS_BYTES);
  scr.part = (double*)(base + SCR_PAIRS_BYTES + SCR_SUM_BYTES);
  const size_t part_cap = SCRATCH_BYTES - SCR_PAIRS_BYTES - SCR_SUM_BYTES;
  scr.max_blocks = (int)std::min<size_t>(1 << 20, part_cap / ((size_t)k1 * k2 * sizeof(double)));
  return gram_impl<float>(ctx, X, Y, n, k1, k2, symmetric, G_host, &scr);
}

int panel_times_small(lightne_ctx* ctx, const float* X, int64_t n, int64_t k1, int64_t ldx, const double* W_host,
                      int64_t k2, float* Out, int normalize_rows, float*, int upper_tri) {
  if (k1 % 16 || k2 % 16 || k1 <= 0 || k2 <= 0 || k1 > 256 || k2 > 16 * PS_MAX_T2 || ldx < k1 || ldx % 4)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel_times_small: bad shape k1=%lld k2=%lld ldx=%lld",
                    (long long)k1, (long long)k2, (long long)ldx);
  if (Out == X && (k1 != k2 || ldx != k1))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel_times_small: in-place needs a square W");
  // column blocks of W small enough for LDS (<= 160 KB per block)
  int64_t kb = k2;
  while ((size_t)k1 * (kb + 4) * sizeof(float) > 150 * 1024) kb -= 16;
  if (kb < k2 && normalize_rows)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel_times_small: row normalisation needs k2 in one block");
  std::vector<float> wf(k1 * k2);
  for (int64_t i = 0; i < k1 * k2; ++i) wf[i] = (float)W_host[i];
  auto* dW = (float*)ctx_buf(ctx, "small_w", wf.size() * sizeof(float));
  if (!dW) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemcpyAsync(dW, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  // wf is a local: the copy must have consumed it before this function returns
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (n <= 0) return LIGHTNE_OK;
  float* dst = Out;
  if (kb < k2 && Out == X) {  // later column blocks must still read the old X
    dst = (float*)ctx_buf(ctx, "ps_tmp", (size_t)n * k2 * sizeof(float));
    if (!dst) return LIGHTNE_ERR_OOM;
  }
  const int64_t nslab = (n + 15) / 16;
  const int64_t waves_per_block = PS_THREADS / 64;
  const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>((nslab + waves_per_block - 1) / waves_per_block,
                                                              (int64_t)ctx->num_cus * 4));
  // an upper-triangular W skips its zero tiles when it is square and in one block
  const bool tri = upper_tri && k1 == k2 && kb == k2;
  for (int64_t j0 = 0; j0 < k2; j0 += kb) {
    const int64_t kk = std::min(kb, k2 - j0);
    const size_t lds = (size_t)k1 * (kk + 4) * sizeof(float);
    const int T2 = (int)(kk / 16);
#define PS_LAUNCH(t, tri_)                                                                                      \
  hipLaunchKernelGGL((k_panel_small<t, tri_>), dim3((unsigned)nblk), dim3(PS_THREADS), lds, ctx->stream, X, ldx, n, \
                     (int)k1, (const float*)(dW + j0), (int)k2, dst + j0, (int64_t)k2, normalize_rows)
#define PS_CASE(t)           \
  case t:                    \
    if (tri)                 \
      PS_LAUNCH(t, true);    \
    else                     \
      PS_LAUNCH(t, false);   \
    break;
    switch (T2) {
      PS_CASE(1) PS_CASE(2) PS_CASE(3) PS_CASE(4) PS_CASE(5) PS_CASE(6) PS_CASE(7) PS_CASE(8)
      PS_CASE(9) PS_CASE(10) PS_CASE(11) PS_CASE(12) PS_CASE(13) PS_CASE(14) PS_CASE(15) PS_CASE(16)
      default:
        return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel_times_small: k2 too large");
    }
#undef PS_CASE
#undef PS_LAUNCH
    LN_LAUNCH(ctx);
  }
  if (dst != Out)
    LN_HIP(ctx, hipMemcpyAsync(Out, dst, (size_t)n * k2 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  return LIGHTNE_OK;
}

int panel_times_small64(lightne_ctx* ctx, const double* X, int64_t n, int64_t k1, const double* W_host, int64_t k2,
                        double* Out) {
  if (k1 % 16 || k2 % 16 || k1 <= 0 || k2 <= 0 || k1 > 256 || k2 > 256 || Out == X)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "panel_times_small64: bad shape or in-place call");
  auto* dW = (double*)ctx_buf(ctx, "small_w64", (size_t)k1 * k2 * sizeof(double));
  if (!dW) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemcpyAsync(dW, W_host, (size_t)k1 * k2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));  // W_host may be a caller's temporary
  if (n <= 0) return LIGHTNE_OK;
  const int64_t nslab = (n + 15) / 16;
  const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>((nslab + 3) / 4, (int64_t)ctx->num_cus * 8));
  for (int c0 = 0; c0 < k2; c0 += 16 * P64_MAX_T) {
    const int nt = (int)std::min<int64_t>(P64_MAX_T, (k2 - c0) / 16);
    const size_t lds = (size_t)k1 * (16 * nt + 1) * sizeof(double);
    hipLaunchKernelGGL(k_panel64, dim3((unsigned)nblk), dim3(256), lds, ctx->stream, X, n, (int)k1,
                       (const double*)dW, (int)k2, c0, nt, Out);
    LN_LAUNCH(ctx);
  }
  return LIGHTNE_OK;
}

int f32_to_f64(lightne_ctx* ctx, const float* in, double* out, int64_t count) {
  if (count <= 0) return LIGHTNE_OK;
  hipLaunchKernelGGL(k_f32_to_f64, dim3(grid_blocks(count, 256)), dim3(256), 0, ctx->stream, in, out, count);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

int f64_to_f32(lightne_ctx* ctx, const double* in, float* out, int64_t count) {
  if (count <= 0) return LIGHTNE_OK;
  hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_blocks(count, 256)), dim3(256), 0, ctx->stream, in, out, count);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

int f32_to_bf16(lightne_ctx* ctx, const float* in, __hip_bfloat16* out, int64_t count) {
  if (count <= 0) return LIGHTNE_OK;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_blocks(count, 256)), dim3(256), 0, ctx->stream, in, out, count);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}

int u_sigma_root_dev(lightne_ctx* ctx, const float* U, const float* S_host, int64_t n, int64_t rank, int64_t dim,
                     int normalize, float* emb) {
  std::vector<float> sq(dim);
  for (int64_t j = 0; j < dim; ++j) sq[j] = S_host ? (float)std::sqrt(S_host[j]) : 1.0f;
  auto* dsq = (float*)ctx_buf(ctx, "usr_sq", dim * sizeof(float));
  if (!dsq) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemcpyAsync(dsq, sq.data(), dim * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));  // sq is a local
  if (n <= 0) return LIGHTNE_OK;
  const int64_t blocks = (n * 64 + 255) / 256;
  hipLaunchKernelGGL(k_u_sigma_root, dim3(grid_blocks(blocks, 1)), dim3(256), 0, ctx->stream, U, (const float*)dsq, n, rank,
                     dim, normalize, emb);
  LN_LAUNCH(ctx);
  return LIGHTNE_OK;
}
