// SpMM argument block, work plan and fused epilogues (shared by spmm.hip and
// its callers).
#pragma once

#include "lightne_internal.h"

enum SpmmEpi {
  EPI_STORE = 0,      // out(XT)  = alpha * op(X)
  EPI_STORE_F32 = 1,  // out(f32) = alpha * op(X)
  EPI_CHEB1 = 2,      // Lx1 = alpha * M X - a ; conv = c0 a + c1 Lx1   (SpectralPropagation.h:114-148)
  EPI_CHEB = 3,       // Lx2 = M X - Lx0 - 2 Lx1 ; conv += c1 Lx2       (:151-197)
  EPI_CHEB_LAST = 4,  // as EPI_CHEB, then y = a - conv, Lx2 not stored  (:151-208)
  EPI_ADJ = 5,        // adjoint step: out = c1 g / (d+1) + c0 (alpha M X - lx1) - lx0
};

enum SpmmImpl {
  IMPL_NONE = 0,  // explicit CSR values
  IMPL_PROP = 1,  // (M x)_i = (1-mu-1/(d_i+1)) x_i - 1/(d_i+1) sum_j x_j  (SpectralPropagation.h:51-75)
  IMPL_AGG = 2,   // ((A+I) x)_i = x_i + sum_j x_j                          (:210-221)
};

struct SpmmPlan {
  void* mem = nullptr;  // one allocation holding every array below
  int64_t n = 0;
  int short_nnz = 0, chunk_nnz = 0;
  // the rows of at most short_nnz entries, grouped by length (the NG rows a
  // wave takes together then take about the same time)
  int64_t nshort = 0;
  int32_t* short_rows = nullptr;
  int64_t nchunks = 0, nmulti = 0, nparts = 0;
  int64_t* chunk_start = nullptr;  // 2 * nchunks: (start, end) pairs
  int32_t* chunk_row = nullptr;
  int32_t* chunk_slot = nullptr;
  int32_t* multi_row = nullptr;
  int64_t* multi_first = nullptr;  // nmulti + 1
};

struct SpmmArgs {
  int64_t n = 0, k = 0;
  // row strides of X and of out in elements (0: k).  A stride above k lets a
  // product run on a column slab of a wider panel; the Chebyshev and adjoint
  // epilogues need contiguous panels (strides k).
  int64_t ldx = 0, ldo = 0;
  const int64_t* rp = nullptr;  // row pointers (CSR or graph offsets)
  const int32_t* col = nullptr;
  const float* val = nullptr;
  const uint32_t* gcol = nullptr;  // graph neighbours (implicit operators)
  int impl_mode = IMPL_NONE;
  float mu = 0.f;
  const void* X = nullptr;  // gathered panel
  void* out = nullptr;
  float alpha = 1.f, c0 = 0.f, c1 = 0.f;
  // EPI_CHEB / EPI_CHEB_LAST accumulator update: 0 adds c1 Lx2 to conv; 1
  // leaves conv alone (the next step adds this step's term from its Lx1); 2
  // adds c_prev Lx1, then c1 Lx2 (the previous step's deferred term first, so
  // the sum is the one the step-by-step update gives)
  int conv_mode = 0;
  float c_prev = 0.f;
  const float* a_in = nullptr;
  const void* lx1 = nullptr;
  const void* lx0 = nullptr;
  float* conv = nullptr;
  void* lx_out = nullptr;
  float* y_out = nullptr;
  // plan (filled by spmm_run)
  int short_nnz = 0, G = 0;
  int64_t nshort = 0;
  const int32_t* short_rows = nullptr;
  int64_t nchunks = 0;
  const int64_t* chunk_start = nullptr;
  const int32_t* chunk_row = nullptr;
  const int32_t* chunk_slot = nullptr;
  int64_t nmulti = 0;
  const int32_t* multi_row = nullptr;
  const int64_t* multi_first = nullptr;
  float* partial = nullptr;  // nparts x k partial sums; NULL: a context buffer
};

// chunk_nnz <= 0: the default chunk length
int spmm_plan_build(lightne_ctx* ctx, const int64_t* rp_dev, int64_t n, SpmmPlan& P, int chunk_nnz = 0);
void spmm_plan_free(SpmmPlan& P);
int spmm_run(lightne_ctx* ctx, const SpmmPlan& P, SpmmArgs a, int xt, int epi, bool impl);
// FP64 form: out(double) = A X(double), explicit CSR values
int spmm_run64(lightne_ctx* ctx, const SpmmPlan& P, SpmmArgs a);

#ifdef __HIPCC__

__device__ __forceinline__ uint32_t f2bf_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return u >> 16;
}

template <typename XT>
__device__ __forceinline__ void load4(const XT* p, float* v);
template <>
__device__ __forceinline__ void load4<float>(const float* p, float* v) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
template <>
__device__ __forceinline__ void load4<__hip_bfloat16>(const __hip_bfloat16* p, float* v) {
  const uint2 t = *reinterpret_cast<const uint2*>(p);
  v[0] = __uint_as_float(t.x << 16);
  v[1] = __uint_as_float(t.x & 0xffff0000u);
  v[2] = __uint_as_float(t.y << 16);
  v[3] = __uint_as_float(t.y & 0xffff0000u);
}

template <typename XT>
__device__ __forceinline__ void store4(XT* p, const float* v);
template <>
__device__ __forceinline__ void store4<float>(float* p, const float* v) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void store4<__hip_bfloat16>(__hip_bfloat16* p, const float* v) {
  uint2 t;
  t.x = f2bf_bits(v[0]) | (f2bf_bits(v[1]) << 16);
  t.y = f2bf_bits(v[2]) | (f2bf_bits(v[3]) << 16);
  *reinterpret_cast<uint2*>(p) = t;
}

template <typename XT, int EPI, bool IMPL>
__device__ __forceinline__ void epilogue(const SpmmArgs& a, int64_t r, int c, float4 acc) {
  float v[4] = {acc.x, acc.y, acc.z, acc.w};
  const int64_t o = r * a.k + c;
  if (IMPL) {
    float xs[4];
    load4<XT>(reinterpret_cast<const XT*>(a.X) + r * a.ldx + c, xs);
    if (a.impl_mode == IMPL_PROP) {
      // SpectralPropagation.h:62,66,72: FP degree = d+1; values formed in double
      const double deg = (double)(float)(a.rp[r + 1] - a.rp[r] + 1);
      const float dg = (float)(1.0 - (double)a.mu - 1.0 / deg);
      const float of = (float)(-1.0 / deg);
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = dg * xs[i] + of * v[i];
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = xs[i] + v[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] *= a.alpha;
  if (EPI == EPI_STORE) {
    store4<XT>(reinterpret_cast<XT*>(a.out) + r * a.ldo + c, v);
  } else if (EPI == EPI_STORE_F32) {
    store4<float>(reinterpret_cast<float*>(a.out) + r * a.ldo + c, v);
  } else if (EPI == EPI_CHEB1) {
    float av[4], cv[4];
    load4<float>(a.a_in + o, av);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = v[i] - av[i];
      cv[i] = a.c0 * av[i] + a.c1 * v[i];
    }
    store4<XT>(reinterpret_cast<XT*>(a.lx_out) + o, v);
    if (a.conv) store4<float>(a.conv + o, cv);
  } else if (EPI == EPI_ADJ) {
    // scaled adjoint of the Chebyshev recurrence: m_{i-1} = -c_{i-1} D~^-1 g
    //   + coef (X m_i) - m_{i+1}, with X m_i = 0.5 M(M m_i) - m_i
    float g[4], l1[4], l0[4] = {0.f, 0.f, 0.f, 0.f};
    load4<float>(a.a_in + o, g);
    load4<XT>(reinterpret_cast<const XT*>(a.lx1) + o, l1);
    if (a.lx0) load4<XT>(reinterpret_cast<const XT*>(a.lx0) + o, l0);
    const float inv_deg = (float)(1.0 / (double)(a.rp[r + 1] - a.rp[r] + 1));
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = a.c1 * (g[i] * inv_deg) + a.c0 * (v[i] - l1[i]) - l0[i];
    store4<XT>(reinterpret_cast<XT*>(a.lx_out) + o, v);
  } else {
    float l1[4], l0[4], cv[4] = {0.f, 0.f, 0.f, 0.f};
    load4<XT>(reinterpret_cast<const XT*>(a.lx1) + o, l1);
    load4<XT>(reinterpret_cast<const XT*>(a.lx0) + o, l0);
    const bool upd = a.conv && a.conv_mode != 1;
    if (upd) load4<float>(a.conv + o, cv);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = (v[i] - l0[i]) + (-2.0f) * l1[i];
      if (a.conv_mode == 2) cv[i] = a.c_prev * l1[i] + cv[i];
      cv[i] = a.c1 * v[i] + cv[i];
    }
    if (EPI == EPI_CHEB) {
      store4<XT>(reinterpret_cast<XT*>(a.lx_out) + o, v);
      if (upd) store4<float>(a.conv + o, cv);  // conv == NULL: plain recurrence
    } else {
      float av[4];
      load4<float>(a.a_in + o, av);
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = av[i] + (-1.0f) * cv[i];
      store4<float>(a.y_out + o, av);
    }
  }
}

#endif  // __HIPCC__
