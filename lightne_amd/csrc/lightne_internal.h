// Internal declarations shared by the HIP translation units of liblightne_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "lightne.h"

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct lightne_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    uint64_t epoch = 0;  // call that last used it
  };
  std::unordered_map<std::string, Buf> bufs;  // grow-only named scratch, kept across calls
  uint64_t epoch = 0;  // numbered top-level library calls
  int depth = 0;       // nesting of library calls in progress
  hipEvent_t ev[16] = {};
  int num_cus = 256;
  size_t lds_per_block = 64 * 1024;  // hipDeviceProp_t::sharedMemPerBlock
  void* scratch = nullptr;  // device address of dense.hip's static scratch array on this device
  void* blas = nullptr;     // rocblas_handle of the wide-rank stage 2 (wide.hip), created on first use
};

struct lightne_table {
  uint64_t nkeys = 0;
  uint64_t* keys = nullptr;  // device, sorted
  uint32_t* cnt = nullptr;   // device
  double* wgh = nullptr;     // device
};

struct SpmmPlan;

struct lightne_csr {
  int64_t n = 0, nnz = 0;
  int64_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  float* val = nullptr;
  bool owns = true;          // false: the arrays are context buffers (internal use)
  SpmmPlan* plan = nullptr;  // nnz-balanced SpMM work split, built on first use
};

int ctx_fail(lightne_ctx* ctx, int code, const char* fmt, ...);
int global_fail(int code, const char* fmt, ...);
void* ctx_buf(lightne_ctx* ctx, const char* name, size_t bytes);  // nullptr on OOM
// Device memory kept across calls is reclaimed only under pressure: when an
// allocation fails inside a scoped call, the cached buffers that call has not
// touched are freed and the allocation is tried once more.
struct CallScope {
  lightne_ctx* c;
  explicit CallScope(lightne_ctx* c_) : c(c_) {
    if (c && c->depth++ == 0) ++c->epoch;
  }
  ~CallScope() {
    if (c) --c->depth;
  }
};
int ctx_malloc(lightne_ctx* ctx, void** p, size_t bytes);  // hipMalloc with reclaim-and-retry; hipError_t
// Sub-buffers carved, 256-B aligned, from the context's shared workspace (the
// buffer named "ws").  Stages whose large temporaries are never live at the
// same time (the sampler's table and sort scratch, the randomized SVD's
// panels, the propagation iterates) reuse the same device memory, so that a
// repeated pipeline neither frees nor maps memory.
struct WsCarve {
  std::vector<std::pair<void**, size_t>> req;
  size_t base = 0;  // first byte of ws this carve may use
  explicit WsCarve(size_t base_ = 0) : base(base_) {}
  void add(void** p, size_t bytes) { req.push_back({p, bytes}); }
  size_t end() const {
    size_t o = base;
    for (auto& r : req) o += (r.second + 255) & ~(size_t)255;
    return o;
  }
  // reserve ws (growing it if needed; contents are lost on growth) and fill
  // the pointers; returns nonzero on OOM
  int commit(lightne_ctx* ctx, size_t reserve = 0);
};

#define LN_HIP(ctx, expr)                                                                        \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return ctx_fail((ctx), LIGHTNE_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr,         \
                      hipGetErrorString(e_));                                                    \
  } while (0)

#define LN_LAUNCH(ctx)                                                                           \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess)                                                                        \
      return ctx_fail((ctx), LIGHTNE_ERR_HIP, "%s:%d launch: %s", __FILE__, __LINE__,            \
                      hipGetErrorString(e_));                                                    \
  } while (0)

#define LN_CHECK(expr)         \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != LIGHTNE_OK) return rc_; \
  } while (0)

// RAII device allocation owned by one call
struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
  void* release() {
    void* q = p;
    p = nullptr;
    return q;
  }
};
int dev_alloc(lightne_ctx* ctx, DevMem& m, size_t bytes, const char* what);

// stage 1 (sparsify.hip): sampler, table, trunc-log filter, CSR.  ctx_owned:
// the CSR arrays are context memory (reused by the next call) and the handle
// does not free them; they then stay clear of the first front_reserve bytes
// of the workspace, which the caller's next stage may carve.  xs (xs_bytes,
// may be NULL): device memory free for stage 1 to use as scratch, e.g. an
// output buffer that is written only after stage 1.
int sparsify_impl(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p, lightne_csr** out,
                  lightne_sparsify_stats* stats, bool ctx_owned, size_t front_reserve, float* xs, size_t xs_bytes);
// Y = A Omega for the very sparse random projection Omega (sparseproj.hip);
// Y is n x k FP32, or FP64 when f64
int sparse_project_sketch(lightne_ctx* ctx, const int64_t* rp, const int32_t* col, const float* val, int64_t n,
                          int64_t k, uint64_t seed, double s_param, void* Y, bool f64);
// full symmetric CSR from upper-triangle storage (sparsify.hip); F's arrays are
// context buffers, F owns nothing
int csr_expand_upper(lightne_ctx* ctx, const lightne_csr* U, lightne_csr* F);
// values of the ne_zhang_et_al matrix on the graph's pattern (nezhang.hip)
int ne_zhang_values(lightne_ctx* ctx, const lightne_graph* g, float* val);

// Blocks for a grid-stride loop over `count` items: one grid dimension may
// hold at most 2^32 - 1 work-items, so large element counts loop.
inline unsigned grid_blocks(int64_t count, int block, int64_t max_blocks = (int64_t)1 << 20) {
  int64_t b = (count + block - 1) / block;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)(b < 1 ? 1 : b);
}

// elapsed seconds between two events recorded on the context stream
double ev_seconds(hipEvent_t a, hipEvent_t b);

// ---------------------------------------------------------------------------
// device primitives (prim.hip)
// ---------------------------------------------------------------------------
// exclusive prefix sum of u32 or u64 counts into u64 offsets; out has n+1
// entries (out[n] = total).  Deterministic.
int scan_u32_to_u64(lightne_ctx* ctx, const uint32_t* in, uint64_t* out, int64_t n);
// stable LSD radix sort of (key, val) pairs on the low `bits` bits of key.
// Sorts on key bits [lo_bit, lo_bit + bits).  The ping-pong buffers (n keys,
// n vals) are k_tmp / v_tmp when given, else context buffers.  keys/vals hold
// the result on return, unless k_res / v_res are given: then they receive
// wherever the last pass left it (no copy back).
int radix_sort_pairs(lightne_ctx* ctx, uint64_t* keys, float* vals, int64_t n, int bits, uint64_t* k_tmp = nullptr,
                     float* v_tmp = nullptr, int lo_bit = 0, uint64_t** k_res = nullptr, float** v_res = nullptr);
int radix_sort_pairs_u32(lightne_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bits);
// digit passes radix_sort_pairs makes over a key of `bits` bits
int radix_sort_passes(int bits);

// ---------------------------------------------------------------------------
// SpMM (spmm.hip)
// ---------------------------------------------------------------------------
enum PanelType { PANEL_F32 = 0, PANEL_BF16 = 1 };

// ---------------------------------------------------------------------------
// dense panels (dense.hip)
// ---------------------------------------------------------------------------
int gaussian_fill(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t count, float* out);
// the same values as gaussian_fill, stored as BF16 (exactly: they are
// BF16-representable)
int gaussian_fill_bf16(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t count, __hip_bfloat16* out);
// the same values as gaussian_fill(count = n k) in an n x k panel of row stride ld
int gaussian_fill_ld(lightne_ctx* ctx, uint64_t seed, uint32_t stream, int64_t n, int64_t k, int64_t ld,
                     float* out);
void gaussian_fill_host(uint64_t seed, uint32_t stream, int64_t count, float* out);
// G(k1 x k2, f64, host) = X^T Y for X: n x k1, Y: n x k2 (row-major f32 device
// panels); symmetric=1 when X == Y (only the upper tiles are computed).
int gram_f64(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2, int symmetric,
             double* G_host);
// Out(n x k2) = X(n x k1) W(k1 x k2) with W a host f64 matrix used as f32.
// Optional row L2-normalisation epilogue (norm > eps).  Out may alias X only
// when k2 == k1 (rows are staged through registers).  upper_tri: W is upper
// triangular (its zero tiles are skipped).
int panel_times_small(lightne_ctx* ctx, const float* X, int64_t n, int64_t k1, int64_t ldx, const double* W_host,
                      int64_t k2, float* Out, int normalize_rows, float* unused, int upper_tri = 0);
// emb = sqrt(S) U[:, :dim] with optional row normalisation (device kernel)
int u_sigma_root_dev(lightne_ctx* ctx, const float* U, const float* S_host, int64_t n, int64_t rank, int64_t dim,
                     int normalize, float* emb);
int f32_to_bf16(lightne_ctx* ctx, const float* in, __hip_bfloat16* out, int64_t count);
// FP64 form of the randomized SVD: Gram of FP64 panels, Out = X W (Out != X),
// and the conversions at its ends
int gram_f64(lightne_ctx* ctx, const double* X, const double* Y, int64_t n, int64_t k1, int64_t k2, int symmetric,
             double* G_host);
// the FP32-panel Gram with every temporary in dense.hip's static scratch array
// (no device allocation); dense_scratch_setup takes the array's device address
// for ctx->device once per device
int gram_f64_noalloc(lightne_ctx* ctx, const float* X, const float* Y, int64_t n, int64_t k1, int64_t k2,
                     int symmetric, double* G_host);
int dense_scratch_setup(lightne_ctx* ctx);

// ---------------------------------------------------------------------------
// stage 2 for sketch widths outside the 16-wide tiling (wide.hip)
// ---------------------------------------------------------------------------
// true when rank k takes the wide path (k > 256 or k not a multiple of 16)
inline bool rsvd_is_wide(int64_t k) { return k > 256 || k % 16 != 0; }
// the randomized SVD with the sketch padded to a multiple of 16 columns:
// out (row-major n x cols, row stride ldo, device) = B V[:, :cols] S^-pw (row
// normalised when asked), or Y = A Omega (cols = k) with random_project_only;
// S receives the k singular values, descending
int rsvd_wide(lightne_ctx* ctx, lightne_csr* A, const lightne_rsvd_params* p, int64_t cols, double pw,
              int normalize, float* out, int64_t ldo, std::vector<double>& S, lightne_rsvd_stats* st);
// workspace front the wide path carves (two padded panels and the BF16 shadow)
size_t rsvd_wide_front(int64_t n, int64_t k, int precision);
void wide_handle_free(lightne_ctx* ctx);
// get_embedding_dense of mat (n x rank, row stride ld) for any rank and dim
int embed_dense_wide(lightne_ctx* ctx, const float* mat, int64_t n, int64_t rank, int64_t ld, int64_t dim,
                     float* emb_out, float* S_host);
int panel_times_small64(lightne_ctx* ctx, const double* X, int64_t n, int64_t k1, const double* W_host, int64_t k2,
                        double* Out);
int f32_to_f64(lightne_ctx* ctx, const float* in, double* out, int64_t count);
int f64_to_f32(lightne_ctx* ctx, const double* in, float* out, int64_t count);

// ---------------------------------------------------------------------------
// host FP64 small linear algebra (linalg_host.cpp)
// ---------------------------------------------------------------------------
#include "linalg_host.h"

// orthonormalise the panel Y (n x k, f32, device) in place by CholeskyQR2 with
// FP64 Gram and Cholesky; optionally returns R (k x k, host) with Y_in = Q R.
int cholqr2(lightne_ctx* ctx, float* Y, int64_t n, int64_t k, double* R_out_host, int* passes_out);
