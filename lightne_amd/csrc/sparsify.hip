// Stage 1: NetSMF path-sampling sparsifier (LightNE/PathEmbed.h:81-316) on the
// device.
//
//  k_edge_src      : source vertex of every directed edge (binary search in
//                    the offsets staged per tile of edges), so the sampler
//                    can run one lane per edge.
//  k_nbr_rows      : every directed edge's target with its degree and first
//                    edge, so a walk step is one random load.
//  k_sample<MODE, PACKED, OFF64, NR> : one lane per directed edge (u, v); its
//                    pbbs::random chain is seeded by fork((u << 32) | v) and
//                    drives exp_rv path samples exactly as PathEmbed.h:128-171;
//                    every sample is aggregated in an open-addressing table
//                    (par_table.h:86-109), linear probing from
//                    hash64_2(key) scaled to the capacity, one 64-bit CAS to
//                    claim a slot.  Slots are 8 bytes when n <= 2^26 (endpoints,
//                    an occupied bit and a 10-bit count whose wraps are added to
//                    a small side table, so counts stay exact), else 16 bytes
//                    {key, count}.
//  k_table_canon + k_canon_fold : every table cluster sorted by key, so the
//                    slot order depends on the key set alone; the occupied
//                    slots per wave quarter of the trunc-log partition, whose
//                    sum is the distinct keys (PathEmbed.h:181-186).
//  k_table_reduce  : -sample 1: distinct keys and total weight
//                    (PathEmbed.h:175-188), deterministic two-level reduction.
//  k_trunc_count / k_trunc_emit : trunc-log (PathEmbed.h:199-218) fused with
//                    the symmetric expansion (PathEmbed.h:278-311): kept keys emit
//                    (row, col) sort keys and float values; per wave quarter
//                    counts, a scan, then each wave writes its run in slot order
//                    (no contended atomics, no block barriers).  Where memory
//                    allows, the count pass also caches every distinct key's
//                    value, so the emission reads no degrees.
//  radix sort + k_csr_finish : stable LSD radix sort on the row bits (on
//                    (row, col) when a cluster was too long to sort), then
//                    row_ptr by boundary detection (no atomics, deterministic).
//  by row ranges : when the sort buffers for every entry do not fit beside
//                    the table and the CSR, the rows are cut into ranges by
//                    the count pass's per-quarter row histograms; each range
//                    is emitted from the table (a scan of its per-quarter
//                    counts, ordered writes), sorted and written into its
//                    slice of the CSR.  The result is the same as one global
//                    sort.
#include "lightne_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr uint64_t EMPTY_KEY = ~0ull;

// wide slot: any n < 2^32
struct __align__(16) Slot {
  unsigned long long key;
  unsigned int cnt;
  unsigned int pad;
};

// packed slot (n <= 2^26), 0 = empty:
//   bits 0..51  (min << 26) | max     bit 52 occupied
//   bit 53      the count wrapped at least once
//   bits 54..63 count mod 1024; an increment adds 1 << 54 and the carry out of
//               bit 63 is lost, so a wrap leaves the key bits intact; the lane
//               whose increment wrapped adds 1024 to the side table
constexpr int PK_VB = 26;
constexpr uint64_t PK_OCC = 1ull << 52;
constexpr uint64_t PK_OVF = 1ull << 53;
constexpr uint64_t PK_ONE = 1ull << 54;
constexpr uint64_t PK_WORD = (PK_OCC << 1) - 1;  // key bits + occupied bit
constexpr uint64_t PK_VMASK = (1ull << PK_VB) - 1;
constexpr uint64_t OVF_CAP = 1ull << 21;  // side table: {word, wrapped count} pairs (32 MiB)

__host__ __device__ __forceinline__ uint64_t hash64_2(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// first probe: the hash scaled to [0, cap)
__device__ __forceinline__ uint64_t slot_of(uint64_t key, uint64_t cap) { return __umul64hi(hash64_2(key), cap); }

__global__ void k_fill_u64(unsigned long long* __restrict__ p, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// side table of wrapped counts: probe for `word`, claim an empty pair
__device__ __forceinline__ bool ovf_add(unsigned long long* ovf, uint64_t word, uint64_t add) {
  uint64_t h = slot_of(word, OVF_CAP);
  for (uint64_t p = 0; p < OVF_CAP; ++p) {
    unsigned long long cur = __hip_atomic_load(&ovf[2 * h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0ull) {
      cur = atomicCAS(&ovf[2 * h], 0ull, (unsigned long long)word);
      if (cur == 0ull) cur = word;
    }
    if (cur == word) {
      atomicAdd(&ovf[2 * h + 1], (unsigned long long)add);
      return true;
    }
    h = (h + 1) & (OVF_CAP - 1);
  }
  return false;
}

__device__ __forceinline__ uint64_t ovf_get(const unsigned long long* __restrict__ ovf, uint64_t word) {
  uint64_t h = slot_of(word, OVF_CAP);
  for (uint64_t p = 0; p < OVF_CAP; ++p) {
    const uint64_t cur = ovf[2 * h];
    if (cur == word) return ovf[2 * h + 1];
    if (cur == 0ull) return 0;
    h = (h + 1) & (OVF_CAP - 1);
  }
  return 0;
}

// read-only view of either table layout
struct TabView {
  const Slot* wide;                // 16-byte slots, or
  const unsigned long long* pk;    // 8-byte slots
  const unsigned long long* ovf;   // side table of the packed layout
  const double* wtab;              // -sample 1 weights (else null)
  uint64_t cap;
};

// a slot as loaded (packed: the word; wide: key and count), so that scans can
// issue several loads before decoding any
struct RawSlot {
  uint64_t a;
  uint32_t b;
};

__device__ __forceinline__ RawSlot raw_load(const TabView& t, uint64_t i) {
  if (t.pk) return RawSlot{t.pk[i], 0u};
  const Slot sl = t.wide[i];
  return RawSlot{sl.key, sl.cnt};
}

// endpoints u <= v and the exact count; false when empty.  PK: the layout
// (8-byte slots) known at compile time
template <bool PK>
__device__ __forceinline__ bool raw_get_t(const TabView& t, const RawSlot& r, uint32_t& u, uint32_t& v,
                                          uint64_t& cnt) {
  if constexpr (PK) {
    const uint64_t w = r.a;
    if (!(w & PK_OCC)) return false;
    u = (uint32_t)((w >> PK_VB) & PK_VMASK);
    v = (uint32_t)(w & PK_VMASK);
    cnt = w >> 54;
    if (w & PK_OVF) cnt += ovf_get(t.ovf, w & PK_WORD);
    return true;
  } else {
    if (r.a == EMPTY_KEY) return false;
    u = (uint32_t)(r.a >> 32);
    v = (uint32_t)r.a;
    cnt = r.b;
    return true;
  }
}

__device__ __forceinline__ bool raw_get(const TabView& t, const RawSlot& r, uint32_t& u, uint32_t& v, uint64_t& cnt) {
  return t.pk ? raw_get_t<true>(t, r, u, v, cnt) : raw_get_t<false>(t, r, u, v, cnt);
}

__device__ __forceinline__ bool tab_get(const TabView& t, uint64_t i, uint32_t& u, uint32_t& v, uint64_t& cnt) {
  return raw_get(t, raw_load(t, i), u, v, cnt);
}

__global__ void k_fill_slots(Slot* __restrict__ tab, uint64_t cap) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < cap; i += stride) {
    Slot s;
    s.key = EMPTY_KEY;
    s.cnt = 0;
    s.pad = 0;
    tab[i] = s;
  }
}

__global__ void k_off32(const int64_t* __restrict__ off, int64_t count, uint32_t* __restrict__ o32) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    o32[i] = (uint32_t)off[i];
}

// largest u in [lo, hi] with off[u] <= j, given off[lo] <= j
__device__ __forceinline__ int64_t edge_row(const int64_t* __restrict__ off, int64_t lo, int64_t hi, int64_t j) {
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Each block takes ES_TILE consecutive edges: two lanes find the rows of its
// first and last edge in the offsets, and when those rows number at most
// ES_ROWS their offsets are staged in LDS and every edge searches there
// (a tile spans about ES_TILE / mean degree rows).  A tile over a longer run
// of rows searches the offsets in memory between the two rows.
constexpr int ES_THREADS = 256, ES_PER = 16, ES_TILE = ES_THREADS * ES_PER, ES_ROWS = 2048;

__global__ __launch_bounds__(ES_THREADS) void k_edge_src(const int64_t* __restrict__ off, int64_t n, int64_t m,
                                                         uint32_t* __restrict__ src) {
  __shared__ int64_t s_off[ES_ROWS];
  __shared__ int64_t s_u[2];
  for (int64_t j0 = (int64_t)blockIdx.x * ES_TILE; j0 < m; j0 += (int64_t)gridDim.x * ES_TILE) {
    const int64_t j1 = j0 + ES_TILE < m ? j0 + ES_TILE : m;
    if (threadIdx.x < 2) s_u[threadIdx.x] = edge_row(off, 0, n - 1, threadIdx.x ? j1 - 1 : j0);
    __syncthreads();
    const int64_t u0 = s_u[0], u1 = s_u[1], rows = u1 - u0 + 1;
    if (rows <= ES_ROWS) {
      for (int64_t i = threadIdx.x; i < rows; i += ES_THREADS) s_off[i] = off[u0 + i];
      __syncthreads();
#pragma unroll 4
      for (int e = 0; e < ES_PER; ++e) {
        const int64_t j = j0 + e * ES_THREADS + threadIdx.x;
        if (j < j1) {
          int lo = 0, hi = (int)rows - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= j)
              lo = mid;
            else
              hi = mid - 1;
          }
          src[j] = (uint32_t)(u0 + lo);
        }
      }
    } else {
      for (int e = 0; e < ES_PER; ++e) {
        const int64_t j = j0 + e * ES_THREADS + threadIdx.x;
        if (j < j1) src[j] = (uint32_t)edge_row(off, u0, u1, j);
      }
    }
    __syncthreads();  // the next tile rewrites s_u and s_off
  }
}

// block-wide sum of three counters, one atomic per counter per block
__device__ __forceinline__ void block_add3(unsigned long long* dst, uint64_t a, uint64_t b, uint64_t c) {
  __shared__ uint64_t s_r[3][4];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    s_r[0][wid] = a;
    s_r[1][wid] = b;
    s_r[2][wid] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    uint64_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_r[threadIdx.x][w];
    if (t) atomicAdd(&dst[threadIdx.x], (unsigned long long)t);
  }
}

// A directed edge's target with that vertex's row: a walk step reads one of
// these (one 16-byte load, one line) instead of a neighbour and then, at the
// next step, the neighbour's row pointers (a second random line).  Built per
// call where the workspace has room and every degree fits 32 bits (2m < 2^32).
struct NbrRow {
  uint32_t v;    // the edge's target
  uint32_t deg;  // its degree
  uint64_t off;  // its first edge
};
static_assert(sizeof(NbrRow) == 16, "one 16-byte load per walk step");

struct SampleArgs {
  // row pointers: 32 bits when 2m < 2^32 (half the bytes of the walk's hot
  // array), else the graph's own 64-bit offsets
  const uint32_t* off32;
  const int64_t* off64;
  const uint32_t* adj;
  const NbrRow* nr;  // neighbour rows (NR instantiations), else null
  uint64_t ith;  // 1: compressed-graph neighbour lookup
  const uint32_t* src;
  int64_t m;
  int64_t j0;  // first directed edge of this launch
  uint64_t seed;
  uint64_t walk_len;
  uint64_t exp_rv;
  double frac;
  double logn;
  Slot* tab;                 // wide layout
  unsigned long long* pk;    // packed layout
  unsigned long long* ovf;   // packed layout: side table of wrapped counts
  double* wtab;
  uint64_t cap;
  unsigned long long* stats;  // 0 samples, 1 steps, 2 attempts, 3 table overflow, 4 bad graph, 5 side table full
};

// random_walk (PathEmbed.h:37-49): `len` steps from x, state advances one
// draw per step.  rand = H(s); next = H(H(s)) = H(rand).  Draw i picks
// neighbour i, or with `ith` = 1 (graphs read from the compressed format)
// neighbour i - 1 unless i is a multiple of 64: the result of
// get_ith_neighbor on such graphs (ligra/encodings/byte_pd_amortized.h:653-688
// decodes i - start - 1 differences after the block's first edge).
// The start vertex's row (o0, d0) comes from the caller, which reads it once
// for all of an edge's samples (every walk from u or v starts there).
template <typename OffT>
__device__ __forceinline__ uint32_t walk(const OffT* __restrict__ off, const uint32_t* __restrict__ adj, uint32_t x,
                                         uint64_t len, uint64_t& s, bool& bad, uint64_t ith, OffT o0, uint64_t d0) {
  for (uint64_t t = 0; t < len; ++t) {
    const OffT o = t ? off[x] : o0;
    const uint64_t d = t ? (uint64_t)(off[x + 1] - o) : d0;
    if (d == 0) {
      bad = true;
      break;
    }
    const uint64_t r = hash64_2(s);
    uint64_t i = r % d;
    i -= ith & (uint64_t)((i & 63) != 0);
    x = adj[(uint64_t)o + i];
    s = hash64_2(r);
  }
  return x;
}

// (from the 32-bit row pointers: the gather's array is half the bytes)
__global__ void k_nbr_rows(const uint32_t* __restrict__ off32, const uint32_t* __restrict__ adj, int64_t m,
                           NbrRow* __restrict__ nr) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t v = adj[j];
    const uint32_t o = off32[v];
    nr[j] = NbrRow{v, off32[v + 1] - o, (uint64_t)o};
  }
}

// random_walk over the neighbour rows: the same draws and picks as walk(),
// the row of each vertex reached coming with it
__device__ __forceinline__ uint32_t walk_nr(const NbrRow* __restrict__ nr, uint32_t x, uint64_t len, uint64_t& s,
                                            bool& bad, uint64_t ith, uint64_t o, uint64_t d) {
  for (uint64_t t = 0; t < len; ++t) {
    if (d == 0) {
      bad = true;
      break;
    }
    const uint64_t r = hash64_2(s);
    uint64_t i = r % d;
    i -= ith & (uint64_t)((i & 63) != 0);
    const NbrRow e = nr[o + i];
    x = e.v;
    o = e.off;
    d = e.deg;
    s = hash64_2(r);
  }
  return x;
}

// par_table::insert_add (par_table.h:86-109): adds c samples (and with
// MODE 1 the weight w) to `key`'s count; 0 when the table is full, -1 when
// the packed layout's side table is full.  Packed counts keep their low 10
// bits in the slot (an add of c * PK_ONE wraps them modulo 1024 without
// touching the key bits) and 1024 per wrap in the side table.
template <int MODE, bool PACKED>
__device__ __forceinline__ int table_add(const SampleArgs& a, uint32_t lo, uint32_t hi, double w, uint64_t c = 1) {
  const uint64_t key = ((uint64_t)lo << 32) + (uint64_t)hi;
  uint64_t h = slot_of(key, a.cap);
  if (PACKED) {
    const uint64_t word = ((uint64_t)lo << PK_VB) | (uint64_t)hi | PK_OCC;
    for (uint64_t p = 0; p < a.cap; ++p) {
      unsigned long long cur = __hip_atomic_load(&a.pk[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == 0ull) {
        cur = atomicCAS(&a.pk[h], 0ull, (unsigned long long)(word | (c & 1023) * PK_ONE | (c >> 10 ? PK_OVF : 0)));
        if (cur == 0ull) {
          int rc = 1;
          if ((c >> 10) && !ovf_add(a.ovf, word, c & ~1023ull)) rc = -1;
          if (MODE == 1) atomicAdd(&a.wtab[h], w);
          return rc;
        }
      }
      if ((cur & PK_WORD) == word) {
        const unsigned long long old = atomicAdd(&a.pk[h], (unsigned long long)(c * PK_ONE));
        int rc = 1;
        const uint64_t wraps = ((old >> 54) + c) >> 10;
        if (wraps) {
          atomicOr(&a.pk[h], (unsigned long long)PK_OVF);
          if (!ovf_add(a.ovf, word, 1024 * wraps)) rc = -1;
        }
        if (MODE == 1) atomicAdd(&a.wtab[h], w);
        return rc;
      }
      h = (h + 1 == a.cap) ? 0 : h + 1;
    }
    return 0;
  }
  for (uint64_t p = 0; p < a.cap; ++p) {
    unsigned long long cur = __hip_atomic_load(&a.tab[h].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == EMPTY_KEY) {
      cur = atomicCAS(&a.tab[h].key, EMPTY_KEY, (unsigned long long)key);
      if (cur == EMPTY_KEY) cur = key;
    }
    if (cur == key) {
      atomicAdd(&a.tab[h].cnt, (unsigned int)c);
      if (MODE == 1) atomicAdd(&a.wtab[h], w);
      return 1;
    }
    h = (h + 1 == a.cap) ? 0 : h + 1;
  }
  return 0;
}

template <int MODE, bool PACKED, bool OFF64, bool NR>
__global__ __launch_bounds__(256) void k_sample(SampleArgs a) {
  const int64_t j = a.j0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const auto* off = [&] {
    if constexpr (OFF64)
      return a.off64;
    else
      return a.off32;
  }();
  uint64_t n_samples = 0, n_steps = 0, n_attempts = 0;
  bool overflow = false, bad = false, side_full = false;
  if (j < a.m) {
    const uint32_t u = a.src[j];
    NbrRow ej{};
    if (NR) ej = a.nr[j];
    const uint32_t v = NR ? ej.v : a.adj[j];
    const uint64_t uv = ((uint64_t)u << 32) | (uint64_t)v;
    uint64_t s = hash64_2(hash64_2(uv + a.seed));  // seed.fork(uv)
    uint64_t e = a.exp_rv;
    if (a.frac > 0.0000000001) {
      const uint64_t r = hash64_2(s);
      e += ((double)r * 0x1p-64 < a.frac) ? 1 : 0;  // map_uint64_t(rand) < frac
      s = hash64_2(r);
    }
    n_attempts = e;
    double p = 1.0, ipw = 1.0;
    // the rows of u and v, where every walk of this edge's samples starts
    const auto ou = off[u];
    const auto ov = NR ? (decltype(ou))ej.off : off[v];
    const uint64_t du_ = (uint64_t)(off[u + 1] - ou), dv_ = NR ? (uint64_t)ej.deg : (uint64_t)(off[v + 1] - ov);
    if (MODE == 1) {
      const double du = (double)du_;
      const double dv = (double)dv_;
      p = (1.0 / du + 1.0 / dv) * a.logn;
      p = p < 1.0 ? p : 1.0;
      ipw = 1.0 / p;
    }
    // -sample 0: the samples of one walk length (r = 1) all hit the edge's own
    // key; their count goes into the table in one add
    uint64_t c_uv = 0;
    for (uint64_t k = 0; k < e; ++k) {
      if (MODE == 1) {
        const uint64_t q = hash64_2(s);
        s = hash64_2(q);
        if (!((double)q * 0x1p-64 < p)) continue;
      }
      uint64_t rr = hash64_2(s);
      const uint64_t r = (uint32_t)(rr % a.walk_len + 1);
      s = hash64_2(rr);
      rr = hash64_2(s);
      const uint32_t kk = (uint32_t)(rr % r + 1);  // path_sample, PathEmbed.h:54
      s = hash64_2(rr);
      if (MODE == 0 && r == 1) {  // walks of no step: the edge itself, added once below
        ++c_uv;
        ++n_samples;
        continue;
      }
      uint32_t x, z;
      if (NR) {
        x = walk_nr(a.nr, u, (uint64_t)kk - 1, s, bad, a.ith, (uint64_t)ou, du_);
        z = walk_nr(a.nr, v, r - kk, s, bad, a.ith, (uint64_t)ov, dv_);
      } else {
        x = walk(off, a.adj, u, (uint64_t)kk - 1, s, bad, a.ith, ou, du_);
        z = walk(off, a.adj, v, r - kk, s, bad, a.ith, ov, dv_);
      }
      n_steps += r - 1;
      const int rc = table_add<MODE, PACKED>(a, min(x, z), max(x, z), ipw);
      if (rc == 0) overflow = true;
      if (rc < 0) side_full = true;
      ++n_samples;
    }
    if (MODE == 0 && c_uv) {
      const int rc = table_add<MODE, PACKED>(a, min(u, v), max(u, v), 1.0, c_uv);
      if (rc == 0) overflow = true;
      if (rc < 0) side_full = true;
    }
  }
  block_add3(a.stats, n_samples, n_steps, n_attempts);
  if (overflow) atomicOr(&a.stats[3], 1ull);
  if (bad) atomicOr(&a.stats[4], 1ull);
  if (side_full) atomicOr(&a.stats[5], 1ull);
}

// ---------------------------------------------------------------------------
// Partition of the table shared by the canonical-order pass and the trunc-log
// passes: trunc_blocks(cap) blocks of trunc_per_block slots each (a multiple
// of TR_WAVES * TR_ITEMS).  Wave w of a block scans its own contiguous quarter
// of the block's slots (wave_span), so every thread's TR_ITEMS slots start
// 32-byte aligned in a table of 8-byte slots, and a wave numbers the occupied
// slots of its quarter in slot order by itself (the value cache's index).
// ---------------------------------------------------------------------------
constexpr int TR_THREADS = 256;
constexpr int TR_ITEMS = 4;  // consecutive slots per thread per step: loads in flight
constexpr int TR_WAVES = TR_THREADS / 64;

// enough blocks that the last round of resident blocks leaves little of the
// device idle on large tables (about seven 256-thread blocks fit a CU)
inline int trunc_blocks(uint64_t cap) { return cap >= (1ull << 30) ? 16384 : 4096; }

inline uint64_t trunc_per_block(uint64_t cap, int nblk) {
  constexpr uint64_t U = (uint64_t)TR_WAVES * TR_ITEMS;
  const uint64_t pb = (cap + nblk - 1) / nblk;
  return (pb + U - 1) / U * U;
}

// ---------------------------------------------------------------------------
// Canonical slot order.  Which slots of a linear-probing table are occupied
// does not depend on the order of the inserts, and neither does which keys a
// cluster (a maximal run of occupied slots) holds: exactly those whose home
// slots lie in it.  Which key sits in which slot of its cluster does depend on
// that order, and concurrent inserts claim slots in a different order on
// every run.  This pass sorts every cluster by key.  The table is then no
// longer a probe layout (nothing looks a key up after it: the passes that
// follow scan the slots), but its slot order depends on the key set alone, so
// every pass that follows slot order (the trunc-log emission, hence the order
// of the entries within a CSR row and of every sum over a row) is the same on
// every run.
//
// The slots of every trunc-log block are cut into windows of CN_WIN owned
// slots (the block's last one shorter), taken by the grid in slot order.  A
// window stages its owned slots, the one before and CN_LA after them
// in LDS, with an occupancy bitmap (one ballot per 64 slots).  Every staged
// occupied slot finds its cluster's head and end from the bitmap (the nearest
// empty slots before and after it) and, when the head is owned, moves to the
// head's slot plus its rank (the number of keys of its cluster below its
// own).  A cluster running past the staged slots raises *err and keeps its
// order (the caller then orders the CSR rows by column instead).
//
// The pass also records per window what the trunc-log passes need first:
// the occupied owned slots in each wave quarter of its trunc-log block (a
// slot's occupancy does not change when its cluster is sorted), summed per
// block by k_canon_fold.
// ---------------------------------------------------------------------------
constexpr int CN_THREADS = 256;
constexpr int CN_WIN = 2048;                      // owned slots per block step
constexpr int CN_LA = 512;                        // staged slots past the owned ones
constexpr int CN_STAGE = 1 + CN_WIN + CN_LA;      // the slot before, the owned ones, the lookahead
constexpr int CN_BM = (CN_STAGE + 63) / 64;       // occupancy bitmap words
static_assert(CN_THREADS % 64 == 0, "one bitmap word per wave and 64 staged slots");
static_assert(CN_WIN % 64 == 0, "owned slots in whole bitmap words");

// occupied owned slots of a window per wave quarter of its trunc-log block
struct CanonWin {
  uint32_t occ[TR_WAVES];
};

struct CanonArgs {
  Slot* wide;
  unsigned long long* pk;
  const unsigned long long* ovf;  // side table of the packed layout
  double* wtab;
  uint64_t cap;
  uint64_t per_block;             // slots per trunc-log block
  uint64_t wpb;                   // windows per trunc-log block
  uint64_t nwin;                  // windows: trunc-log blocks x wpb
  uint32_t* err;
  CanonWin* win;                  // [nwin]
};

// a staged slot holds the word (packed) or the key (wide)
template <bool PK>
__device__ __forceinline__ bool cn_occ(uint64_t w) {
  return PK ? (w & PK_OCC) != 0 : w != EMPTY_KEY;
}
template <bool PK>
__device__ __forceinline__ uint64_t cn_key(uint64_t w) {
  return PK ? (w & (PK_OCC - 1)) : w;
}
__device__ __forceinline__ uint64_t cn_pos(uint64_t i, uint64_t j, uint64_t cap) {
  const uint64_t p = i + j;
  return p >= cap ? p - cap : p;
}

// registers for seven waves per SIMD where the LDS allows seven blocks of
// this size per CU (the packed layout without weights; the other layouts
// stage more per slot and fit fewer)
template <bool PK, bool WT>
constexpr int canon_min_waves() {
  return PK && !WT ? 7 : 1;
}

template <bool PK, bool WT>
__global__ __launch_bounds__(CN_THREADS, (canon_min_waves<PK, WT>())) void k_table_canon(CanonArgs a) {
  __shared__ uint64_t s_w[CN_STAGE];
  __shared__ uint32_t s_c[PK ? 1 : CN_STAGE];
  __shared__ double s_x[WT ? CN_STAGE : 1];
  __shared__ uint64_t s_bm[CN_BM];  // bit k % 64 of word k / 64: staged slot k occupied
  __shared__ uint32_t s_occ[TR_WAVES];
  const uint64_t cap = a.cap;
  const int tid = threadIdx.x;
  const uint64_t q = a.per_block / TR_WAVES;  // slots per wave quarter
  if (tid < TR_WAVES) s_occ[tid] = 0;
  for (uint64_t t = blockIdx.x; t < a.nwin; t += gridDim.x) {
    // window t: the (t % wpb)-th of trunc-log block t / wpb
    const uint64_t b0 = (t / a.wpb) * a.per_block;
    const uint64_t b1 = min(cap, b0 + a.per_block);
    const uint64_t w0 = b0 + (t % a.wpb) * CN_WIN;
    // owned: slots w0 .. w0 + own - 1 (none past the table's end)
    const int own = w0 < b1 ? (int)min((uint64_t)CN_WIN, b1 - w0) : 0;
    if (own > 0) {  // block-uniform
      // staged index k is slot st0 + k (modulo cap): the slot before w0, the
      // owned ones, the lookahead
      const uint64_t st0 = w0 ? w0 - 1 : cap - 1;
      for (int k0 = 0; k0 < CN_STAGE; k0 += CN_THREADS) {  // block-uniform: every lane ballots
        const int k = k0 + tid;
        bool occ = false;
        if (k < CN_STAGE) {
          const uint64_t p = cn_pos(st0, (uint64_t)k % cap, cap);
          uint64_t w;
          if (PK) {
            w = a.pk[p];
          } else {
            const Slot sl = a.wide[p];
            w = sl.key;
            s_c[k] = sl.cnt;
          }
          s_w[k] = w;
          occ = cn_occ<PK>(w);
          if (WT) s_x[k] = a.wtab[p];
        }
        const uint64_t b = __ballot(occ);
        if ((tid & 63) == 0 && k < CN_STAGE) s_bm[k >> 6] = b;
      }
      __syncthreads();
      // occupied owned slots per wave quarter: word i covers staged 1 + 64 i
      // .. 64 + 64 i, slots w0 + 64 i .. (a quarter holds at least TR_ITEMS
      // slots)
      for (int i = tid; 64 * i < own; i += CN_THREADS) {
        uint64_t bits = (s_bm[i] >> 1) | (s_bm[i + 1] << 63);
        const int len = min(64, own - 64 * i);
        if (len < 64) bits &= (1ull << len) - 1;
        uint64_t off = w0 + 64 * (uint64_t)i - b0;  // offset of bit j in the trunc-log block
        int j = 0;
        while (j < len) {
          int wv = 0;  // the quarter of offset off (< per_block = TR_WAVES q)
          while (wv + 1 < TR_WAVES && off >= (uint64_t)(wv + 1) * q) ++wv;
          const uint64_t left = (uint64_t)(wv + 1) * q - off;  // >= 1 slot left in the quarter
          const int jn = left < (uint64_t)(len - j) ? j + (int)left : len;
          const uint64_t m = (jn - j == 64) ? ~0ull : ((1ull << (jn - j)) - 1) << j;
          const uint32_t c = (uint32_t)__popcll(bits & m);
          if (c) atomicAdd(&s_occ[wv], c);
          off += (uint64_t)(jn - j);
          j = jn;
        }
      }
      for (int k = 1 + tid; k < CN_STAGE; k += CN_THREADS) {
        if (!((s_bm[k >> 6] >> (k & 63)) & 1ull)) continue;
        const uint64_t w = s_w[k];
        // the cluster's head: after the nearest empty slot before k (h = 0:
        // every staged slot before k is occupied)
        int h = 0;
        for (int wi = k >> 6; wi >= 0; --wi) {
          uint64_t z = ~s_bm[wi];
          if (wi == (k >> 6)) z &= (1ull << (k & 63)) - 1;  // the slots before k
          if (z) {
            h = 64 * wi + 63 - __clzll((long long)z) + 1;
            break;
          }
        }
        // owned: headed in [w0, w0 + own) (h = 0: headed before w0; staged
        // slots past the table's end wrap to its start, which window 0 owns)
        if (h == 0 || h > own) continue;
        // the cluster's end: its first empty slot after k (staged bits past
        // CN_STAGE read as empty)
        int e = CN_STAGE;
        for (int wi = k >> 6; wi < CN_BM; ++wi) {
          uint64_t z = ~s_bm[wi];
          if (wi == (k >> 6)) z &= ~((2ull << (k & 63)) - 1);  // the slots after k
          if (z) {
            e = 64 * wi + __ffsll((long long)z) - 1;
            break;
          }
        }
        if (e >= CN_STAGE) {  // runs past the staged slots: left as it is
          atomicOr(a.err, 1u);
          continue;
        }
        const uint64_t key = cn_key<PK>(w);
        int rank = 0;
        for (int r = h; r < e; ++r) rank += cn_key<PK>(s_w[r]) < key ? 1 : 0;
        if (h + rank == k) continue;
        const uint64_t qd = cn_pos(st0, (uint64_t)(h + rank), cap);
        if (PK) {
          a.pk[qd] = w;
        } else {
          Slot sl;
          sl.key = w;
          sl.cnt = s_c[k];
          sl.pad = 0;
          a.wide[qd] = sl;
        }
        if (WT) a.wtab[qd] = s_x[k];
      }
    }
    __syncthreads();
    if (tid == 0) {  // the window's record; the counters start again at zero
      CanonWin r;
#pragma unroll
      for (int i = 0; i < TR_WAVES; ++i) {
        r.occ[i] = s_occ[i];
        s_occ[i] = 0;
      }
      a.win[t] = r;
    }
    // the next window's staging barrier orders the reset before its updates
  }
}

// per trunc-log block: the sums of its windows' records (the occupied slots
// of each wave quarter, and of the block)
__global__ __launch_bounds__(256) void k_canon_fold(const CanonWin* __restrict__ win, uint64_t wpb,
                                                    uint32_t* __restrict__ wocc, unsigned long long* __restrict__ bocc) {
  __shared__ uint32_t s_r[TR_WAVES][256 / 64];
  uint32_t v[TR_WAVES] = {};
  const CanonWin* w = win + (uint64_t)blockIdx.x * wpb;
  for (uint64_t i = threadIdx.x; i < wpb; i += 256) {
    const CanonWin r = w[i];
#pragma unroll
    for (int j = 0; j < TR_WAVES; ++j) v[j] += r.occ[j];
  }
#pragma unroll
  for (int j = 0; j < TR_WAVES; ++j) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v[j] += __shfl_xor(v[j], d, 64);
    if ((threadIdx.x & 63) == 0) s_r[j][threadIdx.x >> 6] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    for (int j = 0; j < TR_WAVES; ++j) {
      uint32_t t = 0;
      for (int i = 0; i < 256 / 64; ++i) t += s_r[j][i];
      wocc[(uint64_t)blockIdx.x * TR_WAVES + j] = t;
      b += t;
    }
    bocc[blockIdx.x] = b;
  }
}

// per-block partial (distinct, count sum, weight sum); fixed reduction order
__global__ __launch_bounds__(256) void k_table_reduce(TabView t, unsigned long long* __restrict__ part_u,
                                                      double* __restrict__ part_w) {
  __shared__ unsigned long long s_d[256], s_c[256];
  __shared__ double s_w[256];
  uint64_t d = 0, c = 0;
  double w = 0.0;
  const uint64_t per = (t.cap + gridDim.x - 1) / gridDim.x;
  const uint64_t b0 = (uint64_t)blockIdx.x * per, b1 = min(t.cap, b0 + per);
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    uint32_t u, v;
    uint64_t cnt;
    if (tab_get(t, i, u, v, cnt)) {
      ++d;
      c += cnt;
      if (t.wtab) w += t.wtab[i];
    }
  }
  s_d[threadIdx.x] = d;
  s_c[threadIdx.x] = c;
  s_w[threadIdx.x] = w;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      s_d[threadIdx.x] += s_d[threadIdx.x + st];
      s_c[threadIdx.x] += s_c[threadIdx.x + st];
      s_w[threadIdx.x] += s_w[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part_u[2 * blockIdx.x] = s_d[0];
    part_u[2 * blockIdx.x + 1] = s_c[0];
    part_w[blockIdx.x] = s_w[0];
  }
}

// compact the table into (key, count[, weight]) arrays for the counts API
__global__ __launch_bounds__(256) void k_table_compact(TabView t, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ idx,
                                                       unsigned long long* __restrict__ counter) {
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < t.cap; i0 += (uint64_t)gridDim.x * blockDim.x) {
  const uint64_t i = i0 + threadIdx.x;
  uint32_t u = 0, v = 0;
  uint64_t cnt = 0;
  const bool has = i < t.cap && tab_get(t, i, u, v, cnt);
  const uint64_t bal = __ballot(has);
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0 && bal) base = atomicAdd(counter, (unsigned long long)__popcll(bal));
  base = __shfl(base, 0, 64);
  if (has) {
    const uint64_t pos = base + __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    keys[pos] = ((uint64_t)u << 32) + v;
    idx[pos] = (uint32_t)i;  // slot index (cap <= 2^32 enforced by the host)
  }
  }
}

__global__ void k_gather_counts(TabView t, const uint32_t* __restrict__ idx, int64_t nk, uint32_t* __restrict__ cnt,
                                double* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t s = idx[i];
    uint32_t u, v;
    uint64_t c = 0;
    tab_get(t, s, u, v, c);
    cnt[i] = (uint32_t)c;
    if (w) w[i] = t.wtab ? t.wtab[s] : (double)c;
  }
}

struct TruncArgs {
  TabView t;
  const uint32_t* deg;  // vertex degrees (GA.get_vertex(u).getOutDegree())
  double factor;
  int upper;
  int colbits;
  uint64_t* skey;
  float* sval;
  // value cache, indexed by the rank of an occupied slot among all occupied
  // slots in slot order: the count pass writes the CSR value of every
  // occupied slot of rank below lw_n ((float)log(adj) when kept, else 0) to
  // lw_out; an emission pass given lw_in reads it for those slots instead of
  // the two degrees.  occ_off[block * TR_WAVES + wave]: the rank of the first
  // occupied slot of that wave's quarter of the block.
  float* lw_out;
  const float* lw_in;
  uint64_t lw_n;
  const uint64_t* occ_off;
};

// PathEmbed.h:203-210: adjusted weight of a key of value val (count or
// -sample 1 weight) and endpoint degrees du, dv, evaluated in the reference's
// order (w * factor / d_u / d_v, diagonal weight doubled)
__device__ __forceinline__ double trunc_adj_w(const TruncArgs& a, uint32_t u, uint32_t v, double val, uint32_t du,
                                              uint32_t dv) {
  const double wgh = (u != v) ? val : val * 2;
  return wgh * a.factor / (double)du / (double)dv;
}



__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* s_w) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) s_w[wid] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int w = 0; w < TR_THREADS / 64; ++w) t += s_w[w];
  return t;
}

constexpr int HIST_BUCKETS = 1024;  // row buckets of a wave's histogram

// the calling wave's quarter [w0, w1) of its block's slots; q (a multiple of
// TR_ITEMS) is the same for every wave, so loops over [0, q) take the same
// number of steps across the block (w1 < w0 + q only at the table's end)
struct WaveSpan {
  uint64_t w0, w1, q;
};
__device__ __forceinline__ WaveSpan wave_span(uint64_t cap, uint64_t per_block) {
  const uint64_t q = per_block / TR_WAVES;
  const uint64_t w0 = (uint64_t)blockIdx.x * per_block + (uint64_t)(threadIdx.x >> 6) * q;
  return WaveSpan{w0, min(cap, w0 + q), q};
}

// exclusive prefix of c over the wave's lanes; *tot: the wave's total
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t c, uint32_t* tot) {
  const int lane = threadIdx.x & 63;
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  *tot = __shfl(incl, 63, 64);
  return incl - c;
}

// a thread's TR_ITEMS consecutive slots [ib, ib + TR_ITEMS) of a scan, empty
// past b1: a whole group of 8-byte slots as two 16-byte loads (ib is a
// multiple of TR_ITEMS and the table starts 256-byte aligned)
template <bool PK>
__device__ __forceinline__ void load_group(const TabView& t, uint64_t ib, uint64_t b1, RawSlot* rs) {
  if constexpr (PK) {
    if (ib + TR_ITEMS <= b1) {
      const auto* p = reinterpret_cast<const ulonglong2*>(t.pk + ib);
      const ulonglong2 x0 = p[0], x1 = p[1];
      rs[0] = RawSlot{x0.x, 0u};
      rs[1] = RawSlot{x0.y, 0u};
      rs[2] = RawSlot{x1.x, 0u};
      rs[3] = RawSlot{x1.y, 0u};
      return;
    }
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) rs[j] = ib + j < b1 ? RawSlot{t.pk[ib + j], 0u} : RawSlot{0ull, 0u};
  } else {
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      if (ib + j < b1) {
        const Slot sl = t.wide[ib + j];
        rs[j] = RawSlot{sl.key, sl.cnt};
      } else {
        rs[j] = RawSlot{EMPTY_KEY, 0u};
      }
    }
  }
}

template <bool PK>
__device__ __forceinline__ bool raw_occ(const RawSlot& r) {
  if constexpr (PK)
    return (r.a & PK_OCC) != 0;
  else
    return r.a != EMPTY_KEY;
}

// pass 1: per wave quarter (contiguous slot range) the emitted entries,
// wav_emit[block * TR_WAVES + wave], and per block the kept keys; with HIST
// also every wave's row histogram (HIST_BUCKETS ranges of 2^shift rows),
// stored bucket-major: bh[bucket * (gridDim.x * TR_WAVES) + block * TR_WAVES + wave]
template <bool HIST, bool PK>
__global__ __launch_bounds__(TR_THREADS) void k_trunc_count(TruncArgs a, uint64_t per_block,
                                                            uint32_t* __restrict__ wav_emit,
                                                            uint32_t* __restrict__ blk_kept, int shift, int nbuck,
                                                            uint32_t* __restrict__ bh) {
  __shared__ uint32_t s_w[TR_WAVES];
  __shared__ uint32_t s_h[HIST ? TR_WAVES * HIST_BUCKETS : 1];
  if (HIST) {
    for (int b = threadIdx.x; b < TR_WAVES * HIST_BUCKETS; b += TR_THREADS) s_h[b] = 0;
    __syncthreads();
  }
  const WaveSpan ws = wave_span(a.t.cap, per_block);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t* const wh = s_h + (HIST ? wid * HIST_BUCKETS : 0);  // the wave's histogram
  // occupied rank of the wave's next occupied slot (value cache index)
  uint64_t orun = a.lw_out ? a.occ_off[(uint64_t)blockIdx.x * TR_WAVES + (threadIdx.x >> 6)] : 0;
  uint32_t e = 0, kk = 0;
  for (uint64_t t0 = 0; t0 < ws.q; t0 += 64 * TR_ITEMS) {
    const uint64_t ib = ws.w0 + t0 + (uint64_t)lane * TR_ITEMS;
    RawSlot rs[TR_ITEMS];
    load_group<PK>(a.t, ib, ws.w1, rs);
    // decode the group, then issue all its degree reads (and -sample 1
    // weights) before using any: eight random reads in flight per thread
    uint32_t uu[TR_ITEMS], vv[TR_ITEMS], du[TR_ITEMS], dv[TR_ITEMS];
    uint64_t cc[TR_ITEMS];
    double wv[TR_ITEMS];
    uint32_t occ = 0;
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      if (raw_get_t<PK>(a.t, rs[j], uu[j], vv[j], cc[j])) {
        occ |= 1u << j;
      } else {
        uu[j] = vv[j] = 0;  // an empty slot reads vertex 0's degree, unused
        cc[j] = 0;
      }
    }
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      du[j] = a.deg[uu[j]];
      dv[j] = a.deg[vv[j]];
      wv[j] = (a.t.wtab && ((occ >> j) & 1u)) ? a.t.wtab[ib + j] : (double)cc[j];
    }
    float lwo[TR_ITEMS];
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      lwo[j] = 0.f;
      if (!((occ >> j) & 1u)) continue;
      const uint32_t u = uu[j], v = vv[j];
      const double adj = trunc_adj_w(a, u, v, wv[j], du[j], dv[j]);
      // log(adj) > 0  <=>  adj > 1 for any faithful log
      if (!(adj > 1.0)) continue;
      if (a.lw_out) lwo[j] = (float)log(adj);
      const bool ev = !a.upper && u != v;
      e += 1 + (ev ? 1 : 0);
      ++kk;
      if (HIST) {
        atomicAdd(&wh[u >> shift], 1u);
        if (ev) atomicAdd(&wh[v >> shift], 1u);
      }
    }
    // the value cache: the wave's occupied slots in slot order take
    // consecutive ranks (a wave writes one contiguous run of floats)
    if (a.lw_out) {
      uint32_t wtot;
      uint64_t ci = orun + wave_excl_scan((uint32_t)__popc(occ), &wtot);
#pragma unroll
      for (int j = 0; j < TR_ITEMS; ++j) {
        if (!((occ >> j) & 1u)) continue;
        if (ci < a.lw_n) a.lw_out[ci] = lwo[j];
        ++ci;
      }
      orun += wtot;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) e += __shfl_xor(e, d, 64);
  if (lane == 0) wav_emit[(uint64_t)blockIdx.x * TR_WAVES + wid] = e;
  const uint32_t tk = block_sum_u32(kk, s_w);  // its barriers also order the histograms' updates
  if (threadIdx.x == 0) blk_kept[blockIdx.x] = tk;
  if (HIST) {
    const uint64_t nw = (uint64_t)gridDim.x * TR_WAVES;
    for (int i = threadIdx.x; i < nbuck * TR_WAVES; i += TR_THREADS) {
      const int b = i / TR_WAVES, w = i % TR_WAVES;
      bh[(uint64_t)b * nw + (uint64_t)blockIdx.x * TR_WAVES + w] = s_h[w * HIST_BUCKETS + b];
    }
  }
}

// row histogram totals: bucket b (one block per bucket) sums its nw wave
// quarters' counts, read along the bucket's contiguous row of bh
__global__ __launch_bounds__(256) void k_bh_totals(const uint32_t* __restrict__ bh, int nbuck, int nw,
                                                   unsigned long long* __restrict__ hist) {
  __shared__ unsigned long long s_t[256 / 64];
  const int b = blockIdx.x;
  if (b >= nbuck) return;
  unsigned long long t = 0;
  for (int k = threadIdx.x; k < nw; k += 256) t += bh[(uint64_t)b * nw + k];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
  if ((threadIdx.x & 63) == 0) s_t[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) hist[b] = s_t[0] + s_t[1] + s_t[2] + s_t[3];
}

// per wave quarter: entries of the buckets [lo, hi) (a row range); nw
// quarters
__global__ void k_range_counts(const uint32_t* __restrict__ bh, int lo, int hi, int nw,
                               uint32_t* __restrict__ wav_cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  uint32_t t = 0;
  for (int b = lo; b < hi; ++b) t += bh[(uint64_t)b * nw + k];
  wav_cnt[k] = t;
}

__global__ void k_degrees(const int64_t* __restrict__ off, int64_t n, uint32_t* __restrict__ deg) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    deg[i] = (uint32_t)(off[i + 1] - off[i]);
}

// pass 2: same partition; the entries of rows [r0, r1) (all rows: 0, ~0) at
// the wave quarter's offset (the scan of pass 1's per-quarter counts, or of
// a row range's), as sort keys ((row - r0) << colbits | col) and float
// values, in slot order (a key emits its row-u entry, then its row-v one).
// Every wave writes its own run: no block barrier, no contended atomics.
template <bool PK>
__global__ __launch_bounds__(TR_THREADS) void k_trunc_emit(TruncArgs a, uint64_t per_block, uint32_t r0, uint32_t r1,
                                                           const uint64_t* __restrict__ wav_off) {
  const WaveSpan ws = wave_span(a.t.cap, per_block);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t run = wav_off[(uint64_t)blockIdx.x * TR_WAVES + wid];
  uint64_t orun = a.lw_in ? a.occ_off[(uint64_t)blockIdx.x * TR_WAVES + wid] : 0;
  for (uint64_t t0 = 0; t0 < ws.q; t0 += 64 * TR_ITEMS) {
    const uint64_t ib = ws.w0 + t0 + (uint64_t)lane * TR_ITEMS;
    RawSlot rs[TR_ITEMS];
    load_group<PK>(a.t, ib, ws.w1, rs);
    // decode the group: occupied slots (occ), those with an endpoint in the
    // range (inr) and which rows they emit (eu: row u, ev: row v)
    uint32_t uu[TR_ITEMS], vv[TR_ITEMS];
    uint64_t cc[TR_ITEMS];
    bool eu[TR_ITEMS], ev[TR_ITEMS];
    uint32_t occ = 0, inr = 0;
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      eu[j] = ev[j] = false;
      if (!raw_get_t<PK>(a.t, rs[j], uu[j], vv[j], cc[j])) {
        uu[j] = vv[j] = 0;
        continue;
      }
      occ |= 1u << j;
      eu[j] = uu[j] >= r0 && uu[j] < r1;
      ev[j] = !a.upper && uu[j] != vv[j] && vv[j] >= r0 && vv[j] < r1;
      if (eu[j] || ev[j]) inr |= 1u << j;
    }
    // value cache index of the thread's first occupied slot
    uint64_t ci = 0;
    if (a.lw_in) {
      uint32_t wtot;
      ci = orun + wave_excl_scan((uint32_t)__popc(occ), &wtot);
      orun += wtot;
    }
    // every read of the group before any use: cached values of in-range
    // slots, degrees (and -sample 1 weights) of in-range slots past the cache
    float lw[TR_ITEMS];
    uint32_t du[TR_ITEMS], dv[TR_ITEMS];
    double wv[TR_ITEMS];
    uint32_t cached = 0;
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      const uint64_t cj = ci + (uint64_t)__popc(occ & ((1u << j) - 1u));
      lw[j] = 0.f;
      du[j] = dv[j] = 1;
      wv[j] = 0.0;
      if (!((inr >> j) & 1u)) continue;
      if (a.lw_in && cj < a.lw_n) {
        cached |= 1u << j;
        lw[j] = a.lw_in[cj];
      } else {
        du[j] = a.deg[uu[j]];
        dv[j] = a.deg[vv[j]];
        wv[j] = a.t.wtab ? a.t.wtab[ib + j] : (double)cc[j];
      }
    }
    // trunc-log keeps a key iff log(adj) > 0 (PathEmbed.h:203-217); a kept
    // key's cached value is > 0 (log(adj) >= log(1 + 2^-52) stays positive
    // as a float), a dropped one's is 0
    int c = 0;
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      if (!((inr >> j) & 1u)) continue;
      bool keep;
      if ((cached >> j) & 1u) {
        keep = lw[j] > 0.f;
      } else {
        const double adj = trunc_adj_w(a, uu[j], vv[j], wv[j], du[j], dv[j]);
        keep = adj > 1.0;  // log(adj) > 0 <=> adj > 1 for any faithful log
        if (keep) lw[j] = (float)log(adj);
      }
      if (!keep) {
        eu[j] = ev[j] = false;
        continue;
      }
      c += (eu[j] ? 1 : 0) + (ev[j] ? 1 : 0);
    }
    // the wave's exclusive scan of its lanes' counts
    uint32_t tot;
    uint64_t pos = run + wave_excl_scan((uint32_t)c, &tot);
#pragma unroll
    for (int j = 0; j < TR_ITEMS; ++j) {
      if (!eu[j] && !ev[j]) continue;
      if (eu[j]) {
        a.skey[pos] = ((uint64_t)(uu[j] - r0) << a.colbits) | vv[j];
        a.sval[pos] = lw[j];
        ++pos;
      }
      if (ev[j]) {
        a.skey[pos] = ((uint64_t)(vv[j] - r0) << a.colbits) | uu[j];
        a.sval[pos] = lw[j];
        ++pos;
      }
    }
    run += tot;
  }
}

// row_ptr by boundary detection over the sorted keys; col extraction
// vin / vout: the values when the sort left them elsewhere (else NULL)
// base: offset of these entries in the CSR (a row range's slice): row_ptr
// values are base + i
__global__ void k_csr_finish(const uint64_t* __restrict__ skey, int64_t nnz, int64_t n, int colbits,
                             int64_t* __restrict__ row_ptr, int32_t* __restrict__ col,
                             const float* __restrict__ vin = nullptr, float* __restrict__ vout = nullptr,
                             int64_t base = 0) {
  const uint64_t cmask = (colbits >= 64) ? ~0ull : ((1ull << colbits) - 1);
  const int lane = threadIdx.x & 63;
  // a wave covers 64 consecutive i (blockDim is a multiple of 64), so the
  // previous key comes from the lane below; lane 0 loads its own
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nnz; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = (i < nnz) ? skey[i] : 0ull;
    uint64_t kp = __shfl_up(k, 1, 64);
    if (lane == 0) kp = (i > 0) ? skey[i - 1] : 0ull;
    const int64_t r = (i < nnz) ? (int64_t)(k >> colbits) : n;
    const int64_t rp = (i == 0) ? -1 : (int64_t)(kp >> colbits);
    for (int64_t rr = rp + 1; rr <= r && rr <= n; ++rr) row_ptr[rr] = base + i;
    if (i < nnz) {
      col[i] = (int32_t)(k & cmask);
      if (vin) vout[i] = vin[i];
    }
  }
}

int bits_for(uint64_t n) {
  int b = 0;
  while (b < 63 && ((uint64_t)1 << b) < n) ++b;
  return b;
}

int check_graph(lightne_ctx* ctx, const lightne_graph* g) {
  if (!g || g->n <= 0 || g->m < 0 || !g->offsets || (g->m > 0 && !g->edges))
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "invalid graph descriptor");
  if (g->n >= (int64_t)1 << 32) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "graph has >= 2^32 vertices");
  if (g->ith_rule != LIGHTNE_ITH_PLAIN && g->ith_rule != LIGHTNE_ITH_BYTEPD)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "graph ith_rule must be LIGHTNE_ITH_PLAIN or LIGHTNE_ITH_BYTEPD");
  return LIGHTNE_OK;
}

// Sampler + table.  On success the table (slots, optional weights) occupies
// the start of the context workspace; stats are filled.
int run_sampler(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                lightne_sparsify_stats* st, TabView* tv) {
  LN_CHECK(check_graph(ctx, g));
  if (!p || p->walk_len == 0 || p->negative == 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "walk_len and negative must be >= 1");
  if (p->table_layout < 0 || p->table_layout > 2 || p->csr_parts < 0)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "table_layout must be 0, 1 or 2 and csr_parts >= 0");
  if (p->wide_offsets < 0 || p->wide_offsets > 2)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "wide_offsets must be 0, 1 or 2");
  const bool fits_packed = (uint64_t)g->n <= (1ull << PK_VB);
  if (p->table_layout == 2 && !fits_packed)
    return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "the packed table layout needs n <= 2^26");
  const bool packed = p->table_layout == 2 || (p->table_layout == 0 && fits_packed);
  LN_HIP(ctx, hipSetDevice(ctx->device));
  *st = lightne_sparsify_stats{};
  LN_CHECK(lightne_sample_scalars(g->n, g->m, p, st));
  // Samples: exp_rv per directed edge plus one more with probability frac
  // (PathEmbed.h:142-148), i.e. m*exp_rv + Binomial(m, frac).  Bound the
  // binomial 8 standard deviations above its mean.
  const double m = (double)(uint64_t)g->m;
  const double fr = st->fractional_bits > 0.0000000001 ? st->fractional_bits : 0.0;
  const double bound = m * (double)st->exp_rv_rounded + m * fr + 8.0 * std::sqrt(m * fr * (1.0 - fr)) + 64.0;
  // capacity: load factor <= 0.7 (wide) / 0.8 (packed) even if every sample
  // is a distinct key (the reference sizes by -tablesz / -mem_ratio,
  // 2^ceil(log2(2 table_size)), and exits when that is too small,
  // par_table.h:104-107)
  const double want = bound / (packed ? 0.8 : 0.7) + 1.0;
  uint64_t cap = want > 1.0e12 ? (1ull << 41) : std::max<uint64_t>(((uint64_t)want + 255) & ~255ull, 1024);
  const uint64_t slot_bytes = (packed ? 8 : sizeof(Slot)) + (p->sample ? sizeof(double) : 0);
  if (p->max_table_bytes && cap > p->max_table_bytes / slot_bytes)
    cap = std::max<uint64_t>((p->max_table_bytes / slot_bytes) & ~255ull, 1024);
  if (cap > (1ull << 40))
    return ctx_fail(ctx, LIGHTNE_ERR_OOM, "sampling budget of %.3g samples needs a table of more than 2^40 slots",
                    bound);
  st->table_capacity = cap;
  st->table_layout = packed ? 2 : 1;

  Slot* tab = nullptr;
  unsigned long long* pk = nullptr;
  double* wtab = nullptr;
  // workspace front: the table, then each edge's source vertex (dead after
  // sampling)
  uint32_t* src = nullptr;
  WsCarve wc;
  if (packed)
    wc.add((void**)&pk, cap * 8);
  else
    wc.add((void**)&tab, cap * sizeof(Slot));
  if (p->sample) wc.add((void**)&wtab, cap * sizeof(double));
  wc.add((void**)&src, std::max<int64_t>(g->m, 1) * sizeof(uint32_t));
  LN_CHECK(wc.commit(ctx));
  auto* stats = (unsigned long long*)ctx_buf(ctx, "sample_stats", 8 * sizeof(unsigned long long));
  auto* ovf = packed ? (unsigned long long*)ctx_buf(ctx, "sample_ovf", OVF_CAP * 16) : nullptr;
  // 32-bit row pointers for the walks while every offset fits (2m < 2^32)
  const bool off64 = p->wide_offsets == 1 || (uint64_t)g->m >= ((uint64_t)1 << 32);
  auto* off32 = off64 ? nullptr : (uint32_t*)ctx_buf(ctx, "sample_off32", (size_t)(g->n + 1) * 4);
  if (!stats || (packed && !ovf) || (!off64 && !off32)) return LIGHTNE_ERR_OOM;
  // neighbour rows for the walks (every degree fits 32 bits when 2m < 2^32):
  // behind the table and the sources where the workspace already has the
  // room (it does not grow for them), else in memory of their own for the
  // sampler's run, else the walks read row pointers
  NbrRow* nr = nullptr;
  DevMem nr_mem;
  if (p->wide_offsets == 0 && !off64 && g->m > 0) {
    const auto& wsb = ctx->bufs["ws"];
    const size_t nr_at = wc.end(), nr_bytes = (size_t)g->m * sizeof(NbrRow);
    if (wsb.p && wsb.bytes >= nr_at + nr_bytes)
      nr = (NbrRow*)((char*)wsb.p + nr_at);
    else if (ctx_malloc(ctx, &nr_mem.p, nr_bytes) == hipSuccess)
      nr = (NbrRow*)nr_mem.p;
  }
  st->walk_rows = nr ? 1u : 0u;

  hipStream_t s = ctx->stream;
  LN_HIP(ctx, hipEventRecord(ctx->ev[0], s));
  if (packed) {
    hipLaunchKernelGGL(k_fill_u64, dim3(4096), dim3(256), 0, s, pk, cap);
    LN_LAUNCH(ctx);
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, s, ovf, 2 * OVF_CAP);
    LN_LAUNCH(ctx);
  } else {
    hipLaunchKernelGGL(k_fill_slots, dim3(4096), dim3(256), 0, s, tab, cap);
    LN_LAUNCH(ctx);
  }
  if (wtab) LN_HIP(ctx, hipMemsetAsync(wtab, 0, cap * sizeof(double), s));
  LN_HIP(ctx, hipMemsetAsync(stats, 0, 8 * sizeof(unsigned long long), s));
  if (g->m > 0) {
    hipLaunchKernelGGL(k_edge_src, dim3(grid_blocks(g->m, ES_TILE, (int64_t)1 << 16)), dim3(ES_THREADS), 0, s,
                       g->offsets, g->n, g->m, src);
    LN_LAUNCH(ctx);
    if (!off64) {
      hipLaunchKernelGGL(k_off32, dim3(grid_blocks(g->n + 1, 256)), dim3(256), 0, s, g->offsets, g->n + 1, off32);
      LN_LAUNCH(ctx);
    }
    SampleArgs a;
    a.off32 = off32;
    a.off64 = g->offsets;
    a.adj = g->edges;
    a.nr = nr;
    a.ith = g->ith_rule == LIGHTNE_ITH_BYTEPD ? 1u : 0u;
    a.src = src;
    a.m = g->m;
    a.seed = p->seed;
    a.walk_len = p->walk_len;
    a.exp_rv = st->exp_rv_rounded;
    a.frac = st->fractional_bits;
    a.logn = st->logn;
    a.tab = tab;
    a.pk = pk;
    a.ovf = ovf;
    a.wtab = wtab;
    a.cap = cap;
    a.stats = stats;
    LN_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    if (nr) {
      hipLaunchKernelGGL(k_nbr_rows, dim3(grid_blocks(g->m, 256, (int64_t)1 << 20)), dim3(256), 0, s,
                         (const uint32_t*)off32, g->edges, g->m, nr);
      LN_LAUNCH(ctx);
    }
    // one lane per directed edge; a launch holds at most 2^31 lanes (one grid
    // dimension holds fewer than 2^32 work-items)
    constexpr int64_t LANES = (int64_t)1 << 31;
    for (int64_t j0 = 0; j0 < g->m; j0 += LANES) {
      a.j0 = j0;
      const unsigned nb = (unsigned)((std::min(LANES, g->m - j0) + 255) / 256);
#define LN_SAMPLE(M, P, O, R) hipLaunchKernelGGL((k_sample<M, P, O, R>), dim3(nb), dim3(256), 0, s, a)
      // neighbour rows only with 32-bit offsets (off64 = 0)
      switch ((nr ? 8 : 0) + (p->sample ? 4 : 0) + (packed ? 2 : 0) + (off64 ? 1 : 0)) {
        case 0: LN_SAMPLE(0, false, false, false); break;
        case 1: LN_SAMPLE(0, false, true, false); break;
        case 2: LN_SAMPLE(0, true, false, false); break;
        case 3: LN_SAMPLE(0, true, true, false); break;
        case 4: LN_SAMPLE(1, false, false, false); break;
        case 5: LN_SAMPLE(1, false, true, false); break;
        case 6: LN_SAMPLE(1, true, false, false); break;
        case 7: LN_SAMPLE(1, true, true, false); break;
        case 8: LN_SAMPLE(0, false, false, true); break;
        case 10: LN_SAMPLE(0, true, false, true); break;
        case 12: LN_SAMPLE(1, false, false, true); break;
        case 14: LN_SAMPLE(1, true, false, true); break;
        default: return ctx_fail(ctx, LIGHTNE_ERR_HIP, "sampler: neighbour rows need 32-bit offsets");
      }
#undef LN_SAMPLE
      LN_LAUNCH(ctx);
    }
  } else {
    LN_HIP(ctx, hipEventRecord(ctx->ev[1], s));
  }
  LN_HIP(ctx, hipEventRecord(ctx->ev[2], s));
  unsigned long long hs[8];
  LN_HIP(ctx, hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  st->samples = hs[0];
  st->walk_steps = hs[1];
  st->attempts = hs[2];
  st->t_sample = ev_seconds(ctx->ev[1], ctx->ev[2]);
  if (hs[4]) return ctx_fail(ctx, LIGHTNE_ERR_BAD_GRAPH, "random walk reached a vertex of degree 0 (graph not symmetric?)");
  if (hs[3])
    return ctx_fail(ctx, LIGHTNE_ERR_TABLE_FULL,
                    "hash table full (%llu slots); raise max_table_bytes (the reference exits here, par_table.h:104-107)",
                    (unsigned long long)cap);
  if (hs[5])
    return ctx_fail(ctx, LIGHTNE_ERR_TABLE_FULL,
                    "more than %llu keys were sampled over 1023 times; use the wide table layout",
                    (unsigned long long)OVF_CAP);
  *tv = TabView{tab, pk, ovf, wtab, cap};
  return LIGHTNE_OK;
}

// puts every cluster of the table in canonical order (k_table_canon), over
// the trunc-log partition; *canonical = false when a cluster was too long to
// reorder (the table is a valid one either way).  Also leaves the wave
// quarters' occupied slots in the "tr_wocc" buffer and the blocks' in
// "canon_bocc" (canon_distinct).
int table_canon(lightne_ctx* ctx, const TabView& tv, bool* canonical, double* seconds) {
  *canonical = false;
  const int nblk = trunc_blocks(tv.cap);
  auto* err = (uint32_t*)ctx_buf(ctx, "canon_err", sizeof(uint32_t));
  auto* wocc = (uint32_t*)ctx_buf(ctx, "tr_wocc", (size_t)nblk * TR_WAVES * sizeof(uint32_t));
  auto* bocc = (unsigned long long*)ctx_buf(ctx, "canon_bocc", (size_t)nblk * sizeof(unsigned long long));
  const uint64_t per_block = trunc_per_block(tv.cap, nblk);
  const uint64_t wpb = (per_block + CN_WIN - 1) / CN_WIN;
  const uint64_t nwin = (uint64_t)nblk * wpb;
  auto* win = (CanonWin*)ctx_buf(ctx, "canon_win", nwin * sizeof(CanonWin));
  if (!err || !wocc || !bocc || !win) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemsetAsync(err, 0, sizeof(uint32_t), ctx->stream));
  LN_HIP(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
  CanonArgs a;
  a.wide = const_cast<Slot*>(tv.wide);
  a.pk = const_cast<unsigned long long*>(tv.pk);
  a.ovf = tv.ovf;
  a.wtab = const_cast<double*>(tv.wtab);
  a.cap = tv.cap;
  a.per_block = per_block;
  a.wpb = wpb;
  a.nwin = nwin;
  a.err = err;
  a.win = win;
  const unsigned nb = (unsigned)std::min<uint64_t>(nwin, 16384);
  const hipStream_t s = ctx->stream;
  if (tv.pk) {
    if (tv.wtab)
      hipLaunchKernelGGL((k_table_canon<true, true>), dim3(nb), dim3(CN_THREADS), 0, s, a);
    else
      hipLaunchKernelGGL((k_table_canon<true, false>), dim3(nb), dim3(CN_THREADS), 0, s, a);
  } else {
    if (tv.wtab)
      hipLaunchKernelGGL((k_table_canon<false, true>), dim3(nb), dim3(CN_THREADS), 0, s, a);
    else
      hipLaunchKernelGGL((k_table_canon<false, false>), dim3(nb), dim3(CN_THREADS), 0, s, a);
  }
  LN_LAUNCH(ctx);
  hipLaunchKernelGGL(k_canon_fold, dim3(nblk), dim3(256), 0, s, win, wpb, wocc, bocc);
  LN_LAUNCH(ctx);
  LN_HIP(ctx, hipEventRecord(ctx->ev[7], s));
  uint32_t h = 0;
  LN_HIP(ctx, hipMemcpyAsync(&h, err, sizeof(h), hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  *canonical = h == 0;
  *seconds = ev_seconds(ctx->ev[6], ctx->ev[7]);
  return LIGHTNE_OK;
}

// distinct keys (occupied slots) from the canonical-order pass just run
int canon_distinct(lightne_ctx* ctx, const TabView& tv, uint64_t* distinct) {
  const int nblk = trunc_blocks(tv.cap);
  auto* bocc = (unsigned long long*)ctx_buf(ctx, "canon_bocc", (size_t)nblk * sizeof(unsigned long long));
  if (!bocc) return LIGHTNE_ERR_OOM;
  std::vector<unsigned long long> hb((size_t)nblk);
  LN_HIP(ctx, hipMemcpyAsync(hb.data(), bocc, hb.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t d = 0;
  for (int b = 0; b < nblk; ++b) d += hb[(size_t)b];
  *distinct = d;
  return LIGHTNE_OK;
}

int table_totals(lightne_ctx* ctx, const TabView& tv, uint64_t* distinct, uint64_t* cnt_sum, double* w_sum) {
  const int nb = 2048;
  auto* pu = (unsigned long long*)ctx_buf(ctx, "tred_u", 2 * nb * sizeof(unsigned long long));
  auto* pw = (double*)ctx_buf(ctx, "tred_w", nb * sizeof(double));
  if (!pu || !pw) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_table_reduce, dim3(nb), dim3(256), 0, ctx->stream, tv, pu, pw);
  LN_LAUNCH(ctx);
  std::vector<unsigned long long> hu(2 * nb);
  std::vector<double> hw(nb);
  LN_HIP(ctx, hipMemcpyAsync(hu.data(), pu, hu.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
  LN_HIP(ctx, hipMemcpyAsync(hw.data(), pw, hw.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint64_t d = 0, c = 0;
  double w = 0.0;
  for (int b = 0; b < nb; ++b) {
    d += hu[2 * b];
    c += hu[2 * b + 1];
    w += hw[b];
  }
  *distinct = d;
  *cnt_sum = c;
  *w_sum = w;
  return LIGHTNE_OK;
}

// symmetric expansion of upper-triangle storage: entry e = (i, j, v) emits
// (i, j) and, off the diagonal, (j, i); a diagonal entry's second key is the
// sentinel ~0, which sorts last
__global__ void k_up_emit(const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const float* __restrict__ val, int64_t n, int64_t nnz, int colbits,
                          uint64_t* __restrict__ key, float* __restrict__ kv, unsigned long long* __restrict__ ndiag) {
  __shared__ unsigned long long s_d;
  if (threadIdx.x == 0) s_d = 0;
  __syncthreads();
  unsigned long long d = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = n - 1;  // row of e: largest i with rp[i] <= e
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    const uint64_t i = (uint64_t)lo, j = (uint64_t)(uint32_t)col[e];
    key[2 * e] = (i << colbits) | j;
    kv[2 * e] = val[e];
    key[2 * e + 1] = (i == j) ? ~0ull : ((j << colbits) | i);
    kv[2 * e + 1] = val[e];
    d += (i == j);
  }
  atomicAdd(&s_d, d);
  __syncthreads();
  if (threadIdx.x == 0 && s_d) atomicAdd(ndiag, s_d);
}

}  // namespace

// Bytes this context holds in its cached buffers (reused or reclaimable by
// the call in progress).
static size_t ctx_held_bytes(lightne_ctx* ctx) {
  size_t b = 0;
  for (auto& kv : ctx->bufs) b += kv.second.bytes;
  return b;
}

// Sampler, table, trunc-log filter and CSR.  ctx_owned: the CSR arrays are
// context buffers reused by the next call (the pipeline's own use); else the
// handle owns them.
static int sparsify_sized(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                          lightne_csr** out, lightne_sparsify_stats* stats, bool ctx_owned, size_t front_reserve,
                          float* xs, size_t xs_bytes, int restarts);

int sparsify_impl(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p, lightne_csr** out,
                  lightne_sparsify_stats* stats, bool ctx_owned, size_t front_reserve, float* xs, size_t xs_bytes) {
  return sparsify_sized(ctx, g, p, out, stats, ctx_owned, front_reserve, xs, xs_bytes, 0);
}

// `restarts`: how often the layout was planned again after a workspace growth
// (with -sample 1 the second sampler run may keep a few other keys, below).
// xs: device memory of xs_bytes the caller does not need during this call
// (the pipeline's output buffer), or NULL.
static int sparsify_sized(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                          lightne_csr** out, lightne_sparsify_stats* stats, bool ctx_owned, size_t front_reserve,
                          float* xs, size_t xs_bytes, int restarts) {
  *out = nullptr;
  lightne_sparsify_stats st{};
  TabView tv;
  LN_CHECK(run_sampler(ctx, g, p, &st, &tv));
  st.sampler_runs = 1;
  // canonical slot order: the emission order is then the same on every run,
  // so the CSR rows need sorting on the row bits only (their entries keep the
  // emission order); otherwise rows are sorted by column too
  bool canon = false;
  LN_CHECK(table_canon(ctx, tv, &canon, &st.t_canon));
  hipStream_t s = ctx->stream;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int colbits = std::max(1, bits_for((uint64_t)g->n));
  // blocks of the trunc-log passes (contiguous slot ranges, the partition
  // the canonical-order pass took)
  const int nblk = trunc_blocks(tv.cap);
  // the emission's units are the blocks' wave quarters: entries per quarter
  // and their scan (each quarter's output offset); kept keys per block
  const int nw = nblk * TR_WAVES;
  auto* wav_emit = (uint32_t*)ctx_buf(ctx, "tr_emit", (size_t)nw * sizeof(uint32_t));
  auto* blk_kept = (uint32_t*)ctx_buf(ctx, "tr_kept", nblk * sizeof(uint32_t));
  auto* wav_off = (uint64_t*)ctx_buf(ctx, "tr_off", ((size_t)nw + 1) * sizeof(uint64_t));
  // occupied slots per wave quarter (left by the canonical-order pass) and
  // their scan (the value cache's index)
  auto* wocc = (uint32_t*)ctx_buf(ctx, "tr_wocc", (size_t)nblk * TR_WAVES * sizeof(uint32_t));
  auto* occ_off = (uint64_t*)ctx_buf(ctx, "tr_occ", ((size_t)nblk * TR_WAVES + 1) * sizeof(uint64_t));
  if (!wav_emit || !blk_kept || !wav_off || !wocc || !occ_off) return LIGHTNE_ERR_OOM;
  TruncArgs ta;
  int64_t nnz = 0;
  auto* deg = (uint32_t*)ctx_buf(ctx, "s1_deg", (size_t)g->n * 4);
  if (!deg) return LIGHTNE_ERR_OOM;
  hipLaunchKernelGGL(k_degrees, dim3(grid_blocks(g->n, 256)), dim3(256), 0, s, g->offsets, g->n, deg);
  LN_LAUNCH(ctx);
  // row buckets for a build by row ranges: at most HIST_BUCKETS ranges of
  // 2^shift rows
  const int shift = std::max(0, bits_for((uint64_t)g->n) - 10);
  const int nbuck = (int)(((uint64_t)g->n + ((1ull << shift) - 1)) >> shift);
  static_assert(HIST_BUCKETS == 1 << 10, "n < 2^bits_for(n) rows in 2^10 buckets");
  uint32_t* bh = nullptr;  // per-quarter row histograms, when taken
  // Value cache: the count pass leaves the CSR value of every occupied slot
  // (one float per distinct key, in slot order) in an array right behind the
  // table in the workspace (context-owned builds only), so the emission
  // passes read it instead of two degrees at random places.  Written
  // whenever the workspace holds it beside the table, or as a buffer of its
  // own for a single-sort build that surely fits beside it; the layout chosen
  // after the count keeps it only where it needs no extra memory
  // (value_cache = 1: always, growing the workspace; -1: never).  Where the
  // workspace cannot hold it, the caller's scratch xs holds the values of
  // the first occupied slots it has room for (the rest read the degrees).
  const size_t tab_bytes = al(tv.cap * (tv.pk ? 8 : sizeof(Slot))) + (tv.wtab ? al(tv.cap * sizeof(double)) : 0);
  size_t lw_bytes = 0;  // set by the count pass: one float per distinct key
  const bool lw_force = p->value_cache > 0;
  bool lw_allowed = ctx_owned && p->value_cache >= 0;
  float* lw = nullptr;
  bool lw_in_ws = false;  // else a buffer of its own (single-sort builds only)
  bool lw_ext = false;    // in the caller's scratch, occupied ranks [0, lw_n)
  uint64_t lw_n = 0;
  const bool ext_allowed = xs && p->value_cache >= 0;
  uint64_t ext_n = 0;  // set by the count pass
  // an in-workspace cache the layout cannot keep: the values of its first
  // ext_n occupied slots move to the caller's scratch before anything
  // overwrites them
  auto demote = [&]() -> int {
    if (lw_ext) return LIGHTNE_OK;
    if (lw && lw_in_ws && ext_allowed && ext_n > 0) {
      LN_HIP(ctx, hipMemcpyAsync(xs, lw, ext_n * sizeof(float), hipMemcpyDeviceToDevice, s));
      lw = xs;
      lw_ext = true;
      lw_n = ext_n;
    } else {
      lw = nullptr;
    }
    lw_in_ws = false;
    return LIGHTNE_OK;
  };
  // totals (PathEmbed.h:175-200; the distinct keys from the canonical-order
  // pass, -sample 1 weights by their own reduction), then per-block counts of
  // the kept entries over contiguous slot ranges and their scan; repeated
  // after a workspace regrowth re-ran the (deterministic) sampler
  auto count_pass = [&](bool hist) -> int {
    LN_HIP(ctx, hipEventRecord(ctx->ev[3], s));
    uint64_t distinct, csum;
    double wsum = 0.0;
    if (p->sample) {
      LN_CHECK(table_totals(ctx, tv, &distinct, &csum, &wsum));
    } else {
      // every sample added one to its key's count (the sampler fails
      // otherwise), so the counts sum to the samples drawn
      LN_CHECK(canon_distinct(ctx, tv, &distinct));
      csum = st.samples;
    }
    st.distinct_keys = distinct;
    // PathEmbed.h:187 (sum of slot values) and :199-200 (factor)
    st.M_exact = p->sample ? wsum : (double)csum;
    const double vol = (double)(uint64_t)g->m;
    st.factor = vol * (double)(uint64_t)g->m / st.M_exact / 2 / (double)p->negative;
    LN_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    lw_bytes = al(std::max<uint64_t>(distinct, 1) * sizeof(float));
    ext_n = std::min<uint64_t>(distinct, xs_bytes / sizeof(float));
    ta = TruncArgs{};
    ta.t = tv;
    ta.deg = deg;
    ta.factor = st.factor;
    ta.upper = p->upper;
    ta.colbits = colbits;
    lw = nullptr;
    lw_in_ws = false;
    // the caller's scratch, where it holds the whole cache, costs the
    // workspace nothing (an in-workspace cache shrinks the row ranges)
    const bool ext_whole = ext_allowed && distinct > 0 && ext_n >= distinct;
    if (lw_allowed && (lw_force || !ext_whole)) {
      const size_t nb = 2 * (size_t)distinct + 1;  // nnz <= 2 x kept <= 2 x distinct
      const size_t csr_b = al((g->n + 1) * 8) + 2 * al(nb * 4);
      const size_t ws_b = ctx->bufs["ws"].bytes;
      size_t fb = 0, tb = 0;
      LN_HIP(ctx, hipMemGetInfo(&fb, &tb));
      const size_t mg = ((size_t)3 << 30) + tb / 100;
      const size_t av = fb + ctx_held_bytes(ctx) > mg ? fb + ctx_held_bytes(ctx) - mg : 0;
      const size_t tmp_b = (al(nb * 8) + nb * 4 <= tab_bytes) ? 0 : al(nb * 8) + al(nb * 4);
      const size_t single_b = std::max(tab_bytes + al(nb * 8), front_reserve) + csr_b + tmp_b;
      const bool single = p->csr_parts < 2 && single_b <= av;  // one sort will fit
      if (single) {
        // behind the table when the keys and the CSR still fit behind it,
        // else a buffer of its own when memory holds both
        if (std::max(tab_bytes + lw_bytes + al(nb * 8), front_reserve) + csr_b <= ws_b) {
          lw = (float*)((char*)ctx->bufs["ws"].p + tab_bytes);
          lw_in_ws = true;
        } else if (lw_force || single_b + lw_bytes + tb / 8 <= av) {  // an eighth of the device kept for stages 2-3
          lw = (float*)ctx_buf(ctx, "s1_lw", lw_bytes);
          if (!lw) ctx->err.clear();  // no room after all: the emission reads the degrees
        }
      } else if (tab_bytes + lw_bytes + (ext_allowed && ext_n > 0 ? csr_b : 0) <= ws_b) {
        // row ranges: behind the table (with the caller's scratch at hand,
        // only where the CSR still fits behind it: a cache the layout drops
        // would be copied there after the count)
        lw = (float*)((char*)ctx->bufs["ws"].p + tab_bytes);
        lw_in_ws = true;
      }
    }
    lw_ext = false;
    lw_n = distinct;
    if (!lw && ext_allowed && ext_n > 0) {
      lw = xs;
      lw_ext = true;
      lw_n = ext_n;
    }
    ta.lw_out = lw;
    ta.lw_n = lw_n;
    const uint64_t per_block = trunc_per_block(tv.cap, nblk);
    if (lw) {
      // rank of each wave quarter's first occupied slot
      LN_CHECK(scan_u32_to_u64(ctx, wocc, occ_off, (int64_t)nblk * TR_WAVES));
      ta.occ_off = occ_off;
    }
    if (hist) {
      bh = (uint32_t*)ctx_buf(ctx, "tr_bh", (size_t)nbuck * nw * 4);
      if (!bh) return LIGHTNE_ERR_OOM;
      if (tv.pk)
        hipLaunchKernelGGL((k_trunc_count<true, true>), dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, wav_emit,
                           blk_kept, shift, nbuck, bh);
      else
        hipLaunchKernelGGL((k_trunc_count<true, false>), dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, wav_emit,
                           blk_kept, shift, nbuck, bh);
    } else {
      if (tv.pk)
        hipLaunchKernelGGL((k_trunc_count<false, true>), dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, wav_emit,
                           blk_kept, 0, 0, (uint32_t*)nullptr);
      else
        hipLaunchKernelGGL((k_trunc_count<false, false>), dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, wav_emit,
                           blk_kept, 0, 0, (uint32_t*)nullptr);
    }
    LN_LAUNCH(ctx);
    LN_CHECK(scan_u32_to_u64(ctx, wav_emit, wav_off, nw));
    std::vector<uint32_t> hk(nblk);
    uint64_t htot = 0, otot = distinct;
    LN_HIP(ctx, hipMemcpyAsync(hk.data(), blk_kept, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LN_HIP(ctx, hipMemcpyAsync(&htot, wav_off + nw, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    if (lw)
      LN_HIP(ctx, hipMemcpyAsync(&otot, occ_off + (size_t)nblk * TR_WAVES, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                 s));
    LN_HIP(ctx, hipStreamSynchronize(s));
    if (otot != distinct)
      return ctx_fail(ctx, LIGHTNE_ERR_HIP, "value cache numbered %llu occupied slots, the table holds %llu keys",
                      (unsigned long long)otot, (unsigned long long)distinct);
    nnz = (int64_t)htot;
    st.kept_upper = 0;
    for (uint32_t v : hk) st.kept_upper += v;
    st.nnz = (uint64_t)nnz;
    return LIGHTNE_OK;
  };
  // tables that may need a build by row ranges take the row histograms now
  LN_CHECK(count_pass(p->csr_parts >= 2 || tv.cap >= (1ull << 30)));
  const size_t nz = (size_t)std::max<int64_t>(nnz, 1);
  const size_t csr_bytes = al((g->n + 1) * 8) + 2 * al(nz * 4);
  // Memory: one global sort needs the table, 8-byte keys, the CSR and (unless
  // they fit in the dead table) 12 bytes of ping-pong buffers per entry.  When
  // that exceeds what is free (plus what this context already holds), the CSR
  // is built by row ranges with sort buffers for one range at a time.
  size_t free_b = 0, total_b = 0;
  LN_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
  const size_t margin = ((size_t)3 << 30) + total_b / 100;
  const size_t avail = free_b + ctx_held_bytes(ctx) > margin ? free_b + ctx_held_bytes(ctx) - margin : 0;
  const size_t sort_tmp = (al(nz * 8) + nz * 4 <= tab_bytes) ? 0 : al(nz * 8) + al(nz * 4);
  const size_t single_need = std::max(tab_bytes + al(nz * 8), front_reserve) + csr_bytes + sort_tmp;
  const bool parts = p->csr_parts >= 2 || (nnz > 0 && single_need > avail);
  st.csr_parts = 1;

  DevMem mrp, mcol, mval;
  int64_t* rp = nullptr;
  int32_t* col = nullptr;
  float* val = nullptr;
  if (!parts) {
    // Workspace layout when it is large enough: the table and then the sort
    // keys at the front, the CSR at the back, clear of the front region the
    // next stage carves (front_reserve), so that a repeated pipeline holds one
    // allocation.  Otherwise the keys (and a context-owned CSR) are buffers of
    // their own.
    const size_t ws_bytes = ctx->bufs["ws"].bytes;
    // auto: keep an in-workspace cache only where the keys and the CSR still
    // fit the workspace behind it
    if (lw && lw_in_ws && !lw_force &&
        std::max(tab_bytes + lw_bytes + al(nz * 8), front_reserve) + csr_bytes > ws_bytes)
      LN_CHECK(demote());
    const size_t lwb = lw && lw_in_ws ? lw_bytes : 0;
    const size_t front = std::max(tab_bytes + lwb + al(nz * 8), front_reserve);
    const bool in_ws = ctx_owned && front + csr_bytes <= ws_bytes;
    char* ws = (char*)ctx->bufs["ws"].p;  // the table starts the workspace
    uint64_t* skey = in_ws ? (uint64_t*)(ws + tab_bytes + lwb) : (uint64_t*)ctx_buf(ctx, "s1_skey", nz * 8);
    if (!skey) return LIGHTNE_ERR_OOM;
    if (in_ws) {
      char* back = ws + (ws_bytes - csr_bytes);
      rp = (int64_t*)back;
      col = (int32_t*)(back + al((g->n + 1) * 8));
      val = (float*)(back + al((g->n + 1) * 8) + al(nz * 4));
    } else if (ctx_owned) {
      rp = (int64_t*)ctx_buf(ctx, "s1_rp", (g->n + 1) * 8);
      col = (int32_t*)ctx_buf(ctx, "s1_col", nz * 4);
      val = (float*)ctx_buf(ctx, "s1_val", nz * 4);
      if (!rp || !col || !val) return LIGHTNE_ERR_OOM;
    } else {
      LN_CHECK(dev_alloc(ctx, mval, nz * 4, "sparsifier values"));
      LN_CHECK(dev_alloc(ctx, mrp, (g->n + 1) * 8, "csr row_ptr"));
      LN_CHECK(dev_alloc(ctx, mcol, nz * 4, "csr col"));
      rp = (int64_t*)mrp.p;
      col = (int32_t*)mcol.p;
      val = (float*)mval.p;
    }
    // The sort ping-pongs between (keys, values) and two temporaries, which
    // reuse the dead table when they fit there.  Otherwise they are buffers
    // of their own, alive during the emission too: with an odd number of
    // passes the emission then writes the values into the temporary, so the
    // last pass leaves them in the CSR's value array (no copy).
    const size_t kbytes = al(nz * 8);
    uint64_t* k_tmp = nullptr;
    float* v_tmp = nullptr;
    float* v_emit = val;
    if (kbytes + nz * 4 <= tab_bytes) {
      k_tmp = (uint64_t*)ws;
      v_tmp = (float*)(ws + kbytes);
    } else if (radix_sort_passes(canon ? colbits : 2 * colbits) % 2 == 1) {
      k_tmp = (uint64_t*)ctx_buf(ctx, "rs_keys", nz * 8);
      v_emit = (float*)ctx_buf(ctx, "rs_vals", nz * 4);
      if (!k_tmp || !v_emit) return LIGHTNE_ERR_OOM;
      v_tmp = val;
    }
    ta.skey = skey;
    ta.sval = v_emit;
    ta.lw_out = nullptr;
    ta.lw_in = ta.occ_off ? lw : nullptr;  // numbered by the count pass
    ta.lw_n = lw_n;
    st.value_cache = lw ? (lw_ext ? 2u : 1u) : 0u;
    const uint64_t per_block = trunc_per_block(tv.cap, nblk);
    if (tv.pk)
      hipLaunchKernelGGL(k_trunc_emit<true>, dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, 0u, 0xffffffffu,
                         (const uint64_t*)wav_off);
    else
      hipLaunchKernelGGL(k_trunc_emit<false>, dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, 0u, 0xffffffffu,
                         (const uint64_t*)wav_off);
    LN_LAUNCH(ctx);
    // rows sorted by column: a key's table slot depends on the order concurrent
    // inserts claimed slots, so the column order is what makes the matrix (and
    // every summation over a row) deterministic
    uint64_t* kres = nullptr;
    float* vres = nullptr;
    LN_CHECK(radix_sort_pairs(ctx, skey, v_emit, nnz, canon ? colbits : 2 * colbits, k_tmp, v_tmp,
                              canon ? colbits : 0, &kres, &vres));
    hipLaunchKernelGGL(k_csr_finish, dim3(grid_blocks(nnz + 1, 256)), dim3(256), 0, s, (const uint64_t*)kres, nnz,
                       g->n, colbits, rp, col, vres == val ? nullptr : (const float*)vres, val, (int64_t)0);
    LN_LAUNCH(ctx);
  } else {
    // row histogram of the emitted entries: per wave quarter (taken by the
    // count pass), summed per bucket
    if (!bh) LN_CHECK(count_pass(true));
    auto* hist = (unsigned long long*)ctx_buf(ctx, "tr_hist", HIST_BUCKETS * 8);
    if (!hist) return LIGHTNE_ERR_OOM;
    hipLaunchKernelGGL(k_bh_totals, dim3(nbuck), dim3(256), 0, s, (const uint32_t*)bh, nbuck, nw,
                       hist);
    LN_LAUNCH(ctx);
    std::vector<unsigned long long> hh(HIST_BUCKETS);
    LN_HIP(ctx, hipMemcpyAsync(hh.data(), hist, nbuck * 8, hipMemcpyDeviceToHost, s));
    LN_HIP(ctx, hipStreamSynchronize(s));
    uint64_t hsum = 0;
    for (int b = 0; b < nbuck; ++b) hsum += hh[b];
    if (hsum != (uint64_t)nnz)
      return ctx_fail(ctx, LIGHTNE_ERR_HIP, "row histogram holds %llu entries, the count pass %lld",
                      (unsigned long long)hsum, (long long)nnz);
    // entries per range: forced count, or what the memory left beside the
    // table (and value cache) and the CSR allows (24 bytes per entry: keys,
    // values, ping-pong).  Context-owned, the workspace is rebuilt as
    // [table | cache | ranges | CSR]; otherwise the range buffers use the
    // workspace's room behind the table or memory of their own.
    const size_t ws_now = ctx->bufs["ws"].bytes;
    // ranges of whole buckets, at most 2^24 rows each, sized for the room
    // left beside the table, the value cache (lwb bytes) and the CSR
    std::vector<std::pair<int, int>> rng;  // bucket ranges
    std::vector<uint64_t> rcnt;
    // Where a range's sort buffers (keys, key temporary: 8 bytes per entry;
    // values, value temporary: 4) live.  The CSR slices of the ranges after
    // it are still unwritten, so its col and val arrays have a free tail of
    // 4 bytes per later entry each:
    //   RP_OWN   all four in the range buffer region (24 B per entry there);
    //   RP_KEYS  keys in the col tail, key temporary in the val tail, values
    //            in the range buffer region (8 B per entry there);
    //   RP_TAILS keys and values in the col tail, temporaries in the val tail.
    // Early ranges, with many entries after them, can thus be larger than
    // the range buffer region alone allows: fewer ranges, fewer table scans.
    enum { RP_OWN = 0, RP_KEYS = 1, RP_TAILS = 2 };
    std::vector<int> rplace;
    size_t pe = 1, part_bytes = 0;
    // in_ws: the ranges must fit the workspace as it is
    auto plan = [&](size_t lwb, bool in_ws) -> int {
      size_t part_room;
      // the pipeline's later stages allocate outside the workspace: per-row
      // SpMM plan scratch (about 40 bytes per row), chunk lists, the graph
      // plan's partial sums; a repeated call must find them free again, or
      // it evicts cached buffers and waits for freed memory to be wiped
      const size_t later = ctx_owned ? 80 * ((size_t)g->n + 1) + ((size_t)4 << 30) : 0;
      if (ctx_owned)
        part_room = avail > tab_bytes + lwb + csr_bytes + later ? avail - tab_bytes - lwb - csr_bytes - later : 0;
      else
        part_room = std::max(ws_now - tab_bytes, avail > ws_now + csr_bytes ? avail - ws_now - csr_bytes : 0);
      if (in_ws) {
        const size_t fixed = std::max(tab_bytes + lwb, front_reserve) + csr_bytes;
        part_room = std::min(part_room, ws_now > fixed ? ws_now - fixed : 0);
      }
      const bool forced = p->csr_parts >= 2;
      // at least 4 ranges: memory for later stages
      const uint64_t quarter = ((uint64_t)nnz + 3) / 4;
      // entries a range may hold in the range buffer region alone (RP_OWN),
      // with its keys in the CSR tails (RP_KEYS), or in the tails alone
      // (RP_TAILS), when `rem` entries (this range's and the later ones')
      // are still to be built; 256-byte alignment of every array
      const uint64_t own_cap = part_room > 1024 ? (part_room - 1024) / 24 : 0;
      auto cap_of = [&](uint64_t rem, int* place) -> uint64_t {
        uint64_t best = own_cap;
        *place = RP_OWN;
        if (!forced) {
          const uint64_t keys_cap = std::min<uint64_t>(4 * rem > 768 ? (4 * rem - 768) / 12 : 0,
                                                       part_room > 512 ? (part_room - 512) / 8 : 0);
          const uint64_t tails_cap = 4 * rem > 1024 ? (4 * rem - 1024) / 16 : 0;
          if (keys_cap > best) {
            best = keys_cap;
            *place = RP_KEYS;
          }
          if (tails_cap > best) {
            best = tails_cap;
            *place = RP_TAILS;
          }
        }
        return std::min<uint64_t>(best, quarter);
      };
      rng.clear();
      rcnt.clear();
      rplace.clear();
      // ranges of at most 2^24 rows: (row - r0) << colbits | col then has at
      // most 24 + 26 = 50 bits, six radix passes (two of 9 bits)
      const int max_buckets = std::max(1, (1 << 24) >> shift);
      const uint64_t forced_cap = forced ? ((uint64_t)nnz + p->csr_parts - 1) / p->csr_parts : 0;
      uint64_t rem = (uint64_t)nnz;
      int lo = 0, place = RP_OWN;
      uint64_t acc = 0;
      uint64_t cap_e = forced ? forced_cap : cap_of(rem, &place);
      for (int b = 0; b < nbuck; ++b) {
        if (b > lo && (acc + hh[b] > std::max<uint64_t>(cap_e, 1) || b - lo + 1 > max_buckets)) {
          rng.push_back({lo, b});
          rcnt.push_back(acc);
          rplace.push_back(place);
          rem -= acc;
          lo = b;
          acc = 0;
          cap_e = forced ? forced_cap : cap_of(rem, &place);
        }
        acc += hh[b];
      }
      rng.push_back({lo, nbuck});
      rcnt.push_back(acc);
      rplace.push_back(place);
      // each range takes the placement that needs the least of the range
      // buffer region among those its CSR tails allow (a range the plan made
      // larger than its tails hold, a single dense bucket, falls back to its
      // own region); the region is sized for the largest need
      part_bytes = 0;
      uint64_t pmax = 1, done = 0;
      for (size_t r = 0; r < rng.size(); ++r) {
        const uint64_t e = std::max<uint64_t>(rcnt[r], 1);
        done += rcnt[r];
        const uint64_t tail = 4 * ((uint64_t)nnz - done);  // bytes free in each of col and val after it
        if (al(e * 8) + al(e * 4) + 256 <= tail)
          rplace[r] = RP_TAILS;
        else if (al(e * 8) + 256 <= tail)
          rplace[r] = RP_KEYS;
        else
          rplace[r] = RP_OWN;
        const size_t need = rplace[r] == RP_OWN ? 2 * al(e * 8) + 2 * al(e * 4)
                            : rplace[r] == RP_KEYS ? 2 * al(e * 4)
                                                   : 0;
        part_bytes = std::max(part_bytes, need);
        pmax = std::max(pmax, e);
      }
      pe = (size_t)pmax;
      if (!forced && part_bytes > part_room)
        return ctx_fail(ctx, LIGHTNE_ERR_OOM,
                        "sparsifier of %lld entries: %.1f GB of table and %.1f GB of CSR leave %.1f GB for sort "
                        "buffers, the row ranges need %.1f GB",
                        (long long)nnz, tab_bytes / 1e9, csr_bytes / 1e9, part_room / 1e9, part_bytes / 1e9);
      return LIGHTNE_OK;
    };
    // keep the value cache only inside the workspace and where the plan with
    // it fits the workspace as it is (no regrowth, which would re-run the
    // sampler)
    if (lw && !lw_in_ws && !lw_ext) lw = nullptr;
    bool keep_lw = ctx_owned && ((lw && lw_in_ws) || (lw_force && lw_allowed));
    if (keep_lw) {
      const int rc = plan(lw_bytes, !lw_force);
      if (rc != LIGHTNE_OK) {
        if (lw_force) return rc;
        keep_lw = false;
      } else if (!lw_force && std::max(tab_bytes + lw_bytes + part_bytes, front_reserve) + csr_bytes > ws_now) {
        keep_lw = false;
      }
    }
    if (!keep_lw) {
      LN_CHECK(demote());
      lw_allowed = false;  // a regrowth's count pass must not place it either
      ctx->err.clear();
      LN_CHECK(plan(0, false));
      // A repeated call finds the workspace an earlier call sized: ranges that
      // fit it as it is avoid a regrowth (a second sampler run and the wait
      // for the freed memory's wipe) that a plan made from a slightly
      // different free-memory reading would force.  Taken unless it needs
      // half as many ranges again.
      if (ctx_owned && p->csr_parts < 2 && ws_now > 0) {
        const auto rng_free = rng;
        const auto rcnt_free = rcnt;
        const auto rplace_free = rplace;
        const size_t pe_free = pe, part_free = part_bytes;
        const bool fits = plan(0, true) == LIGHTNE_OK;
        ctx->err.clear();
        if (!(fits && rng.size() <= rng_free.size() + rng_free.size() / 2 + 1 &&
              std::max(tab_bytes + part_bytes, front_reserve) + csr_bytes <= ws_now)) {
          rng = rng_free;
          rcnt = rcnt_free;
          rplace = rplace_free;
          pe = pe_free;
          part_bytes = part_free;
        }
      }
    }
    const size_t lwb = keep_lw ? lw_bytes : 0;
    st.csr_parts = (uint32_t)rng.size();
    // Context-owned: [table | value cache | range buffers | ... | CSR] in the
    // workspace, which grows (and the deterministic sampler runs again) when
    // it is too small.  Otherwise the range buffers follow the table when
    // they fit and the CSR arrays belong to the handle.
    char* pbuf = nullptr;
    if (ctx_owned) {
      const size_t need = std::max(tab_bytes + lwb + part_bytes, front_reserve) + csr_bytes;
      if (ctx->bufs["ws"].bytes < need) {
        if (p->csr_parts < 2 && need > avail)
          return ctx_fail(ctx, LIGHTNE_ERR_OOM, "stage 1 needs a %.1f GB workspace, %.1f GB are available",
                          need / 1e9, avail / 1e9);
        // -sample 1: a little room for the re-run below to keep a few more keys
        const size_t slack = p->sample ? need / 1000 + ((size_t)64 << 20) : 0;
        if (!ctx_buf(ctx, "ws", need + slack)) return LIGHTNE_ERR_OOM;
        const lightne_sparsify_stats keep = st;
        LN_CHECK(run_sampler(ctx, g, p, &st, &tv));
        st.sampler_runs = keep.sampler_runs + 1;
        LN_CHECK(table_canon(ctx, tv, &canon, &st.t_canon));
        LN_CHECK(count_pass(true));  // places the cache behind the table iff keep_lw
        if (keep_lw && !(lw && lw_in_ws))
          return ctx_fail(ctx, LIGHTNE_ERR_HIP, "stage 1: the value cache did not fit the grown workspace");
        // the key -> count multiset is exact (integer counts), so the keys and
        // their number are the first run's
        if (st.distinct_keys != keep.distinct_keys || st.samples != keep.samples)
          return ctx_fail(ctx, LIGHTNE_ERR_HIP, "the sampler gave a different table on its second run");
        if (p->sample) {
          // -sample 1: the FP64 weights are sums in the order of the atomic
          // adds, so M_exact and a key within an ulp of the trunc-log
          // threshold (adj > 1) can differ between runs, and with them the
          // nonzeros per row range; plan the layout again from a fresh run on
          // the grown workspace
          if (restarts >= 2)
            return ctx_fail(ctx, LIGHTNE_ERR_HIP, "stage 1: the workspace plan did not settle after growing");
          return sparsify_sized(ctx, g, p, out, stats, ctx_owned, front_reserve, xs, xs_bytes, restarts + 1);
        }
        if (st.nnz != keep.nnz)
          return ctx_fail(ctx, LIGHTNE_ERR_HIP, "the sampler gave a different sparsifier on its second run");
        st.csr_parts = keep.csr_parts;
      }
      char* ws = (char*)ctx->bufs["ws"].p;
      const size_t ws_bytes = ctx->bufs["ws"].bytes;
      pbuf = ws + tab_bytes + lwb;
      char* back = ws + (ws_bytes - csr_bytes);
      rp = (int64_t*)back;
      col = (int32_t*)(back + al((g->n + 1) * 8));
      val = (float*)(back + al((g->n + 1) * 8) + al(nz * 4));
    } else {
      const size_t ws_bytes = ctx->bufs["ws"].bytes;
      pbuf = tab_bytes + part_bytes <= ws_bytes ? (char*)ctx->bufs["ws"].p + tab_bytes
                                                : (char*)ctx_buf(ctx, "s1_part", part_bytes);
      if (!pbuf) return LIGHTNE_ERR_OOM;
      LN_CHECK(dev_alloc(ctx, mval, nz * 4, "sparsifier values"));
      LN_CHECK(dev_alloc(ctx, mrp, (g->n + 1) * 8, "csr row_ptr"));
      LN_CHECK(dev_alloc(ctx, mcol, nz * 4, "csr col"));
      rp = (int64_t*)mrp.p;
      col = (int32_t*)mcol.p;
      val = (float*)mval.p;
    }
    ta.lw_out = nullptr;
    ta.lw_in = ta.occ_off ? lw : nullptr;  // numbered by the count pass
    ta.lw_n = lw_n;
    st.value_cache = lw ? (lw_ext ? 2u : 1u) : 0u;
    int64_t base = 0;
    for (size_t r = 0; r < rng.size(); ++r) {
      const uint64_t r0 = (uint64_t)rng[r].first << shift;
      const uint64_t r1 = std::min<uint64_t>((uint64_t)rng[r].second << shift, (uint64_t)g->n);
      const int64_t cnt = (int64_t)rcnt[r];
      // this range's sort buffers (the plan checked that they fit)
      const size_t e = (size_t)std::max<int64_t>(cnt, 1);
      auto up = [&](const void* q) { return (char*)(((uintptr_t)q + 255) & ~(uintptr_t)255); };
      uint64_t *pkey, *ktmp;
      float *pval, *vtmp;
      if (rplace[r] == RP_OWN) {
        pkey = (uint64_t*)pbuf;
        ktmp = (uint64_t*)(pbuf + al(e * 8));
        pval = (float*)(pbuf + 2 * al(e * 8));
        vtmp = (float*)(pbuf + 2 * al(e * 8) + al(e * 4));
      } else {
        char* ctail = up(col + base + cnt);
        char* vtail = up(val + base + cnt);
        pkey = (uint64_t*)ctail;
        ktmp = (uint64_t*)vtail;
        if (rplace[r] == RP_KEYS) {
          pval = (float*)pbuf;
          vtmp = (float*)(pbuf + al(e * 4));
        } else {
          pval = (float*)(ctail + al(e * 8));
          vtmp = (float*)(vtail + al(e * 8));
        }
      }
      ta.skey = pkey;
      ta.sval = pval;
      uint64_t* kres = pkey;
      float* vres = pval;
      if (cnt > 0) {
        const uint64_t per_block = trunc_per_block(tv.cap, nblk);
        hipLaunchKernelGGL(k_range_counts, dim3((nw + 255) / 256), dim3(256), 0, s, (const uint32_t*)bh,
                           rng[r].first, rng[r].second, nw, wav_emit);
        LN_LAUNCH(ctx);
        LN_CHECK(scan_u32_to_u64(ctx, wav_emit, wav_off, nw));
        uint64_t got = 0;
        LN_HIP(ctx, hipMemcpyAsync(&got, wav_off + nw, 8, hipMemcpyDeviceToHost, s));
        LN_HIP(ctx, hipStreamSynchronize(s));
        if ((int64_t)got != cnt)
          return ctx_fail(ctx, LIGHTNE_ERR_HIP, "row range %llu..%llu holds %llu entries, the histogram %lld",
                          (unsigned long long)r0, (unsigned long long)r1, (unsigned long long)got, (long long)cnt);
        if (tv.pk)
          hipLaunchKernelGGL(k_trunc_emit<true>, dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, (uint32_t)r0,
                             (uint32_t)r1, (const uint64_t*)wav_off);
        else
          hipLaunchKernelGGL(k_trunc_emit<false>, dim3(nblk), dim3(TR_THREADS), 0, s, ta, per_block, (uint32_t)r0,
                             (uint32_t)r1, (const uint64_t*)wav_off);
        LN_LAUNCH(ctx);
        LN_CHECK(radix_sort_pairs(ctx, pkey, pval, cnt, bits_for(r1 - r0) + (canon ? 0 : colbits), ktmp, vtmp,
                                  canon ? colbits : 0, &kres, &vres));
      }
      hipLaunchKernelGGL(k_csr_finish, dim3(grid_blocks(cnt + 1, 256)), dim3(256), 0, s, (const uint64_t*)kres, cnt,
                         (int64_t)(r1 - r0), colbits, rp + r0, col + base, (const float*)vres, val + base, base);
      LN_LAUNCH(ctx);
      base += cnt;
    }
  }
  LN_HIP(ctx, hipEventRecord(ctx->ev[5], s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  auto* csr = new lightne_csr();
  csr->n = g->n;
  csr->nnz = nnz;
  csr->owns = !ctx_owned;
  csr->row_ptr = ctx_owned ? rp : (int64_t*)mrp.release();
  csr->col = ctx_owned ? col : (int32_t*)mcol.release();
  csr->val = ctx_owned ? val : (float*)mval.release();
  st.t_table = ev_seconds(ctx->ev[3], ctx->ev[4]);
  st.t_csr = ev_seconds(ctx->ev[4], ctx->ev[5]);
  *out = csr;
  st.row_order = canon ? 1u : 2u;
  if (stats) *stats = st;
  return LIGHTNE_OK;
}


// Full symmetric CSR from upper-triangle storage (the matrix the reference's
// SPARSE_MATRIX_TYPE_SYMMETRIC / SPARSE_FILL_MODE_UPPER descriptor denotes,
// mklredsvd.cc:95-98): A = U + U^T - diag(U).  The arrays are context
// buffers ("up_*"), valid until the next expansion.
int csr_expand_upper(lightne_ctx* ctx, const lightne_csr* U, lightne_csr* F) {
  const int64_t n = U->n, nnz = U->nnz;
  const int colbits = std::max(1, bits_for((uint64_t)n + 1));  // 2^colbits > n: no key equals the sentinel
  if (2 * colbits > 64) return ctx_fail(ctx, LIGHTNE_ERR_BAD_ARG, "upper expansion: n too large");
  const size_t m2 = (size_t)std::max<int64_t>(2 * nnz, 1);
  auto* key = (uint64_t*)ctx_buf(ctx, "up_key", m2 * 8);
  auto* kv = (float*)ctx_buf(ctx, "up_val", m2 * 4);
  auto* rp = (int64_t*)ctx_buf(ctx, "up_rp", (n + 1) * 8);
  auto* col = (int32_t*)ctx_buf(ctx, "up_col", m2 * 4);
  auto* nd = (unsigned long long*)ctx_buf(ctx, "up_ndiag", 8);
  if (!key || !kv || !rp || !col || !nd) return LIGHTNE_ERR_OOM;
  hipStream_t s = ctx->stream;
  LN_HIP(ctx, hipMemsetAsync(nd, 0, 8, s));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_up_emit, dim3(grid_blocks(nnz, 256)), dim3(256), 0, s, U->row_ptr, U->col, U->val, n, nnz,
                       colbits, key, kv, nd);
    LN_LAUNCH(ctx);
    LN_CHECK(radix_sort_pairs(ctx, key, kv, 2 * nnz, 2 * colbits));
  }
  unsigned long long ndiag = 0;
  LN_HIP(ctx, hipMemcpyAsync(&ndiag, nd, 8, hipMemcpyDeviceToHost, s));
  LN_HIP(ctx, hipStreamSynchronize(s));
  const int64_t nf = 2 * nnz - (int64_t)ndiag;  // the sentinels sort past position nf
  hipLaunchKernelGGL(k_csr_finish, dim3(grid_blocks(nf + 1, 256)), dim3(256), 0, s, (const uint64_t*)key, nf, n,
                     colbits, rp, col, nullptr, nullptr);
  LN_LAUNCH(ctx);
  LN_HIP(ctx, hipStreamSynchronize(s));
  F->n = n;
  F->nnz = nf;
  F->row_ptr = rp;
  F->col = col;
  F->val = kv;  // the first nf sorted values
  F->owns = false;
  F->plan = nullptr;
  return LIGHTNE_OK;
}

extern "C" {

int lightne_sample_scalars(int64_t n, int64_t m, const lightne_sparsify_params* p, lightne_sparsify_stats* out) {
  if (!p || !out || n <= 0 || m < 0) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_sample_scalars: bad argument");
  // PathEmbed.h:89 logn = log(n); :99 table_size; :108-111 M, 1/lambda, rounding
  const double logn = log((double)(uint64_t)n);
  const uint64_t tn = p->walk_len * (uint64_t)n;
  uint64_t ts = (uint64_t)((double)tn * logn * (double)p->mem_ratio);
  if (p->tablesz > ts) ts = p->tablesz;
  const double M = (double)tn * logn * (double)p->sample_ratio;
  // a graph without edges draws nothing (the reference divides by m = 0)
  const double inv_lambda = m > 0 ? M / (double)(uint64_t)m : 0.0;
  if (!(inv_lambda < 1.8e19))
    return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_sample_scalars: %g samples per edge is out of range", inv_lambda);
  const uint64_t e = (uint64_t)inv_lambda;
  out->logn = logn;
  out->M = M;
  out->one_over_lambda = inv_lambda;
  out->exp_rv_rounded = e;
  out->fractional_bits = inv_lambda - (double)e;
  out->table_size = ts;
  return LIGHTNE_OK;
}

int lightne_sparsify_counts(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p,
                            lightne_table** out, lightne_sparsify_stats* stats) {
  if (!ctx || !out) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_sparsify_counts: NULL argument");
  CallScope scope(ctx);
  *out = nullptr;
  lightne_sparsify_stats st{};
  TabView tv;
  LN_CHECK(run_sampler(ctx, g, p, &st, &tv));
  st.sampler_runs = 1;
  const uint64_t cap = tv.cap;
  // the compacted table addresses slots with 32-bit indices
  if (cap > (1ull << 32))
    return ctx_fail(ctx, LIGHTNE_ERR_OOM, "lightne_sparsify_counts: table of %llu slots exceeds 2^32",
                    (unsigned long long)cap);
  uint64_t distinct, csum;
  double wsum;
  LN_CHECK(table_totals(ctx, tv, &distinct, &csum, &wsum));
  st.distinct_keys = distinct;
  st.M_exact = p->sample ? wsum : (double)csum;
  DevMem mk, mc, mw, mi;
  const uint64_t nk = distinct;
  LN_CHECK(dev_alloc(ctx, mk, std::max<uint64_t>(nk, 1) * 8, "table keys"));
  LN_CHECK(dev_alloc(ctx, mc, std::max<uint64_t>(nk, 1) * 4, "table counts"));
  LN_CHECK(dev_alloc(ctx, mw, std::max<uint64_t>(nk, 1) * 8, "table weights"));
  LN_CHECK(dev_alloc(ctx, mi, std::max<uint64_t>(nk, 1) * 4, "table index"));
  auto* counter = (unsigned long long*)ctx_buf(ctx, "compact_ctr", 8);
  if (!counter) return LIGHTNE_ERR_OOM;
  LN_HIP(ctx, hipMemsetAsync(counter, 0, 8, ctx->stream));
  hipLaunchKernelGGL(k_table_compact, dim3(grid_blocks((int64_t)cap, 256)), dim3(256), 0, ctx->stream, tv,
                     (uint64_t*)mk.p, (uint32_t*)mi.p, counter);
  LN_LAUNCH(ctx);
  LN_CHECK(radix_sort_pairs_u32(ctx, (uint64_t*)mk.p, (uint32_t*)mi.p, (int64_t)nk, 32 + bits_for((uint64_t)g->n)));
  if (nk) {
    hipLaunchKernelGGL(k_gather_counts, dim3(grid_blocks((int64_t)nk, 256)), dim3(256), 0, ctx->stream, tv,
                       (const uint32_t*)mi.p, (int64_t)nk, (uint32_t*)mc.p, (double*)mw.p);
    LN_LAUNCH(ctx);
  }
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  auto* t = new lightne_table();
  t->nkeys = nk;
  t->keys = (uint64_t*)mk.release();
  t->cnt = (uint32_t*)mc.release();
  t->wgh = (double*)mw.release();
  *out = t;
  if (stats) *stats = st;
  return LIGHTNE_OK;
}

int lightne_table_info(const lightne_table* t, uint64_t* nkeys) {
  if (!t || !nkeys) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_table_info: NULL argument");
  *nkeys = t->nkeys;
  return LIGHTNE_OK;
}

int lightne_table_copy(lightne_ctx* ctx, const lightne_table* t, uint64_t* keys, uint32_t* counts, double* weights) {
  if (!ctx || !t) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_table_copy: NULL argument");
  if (t->nkeys == 0) return LIGHTNE_OK;
  if (keys) LN_HIP(ctx, hipMemcpyAsync(keys, t->keys, t->nkeys * 8, hipMemcpyDefault, ctx->stream));
  if (counts) LN_HIP(ctx, hipMemcpyAsync(counts, t->cnt, t->nkeys * 4, hipMemcpyDefault, ctx->stream));
  if (weights) LN_HIP(ctx, hipMemcpyAsync(weights, t->wgh, t->nkeys * 8, hipMemcpyDefault, ctx->stream));
  LN_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LIGHTNE_OK;
}

int lightne_table_free(lightne_table* t) {
  if (!t) return LIGHTNE_OK;
  if (t->keys) (void)hipFree(t->keys);
  if (t->cnt) (void)hipFree(t->cnt);
  if (t->wgh) (void)hipFree(t->wgh);
  delete t;
  return LIGHTNE_OK;
}

int lightne_sparsify(lightne_ctx* ctx, const lightne_graph* g, const lightne_sparsify_params* p, lightne_csr** out,
                     lightne_sparsify_stats* stats) {
  if (!ctx || !out) return global_fail(LIGHTNE_ERR_BAD_ARG, "lightne_sparsify: NULL argument");
  CallScope scope(ctx);
  return sparsify_impl(ctx, g, p, out, stats, false, 0, nullptr, 0);
}

}  // extern "C"
