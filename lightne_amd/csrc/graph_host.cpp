// Host graph utilities: Ligra AdjacencyGraph text IO (ligra/graph_io.h:185-231,
// util/x2adj.py:81-119), a binary CSR cache, and seeded synthetic generators
// (Chung-Lu, R-MAT).
// Generators are counter-based (every draw is a hash of (seed, edge, draw)),
// so the graph does not depend on the thread count.
#include "lightne_internal.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
// uniform double in [0, 1) from draw `i` of stream `s`
inline double unif(uint64_t seed, uint64_t s, uint64_t i) {
  const uint64_t h = mix64(mix64(seed ^ 0x9e3779b97f4a7c15ull) + mix64(s * 0x100000001b3ull + i));
  return (double)(h >> 11) * 0x1p-53;
}

unsigned nthreads() {
  unsigned t = std::thread::hardware_concurrency();
  if (const char* e = std::getenv("LIGHTNE_HOST_THREADS")) t = (unsigned)std::max(1, atoi(e));
  return std::max(1u, std::min(t, 64u));
}

template <class F>
void par_for(int64_t n, F f) {
  const unsigned T = nthreads();
  if (T == 1 || n < 2) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t) {
    const int64_t b = n * t / T, e = n * (t + 1) / T;
    th.emplace_back([=] { f(b, e); });
  }
  for (auto& x : th) x.join();
}

// undirected edge list (u, v), u != v -> symmetric CSR with sorted, deduplicated
// neighbour lists
int build_csr(int64_t n, const std::vector<uint32_t>& eu, const std::vector<uint32_t>& ev, int64_t* m_out,
              int64_t** offsets_host, uint32_t** edges_host) {
  const int64_t E = (int64_t)eu.size();
  std::vector<std::atomic<int64_t>> deg(n);
  for (auto& d : deg) d.store(0, std::memory_order_relaxed);
  par_for(E, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      if (eu[i] == ev[i]) continue;
      deg[eu[i]].fetch_add(1, std::memory_order_relaxed);
      deg[ev[i]].fetch_add(1, std::memory_order_relaxed);
    }
  });
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i].load(std::memory_order_relaxed);
  std::vector<uint32_t> adj(std::max<int64_t>(off[n], 1));
  std::vector<std::atomic<int64_t>> pos(n);
  for (int64_t i = 0; i < n; ++i) pos[i].store(off[i], std::memory_order_relaxed);
  par_for(E, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      if (eu[i] == ev[i]) continue;
      adj[pos[eu[i]].fetch_add(1, std::memory_order_relaxed)] = ev[i];
      adj[pos[ev[i]].fetch_add(1, std::memory_order_relaxed)] = eu[i];
    }
  });
  std::vector<int64_t> newdeg(n);
  par_for(n, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      auto* s = adj.data() + off[u];
      auto* t = adj.data() + off[u + 1];
      std::sort(s, t);
      newdeg[u] = std::unique(s, t) - s;
    }
  });
  auto* o = (int64_t*)std::malloc((n + 1) * sizeof(int64_t));
  if (!o) return global_fail(LIGHTNE_ERR_OOM, "graph: host allocation failed");
  o[0] = 0;
  for (int64_t i = 0; i < n; ++i) o[i + 1] = o[i] + newdeg[i];
  auto* a = (uint32_t*)std::malloc(std::max<int64_t>(o[n], 1) * sizeof(uint32_t));
  if (!a) {
    std::free(o);
    return global_fail(LIGHTNE_ERR_OOM, "graph: host allocation failed");
  }
  par_for(n, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) std::memcpy(a + o[u], adj.data() + off[u], newdeg[u] * sizeof(uint32_t));
  });
  *m_out = o[n];
  *offsets_host = o;
  *edges_host = a;
  return LIGHTNE_OK;
}

inline const char* skip_ws(const char* p, const char* end) {
  while (p < end && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
  return p;
}

}  // namespace

extern "C" {

int lightne_graph_read_adj(const char* path, int64_t* n_out, int64_t* m_out, int64_t** offsets_host,
                           uint32_t** edges_host) {
  if (!path || !n_out || !m_out || !offsets_host || !edges_host)
    return global_fail(LIGHTNE_ERR_BAD_ARG, "read_adj: NULL argument");
  int fd = open(path, O_RDONLY);
  if (fd < 0) return global_fail(LIGHTNE_ERR_IO, "read_adj: cannot open %s", path);
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return global_fail(LIGHTNE_ERR_IO, "read_adj: cannot stat %s", path);
  }
  const size_t len = (size_t)sb.st_size;
  const char* data = len ? (const char*)mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
  close(fd);
  if (len && data == MAP_FAILED) return global_fail(LIGHTNE_ERR_IO, "read_adj: mmap failed for %s", path);
  const char* p = data;
  const char* end = data + len;
  p = skip_ws(p, end);
  const char* hdr = "AdjacencyGraph";
  if ((size_t)(end - p) < strlen(hdr) || strncmp(p, hdr, strlen(hdr)) != 0) {
    if (len) munmap((void*)data, len);
    return global_fail(LIGHTNE_ERR_IO, "read_adj: %s is not an AdjacencyGraph file", path);
  }
  p += strlen(hdr);
  auto next_num = [&](uint64_t& v) -> bool {
    p = skip_ws(p, end);
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') x = x * 10 + (uint64_t)(*p++ - '0');
    v = x;
    return true;
  };
  uint64_t n = 0, m = 0;
  if (!next_num(n) || !next_num(m) || n == 0) {
    if (len) munmap((void*)data, len);
    return global_fail(LIGHTNE_ERR_IO, "read_adj: bad header in %s", path);
  }
  auto* off = (int64_t*)std::malloc((n + 1) * sizeof(int64_t));
  auto* adj = (uint32_t*)std::malloc(std::max<uint64_t>(m, 1) * sizeof(uint32_t));
  if (!off || !adj) {
    std::free(off);
    std::free(adj);
    if (len) munmap((void*)data, len);
    return global_fail(LIGHTNE_ERR_OOM, "read_adj: host allocation failed");
  }
  bool ok = true;
  for (uint64_t i = 0; i < n && ok; ++i) {
    uint64_t v;
    ok = next_num(v);
    off[i] = (int64_t)v;
  }
  off[n] = (int64_t)m;  // graph_io.h: offsets[n] = m
  for (uint64_t i = 0; i < m && ok; ++i) {
    uint64_t v;
    ok = next_num(v);
    if (ok && v >= n) ok = false;
    adj[i] = (uint32_t)v;
  }
  if (len) munmap((void*)data, len);
  for (uint64_t i = 0; i < n && ok; ++i) ok = off[i] <= off[i + 1];
  if (!ok) {
    std::free(off);
    std::free(adj);
    return global_fail(LIGHTNE_ERR_IO, "read_adj: truncated or inconsistent file %s", path);
  }
  *n_out = (int64_t)n;
  *m_out = (int64_t)m;
  *offsets_host = off;
  *edges_host = adj;
  return LIGHTNE_OK;
}

int lightne_graph_write_adj(const char* path, int64_t n, const int64_t* off, const uint32_t* adj) {
  if (!path || n <= 0 || !off) return global_fail(LIGHTNE_ERR_BAD_ARG, "write_adj: bad argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return global_fail(LIGHTNE_ERR_IO, "write_adj: cannot open %s", path);
  std::vector<char> buf(1 << 20);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  std::fprintf(f, "AdjacencyGraph\n%lld\n%lld\n", (long long)n, (long long)off[n]);
  for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%lld\n", (long long)off[i]);
  for (int64_t j = 0; j < off[n]; ++j) std::fprintf(f, "%u\n", adj[j]);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  return bad ? global_fail(LIGHTNE_ERR_IO, "write_adj: write error on %s", path) : LIGHTNE_OK;
}

// Binary CSR: "LNECSR01", int64 n, int64 m, int64 offsets[n+1], uint32 edges[m]
static const char kBinMagic[8] = {'L', 'N', 'E', 'C', 'S', 'R', '0', '1'};

int lightne_graph_write_bin(const char* path, int64_t n, const int64_t* off, const uint32_t* adj) {
  if (!path || n <= 0 || !off || (off[n] > 0 && !adj)) return global_fail(LIGHTNE_ERR_BAD_ARG, "write_bin: bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return global_fail(LIGHTNE_ERR_IO, "write_bin: cannot open %s", path);
  const int64_t hdr[2] = {n, off[n]};
  bool ok = std::fwrite(kBinMagic, 1, 8, f) == 8 && std::fwrite(hdr, 8, 2, f) == 2 &&
            std::fwrite(off, 8, (size_t)(n + 1), f) == (size_t)(n + 1) &&
            std::fwrite(adj, 4, (size_t)off[n], f) == (size_t)off[n];
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LIGHTNE_OK : global_fail(LIGHTNE_ERR_IO, "write_bin: write error on %s", path);
}

int lightne_graph_read_bin(const char* path, int64_t* n_out, int64_t* m_out, int64_t** offsets_host,
                           uint32_t** edges_host) {
  if (!path || !n_out || !m_out || !offsets_host || !edges_host)
    return global_fail(LIGHTNE_ERR_BAD_ARG, "read_bin: NULL argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return global_fail(LIGHTNE_ERR_IO, "read_bin: cannot open %s", path);
  char magic[8];
  int64_t hdr[2];
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kBinMagic, 8) != 0 || std::fread(hdr, 8, 2, f) != 2) {
    std::fclose(f);
    return global_fail(LIGHTNE_ERR_IO, "read_bin: %s is not a binary CSR file", path);
  }
  const int64_t n = hdr[0], m = hdr[1];
  struct stat sb;
  const bool sized = fstat(fileno(f), &sb) == 0 && n > 0 && n < ((int64_t)1 << 32) && m >= 0 &&
                     (uint64_t)sb.st_size == 24 + 8 * (uint64_t)(n + 1) + 4 * (uint64_t)m;
  if (!sized) {
    std::fclose(f);
    return global_fail(LIGHTNE_ERR_IO, "read_bin: header of %s does not match its size", path);
  }
  auto* off = (int64_t*)std::malloc((size_t)(n + 1) * sizeof(int64_t));
  auto* adj = (uint32_t*)std::malloc((size_t)std::max<int64_t>(m, 1) * sizeof(uint32_t));
  bool ok = off && adj && std::fread(off, 8, (size_t)(n + 1), f) == (size_t)(n + 1) &&
            std::fread(adj, 4, (size_t)m, f) == (size_t)m;
  std::fclose(f);
  ok = ok && off[0] == 0 && off[n] == m;
  for (int64_t i = 0; i < n && ok; ++i) ok = off[i] <= off[i + 1];
  for (int64_t j = 0; j < m && ok; ++j) ok = adj[j] < (uint64_t)n;
  if (!ok) {
    std::free(off);
    std::free(adj);
    return global_fail(LIGHTNE_ERR_IO, "read_bin: truncated or inconsistent file %s", path);
  }
  *n_out = n;
  *m_out = m;
  *offsets_host = off;
  *edges_host = adj;
  return LIGHTNE_OK;
}

int lightne_gen_chung_lu(int64_t n, double beta, double mean_deg, double max_deg, uint64_t seed, int64_t* m_out,
                         int64_t** offsets_host, uint32_t** edges_host) {
  if (n < 2 || n >= (int64_t)1 << 32 || !(beta > 1.0) || !(mean_deg > 0.0) || !(max_deg >= mean_deg) || !m_out ||
      !offsets_host || !edges_host)
    return global_fail(LIGHTNE_ERR_BAD_ARG, "gen_chung_lu: bad argument");
  // expected degrees w_i = w_0 (1 + i/i0)^-g, g = 1/(beta-1).  Duplicate
  // (u,v) draws and self loops are dropped, which lowers the realised mean and
  // maximum; the model targets are calibrated (3 rounds) so that the realised
  // 2m/n and max degree approach (mean_deg, max_deg).
  const double g = 1.0 / (beta - 1.0);
  // fixed 64K-index blocks summed in block order: independent of the thread count
  const int64_t nblk = (n + 65535) / 65536;
  std::vector<double> part(nblk);
  auto model_mean = [&](double i0, double w0) {
    par_for(nblk, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) {
        double s = 0.0;
        for (int64_t i = k * 65536; i < std::min<int64_t>(n, (k + 1) * 65536); ++i)
          s += std::pow(1.0 + (double)i / i0, -g);
        part[k] = s;
      }
    });
    double s = 0.0;
    for (double v : part) s += v;
    return w0 * s / (double)n;
  };
  double tgt_mean = mean_deg, tgt_max = max_deg;
  std::vector<uint32_t> eu, ev;
  int rc = LIGHTNE_OK;
  for (int round = 0; round < 4; ++round) {
    double lo = 1e-9, hi = 1e12;
    for (int it = 0; it < 60; ++it) {
      const double mid = std::sqrt(lo * hi);
      if (model_mean(mid, tgt_max) < tgt_mean)
        lo = mid;
      else
        hi = mid;
    }
    const double i0 = std::sqrt(lo * hi);
    std::vector<double> cdf(n + 1, 0.0);
    for (int64_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + tgt_max * std::pow(1.0 + (double)i / i0, -g);
    const double W = cdf[n];
    const int64_t E = (int64_t)std::llround(W / 2.0);
    eu.assign(E, 0);
    ev.assign(E, 0);
    auto pick = [&](double r) -> uint32_t {
      const double x = r * W;
      int64_t k = std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin() - 1;
      return (uint32_t)std::min<int64_t>(std::max<int64_t>(k, 0), n - 1);
    };
    par_for(E, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        eu[i] = pick(unif(seed, 1, 2 * (uint64_t)i));
        ev[i] = pick(unif(seed, 1, 2 * (uint64_t)i + 1));
      }
    });
    if (round == 3) break;
    int64_t m = 0;
    int64_t* o = nullptr;
    uint32_t* a = nullptr;
    rc = build_csr(n, eu, ev, &m, &o, &a);
    if (rc != LIGHTNE_OK) return rc;
    int64_t dmax = 0;
    for (int64_t i = 0; i < n; ++i) dmax = std::max<int64_t>(dmax, o[i + 1] - o[i]);
    std::free(o);
    std::free(a);
    const double real_mean = (double)m / (double)n;
    tgt_mean *= mean_deg / std::max(real_mean, 1e-9);
    tgt_max *= max_deg / std::max<double>((double)dmax, 1.0);
    tgt_max = std::max(tgt_max, tgt_mean);
  }
  return build_csr(n, eu, ev, m_out, offsets_host, edges_host);
}

int lightne_gen_rmat(int32_t scale, double edge_factor, double a, double b, double c, uint64_t seed, int64_t* n_out,
                     int64_t* m_out, int64_t** offsets_host, uint32_t** edges_host) {
  if (scale < 1 || scale > 31 || !(edge_factor > 0) || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 || !n_out ||
      !m_out || !offsets_host || !edges_host)
    return global_fail(LIGHTNE_ERR_BAD_ARG, "gen_rmat: bad argument");
  const int64_t n = (int64_t)1 << scale;
  const int64_t E = (int64_t)std::llround(edge_factor * (double)n);
  std::vector<uint32_t> eu(E), ev(E);
  const uint32_t mask = (uint32_t)(n - 1);
  // vertex relabelling: a bijection of [0, 2^scale) (odd multiplier, xor-shift)
  const uint64_t mul = (mix64(seed + 7) | 1ull) & mask;
  auto relabel = [&](uint32_t x) -> uint32_t {
    uint64_t y = x;
    y = (y * (mul | 1ull)) & mask;
    y ^= y >> ((scale + 1) / 2);
    y = (y * 0x9E3779B1ull) & mask;
    y ^= y >> ((scale + 2) / 3);
    return (uint32_t)y;
  };
  par_for(E, [&](int64_t bb, int64_t ee) {
    for (int64_t i = bb; i < ee; ++i) {
      uint32_t u = 0, v = 0;
      for (int l = 0; l < scale; ++l) {
        const double r = unif(seed, 2, (uint64_t)i * 64 + (uint64_t)l);
        const uint32_t bit = 1u << (scale - 1 - l);
        if (r < a) {
        } else if (r < a + b) {
          v |= bit;
        } else if (r < a + b + c) {
          u |= bit;
        } else {
          u |= bit;
          v |= bit;
        }
      }
      eu[i] = relabel(u);
      ev[i] = relabel(v);
    }
  });
  *n_out = n;
  return build_csr(n, eu, ev, m_out, offsets_host, edges_host);
}

}  // extern "C"
