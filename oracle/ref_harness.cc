// ref_harness.cc -- golden-vector harness built from the reference's own
// headers (TEST INFRASTRUCTURE, output only under oracle/_ref/ and the
// fixtures that tests/golden/make_golden.py writes from it).
//
// What comes from the reference tree (compiled here, not copied):
//   * gbbs_io::read_unweighted_symmetric_graph  (ligra/graph_io.h:185-252)
//   * symmetric_graph / symmetric_vertex accessors and map_edges
//     (ligra/graph.h:113-118, ligra/vertex.h:296-333,445-470)
//   * pbbs::random and pbbs::hash64_2           (pbbslib/random.h:38-51,
//                                                pbbslib/utilities.h:154-159)
//   * make_par_table / par_table::insert_add / map_table
//                                               (LightNE/par_table.h:60-109,163-170)
//   * pbbs::reduce                              (pbbslib/sequence_ops.h:66-83)
// LightNE/PathEmbed.h itself cannot be included: it pulls in mkl.h through
// MKLSVD.h, which this image does not have.  The per-edge loop and the walk
// (PathEmbed.h:37-58,128-171) and the trunc-log map (PathEmbed.h:199-218) are
// therefore restated below, driven entirely by the reference primitives above.
//
// With a tenth argument "-c" the graph is read by
// gbbs_io::read_compressed_symmetric_graph (ligra/graph_io.h:344-377), so the
// walks go through the compressed vertex's get_ith_out_neighbor
// (ligra/encodings/byte_pd_amortized.h:653-688) and map_edges through its
// block decoder; the harness then also writes out.bin.csr:
//   u64 n, u64 m, u64 offsets[n+1] (prefix sums of the degrees),
//   u32 decoded[m] (mapOutNghIdx: edge j of u at offsets[u] + j),
//   u32 ith[m]     (get_ith_out_neighbor(u, j) at offsets[u] + j).
//
// usage: ref_harness graph seed walk_len sample sample_ratio mem_ratio
//                    tablesz negative out.bin [-c]
// out.bin: u64 magic 'LNEGOLD1', u64 nkeys, u64 keys[nkeys] (sorted),
//          f64 vals[nkeys], f64 trunclog[nkeys] (NaN where erased),
//          f64 {logn, M, one_over_lambda, frac, M_exact, factor},
//          u64 {exp_rv_rounded, table_size, HT.m}
#include "ligra.h"
#include "pbbslib/random.h"
#include "pbbslib/sequence_ops.h"
#include "pbbslib/monoid.h"
#include "par_table.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

namespace {

template <class Graph>
uintE walk_steps(Graph& GA, uintE u, size_t len, pbbs::random& s) {
  for (size_t i = 0; i < len; i++) {
    auto vtx = GA.get_vertex(u);
    uintE d = vtx.getOutDegree();
    size_t j = s.rand() % d;
    u = std::get<0>(vtx.get_ith_out_neighbor(u, j));
    s = s.next();
  }
  return u;
}

}  // namespace

template <class Graph>
int run(Graph& GA, char** argv) {
  size_t seed = std::stoull(argv[2]);
  size_t walk_len = std::stoull(argv[3]);
  bool sample = std::stoi(argv[4]) != 0;
  float sample_ratio = static_cast<float>(std::stod(argv[5]));
  float mem_ratio = static_cast<float>(std::stod(argv[6]));
  size_t table_size = std::stoull(argv[7]);
  size_t negative = std::stoull(argv[8]);

  using K = size_t;
  using V = double;
  using T = std::tuple<K, V>;
  size_t n = GA.n, m = GA.m;
  double logn = log(n);
  auto empty = std::make_tuple(std::numeric_limits<size_t>::max(), static_cast<V>(0));
  table_size = std::max(static_cast<size_t>(walk_len * GA.n * logn * mem_ratio), table_size);
  auto hash_f = [&](size_t k) { return pbbs::hash64_2(k); };
  auto HT = make_par_table(table_size, empty, hash_f, 2);
  auto M = walk_len * GA.n * logn * sample_ratio;
  double one_over_lambda = static_cast<double>(M) / static_cast<double>(m);
  size_t exp_rv_rounded = static_cast<size_t>(one_over_lambda);
  double fractional_bits = one_over_lambda - static_cast<double>(exp_rv_rounded);
  pbbs::random base(seed);

  auto emit = [&](uintE x, uintE z, V w) {
    size_t key = (static_cast<size_t>(std::min(x, z)) << 32UL) + static_cast<size_t>(std::max(x, z));
    HT.insert_add(std::make_tuple(key, w));
  };
  auto one_sample = [&](uintE u, uintE v, pbbs::random& s) {
    uintE r = s.rand() % walk_len + 1;
    s = s.next();
    uintE k = s.rand() % r + 1;
    s = s.next();
    uintE x = walk_steps(GA, u, k - 1, s);
    uintE z = walk_steps(GA, v, r - k, s);
    return std::make_pair(x, z);
  };
  auto edge_f = [&](const uintE& u, const uintE& v, const pbbs::empty&) {
    size_t uv = (static_cast<size_t>(u) << 32UL) | static_cast<size_t>(v);
    auto s = base.fork(uv);
    size_t e = exp_rv_rounded;
    if (fractional_bits > 0.0000000001) {
      double q = static_cast<double>(s.rand()) / 18446744073709551616.0;
      e += static_cast<double>(q < fractional_bits);
      s = s.next();
    }
    if (!sample) {
      for (size_t k = 0; k < e; k++) {
        auto xz = one_sample(u, v, s);
        emit(xz.first, xz.second, static_cast<V>(1));
      }
    } else {
      double p = static_cast<double>(1.0 / GA.get_vertex(u).getOutDegree() +
                                     1.0 / GA.get_vertex(v).getOutDegree()) * logn;
      p = std::min(p, 1.0);
      double ipw = 1.0 / p;
      for (size_t k = 0; k < e; k++) {
        double q = static_cast<double>(s.rand()) / 18446744073709551616.0;
        s = s.next();
        if (q < p) {
          auto xz = one_sample(u, v, s);
          emit(xz.first, xz.second, static_cast<V>(ipw));
        }
      }
    }
  };
  GA.map_edges(edge_f);

  auto wgh_seq = pbbslib::make_sequence<V>(HT.m, [&](size_t i) {
    return std::get<0>(HT.table[i]) != std::numeric_limits<size_t>::max() ? std::get<1>(HT.table[i])
                                                                          : static_cast<V>(0);
  });
  V M_exact = pbbs::reduce(wgh_seq, pbbs::addm<V>());
  V vol = GA.m;
  V factor = vol * m / M_exact / 2 / negative;

  std::vector<T> before;
  for (size_t i = 0; i < HT.m; ++i)
    if (std::get<0>(HT.table[i]) != std::numeric_limits<size_t>::max()) before.push_back(HT.table[i]);
  std::sort(before.begin(), before.end(),
            [](const T& a, const T& b) { return std::get<0>(a) < std::get<0>(b); });

  auto map_f = [&](T kv) {
    uintE u = (uintE)(std::get<0>(kv) >> 32UL);
    uintE v = (uintE)(std::get<0>(kv));
    uintE d_u = GA.get_vertex(u).getOutDegree();
    uintE d_v = GA.get_vertex(v).getOutDegree();
    V wgh = (u != v) ? std::get<1>(kv) : std::get<1>(kv) * 2;
    V adjusted_wgh = wgh * factor / d_u / d_v;
    V log_wgh = log(adjusted_wgh);
    if (log_wgh > 0.0) return std::make_tuple(std::get<0>(kv), log_wgh);
    return empty;
  };
  HT.map_table(map_f);
  std::vector<T> after;
  for (size_t i = 0; i < HT.m; ++i)
    if (std::get<0>(HT.table[i]) != std::numeric_limits<size_t>::max()) after.push_back(HT.table[i]);
  std::sort(after.begin(), after.end(),
            [](const T& a, const T& b) { return std::get<0>(a) < std::get<0>(b); });

  FILE* f = std::fopen(argv[9], "wb");
  if (!f) return 3;
  uint64_t magic = 0x31444c4f47454e4cULL;  // "LNEGOLD1"
  uint64_t nk = before.size();
  std::fwrite(&magic, 8, 1, f);
  std::fwrite(&nk, 8, 1, f);
  for (auto& kv : before) { uint64_t k = std::get<0>(kv); std::fwrite(&k, 8, 1, f); }
  for (auto& kv : before) { double v = std::get<1>(kv); std::fwrite(&v, 8, 1, f); }
  size_t j = 0;
  for (auto& kv : before) {
    double t = std::numeric_limits<double>::quiet_NaN();
    if (j < after.size() && std::get<0>(after[j]) == std::get<0>(kv)) { t = std::get<1>(after[j]); ++j; }
    std::fwrite(&t, 8, 1, f);
  }
  double sc[6] = {logn, static_cast<double>(M), one_over_lambda, fractional_bits, M_exact, factor};
  std::fwrite(sc, 8, 6, f);
  uint64_t su[3] = {exp_rv_rounded, table_size, HT.m};
  std::fwrite(su, 8, 3, f);
  std::fclose(f);
  std::printf("n=%zu m=%zu nkeys=%zu kept=%zu M_exact=%.17g factor=%.17g\n", n, m, before.size(),
              after.size(), M_exact, factor);
  return 0;
}

template <class Graph>
int dump_csr(Graph& GA, const char* path) {
  const size_t n = GA.n, m = GA.m;
  std::vector<uint64_t> off(n + 1, 0);
  for (size_t u = 0; u < n; ++u) off[u + 1] = off[u] + GA.get_vertex(u).getOutDegree();
  std::vector<uint32_t> dec(m, 0xffffffffu), ith(m, 0xffffffffu);
  for (size_t u = 0; u < n; ++u) {
    auto f = [&](const uintE& src, const uintE& v, const pbbs::empty&, const uintT& j) { dec[off[src] + j] = v; };
    GA.get_vertex(u).mapOutNghIdx(u, f, false);
    const size_t d = off[u + 1] - off[u];
    for (size_t j = 0; j < d; ++j) ith[off[u] + j] = std::get<0>(GA.get_vertex(u).get_ith_out_neighbor(u, j));
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return 3;
  const uint64_t hdr[2] = {n, m};
  std::fwrite(hdr, 8, 2, f);
  std::fwrite(off.data(), 8, n + 1, f);
  std::fwrite(dec.data(), 4, m, f);
  std::fwrite(ith.data(), 4, m, f);
  std::fclose(f);
  return 0;
}

int main(int argc, char** argv) {
  const bool compressed = argc == 11 && std::string(argv[10]) == "-c";
  if (argc != 10 && !compressed) {
    std::fprintf(stderr, "usage: %s graph seed walk_len sample sample_ratio mem_ratio tablesz negative out [-c]\n",
                 argv[0]);
    return 2;
  }
  if (compressed) {
    auto GA = gbbs_io::read_compressed_symmetric_graph<pbbslib::empty>(argv[1], true, false);
    const int rc = run(GA, argv);
    return rc ? rc : dump_csr(GA, (std::string(argv[9]) + ".csr").c_str());
  }
  auto GA = gbbs_io::read_unweighted_symmetric_graph(argv[1], false);
  return run(GA, argv);
}
