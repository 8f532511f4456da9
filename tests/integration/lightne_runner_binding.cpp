// The binding a maintainer adds to the reference runner (LightNE/LightNE.cc:29-122)
// to run its three stages through liblightne_amd, written against include/lightne.h
// only.  LightNE_mkl_runner's Ligra graph reaches this function as its CSR:
//   offsets[i] = sum of GA.get_vertex(j).getOutDegree() for j < i,
//   edges      = the neighbour lists in the graph's own order (ligra/vertex.h:296-300).
// main() stands in for Ligra's loader: it reads an AdjacencyGraph file with
// lightne_graph_read_adj (ligra/graph_io.h:185-252) and runs the stages with the
// reference's defaults, writing -ne_out / -pro_out as save() does (LightNE.cc:18-27).
// Build: g++ -std=c++17 -D__HIP_PLATFORM_AMD__ -I<repo>/include -I/opt/rocm/include
//        lightne_runner_binding.cpp -L<repo>/lightne_amd -llightne_amd -L/opt/rocm/lib -lamdhip64
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lightne.h"

struct RunnerArgs {  // the flags of LightNE.cc:31-50 this binding forwards
  size_t walk_len = 10, rank = 256, dim = 128, order = 10, negative = 1, table_size = 0;
  bool upper = false, sample = false, normalize = false;
  float sample_ratio = 1.0f, mem_ratio = 1.0f;
  uint64_t seed = 1;  // the reference seeds with time(0) (PathEmbed.h:499)
  std::string ne_out, pro_out;
};

static bool save(const std::string& file, const float* dev, size_t n, size_t d) {
  std::vector<float> host(n * d);
  if (hipMemcpy(host.data(), dev, n * d * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return false;
  FILE* f = std::fopen(file.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(host.data(), sizeof(float), n * d, f) == n * d;
  std::fclose(f);
  return ok;
}

// LightNE_mkl_runner with the three stages on the device; returns the running
// time in seconds (graph upload excluded, as the reference's graph load is),
// or -1 with the library's message on stderr
double lightne_gpu_runner(int64_t n, int64_t m, const int64_t* offsets_host, const uint32_t* edges_host,
                          const RunnerArgs& a) {
  lightne_ctx* ctx = nullptr;
  if (lightne_ctx_create(0, nullptr, &ctx) != LIGHTNE_OK) {
    std::fprintf(stderr, "%s\n", lightne_global_error());
    return -1;
  }
  int64_t* off_d = nullptr;
  uint32_t* adj_d = nullptr;
  float *emb = nullptr, *pro = nullptr;
  double secs = -1;
  if (hipMalloc(&off_d, (n + 1) * sizeof(int64_t)) != hipSuccess ||
      hipMalloc(&adj_d, (m > 0 ? m : 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMalloc(&emb, n * a.dim * sizeof(float)) != hipSuccess ||
      hipMalloc(&pro, n * a.dim * sizeof(float)) != hipSuccess ||
      hipMemcpy(off_d, offsets_host, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
      (m > 0 && hipMemcpy(adj_d, edges_host, m * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)) {
    std::fprintf(stderr, "device allocation or upload failed\n");
  } else {
    lightne_graph g{};
    g.n = n;
    g.m = m;
    g.offsets = off_d;
    g.edges = adj_d;
    g.ith_rule = LIGHTNE_ITH_PLAIN;
    lightne_sparsify_params sp{};  // zero-initialised: the automatic layouts
    sp.walk_len = a.walk_len;
    sp.negative = a.negative;
    sp.sample = a.sample;
    sp.upper = a.upper;
    sp.sample_ratio = a.sample_ratio;
    sp.mem_ratio = a.mem_ratio;
    sp.tablesz = a.table_size;
    sp.seed = a.seed;
    lightne_rsvd_params rp{};
    rp.rank = (int64_t)a.rank;
    rp.power_iters = 1;
    rp.precision = LIGHTNE_FP32;
    rp.upper = a.upper;
    rp.seed = 777;  // the reference's VSL seed (mklredsvd.cc:141)
    std::vector<float> S(a.rank);
    lightne_netsmf_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    // NetSMF (PathEmbed.h:494-563): generate_trunc_log_matrix_v2 + MKLRedSVD + compute_u_sigma_root
    if (lightne_netsmf(ctx, &g, &sp, &rp, (int64_t)a.dim, a.normalize, emb, S.data(), &st) != LIGHTNE_OK) {
      std::fprintf(stderr, "%s\n", lightne_last_error(ctx));  // the reference would exit(0) here
    } else if (!a.ne_out.empty() && !save(a.ne_out, emb, n, a.dim)) {
      std::fprintf(stderr, "cannot write %s\n", a.ne_out.c_str());
    } else {
      bool ok = true;
      if (!a.pro_out.empty()) {
        // chebyshev_expansion(emb, GA, dim, dim, order, 0.2, 0.5) (LightNE.cc:114)
        lightne_prop_params pp{};
        pp.order = (int64_t)a.order;
        pp.mu = 0.2f;
        pp.theta = 0.5f;
        pp.precision = LIGHTNE_FP32;
        lightne_prop_stats ps{};
        ok = lightne_chebyshev_expansion(ctx, &g, emb, (int64_t)a.dim, &pp, pro, &ps) == LIGHTNE_OK;
        if (!ok) std::fprintf(stderr, "%s\n", lightne_last_error(ctx));
        ok = ok && save(a.pro_out, pro, n, a.dim);
      }
      if (ok) {
        secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("### Running Time: %g\n", secs);
      }
    }
  }
  (void)hipFree(off_d);
  (void)hipFree(adj_d);
  (void)hipFree(emb);
  (void)hipFree(pro);
  lightne_ctx_destroy(ctx);
  return secs;
}

int main(int argc, char** argv) {
  RunnerArgs a;
  std::string graph;
  for (int i = 1; i < argc; ++i) {
    const std::string f = argv[i];
    const char* v = i + 1 < argc ? argv[i + 1] : "";
    if (f == "-walklen") a.walk_len = std::strtoull(v, nullptr, 10), ++i;
    else if (f == "-rank") a.rank = std::strtoull(v, nullptr, 10), ++i;
    else if (f == "-dim") a.dim = std::strtoull(v, nullptr, 10), ++i;
    else if (f == "-order") a.order = std::strtoull(v, nullptr, 10), ++i;
    else if (f == "-sample_ratio") a.sample_ratio = std::strtof(v, nullptr), ++i;
    else if (f == "-sample") a.sample = std::atoi(v) != 0, ++i;
    else if (f == "-seed") a.seed = std::strtoull(v, nullptr, 10), ++i;
    else if (f == "-ne_out") a.ne_out = v, ++i;
    else if (f == "-pro_out") a.pro_out = v, ++i;
    else graph = f;
  }
  int64_t n = 0, m = 0, *off = nullptr;
  uint32_t* edg = nullptr;
  if (lightne_graph_read_adj(graph.c_str(), &n, &m, &off, &edg) != LIGHTNE_OK) {
    std::fprintf(stderr, "%s\n", lightne_global_error());
    return 1;
  }
  const double t = lightne_gpu_runner(n, m, off, edg, a);
  lightne_host_free(off);
  lightne_host_free(edg);
  return t < 0 ? 1 : 0;
}
