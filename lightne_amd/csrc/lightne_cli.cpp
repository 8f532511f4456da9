// lightne -- command-line driver with the flags and log lines of the
// reference runner (LightNE/LightNE.cc:29-124) and of Ligra's
// generate_symmetric_main (ligra/ligra.h:494-505,642-667), running the three
// stages through the liblightne_amd C-ABI on one GPU.
//
//   lightne [flags] -s <graph.adj>       AdjacencyGraph text, or a binary CSR
//                                         file (recognised by its magic)
//   lightne [flags] -s -c <graph.cmp>    Ligra compressed format, decoded on
//                                         the device (-m is accepted; the file
//                                         is read whole either way)
//
// Flag parsing follows ligra/parse_command_line.h:32-113: exact string match,
// the value is the next argv entry, the graph file is the LAST argument.
// Reference flags (defaults as in LightNE.cc:31-50):
//   -walksperedge 100 (unused, as in the reference)  -walklen 10  -tablesz 0
//   -rank 256  -dim 128  -order 10  -negative 1  -upper 0  -sample 0
//   -analyze 0  -normalize 0  -sample_ratio 1.0  -mem_ratio 1.0
//   -ne_method netsmf|ne_zhang_et_al  -ne_out FILE  -pro_out FILE
//   -step_coeff 1,..,1 (parsed, unused)  -random_project_only 0
//   -sparse_project 0  -sparse_project_s 100  -s  -m  -c  -rounds 3
// Added flags:
//   -seed S        pbbs::random seed of the sampler (reference: time(0))
//   -svd_seed S    Gaussian sketch seed (reference: VSL seed 777)
//   -power_iters q randomized-SVD power iterations (1 = reference schedule)
//   -precision fp32|bf16|bf16_power|fp64   SpMM panels / propagation iterates
//                          (fp64: randomized SVD in FP64, propagation FP32)
//   -mu X -theta X propagation parameters (reference hard-codes 0.2, 0.5)
//   -ne_in FILE    read the stage-2 embedding (raw FP32 n x dim) instead of
//                  computing it, so stage 3 can run alone
//   -device D      HIP device ordinal
//   -ith_rule plain|bytepd   neighbour lookup of the walks; bytepd (the
//                  default with -c) reproduces the reference's lookup on
//                  compressed graphs
//   -gen SPEC      synthetic graph instead of a file:
//                  chung_lu:n,beta,mean,max,seed | rmat:scale,ef,a,b,c,seed
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <sstream>
#include <string>
#include <vector>

#include "lightne.h"

namespace {

struct Args {
  int argc;
  char** argv;
  bool has(const std::string& o) const {
    for (int i = 1; i < argc - 1; ++i)
      if (o == argv[i]) return true;
    return false;
  }
  std::string str(const std::string& o, const std::string& def) const {
    for (int i = 1; i < argc - 1; ++i)
      if (o == argv[i]) return argv[i + 1];
    return def;
  }
  long lng(const std::string& o, long def) const {
    for (int i = 1; i < argc - 1; ++i)
      if (o == argv[i]) return atol(argv[i + 1]);
    return def;
  }
  double dbl(const std::string& o, double def) const {
    for (int i = 1; i < argc - 1; ++i)
      if (o == argv[i]) {
        double v;
        if (sscanf(argv[i + 1], "%lf", &v) != 1) {
          fprintf(stderr, "bad value for %s\n", o.c_str());
          exit(2);
        }
        return v;
      }
    return def;
  }
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] void die(const char* what, lightne_ctx* ctx = nullptr) {
  fprintf(stderr, "lightne: %s: %s\n", what, ctx ? lightne_last_error(ctx) : lightne_global_error());
  exit(1);
}

#define CK(expr, ctx)                \
  do {                               \
    if ((expr) != LIGHTNE_OK) die(#expr, ctx); \
  } while (0)

#define HK(expr)                                                               \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "lightne: %s: %s\n", #expr, hipGetErrorString(e_));      \
      exit(1);                                                                 \
    }                                                                          \
  } while (0)

// LightNE.cc:18-27: raw FP32 row-major n x d, no header
void save(const std::string& file, const float* dev, int64_t n, int64_t d) {
  const double t0 = now();
  std::vector<float> h((size_t)n * d);
  HK(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  if (FILE* f = fopen(file.c_str(), "wb")) {
    fwrite(h.data(), sizeof(float), h.size(), f);
    fclose(f);
  } else {
    fprintf(stderr, "lightne: cannot write %s\n", file.c_str());
    exit(1);
  }
  printf("save time: %.6f\n", now() - t0);
}

bool gen_graph(const std::string& spec, int64_t* n, int64_t* m, int64_t** off, uint32_t** adj) {
  const auto c = spec.find(':');
  if (c == std::string::npos) return false;
  const std::string kind = spec.substr(0, c);
  std::vector<double> v;
  std::stringstream ss(spec.substr(c + 1));
  std::string tok;
  while (std::getline(ss, tok, ',')) v.push_back(atof(tok.c_str()));
  if (kind == "chung_lu" && v.size() == 5) {
    *n = (int64_t)v[0];
    return lightne_gen_chung_lu(*n, v[1], v[2], v[3], (uint64_t)v[4], m, off, adj) == LIGHTNE_OK;
  }
  if (kind == "rmat" && v.size() == 6)
    return lightne_gen_rmat((int32_t)v[0], v[1], v[2], v[3], v[4], (uint64_t)v[5], n, m, off, adj) == LIGHTNE_OK;
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  Args P{argc, argv};
  if (argc < 2) {
    printf("usage: %s [flags] -s <inFile>\n", argv[0]);
    return 2;
  }
  const std::string gen = P.str("-gen", "");
  const char* iFile = argv[argc - 1];
  if (gen.empty() && !P.has("-s")) {
    printf("# The application expects the input graph to be symmetric (-s flag).\n");
    printf("# Please run on a symmetric input.\n");
  }
  const bool compressed = P.has("-c");
  const size_t walks_per_edge = (size_t)P.lng("-walksperedge", 100);
  const size_t walk_len = (size_t)P.lng("-walklen", 10);
  const size_t table_size = (size_t)P.lng("-tablesz", 0);
  const size_t rank = (size_t)P.lng("-rank", 256);
  const size_t dim = (size_t)P.lng("-dim", 128);
  const size_t order = (size_t)P.lng("-order", 10);
  const size_t negative = (size_t)P.lng("-negative", 1);
  const bool upper = P.lng("-upper", 0) != 0;
  const bool sample = P.lng("-sample", 0) != 0;
  const bool analyze = P.lng("-analyze", 0) != 0;
  const bool normalize = P.lng("-normalize", 0) != 0;
  const float sample_ratio = (float)P.dbl("-sample_ratio", 1.0);
  const float mem_ratio = (float)P.dbl("-mem_ratio", 1.0);
  const std::string ne_method = P.str("-ne_method", "netsmf");
  const std::string ne_out = P.str("-ne_out", "");
  const std::string pro_out = P.str("-pro_out", "");
  const std::string ne_in = P.str("-ne_in", "");
  const std::string step_coeff_str = P.str("-step_coeff", "1,1,1,1,1,1,1,1,1,1");
  const bool random_project_only = P.lng("-random_project_only", 0) != 0;
  const bool sparse_project = P.lng("-sparse_project", 0) != 0;
  const float sparse_project_s = (float)P.dbl("-sparse_project_s", 100.0);
  const size_t rounds = (size_t)P.lng("-rounds", 3);
  const uint64_t seed = (uint64_t)P.lng("-seed", (long)time(nullptr));
  const uint64_t svd_seed = (uint64_t)P.lng("-svd_seed", 777);
  const int power_iters = (int)P.lng("-power_iters", 1);
  const std::string precision = P.str("-precision", "fp32");
  const float mu = (float)P.dbl("-mu", 0.2);
  const float theta = (float)P.dbl("-theta", 0.5);
  const int device = (int)P.lng("-device", 0);

  std::vector<float> step_coeff;
  {
    std::stringstream ss(step_coeff_str);
    std::string t;
    while (std::getline(ss, t, ',')) step_coeff.push_back((float)atof(t.c_str()));
  }
  if (ne_method != "netsmf" && ne_method != "ne_zhang_et_al") {
    fprintf(stderr, "lightne: -ne_method must be netsmf or ne_zhang_et_al\n");
    return 2;
  }
  if (precision != "fp32" && precision != "bf16" && precision != "bf16_power" && precision != "fp64") {
    fprintf(stderr, "lightne: -precision must be fp32, bf16, bf16_power or fp64\n");
    return 2;
  }

  // -c graphs keep the reference's i-th-neighbour lookup for compressed input
  // unless -ith_rule plain is given
  const std::string ith_rule = P.str("-ith_rule", compressed ? "bytepd" : "plain");
  if (ith_rule != "plain" && ith_rule != "bytepd") {
    fprintf(stderr, "lightne: -ith_rule must be plain or bytepd\n");
    return 2;
  }

  // graph load (excluded from the running time, as in the reference): host
  // reads first, so a missing file is reported before any device work
  int64_t n = 0, m = 0;
  int64_t* off_h = nullptr;
  uint32_t* adj_h = nullptr;
  std::vector<char> img;  // -c: the whole file, decoded on the device
  if (!gen.empty()) {
    if (!gen_graph(gen, &n, &m, &off_h, &adj_h)) die("graph generation (-gen)");
  } else if (compressed) {
    FILE* f = fopen(iFile, "rb");
    if (!f) {
      fprintf(stderr, "lightne: reading the graph: cannot open %s\n", iFile);
      return 1;
    }
    fseek(f, 0, SEEK_END);
    const long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    img.resize(len > 0 ? (size_t)len : 0);
    const bool ok = len >= 24 && fread(img.data(), 1, img.size(), f) == img.size();
    fclose(f);
    if (!ok) {
      fprintf(stderr, "lightne: reading the graph: %s is not a compressed graph file\n", iFile);
      return 1;
    }
    memcpy(&n, img.data(), 8);
    memcpy(&m, img.data() + 8, 8);
    if (n <= 0 || m < 0) {
      fprintf(stderr, "lightne: reading the graph: bad header in %s\n", iFile);
      return 1;
    }
  } else {
    char magic[8] = {0};
    FILE* f = fopen(iFile, "rb");
    const bool bin = f && fread(magic, 1, 8, f) == 8 && memcmp(magic, "LNECSR01", 8) == 0;
    if (f) fclose(f);
    const int rc = bin ? lightne_graph_read_bin(iFile, &n, &m, &off_h, &adj_h)
                       : lightne_graph_read_adj(iFile, &n, &m, &off_h, &adj_h);
    if (rc != LIGHTNE_OK) die("reading the graph");
  }
  lightne_ctx* ctx = nullptr;
  if (lightne_ctx_create(device, nullptr, &ctx) != LIGHTNE_OK) die("lightne_ctx_create");
  int64_t* off_d = nullptr;
  uint32_t* adj_d = nullptr;
  HK(hipMalloc(&off_d, (n + 1) * sizeof(int64_t)));
  HK(hipMalloc(&adj_d, std::max<int64_t>(m, 1) * sizeof(uint32_t)));
  if (compressed && gen.empty()) {
    void* img_d = nullptr;
    HK(hipMalloc(&img_d, img.size()));
    HK(hipMemcpy(img_d, img.data(), img.size(), hipMemcpyHostToDevice));
    if (lightne_graph_decode_compressed(ctx, img_d, img.size(), n, m, off_d, adj_d) != LIGHTNE_OK)
      die("decoding the compressed graph", ctx);
    HK(hipFree(img_d));
    std::vector<char>().swap(img);
  } else {
    HK(hipMemcpy(off_d, off_h, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (m) HK(hipMemcpy(adj_d, adj_h, m * sizeof(uint32_t), hipMemcpyHostToDevice));
    lightne_host_free(off_h);
    lightne_host_free(adj_h);
  }
  lightne_graph g{n, m, off_d, adj_d, ith_rule == "bytepd" ? LIGHTNE_ITH_BYTEPD : LIGHTNE_ITH_PLAIN, 0};

  printf("### Application: LightNE_mkl\n");
  printf("### Graph: %s\n", gen.empty() ? iFile : gen.c_str());
  printf("### Threads: 1 GPU (%s)\n", lightne_version());
  printf("### n: %lld\n", (long long)n);
  printf("### m: %lld\n", (long long)m);
  printf("### Params: \n");
  printf("###  -walksperedge = %zu\n", walks_per_edge);
  printf("###  -walklen = %zu\n", walk_len);
  printf("###  -tablesz = %zu\n", table_size);
  printf("###  -rank = %zu\n", rank);
  printf("###  -dim = %zu\n", dim);
  printf("###  -order = %zu\n", order);
  printf("###  -negative = %zu\n", negative);
  printf("###  -upper = %s\n", upper ? "true" : "false");
  printf("###  -sample = %s\n", sample ? "true" : "false");
  printf("###  -analyze = %s\n", analyze ? "true" : "false");
  printf("###  -ne_method = %s\n", ne_method.c_str());
  printf("###  -normalize = %d\n", (int)normalize);
  printf("###  -sample_ratio = %g\n", sample_ratio);
  printf("###  -mem_ratio = %g\n", mem_ratio);
  printf("###  -random_project_only = %d\n", (int)random_project_only);
  printf("###  -sparse_project = %d\n", (int)sparse_project);
  printf("###  -sparse_project_s = %g\n", sparse_project_s);
  printf("###  -ne_out = %s\n", ne_out.c_str());
  printf("###  -pro_out = %s\n", pro_out.c_str());
  printf("###  -step_coeff = ");
  for (float c : step_coeff) printf("%g ", c);
  printf("\n###  -seed = %llu -svd_seed = %llu -power_iters = %d -precision = %s -mu = %g -theta = %g\n",
         (unsigned long long)seed, (unsigned long long)svd_seed, power_iters, precision.c_str(), mu, theta);
  printf("### ------------------------------------\n");
  fflush(stdout);

  float* emb = nullptr;
  float* pro = nullptr;
  HK(hipMalloc(&emb, (size_t)n * dim * sizeof(float)));
  double total = 0.0;
  for (size_t r = 0; r < rounds; ++r) {
    const double t0 = now();
    const int prec = precision == "bf16"         ? LIGHTNE_BF16
                     : precision == "bf16_power" ? LIGHTNE_BF16_POWER
                     : precision == "fp64"       ? LIGHTNE_FP64
                                                 : LIGHTNE_FP32;
    if (!ne_in.empty()) {
      std::vector<float> h((size_t)n * dim);
      FILE* f = fopen(ne_in.c_str(), "rb");
      if (!f || fread(h.data(), sizeof(float), h.size(), f) != h.size()) {
        fprintf(stderr, "lightne: cannot read %zu floats from %s\n", h.size(), ne_in.c_str());
        return 1;
      }
      fclose(f);
      HK(hipMemcpy(emb, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    } else if (ne_method == "ne_zhang_et_al") {
      // PathEmbed.h:410-463 (LightNE.cc:106)
      lightne_rsvd_params rp{};
      rp.rank = (int64_t)rank;
      rp.power_iters = power_iters;
      rp.precision = prec;
      rp.seed = svd_seed;
      lightne_rsvd_stats st{};
      std::vector<float> S(rank);
      const double ts = now();
      CK(lightne_ne_zhang(ctx, &g, &rp, (int64_t)dim, emb, S.data(), &st), ctx);
      printf("# svd time: %.6f\n", now() - ts);
    } else {
      lightne_sparsify_params sp{};
      sp.walk_len = walk_len;
      sp.negative = negative;
      sp.sample = sample;
      sp.upper = upper;
      sp.sample_ratio = sample_ratio;
      sp.mem_ratio = mem_ratio;
      sp.tablesz = table_size;
      sp.seed = seed;
      lightne_rsvd_params rp{};
      rp.rank = (int64_t)rank;
      rp.power_iters = power_iters;
      rp.precision = prec;
      rp.random_project_only = random_project_only;
      rp.upper = upper;
      rp.seed = svd_seed;
      rp.sparse_project = sparse_project;
      rp.sparse_project_s = sparse_project_s;
      lightne_netsmf_stats st{};
      std::vector<float> S(rank);
      printf("# seed = %llu\n", (unsigned long long)seed);
      CK(lightne_netsmf(ctx, &g, &sp, &rp, (int64_t)dim, normalize, emb, S.data(), &st), ctx);
      const auto& s1 = st.sparsify;
      printf("# logn = %g sample_ratio = %g mem_ratio = %g\n", s1.logn, sample_ratio, mem_ratio);
      printf("# table_size = %llu\n", (unsigned long long)s1.table_size);
      printf("# actual table.size = %llu\n", (unsigned long long)s1.table_capacity);
      printf("# total samples = %g\n", s1.M);
      printf("# sample param = %d\n", (int)sample);
      printf("# 1/lambda = %g\n", s1.one_over_lambda);
      printf("# exp_rv_rounded = %llu\n", (unsigned long long)s1.exp_rv_rounded);
      printf("# fractional_bits = %g\n", s1.fractional_bits);
      printf("# count time: %.6f\n", s1.t_sample);
      printf("# tot wgh (samples) = %.17g\n", s1.M_exact);
      printf("# factor = %.17g\n", s1.factor);
      printf("# nnz in upper triangle of sparsifier = %llu\n", (unsigned long long)s1.kept_upper);
      printf("# nnz in sparsifier = %llu\n", (unsigned long long)s1.nnz);
      printf("# scan time: %.6f\n", s1.t_table + s1.t_csr);
      printf("# generate matrix time: %.6f\n", st.t_sparsify);
      if (!random_project_only)
        printf("# computing u*sqrt(sigma), sigma(0)=%g, sigma(%zu)=%g\n", S[0], dim - 1, S[dim - 1]);
      printf("# svd time: %.6f\n", st.t_svd);
    }
    if (!ne_out.empty()) {
      printf("dump network embedding to %s\n", ne_out.c_str());
      save(ne_out, emb, n, (int64_t)dim);
    }
    if (!pro_out.empty()) {
      if (!pro) HK(hipMalloc(&pro, (size_t)n * dim * sizeof(float)));
      const double tp = now();
      lightne_prop_params pp{};
      pp.order = (int64_t)order;
      pp.mu = mu;
      pp.theta = theta;
      pp.precision = prec == LIGHTNE_FP64 ? LIGHTNE_FP32 : prec;
      lightne_prop_stats ps{};
      CK(lightne_chebyshev_expansion(ctx, &g, emb, (int64_t)dim, &pp, pro, &ps), ctx);
      printf("proX time: %.6f\n", now() - tp);
      printf("dump proX embedding to %s\n", pro_out.c_str());
      save(pro_out, pro, n, (int64_t)dim);
    }
    const double tt = now() - t0;
    printf("### Running Time: %.6f\n", tt);
    fflush(stdout);
    total += tt;
  }
  printf("# time per iter: %.6f\n", total / (double)std::max<size_t>(rounds, 1));
  (void)hipFree(emb);
  if (pro) (void)hipFree(pro);
  (void)hipFree(off_d);
  (void)hipFree(adj_d);
  lightne_ctx_destroy(ctx);
  return 0;
}
