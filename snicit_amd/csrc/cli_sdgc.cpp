// snicit_sdgc -- drop-in for the reference's SDGC executable (main/SDGC.cu:11-114).
//
// Same flags: -m/--mode {SNICIT,BF,SNIG,XY} (default SNICIT), -n N (1024), -l L (120),
// -t threshold (30), -b batch (60000, must divide the 60000 inputs); same dataset paths
// relative to the working directory (main/SDGC.cu:58-62); same stdout lines
// (src/SDGC/SNICIT/SNICIT.hpp:659-664, src/SDGC/BF/BF.hpp:219, src/SDGC/SNIG/SNIG.hpp:304,
// src/SDGC/XY/XY.hpp:500-501) and the identity-level golden check (utility.hpp:243-266).
// Extra flags: --synthetic (seeded SDGC-like net and blob images instead of the files;
// the golden categories then come from a plain BF run), --seed S, --inputs K, --dataset DIR,
// --hash-reorder (renumber the neurons of the weight and input files by the reference's
// bucket hash before the run, as its SNICIT mode does: src/SDGC/SNICIT/SNICIT.hpp:99-139).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "snicit.h"

namespace {

[[noreturn]] void die(const std::string& m) {
  std::cerr << "snicit_sdgc: " << m << std::endl;
  std::exit(1);
}

void check(int rc, const char* what) {
  if (rc != SNICIT_OK) die(std::string(what) + ": " + snicit_last_error());
}

// Weight file n{N}-l{l}.tsv: "i\tj\tv" = input i -> output j, 1-based (utility.hpp:109-113).
void read_layer(const std::string& path, int N, std::vector<int32_t>& cols, std::vector<float>& vals,
                int l, int K) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  std::vector<int> fill(N, 0);
  int i, j;
  float v;
  while (f >> i >> j >> v) {
    if (i < 1 || i > N || j < 1 || j > N) die("index out of range in " + path);
    int o = j - 1;
    if (fill[o] >= K) die("row with more than 32 nonzeros in " + path);
    size_t at = ((size_t)l * N + o) * K + fill[o]++;
    cols[at] = i - 1;
    vals[at] = v;
  }
  for (int o = 0; o < N; ++o)
    for (int r = fill[o]; r < K; ++r) {  // pad with value-0 slots
      size_t at = ((size_t)l * N + o) * K + r;
      cols[at] = fill[o] ? cols[((size_t)l * N + o) * K] : 0;
      vals[at] = 0.f;
    }
}

// Inputs "sample\tneuron\tvalue", 1-based (read_input_xy, utility.hpp:268-293).
void read_inputs(const std::string& path, int N, int64_t num, std::vector<float>& y) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  y.assign((size_t)num * N, 0.f);
  long long b, n;
  float v;
  while (f >> b >> n >> v)
    if (b >= 1 && b <= num && n >= 1 && n <= N) y[(size_t)(b - 1) * N + (n - 1)] = v;
}

// The reference's neuron renumbering (src/SDGC/reorder/hash.h:14-21, buckets per N at
// src/SDGC/SNICIT/SNICIT.hpp:99-104): neuron v = q * buckets + r moves to r * (N / buckets) + q,
// so neurons that share a residue modulo the bucket count become neighbours.  It is a
// permutation of [0, N) when buckets divides N, applied alike to the input and output index of
// every weight and to the input features.  Categories do not depend on neuron order, so the run
// must classify as without it (the golden check is on categories); the final activations are the
// same values renumbered, up to the FP32 rounding of the row sums' changed input order.
std::vector<int32_t> hash_permutation(int N) {
  const std::map<int, int> buckets = {{65536, 4096}, {16384, 1024}, {4096, 256}, {1024, 64}};
  const int nb = buckets.at(N), width = N / nb;
  std::vector<int32_t> to(N);
  for (int v = 0; v < N; ++v) to[v] = v / nb + (v % nb) * width;
  return to;
}

// Renumber an ELL net ([L][N][K], row o holds output neuron o) and the row-major inputs.
void apply_permutation(const std::vector<int32_t>& to, int L, int K, std::vector<int32_t>& cols,
                       std::vector<float>& vals, std::vector<float>& y, int64_t num) {
  const size_t N = to.size(), row = (size_t)K;
  std::vector<int32_t> c2(cols.size());
  std::vector<float> v2(vals.size());
  for (int l = 0; l < L; ++l)
    for (size_t o = 0; o < N; ++o) {
      const size_t src = ((size_t)l * N + o) * row, dst = ((size_t)l * N + to[o]) * row;
      for (size_t r = 0; r < row; ++r) {
        c2[dst + r] = to[cols[src + r]];
        v2[dst + r] = vals[src + r];
      }
    }
  cols.swap(c2);
  vals.swap(v2);
  std::vector<float> y2(y.size());
  for (int64_t b = 0; b < num; ++b)
    for (size_t n = 0; n < N; ++n) y2[(size_t)b * N + to[n]] = y[(size_t)b * N + n];
  y.swap(y2);
}

// Golden categories: one 1-based sample id per line (read_golden, utility.hpp:243-256).
std::vector<int32_t> read_golden(const std::string& path, int64_t num) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  std::vector<int32_t> g(num, 0);
  long long id;
  while (f >> id)
    if (id >= 1 && id <= num) g[id - 1] = 1;
  return g;
}

}  // namespace

int main(int argc, char** argv) {
  std::string mode = "SNICIT", dataset = "../dataset/SDGC";
  int N = 1024, L = 120, t = 30;
  int64_t batch = 60000, inputs = 60000;
  bool synthetic = false, hash_reorder = false;
  uint64_t seed = 2023;
  for (int a = 1; a < argc; ++a) {
    std::string s = argv[a];
    auto next = [&]() -> std::string {
      if (a + 1 >= argc) die("missing value for " + s);
      return argv[++a];
    };
    if (s == "-m" || s == "--mode") mode = next();
    else if (s == "-n" || s == "--num_neurons") N = std::atoi(next().c_str());
    else if (s == "-l" || s == "--num_layers") L = std::atoi(next().c_str());
    else if (s == "-t" || s == "--threshold") t = std::atoi(next().c_str());
    else if (s == "-b" || s == "--input_batch_size") batch = std::atoll(next().c_str());
    else if (s == "--inputs") inputs = std::atoll(next().c_str());
    else if (s == "--synthetic") synthetic = true;
    else if (s == "--seed") seed = std::strtoull(next().c_str(), nullptr, 10);
    else if (s == "--dataset") dataset = next();
    else if (s == "--hash-reorder") hash_reorder = true;
    else if (s == "-h" || s == "--help") {
      std::cout << "usage: snicit_sdgc [-m SNICIT|BF|SNIG|XY] [-n N] [-l L] [-t T] [-b B] "
                   "[--synthetic] [--seed S] [--inputs K] [--dataset DIR] [--hash-reorder]\n";
      return 0;
    } else die("unknown option " + s);
  }
  // bias per N (main/SDGC.cu:49-54)
  std::map<int, float> bias_map = {{65536, -0.45f}, {16384, -0.4f}, {4096, -0.35f}, {1024, -0.3f}};
  if (!bias_map.count(N)) die("N must be one of 1024, 4096, 16384, 65536");
  const float bias = bias_map[N];
  std::map<std::string, int> modes = {{"SNICIT", SNICIT_MODE_SNICIT}, {"BF", SNICIT_MODE_BF},
                                      {"SNIG", SNICIT_MODE_SNIG}, {"XY", SNICIT_MODE_XY}};
  if (!modes.count(mode)) die("Error mode. Please correct your mode name");
  if (batch <= 0 || inputs % batch) die("batch must be a factor of the number of inputs");
  std::cout << "Current mode: " << mode << std::endl;

  // every file is read and checked before the device is touched
  const int K = 32;
  const bool generate_on_device = synthetic && !hash_reorder;
  std::vector<float> y0;
  std::vector<int32_t> golden, cols;
  std::vector<float> vals;
  if (generate_on_device) {
    y0.resize((size_t)inputs * N);
    check(snicit_gen_images(N, seed + 1, 0, inputs, y0.data()), "images");
  } else {
    cols.resize((size_t)L * N * K);
    vals.assign((size_t)L * N * K, 0.0625f);
    if (synthetic) {
      check(snicit_gen_sdgc_ell(N, L, seed, cols.data()), "net");
      y0.resize((size_t)inputs * N);
      check(snicit_gen_images(N, seed + 1, 0, inputs, y0.data()), "images");
    } else {
      const std::string wdir = dataset + "/tsv_weights/neuron" + std::to_string(N);
      for (int l = 0; l < L; ++l)
        read_layer(wdir + "/n" + std::to_string(N) + "-l" + std::to_string(l + 1) + ".tsv", N,
                   cols, vals, l, K);
      read_inputs(dataset + "/MNIST/sparse-images-" + std::to_string(N) + ".tsv", N, inputs, y0);
      golden = read_golden(dataset + "/MNIST/neuron" + std::to_string(N) + "-l" +
                               std::to_string(L) + "-categories.tsv",
                           inputs);
    }
    if (hash_reorder) {
      apply_permutation(hash_permutation(N), L, K, cols, vals, y0, inputs);
      std::cout << "Neurons renumbered by the bucket hash (" << N << " neurons)" << std::endl;
    }
  }

  snicit_ctx* ctx = nullptr;
  check(snicit_ctx_create(0, &ctx), "context");
  snicit_net* net = nullptr;
  if (generate_on_device) {
    check(snicit_net_generate_sdgc(ctx, N, L, seed, 0.0625f, bias, 32.f, &net), "net");
  } else {
    check(snicit_net_create_ell(ctx, N, L, K, cols.data(), vals.data(), SNICIT_BIAS_SCALAR, &bias,
                                32.f, &net),
          "net");
    std::vector<int32_t>().swap(cols);
    std::vector<float>().swap(vals);
  }
  snicit_run_cfg cfg;
  snicit_run_cfg_default(&cfg);
  cfg.mode = modes[mode];
  cfg.threshold = t;
  std::vector<int32_t> cat(inputs);
  snicit_stats st{};
  check(snicit_infer(net, &cfg, y0.data(), inputs, batch, cat.data(), &st), "infer");
  if (synthetic) {
    snicit_run_cfg bf;
    snicit_run_cfg_default(&bf);
    bf.mode = SNICIT_MODE_BF;
    golden.resize(inputs);
    snicit_stats sb{};
    check(snicit_infer(net, &bf, y0.data(), inputs, batch, golden.data(), &sb), "golden");
  }
  const int post_layers = mode == "SNICIT" ? L - t : L;
  if (mode == "SNICIT") {
    std::cout << "pre-conv Time [SNICIT] = " << st.ms_pre << "ms" << std::endl;
    std::cout << "cluster-based conversion Time [overhead] = " << st.ms_cluster << "ms" << std::endl;
    std::cout << "post-conv Time [SNICIT] = " << st.ms_post << "ms" << std::endl;
    std::cout << "recover Time [SNICIT] = " << st.ms_recover << "ms" << std::endl;
    std::cout << "SNICIT info: runtime " << st.ms_total << "ms avgpost "
              << st.ms_post / (post_layers > 0 ? post_layers : 1) << "ms" << std::endl;
  } else if (mode == "XY") {
    std::cout << "XY info: runtime " << st.ms_total << "ms avgpost "
              << st.ms_total / (L > 0 ? L : 1) << "ms" << std::endl;
  } else {
    std::cout << mode << " info: runtime " << st.ms_total << " ms" << std::endl;
  }
  int64_t diff = 0;
  for (int64_t i = 0; i < inputs; ++i) diff += (cat[i] != 0) != (golden[i] != 0);
  std::cout << "\nNumber of different categories: " << diff << std::endl;
  std::cout << (diff == 0 ? "CHALLENGE PASSED" : "CHALLENGE FAILED") << std::endl;
  snicit_net_destroy(net);
  snicit_ctx_destroy(ctx);
  return diff == 0 ? 0 : 3;
}
