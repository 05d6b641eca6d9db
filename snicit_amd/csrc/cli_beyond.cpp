// snicit_beyond -- drop-in for the reference's beyond executable (main/beyond.cu:10-150).
//
// Same flags: -m/--mode {SNICIT,BF,SNIG} (default SNICIT), -k/--benchmark {A,B,C,D} (default
// A), -n/--num_input (10000), -b/--batch_size (10000), -t/--threshold (default per net:
// 8/8/6/6, main/beyond.cu:65,75,85,95); same dataset paths relative to the working directory
// (main/beyond.cu:57-95) and file formats (src/beyond/SNICIT/SNICIT.hpp:175-365):
//   tsv_weights/<net>/l1-dense.tsv       input layer, one value per line, input-major
//   tsv_weights/<net>/l<i>-sparse.tsv     hidden layer, "in<TAB>out<TAB>value", 0-based
//   tsv_weights/<net>/l<L+2>-dense.tsv    output layer, input-major
//   tsv_biases/<net>/...                  one bias per line, same names
//   MNIST/t10k-images-idx3-ubyte, t10k-labels-idx1-ubyte; CIFAR-10/cifar-input.txt,
//   cifar-label.bin (CIFAR nets have no input layer: the hidden files start at l1)
// Same result line: "<MODE> info: accuracy <p>% runtime <ms>ms avgpost <ms>ms"
// (src/beyond/SNICIT/SNICIT.hpp:635-636).  Extra flags: --dataset DIR, --synthetic (seeded
// inputs with 20% of features set to U[0,1) values when the image files are absent; the
// labels then come from a BF run of the same net), --seed S.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "snicit.h"

namespace {

[[noreturn]] void die(const std::string& m) {
  std::cerr << "snicit_beyond: " << m << std::endl;
  std::exit(1);
}

void check(int rc, const char* what) {
  if (rc != SNICIT_OK) die(std::string(what) + ": " + snicit_last_error());
}

std::vector<float> read_values(const std::string& path) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  std::vector<float> v;
  std::string line;
  for (int64_t no = 1; std::getline(f, line); ++no) {
    if (line.empty()) continue;
    char* end = nullptr;
    const float x = std::strtof(line.c_str(), &end);  // std::stof's conversion, without throwing
    if (end == line.c_str()) die("not a number at line " + std::to_string(no) + " of " + path);
    v.push_back(x);
  }
  return v;
}

// Hidden layer file: lines "in out value" grouped by input row (CSR over input neurons).
void read_sparse(const std::string& path, int N, std::vector<int64_t>& row_ptr,
                 std::vector<int32_t>& cols, std::vector<float>& vals) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  const int64_t base = (int64_t)cols.size();
  std::vector<int64_t> count(N, 0);
  int i, j;
  float v;
  int last = -1;
  while (f >> i >> j >> v) {
    if (i < 0 || i >= N || j < 0 || j >= N) die("index out of range in " + path);
    if (i < last) die("entries not grouped by input row in " + path);
    last = i;
    ++count[i];
    cols.push_back(j);
    vals.push_back(v);
  }
  int64_t off = base;
  for (int r = 0; r < N; ++r) {
    row_ptr.push_back(off);
    off += count[r];
  }
  row_ptr.push_back(off);
}

// idx files: big-endian 32-bit header fields, then bytes.
bool read_idx(const std::string& path, int header_words, std::vector<uint8_t>& data) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.seekg(4 * header_words);
  data.assign(std::istreambuf_iterator<char>(f), {});
  return true;
}

// Reads an input-major [rows_in][cols] dense layer and returns it as [cols][rows_in]; the
// file must hold exactly rows_in * cols values.
std::vector<float> read_dense_transposed(const std::string& path, int rows_in, int cols) {
  const std::vector<float> a = read_values(path);
  if ((int64_t)a.size() != (int64_t)rows_in * cols)
    die("dense layer size mismatch in " + path + ": " + std::to_string(a.size()) +
        " values, expected " + std::to_string((int64_t)rows_in * cols));
  std::vector<float> t((size_t)rows_in * cols);
  for (int i = 0; i < rows_in; ++i)
    for (int k = 0; k < cols; ++k) t[(size_t)k * rows_in + i] = a[(size_t)i * cols + k];
  return t;
}

}  // namespace

int main(int argc, char** argv) {
  std::string mode = "SNICIT", bench = "A", dataset = "../dataset/beyond";
  int64_t num_input = 10000, batch = 10000;
  int threshold = -1;
  bool synthetic = false;
  uint64_t seed = 7;
  for (int a = 1; a < argc; ++a) {
    std::string s = argv[a];
    auto next = [&]() -> std::string {
      if (a + 1 >= argc) die("missing value for " + s);
      return argv[++a];
    };
    if (s == "-m" || s == "--mode") mode = next();
    else if (s == "-k" || s == "--benchmark") bench = next();
    else if (s == "-n" || s == "--num_input") num_input = std::atoll(next().c_str());
    else if (s == "-b" || s == "--batch_size") batch = std::atoll(next().c_str());
    else if (s == "-t" || s == "--threshold") threshold = std::atoi(next().c_str());
    else if (s == "--dataset") dataset = next();
    else if (s == "--synthetic") synthetic = true;
    else if (s == "--seed") seed = std::strtoull(next().c_str(), nullptr, 10);
    else if (s == "-h" || s == "--help") {
      std::cout << "usage: snicit_beyond [-m SNICIT|BF|SNIG] [-k A|B|C|D] [-n N] [-b B] [-t T] "
                   "[--dataset DIR] [--synthetic] [--seed S]\n";
      return 0;
    } else die("unknown option " + s);
  }
  // benchmark table (main/beyond.cu:57-95)
  std::string net_dir;
  int N = 0, L = 0, t_default = 0;
  bool cifar = false;
  if (bench == "A") net_dir = "n128-l18-acc94.94", N = 128, L = 18, t_default = 8;
  else if (bench == "B") net_dir = "n256-l18-acc96.88", N = 256, L = 18, t_default = 8;
  else if (bench == "C") net_dir = "n256-l12-acc95.61", N = 256, L = 12, t_default = 6;
  else if (bench == "D") net_dir = "n256-l12-acc75.86", N = 256, L = 12, t_default = 6, cifar = true;
  else die("Error benchmark. Please correct your benchmark name");
  if (threshold < 0) threshold = t_default;
  if (mode != "SNICIT" && mode != "BF" && mode != "SNIG") die("mode must be SNICIT, BF or SNIG");
  if (batch <= 0 || num_input % batch) die("batch_size must divide num_input");
  std::cout << "Current mode: " << mode << std::endl;
  std::cout << "Benchmark: " << bench << std::endl;

  const std::string wdir = dataset + "/networks/tsv_weights/" + net_dir + "/";
  const std::string bdir = dataset + "/networks/tsv_biases/" + net_dir + "/";
  const int off = cifar ? 1 : 2;  // file number of the first hidden layer
  const int Nin = cifar ? 0 : 784, Nout = 10;
  std::vector<float> w_in, b_in;
  if (!cifar) {
    w_in = read_dense_transposed(wdir + "l1-dense.tsv", Nin, N);
    b_in = read_values(bdir + "l1-dense.tsv");
    if ((int)b_in.size() != N) die("input layer bias size mismatch");
  }
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> cols;
  std::vector<float> vals, bias;
  for (int h = 0; h < L; ++h) {
    const std::string name = "l" + std::to_string(h + off) + "-sparse.tsv";
    read_sparse(wdir + name, N, row_ptr, cols, vals);
    std::vector<float> b = read_values(bdir + name);
    if ((int)b.size() != N) die("bias size mismatch in " + name);
    bias.insert(bias.end(), b.begin(), b.end());
  }
  const std::string out_name = "l" + std::to_string(L + off) + "-dense.tsv";
  std::vector<float> w_out = read_dense_transposed(wdir + out_name, N, Nout);
  std::vector<float> b_out = read_values(bdir + out_name);
  if ((int)b_out.size() != Nout) die("output layer bias size mismatch");

  // inputs and labels
  const int width = cifar ? N : Nin;
  std::vector<float> x((size_t)num_input * width);
  std::vector<int32_t> labels;
  bool have_inputs = false;
  if (cifar) {
    std::vector<float> v;
    std::ifstream f(dataset + "/CIFAR-10/cifar-input.txt");
    if (f) {
      v = read_values(dataset + "/CIFAR-10/cifar-input.txt");
      if ((int64_t)v.size() >= (int64_t)num_input * width) {
        std::copy(v.begin(), v.begin() + (size_t)num_input * width, x.begin());
        have_inputs = true;
      }
    }
    std::vector<uint8_t> lb;
    // CIFAR-10 binary records: one label byte then 3072 pixel bytes; the label of sample i is
    // byte i * 3073 (src/beyond/SNICIT/SNICIT.hpp:474-477)
    constexpr int64_t kCifarRecord = 3073;
    if (read_idx(dataset + "/CIFAR-10/cifar-label.bin", 0, lb) && num_input > 0 &&
        (int64_t)lb.size() >= (num_input - 1) * kCifarRecord + 1) {
      labels.resize(num_input);
      for (int64_t i = 0; i < num_input; ++i) labels[i] = lb[(size_t)(i * kCifarRecord)];
    }
  } else {
    std::vector<uint8_t> img, lb;
    if (read_idx(dataset + "/MNIST/t10k-images-idx3-ubyte", 4, img) &&
        (int64_t)img.size() >= num_input * width) {
      for (size_t i = 0; i < (size_t)num_input * width; ++i) x[i] = img[i] / 255.0f;
      have_inputs = true;
    }
    if (read_idx(dataset + "/MNIST/t10k-labels-idx1-ubyte", 2, lb) && (int64_t)lb.size() >= num_input)
      labels.assign(lb.begin(), lb.begin() + num_input);
  }
  if (!have_inputs) {
    if (!synthetic) die("input images not found under " + dataset + " (use --synthetic)");
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> u(0.f, 1.f);
    for (auto& v : x) v = u(rng) < 0.2f ? u(rng) : 0.f;
    labels.clear();  // the synthetic inputs have no labels: compare against BF below
  }

  snicit_ctx* ctx = nullptr;
  check(snicit_ctx_create(0, &ctx), "context");
  snicit_net* net = nullptr;
  check(snicit_net_create_csr(ctx, N, L, row_ptr.data(), cols.data(), vals.data(), 1,
                              SNICIT_BIAS_PER_NEURON, bias.data(), 1.0f, &net),
        "net");
  check(snicit_net_set_io(net, Nin, cifar ? nullptr : w_in.data(), cifar ? nullptr : b_in.data(),
                          1.0f, Nout, w_out.data(), b_out.data()),
        "dense layers");
  snicit_run_cfg cfg;
  snicit_run_cfg_default(&cfg);
  cfg.mode = mode == "SNICIT" ? SNICIT_MODE_SNICIT : mode == "SNIG" ? SNICIT_MODE_SNIG : SNICIT_MODE_BF;
  cfg.threshold = threshold;
  cfg.cluster = SNICIT_CLUSTER_LOSSY;
  std::vector<int32_t> cat(num_input);
  snicit_stats st;
  check(snicit_infer(net, &cfg, x.data(), num_input, batch, cat.data(), &st), "inference");
  if (labels.empty()) {
    snicit_run_cfg bf = cfg;
    bf.mode = SNICIT_MODE_BF;
    labels.resize(num_input);
    snicit_stats st_bf;
    check(snicit_infer(net, &bf, x.data(), num_input, batch, labels.data(), &st_bf), "BF labels");
  }
  int64_t hit = 0;
  for (int64_t i = 0; i < num_input; ++i) hit += cat[i] == labels[i];
  const double post_layers = cfg.mode == SNICIT_MODE_SNICIT ? L - threshold : L;
  const double post_ms = cfg.mode == SNICIT_MODE_SNICIT ? st.ms_post : st.ms_total;
  std::cout << mode << " info: accuracy " << 100.0 * (double)hit / (double)num_input << "%"
            << " runtime " << st.ms_total << "ms avgpost " << post_ms / post_layers << "ms"
            << std::endl;
  snicit_net_destroy(net);
  snicit_ctx_destroy(ctx);
  return 0;
}
