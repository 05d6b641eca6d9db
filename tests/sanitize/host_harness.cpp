// Runs the library's host-only code (the seeded generators of gen.cpp and the panel packer
// of net.cpp) on small shapes, for AddressSanitizer and UndefinedBehaviorSanitizer builds.
// Exits 0 when every structural check holds.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "../../snicit_amd/csrc/internal.hpp"

using namespace snicit;

static int failures = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

// Every row of the packed layer lists, in its panel's slots, exactly its own inputs with its
// own weights (other slots weight 0), and every output row appears once.
static void check_packed(int32_t N, int32_t K, const std::vector<int32_t>& cols,
                         const std::vector<float>& vals, const HostLayerPanels& hp) {
  const int64_t P = (int64_t)hp.pbeg.size() - 1;
  CHECK(P >= 1 && hp.pbeg[0] == 0 && hp.pbeg[P] == N);
  CHECK((int64_t)hp.pcols.size() == P * kSlots && (int64_t)hp.outs.size() == N);
  std::vector<int> seen(N, 0);
  for (int64_t p = 0; p < P; ++p) {
    for (int32_t j = hp.pbeg[p]; j < hp.pbeg[p + 1]; ++j) {
      const int32_t o = hp.outs[j];
      CHECK(o >= 0 && o < N);
      if (o < 0 || o >= N) continue;
      ++seen[o];
      double want = 0.0, got = 0.0;
      for (int32_t r = 0; r < K; ++r) want += vals[(size_t)o * K + r] * (1.0 + cols[(size_t)o * K + r]);
      for (int r = 0; r < kSlots; ++r)
        got += hp.w[(size_t)j * kSlots + r] * (1.0 + hp.pcols[(size_t)p * kSlots + r]);
      CHECK(std::abs(want - got) <= 1e-3 * (1.0 + std::abs(want)));
    }
  }
  for (int32_t o = 0; o < N; ++o) CHECK(seen[o] == 1);
}

int main() {
  // SDGC-like generator: panels and per-row lists agree
  for (int32_t N : {1024, 4096}) {
    for (int32_t l : {0, 1, 7, 13}) {
      std::vector<int32_t> cols((size_t)N * 32);
      sdgc_layer_ell(N, l, 2023, cols.data());
      HostLayerPanels hp;
      sdgc_layer_panels(N, l, 2023, hp);
      std::vector<float> vals(cols.size(), 0.0625f);
      for (int32_t o = 0; o < N; ++o)
        for (int r = 0; r < 32; ++r) CHECK(cols[(size_t)o * 32 + r] >= 0 && cols[(size_t)o * 32 + r] < N);
      for (int32_t grouping : {SNICIT_GROUP_NONE, SNICIT_GROUP_CONSECUTIVE, SNICIT_GROUP_SEARCH}) {
        HostLayerPanels packed;
        pack_layer_ell(N, 32, cols.data(), vals.data(), packed, grouping);
        check_packed(N, 32, cols, vals, packed);
      }
    }
    // images
    ImageGeom g = image_geometry(N, 60000);
    std::vector<float> row(N);
    for (int64_t s = 0; s < 64; ++s) {
      SampleParams sp = image_sample_params(g, 60000, s);
      image_render(g, sp, N, row.data());
      for (float v : row) CHECK(v == 0.f || v == 1.f);
    }
  }
  // unstructured rows, dense layers, features
  {
    const int32_t N = 512, K = 48;
    std::vector<int32_t> cols((size_t)N * K);
    std::vector<float> vals((size_t)N * K);
    unstructured_layer(N, K, 3, 7, cols.data(), vals.data());
    for (int32_t o = 0; o < N; ++o) {
      std::set<int32_t> s(cols.begin() + (size_t)o * K, cols.begin() + (size_t)(o + 1) * K);
      CHECK((int32_t)s.size() == K);
    }
    std::vector<float> w((size_t)64 * 100);
    dense_layer(64, 100, 7, 1, w.data());
    std::vector<float> f((size_t)9 * 300);
    uniform_features(300, 5, 11, 9, f.data());
    for (float v : f) CHECK(v >= 0.f && v < 1.f);
  }
  // the panel packer on rows that share inputs, short rows (merged panels), repeated inputs
  // and equal weight rows
  for (int32_t K : {32, 20, 8, 3, 1}) {
    const int32_t N = 256;
    std::vector<int32_t> cols((size_t)N * K);
    std::vector<float> vals((size_t)N * K);
    Rng r{(uint64_t)K * 977u};
    for (int32_t o = 0; o < N; ++o) {
      const int32_t g = o / 4;  // four rows share an input set
      Rng rg{(uint64_t)g * 131u + (uint64_t)K};
      for (int32_t q = 0; q < K; ++q) {
        cols[(size_t)o * K + q] = (int32_t)(rg.next() % N);
        vals[(size_t)o * K + q] = (o % 3 == 0) ? 0.5f : (float)((r.next() % 1000) / 250.0 - 2.0);
      }
    }
    for (int32_t grouping : {SNICIT_GROUP_NONE, SNICIT_GROUP_CONSECUTIVE, SNICIT_GROUP_SEARCH}) {
      HostLayerPanels packed;
      pack_layer_ell(N, K, cols.data(), vals.data(), packed, grouping);
      check_packed(N, K, cols, vals, packed);
    }
  }
  // grouping properties on random short-row layers: every grouping keeps each row's entries
  // (check_packed), and the search never packs into more panels than the consecutive rule,
  // which never packs into more than one input set per panel
  for (int trial = 0; trial < 24; ++trial) {
    Rng r{(uint64_t)trial * 7919u + 17u};
    const int32_t N = 256 + (int32_t)(r.next() % 768);
    const int32_t K = 1 + (int32_t)(r.next() % 31);
    const int32_t span = 1 + (int32_t)(r.next() % 64);  // inputs drawn near a row-specific base
    std::vector<int32_t> cols((size_t)N * K);
    std::vector<float> vals((size_t)N * K);
    for (int32_t o = 0; o < N; ++o) {
      const int32_t base = (int32_t)(r.next() % N);
      for (int32_t q = 0; q < K; ++q) {
        cols[(size_t)o * K + q] = (base + (int32_t)(r.next() % (uint64_t)(span + K))) % N;
        vals[(size_t)o * K + q] = (float)((int32_t)(r.next() % 2001) - 1000) / 512.0f;
      }
    }
    size_t panels[3];
    for (int32_t grouping : {SNICIT_GROUP_NONE, SNICIT_GROUP_CONSECUTIVE, SNICIT_GROUP_SEARCH}) {
      HostLayerPanels packed;
      pack_layer_ell(N, K, cols.data(), vals.data(), packed, grouping);
      check_packed(N, K, cols, vals, packed);
      panels[grouping] = packed.pbeg.size() - 1;
    }
    CHECK(panels[SNICIT_GROUP_SEARCH] <= panels[SNICIT_GROUP_CONSECUTIVE]);
    CHECK(panels[SNICIT_GROUP_CONSECUTIVE] <= panels[SNICIT_GROUP_NONE]);
  }
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host harness ok\n");
  return 0;
}
