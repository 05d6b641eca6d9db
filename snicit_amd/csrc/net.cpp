// Net ingestion: per-row ELL -> fan-in panels, resident in HBM.
//
// Replaces the reference's weight packing (CSR slabs, src/SDGC/base.hpp:89-141 and
// src/SDGC/utility.hpp:85-157) and the XY offline layout (flat value/row_access arrays,
// src/SDGC/SNICIT/SNICIT.hpp:127-183).  Rows of a layer whose sorted input sets are
// identical become one panel: the kernel gathers the panel's 32 inputs once per sample
// and produces every row of the panel from registers.
#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "internal.hpp"

namespace snicit {

namespace {

struct KeyHash {
  size_t operator()(const std::vector<int32_t>& k) const {
    uint64_t h = 0x84222325CBF29CE4ULL;
    for (int32_t v : k) h = sm_mix(h ^ (uint64_t)(uint32_t)v);
    return (size_t)h;
  }
};

}  // namespace

void pack_layer_ell(int32_t N, int32_t K, const int32_t* cols, const float* vals,
                    HostLayerPanels& out) {
  // Sort each row by input neuron; rows with the same input set share a panel.
  std::vector<std::vector<int32_t>> keys(N);
  std::vector<std::vector<float>> wrow(N);
  std::vector<std::pair<int32_t, float>> tmp(K);
  for (int32_t o = 0; o < N; ++o) {
    for (int32_t r = 0; r < K; ++r) {
      int32_t c = cols[(size_t)o * K + r];
      SNICIT_REQUIRE(c >= 0 && c < N, "ELL column out of range at row " + std::to_string(o));
      tmp[r] = {c, vals[(size_t)o * K + r]};
    }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) {
                       return a.first < b.first;
                     });
    keys[o].resize(kSlots);
    wrow[o].resize(kSlots);
    for (int32_t r = 0; r < kSlots; ++r) {
      // Pad short rows with value-0 slots on the last column (a zero product leaves the
      // fused multiply-add chain unchanged).
      int32_t rr = r < K ? r : K - 1;
      keys[o][r] = tmp[rr].first;
      wrow[o][r] = r < K ? tmp[rr].second : 0.f;
    }
  }
  std::unordered_map<std::vector<int32_t>, std::vector<int32_t>, KeyHash> groups;
  std::vector<const std::vector<int32_t>*> order;
  for (int32_t o = 0; o < N; ++o) {
    auto it = groups.find(keys[o]);
    if (it == groups.end()) {
      it = groups.emplace(keys[o], std::vector<int32_t>()).first;
      order.push_back(&it->first);
    }
    it->second.push_back(o);
  }
  out.pcols.clear();
  out.pbeg.clear();
  out.outs.clear();
  out.w.clear();
  out.outs.reserve(N);
  out.w.reserve((size_t)N * kSlots);
  for (const std::vector<int32_t>* key : order) {
    const std::vector<int32_t>& rows = groups[*key];
    for (size_t b = 0; b < rows.size(); b += kMaxPanelRows) {
      out.pbeg.push_back((int32_t)out.outs.size());
      out.pcols.insert(out.pcols.end(), key->begin(), key->end());
      size_t e = std::min(rows.size(), b + kMaxPanelRows);
      for (size_t i = b; i < e; ++i) {
        out.outs.push_back(rows[i]);
        out.w.insert(out.w.end(), wrow[rows[i]].begin(), wrow[rows[i]].end());
      }
    }
  }
  out.pbeg.push_back((int32_t)out.outs.size());
}

}  // namespace snicit
