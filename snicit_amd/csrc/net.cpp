// Net ingestion: per-row ELL -> fan-in panels, resident in HBM.
//
// Replaces the reference's weight packing (CSR slabs, src/SDGC/base.hpp:89-141 and
// src/SDGC/utility.hpp:85-157) and the XY offline layout (flat value/row_access arrays,
// src/SDGC/SNICIT/SNICIT.hpp:127-183).  Rows of a layer whose sorted input sets are
// identical become one panel: the kernel gathers the panel's 32 inputs once per sample
// and produces every row of the panel from registers.  Rows with fewer than 32 inputs
// also share a panel when their input sets together fit its 32 slots.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <iterator>
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

// Distinct input neurons of a padded key (short rows repeat their last column), or an empty
// vector when a row lists one input twice: such groups are never merged.
std::vector<int32_t> distinct_inputs(const std::vector<int32_t>& key, int32_t K) {
  std::vector<int32_t> d(key.begin(), key.begin() + std::min<int32_t>(K, kSlots));
  for (size_t i = 1; i < d.size(); ++i)
    if (d[i] == d[i - 1]) return {};
  return d;
}

// SNICIT_PANEL_MERGE=0 keeps one input set per panel (the layout without merging).
bool merge_short_rows_enabled() {
  const char* e = std::getenv("SNICIT_PANEL_MERGE");
  return !(e && std::strcmp(e, "0") == 0);
}

// Inspector-style grouping (the role of the XY offline layout search,
// src/SDGC/inspector/gpu_block_scheduler.h:58-171): consecutive groups of rows whose input
// sets together hold at most 32 distinct inputs, and at most 32 rows, share one panel.  The
// panel's slots are the sorted union; a row's weight is 0 in the slots it does not use.
// Each row's own entries stay in ascending input order, and a zero-weight step
// fma(0, x, acc) returns acc exactly (finite x, acc never -0), so every row's chain gives
// the bits of its own sorted chain.  Returns false (nothing written) when no two groups
// merge, which is the case for full 32-input rows.
bool merge_groups(const std::vector<std::vector<int32_t>>& keys,
                  const std::vector<std::vector<float>>& wrow,
                  std::unordered_map<std::vector<int32_t>, std::vector<int32_t>, KeyHash>& groups,
                  const std::vector<const std::vector<int32_t>*>& order, HostLayerPanels& out,
                  int32_t K) {
  struct Panel {
    std::vector<int32_t> uni;   // sorted distinct inputs
    std::vector<int32_t> rows;  // output rows
  };
  std::vector<Panel> panels;
  bool merged = false;
  for (const std::vector<int32_t>* key : order) {
    const std::vector<int32_t>& rows = groups[*key];
    const std::vector<int32_t> d = distinct_inputs(*key, K);
    for (size_t b = 0; b < rows.size(); b += kMaxPanelRows) {
      const size_t e = std::min(rows.size(), b + kMaxPanelRows);
      if (!d.empty() && !panels.empty() && !panels.back().uni.empty()) {
        Panel& cur = panels.back();
        std::vector<int32_t> u;
        std::set_union(cur.uni.begin(), cur.uni.end(), d.begin(), d.end(), std::back_inserter(u));
        if (u.size() <= (size_t)kSlots && cur.rows.size() + (e - b) <= (size_t)kMaxPanelRows) {
          cur.uni.swap(u);
          cur.rows.insert(cur.rows.end(), rows.begin() + b, rows.begin() + e);
          merged = true;
          continue;
        }
      }
      Panel np;
      np.uni = d;  // empty: a row with a repeated input, kept on its own key
      np.rows.assign(rows.begin() + b, rows.begin() + e);
      panels.push_back(std::move(np));
    }
  }
  if (!merged) return false;
  for (const Panel& pn : panels) {
    out.pbeg.push_back((int32_t)out.outs.size());
    if (pn.uni.empty()) {  // unmerged group: its own padded key and weights
      const std::vector<int32_t>& key = keys[pn.rows[0]];
      out.pcols.insert(out.pcols.end(), key.begin(), key.end());
      for (int32_t o : pn.rows) {
        out.outs.push_back(o);
        out.w.insert(out.w.end(), wrow[o].begin(), wrow[o].end());
      }
      continue;
    }
    std::vector<int32_t> slots(pn.uni);
    slots.resize(kSlots, pn.uni.back());  // pad with the last input, weight 0
    out.pcols.insert(out.pcols.end(), slots.begin(), slots.end());
    for (int32_t o : pn.rows) {
      out.outs.push_back(o);
      std::vector<float> wv(kSlots, 0.f);
      // the row's entries (keys[o][0..K), distinct and sorted) land on their union slots
      size_t u = 0;
      for (int32_t r = 0; r < std::min<int32_t>(K, kSlots); ++r) {
        while (pn.uni[u] != keys[o][r]) ++u;
        wv[u] = wrow[o][r];
      }
      out.w.insert(out.w.end(), wv.begin(), wv.end());
    }
  }
  out.pbeg.push_back((int32_t)out.outs.size());
  return true;
}

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
  if (K < kSlots && merge_short_rows_enabled() && merge_groups(keys, wrow, groups, order, out, K))
    return;
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
