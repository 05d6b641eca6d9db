// Net ingestion: per-row ELL -> fan-in panels, resident in HBM.
//
// Replaces the reference's weight packing (CSR slabs, src/SDGC/base.hpp:89-141 and
// src/SDGC/utility.hpp:85-157) and the XY offline layout (flat value/row_access arrays,
// src/SDGC/SNICIT/SNICIT.hpp:127-183).  Rows of a layer whose sorted input sets are
// identical become one panel: the kernel gathers the panel's 32 inputs once per sample
// and produces every row of the panel from registers.  Rows with fewer than 32 inputs
// also share a panel when their input sets together fit its 32 slots.
#include <algorithm>
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

struct Panel {
  std::vector<int32_t> uni;   // sorted distinct inputs (empty: a group kept on its own key)
  std::vector<int32_t> rows;  // output rows
};

// A group's rows in chunks of at most kMaxPanelRows, each with the group's distinct inputs.
struct Unit {
  std::vector<int32_t> d;
  std::vector<int32_t> rows;
};

size_t union_size(const std::vector<int32_t>& a, const std::vector<int32_t>& b) {
  size_t n = 0, i = 0, j = 0;
  while (i < a.size() || j < b.size()) {
    if (j == b.size() || (i < a.size() && a[i] < b[j])) ++i;
    else if (i == a.size() || b[j] < a[i]) ++j;
    else ++i, ++j;
    ++n;
  }
  return n;
}

void absorb(Panel& pn, const Unit& u) {
  std::vector<int32_t> m;
  std::set_union(pn.uni.begin(), pn.uni.end(), u.d.begin(), u.d.end(), std::back_inserter(m));
  pn.uni.swap(m);
  pn.rows.insert(pn.rows.end(), u.rows.begin(), u.rows.end());
}

bool fits(const Panel& pn, const Unit& u) {
  return !u.d.empty() && pn.rows.size() + u.rows.size() <= (size_t)kMaxPanelRows &&
         union_size(pn.uni, u.d) <= (size_t)kSlots;
}

// Consecutive grouping: each unit joins the panel before it when the union of their input
// sets and their rows still fit one panel.
std::vector<Panel> group_consecutive(const std::vector<Unit>& units) {
  std::vector<Panel> panels;
  for (const Unit& u : units) {
    if (!panels.empty() && !panels.back().uni.empty() && fits(panels.back(), u)) {
      absorb(panels.back(), u);
      continue;
    }
    Panel np;
    np.uni = u.d;
    np.rows = u.rows;
    panels.push_back(std::move(np));
  }
  return panels;
}

// Search grouping, after the inspector's greedy maximum-reuse schedule
// (src/SDGC/inspector/gpu_block_scheduler.h:58-86, which merges column blocks whose row
// accesses overlap, at any distance): a panel is seeded with the first unplaced unit, then
// repeatedly takes the unplaced unit that shares the most inputs with the panel's union
// (found through an inverted index from input neuron to units; ties go to the earlier unit)
// while union and rows still fit; what room is left is filled with the next unplaced units
// in order that fit (looking up to kFillWindow units ahead), so the search never packs worse
// than the consecutive rule.  Inputs listed by more than kHotInput units do not score
// candidates (they discriminate little and would make the search quadratic).
std::vector<Panel> group_search(const std::vector<Unit>& units, int32_t N) {
  constexpr size_t kHotInput = 4096;
  constexpr size_t kFillWindow = 256;
  std::vector<std::vector<int32_t>> index(N);
  for (size_t u = 0; u < units.size(); ++u)
    for (int32_t x : units[u].d) index[x].push_back((int32_t)u);
  std::vector<uint8_t> placed(units.size(), 0);
  // score[v]: inputs of unit v found in the panel's union so far, counted incrementally over
  // the inputs each absorbed unit adds.  Without skipped hot inputs the count is exact and
  // |union of the panel and v| = |union| + |v| - score[v]; with them that is an upper bound,
  // and the exact merge (fits) decides.
  std::vector<int32_t> score(units.size(), 0);
  std::vector<int32_t> touched;
  std::vector<int32_t> fresh;
  std::vector<Panel> panels;
  size_t next = 0;  // first unit that may be unplaced
  const auto rows_fit = [](const Panel& pn, const Unit& u) {
    return pn.rows.size() + u.rows.size() <= (size_t)kMaxPanelRows;
  };
  for (size_t seed = 0; seed < units.size(); ++seed) {
    if (placed[seed]) continue;
    placed[seed] = 1;
    Panel pn;
    pn.uni = units[seed].d;
    pn.rows = units[seed].rows;
    if (!pn.uni.empty()) {
      bool hot = false;
      fresh = pn.uni;
      for (;;) {
        for (int32_t x : fresh) {
          if (index[x].size() > kHotInput) {
            hot = true;
            continue;
          }
          for (int32_t v : index[x]) {
            if (placed[v]) continue;
            if (score[v]++ == 0) touched.push_back(v);
          }
        }
        int32_t best = -1;
        for (int32_t v : touched) {
          if (placed[v] || !rows_fit(pn, units[v])) continue;
          if (best >= 0 && (score[v] < score[best] || (score[v] == score[best] && v > best)))
            continue;
          if (pn.uni.size() + units[v].d.size() - (size_t)score[v] <= (size_t)kSlots ||
              (hot && fits(pn, units[v])))
            best = v;
        }
        if (best < 0) break;
        placed[best] = 1;
        fresh.clear();
        std::set_difference(units[best].d.begin(), units[best].d.end(), pn.uni.begin(),
                            pn.uni.end(), std::back_inserter(fresh));
        absorb(pn, units[best]);
      }
      for (int32_t v : touched) score[v] = 0;
      touched.clear();
      // Units the search left share no counted input with the union (untouched) or did not fit
      // it (touched; the union only grows), so until the fill absorbs a unit, one joins only
      // when its inputs fit beside the union's.
      bool filled = false;
      while (next < units.size() && placed[next]) ++next;
      for (size_t v = next, seen = 0; v < units.size() && seen < kFillWindow; ++v) {
        if (placed[v]) continue;
        ++seen;
        const Unit& u = units[v];
        if (u.d.empty() || !rows_fit(pn, u)) continue;
        if (pn.uni.size() + u.d.size() <= (size_t)kSlots || ((hot || filled) && fits(pn, u))) {
          placed[v] = 1;
          absorb(pn, u);
          filled = true;
        }
      }
    }
    panels.push_back(std::move(pn));
  }
  return panels;
}

// Inspector-style grouping (the role of the XY offline layout search,
// src/SDGC/inspector/gpu_block_scheduler.h:58-171): groups of rows whose input sets together
// hold at most 32 distinct inputs, and at most 32 rows, share one panel (grouping 1: only
// consecutive groups; 2: the search above).  The panel's slots are the sorted union; a row's
// weight is 0 in the slots it does not use.  Each row's own entries stay in ascending input
// order, and a zero-weight step fma(0, x, acc) returns acc exactly (finite x, acc never -0),
// so every row's chain gives the bits of its own sorted chain.  Returns false (nothing
// written) when no two groups merge, which is the case for full 32-input rows.
bool merge_groups(const std::vector<std::vector<int32_t>>& keys,
                  const std::vector<std::vector<float>>& wrow,
                  std::unordered_map<std::vector<int32_t>, std::vector<int32_t>, KeyHash>& groups,
                  const std::vector<const std::vector<int32_t>*>& order, HostLayerPanels& out,
                  int32_t N, int32_t K, int32_t grouping) {
  std::vector<Unit> units;
  for (const std::vector<int32_t>* key : order) {
    const std::vector<int32_t>& rows = groups[*key];
    const std::vector<int32_t> d = distinct_inputs(*key, K);
    for (size_t b = 0; b < rows.size(); b += kMaxPanelRows) {
      Unit u;
      u.d = d;  // empty: a row with a repeated input, kept on its own key
      u.rows.assign(rows.begin() + b, rows.begin() + std::min(rows.size(), b + kMaxPanelRows));
      units.push_back(std::move(u));
    }
  }
  const std::vector<Panel> panels =
      grouping == SNICIT_GROUP_SEARCH ? group_search(units, N) : group_consecutive(units);
  if (panels.size() == units.size()) return false;
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
                    HostLayerPanels& out, int32_t grouping) {
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
  if (K < kSlots && grouping != SNICIT_GROUP_NONE &&
      merge_groups(keys, wrow, groups, order, out, N, K, grouping))
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
