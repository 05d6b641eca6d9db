// C-ABI of libsnicit: contexts, nets, plans and the layer loop.
//
// The layer loop follows SNICIT::_infer (src/SDGC/SNICIT/SNICIT.hpp:205-692) and
// BF::_infer (src/SDGC/BF/BF.hpp:157-220) in what it computes, with every piece of
// bookkeeping on the device: live-row lists, the centroid pool, assignment and the
// non-empty delta rows are built by kernels, counts stay in device memory, and each
// stage (pre / cluster / post / recover) can be replayed as one captured hipGraph
// because no launch size depends on a host read-back.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>

#include "plan.hpp"

using namespace snicit;


namespace {

// Upload one layer's host panels into the net's concatenated arrays.  Row-major weights
// serve every panel kernel; SNICIT_KERNEL=mfma stores them slot-major for panel_mfma_kernel.
int32_t max_rows(const HostLayerPanels& hp) {
  int32_t m = 0;
  for (size_t p = 0; p + 1 < hp.pbeg.size(); ++p) m = std::max(m, hp.pbeg[p + 1] - hp.pbeg[p]);
  return m;
}

// SNICIT_KERNEL=auto|valu|mfma selects the panel kernel family when a net is built (all
// three give the same bits).  auto (0): plain-epilogue layers on the matrix cores with the
// weights staged in LDS (panel_plain_mfma_kernel), centroid and delta epilogues on the
// vector ALU.  valu (1): the vector-ALU kernel for every epilogue.  mfma (2): slot-major
// weights held in registers (panel_mfma_kernel, ~170 VGPRs, measured slower).
int kernel_choice_from_env() {
  const char* e = std::getenv("SNICIT_KERNEL");
  if (!e || !*e || std::strcmp(e, "auto") == 0) return 0;
  if (std::strcmp(e, "valu") == 0) return 1;
  if (std::strcmp(e, "mfma") == 0) return 2;
  throw Error(SNICIT_E_INVALID, std::string("SNICIT_KERNEL must be auto, valu or mfma, got ") + e);
}

// The layer with each panel's rows reordered so that rows with equal weights (bit for bit)
// are adjacent, first occurrences in their original order; returns `out`.
const HostLayerPanels& group_equal_rows(const HostLayerPanels& hp, HostLayerPanels& out) {
  out.pcols = hp.pcols;
  out.pbeg = hp.pbeg;
  out.outs.resize(hp.outs.size());
  out.w.resize(hp.w.size());
  const int64_t P = (int64_t)hp.pbeg.size() - 1;
  std::vector<int32_t> order;
  for (int64_t p = 0; p < P; ++p) {
    const int32_t b = hp.pbeg[p], G = hp.pbeg[p + 1] - b;
    order.clear();
    std::vector<bool> used(G, false);
    for (int32_t g = 0; g < G; ++g) {
      if (used[g]) continue;
      for (int32_t h = g; h < G; ++h)
        if (!used[h] && std::memcmp(&hp.w[(size_t)(b + h) * kSlots],
                                    &hp.w[(size_t)(b + g) * kSlots], kSlots * 4) == 0) {
          used[h] = true;
          order.push_back(h);
        }
    }
    for (int32_t i = 0; i < G; ++i) {
      out.outs[b + i] = hp.outs[b + order[i]];
      std::memcpy(&out.w[(size_t)(b + i) * kSlots], &hp.w[(size_t)(b + order[i]) * kSlots],
                  kSlots * 4);
    }
  }
  return out;
}

struct NetBuilder {
  Net* net;
  int kernel_choice = kernel_choice_from_env();
  std::vector<int32_t> pcols, pbeg, outs;
  std::vector<float> w;
  std::vector<uint32_t> psame;
  bool weights_given = true;
  int64_t w_floats = 0;
  // wide-ELL layers
  std::vector<int32_t> ecols;
  std::vector<float> evals;
  std::vector<int32_t> erng;
  std::vector<int2> eent;
  void add_ell(const int32_t* cols, const float* vals, int32_t K) {
    const int32_t Kp = (K + 15) / 16 * 16;
    const int32_t N = net->N;
    net->panel_off.push_back(net->panel_off.back());
    pbeg.push_back(0);
    net->mfma.push_back(0);
    net->dup_rows.push_back(0);
    net->few_rows.push_back(0);
    net->w_off.push_back(w_floats);
    net->ell_K.push_back(Kp);
    net->ell_off.push_back((int64_t)ecols.size());
    net->erng_off.push_back((int64_t)erng.size());
    net->eent_off.push_back((int64_t)eent.size());
    const int64_t stream0 = (int64_t)eent.size();
    const int32_t nch = ell_chunks(N);
    for (int32_t o = 0; o < N; ++o) outs.push_back(o);  // keeps the [L][N] row-id layout
    std::vector<std::pair<int32_t, float>> tmp(K);
    for (int32_t o = 0; o < N; ++o) {
      for (int32_t r = 0; r < K; ++r) {
        const int32_t c = cols[(size_t)o * K + r];
        SNICIT_REQUIRE(c >= 0 && c < N, "ELL column out of range at row " + std::to_string(o));
        tmp[r] = {c, vals[(size_t)o * K + r]};
      }
      // ascending input neuron, stable for ties (the oracle's order)
      std::stable_sort(tmp.begin(), tmp.end(),
                       [](const std::pair<int32_t, float>& x, const std::pair<int32_t, float>& y) {
                         return x.first < y.first;
                       });
      for (int32_t r = 0; r < Kp; ++r) {
        ecols.push_back(r < K ? tmp[r].first : tmp[K - 1].first);
        evals.push_back(r < K ? tmp[r].second : 0.f);
      }
      // the row without its trailing value-0 entries (padding, here or from a CSR row
      // shorter than the longest), chunk by chunk, each chunk's run padded to whole groups
      int32_t end = K;
      while (end > 0 && tmp[end - 1].second == 0.f) --end;
      int32_t q = 0;
      for (int32_t c = 0; c < nch; ++c) {
        const int64_t at = (int64_t)eent.size() - stream0;
        SNICIT_REQUIRE(at < (int64_t(1) << 31), "wide layer too large for its entry stream");
        erng.push_back((int32_t)at);
        int32_t last = -1;
        int32_t n = 0;
        for (; q < end && tmp[q].first < (c + 1) * kEllChunk; ++q, ++n) {
          int2 e;
          e.x = tmp[q].first - c * kEllChunk;
          std::memcpy(&e.y, &tmp[q].second, 4);
          eent.push_back(e);
          last = e.x;
        }
        for (; n % kEllGroup != 0; ++n) eent.push_back(int2{last, 0});
      }
      erng.push_back((int32_t)((int64_t)eent.size() - stream0));
    }
  }
  void add(const HostLayerPanels& hp_in) {
    const int64_t P = (int64_t)hp_in.pbeg.size() - 1;
    net->ell_K.push_back(0);
    net->ell_off.push_back((int64_t)ecols.size());
    net->erng_off.push_back((int64_t)erng.size());
    net->eent_off.push_back((int64_t)eent.size());
    net->panel_off.push_back(net->panel_off.back() + P);
    HostLayerPanels grouped;
    const HostLayerPanels& hp = weights_given ? group_equal_rows(hp_in, grouped) : hp_in;
    bool all_dup = P > 0, all_few = P > 0;
    for (int64_t p = 0; p < P; ++p) {
      const int32_t b = hp.pbeg[p], G = hp.pbeg[p + 1] - b;
      uint32_t m = 0;
      for (int32_t g = 1; g < G && g < 32; ++g) {
        // generated nets hold one weight value everywhere
        const bool same = !weights_given ||
                          std::memcmp(&hp.w[(size_t)(b + g) * kSlots],
                                      &hp.w[(size_t)(b + g - 1) * kSlots], kSlots * 4) == 0;
        if (same) m |= 1u << g;
      }
      psame.push_back(m);
      const uint32_t full = G >= 32 ? 0xFFFFFFFEu : ((1u << G) - 1u) & ~1u;
      all_dup = all_dup && m == full;
      all_few = all_few && G <= 32 && __builtin_popcount(full & ~m) <= 1;  // rows 1.. that differ
    }
    net->dup_rows.push_back(all_dup ? 1 : 0);
    net->few_rows.push_back(all_few ? 1 : 0);
    pcols.insert(pcols.end(), hp.pcols.begin(), hp.pcols.end());
    pbeg.insert(pbeg.end(), hp.pbeg.begin(), hp.pbeg.end());
    outs.insert(outs.end(), hp.outs.begin(), hp.outs.end());
    const bool mfma = kernel_choice == 2 && P > 0 && max_rows(hp) <= 32;
    net->mfma.push_back(mfma ? 1 : 0);
    net->w_off.push_back(w_floats);
    const int64_t nf = mfma ? P * kSlots * 32 : (int64_t)net->N * kSlots;
    w_floats += nf;
    if (!weights_given) return;
    if (!mfma) {
      w.insert(w.end(), hp.w.begin(), hp.w.end());
      return;
    }
    const size_t at = w.size();
    w.resize(at + (size_t)nf, 0.f);
    for (int64_t p = 0; p < P; ++p)
      for (int32_t g = 0; g < hp.pbeg[p + 1] - hp.pbeg[p]; ++g)
        for (int r = 0; r < kSlots; ++r)
          w[at + (size_t)p * kSlots * 32 + (size_t)r * 32 + g] =
              hp.w[(size_t)(hp.pbeg[p] + g) * kSlots + r];
  }
};

void upload_net(Net* net, NetBuilder& nb, float fill_weight) {
  hipStream_t s = net->ctx->stream;
  net->plain_mfma = nb.kernel_choice == 0;
  const int64_t Ptot = net->panel_off.back();
  net->d_pcols = dev_alloc<int32_t>((size_t)Ptot * kSlots, &net->device_bytes);
  net->d_pbeg = dev_alloc<int32_t>((size_t)Ptot + net->L, &net->device_bytes);
  net->d_outs = dev_alloc<int32_t>((size_t)net->L * net->N, &net->device_bytes);
  net->d_w = dev_alloc<float>((size_t)nb.w_floats, &net->device_bytes);
  net->d_psame = dev_alloc<uint32_t>(std::max<size_t>(1, nb.psame.size()), &net->device_bytes);
  if (!nb.psame.empty())
    SNICIT_HIP(hipMemcpy(net->d_psame, nb.psame.data(), nb.psame.size() * 4,
                         hipMemcpyHostToDevice));
  SNICIT_HIP(hipMemcpy(net->d_pcols, nb.pcols.data(), nb.pcols.size() * 4, hipMemcpyHostToDevice));
  SNICIT_HIP(hipMemcpy(net->d_pbeg, nb.pbeg.data(), nb.pbeg.size() * 4, hipMemcpyHostToDevice));
  SNICIT_HIP(hipMemcpy(net->d_outs, nb.outs.data(), nb.outs.size() * 4, hipMemcpyHostToDevice));
  if (nb.weights_given) {
    SNICIT_HIP(hipMemcpy(net->d_w, nb.w.data(), nb.w.size() * 4, hipMemcpyHostToDevice));
  } else {
    // generated nets: every panel has 32 rows, so both layouts hold only `fill_weight`
    launch_fill_f32(net->d_w, fill_weight, nb.w_floats, s);
  }
  if (!nb.ecols.empty()) {
    net->d_ecols = dev_alloc<int32_t>(nb.ecols.size(), &net->device_bytes);
    net->d_evals = dev_alloc<float>(nb.evals.size(), &net->device_bytes);
    SNICIT_HIP(hipMemcpy(net->d_ecols, nb.ecols.data(), nb.ecols.size() * 4, hipMemcpyHostToDevice));
    SNICIT_HIP(hipMemcpy(net->d_evals, nb.evals.data(), nb.evals.size() * 4, hipMemcpyHostToDevice));
    // one group of slack: the kernel loads the group after a row's last one
    net->d_eent = dev_alloc<int2>(nb.eent.size() + kEllGroup, &net->device_bytes);
    SNICIT_HIP(hipMemcpy(net->d_eent, nb.eent.data(), nb.eent.size() * 8, hipMemcpyHostToDevice));
    SNICIT_HIP(hipMemset(net->d_eent + nb.eent.size(), 0, kEllGroup * 8));
    net->d_erng = dev_alloc<int32_t>(nb.erng.size(), &net->device_bytes);
    SNICIT_HIP(hipMemcpy(net->d_erng, nb.erng.data(), nb.erng.size() * 4, hipMemcpyHostToDevice));
  }
  if (net->bias_kind == SNICIT_BIAS_PER_NEURON) {
    net->d_bias = dev_alloc<float>((size_t)net->L * net->N, &net->device_bytes);
    SNICIT_HIP(hipMemcpy(net->d_bias, net->bias_host.data(), net->bias_host.size() * 4,
                         hipMemcpyHostToDevice));
  }
  SNICIT_HIP(hipStreamSynchronize(s));
}

void net_set_bias(Net* net, int32_t bias_kind, const float* bias) {
  SNICIT_REQUIRE(bias != nullptr, "bias must not be null");
  net->bias_kind = bias_kind;
  size_t nb = bias_kind == SNICIT_BIAS_SCALAR      ? 1
              : bias_kind == SNICIT_BIAS_PER_LAYER ? (size_t)net->L
              : bias_kind == SNICIT_BIAS_PER_NEURON ? (size_t)net->L * net->N
                                                     : 0;
  SNICIT_REQUIRE(nb > 0, "bad bias_kind");
  net->bias_host.assign(bias, bias + nb);
  net->skip_dead = true;
  for (float b : net->bias_host)
    if (!(b <= 0.f)) net->skip_dead = false;
}

void destroy_net(Net* net) {
  if (net->h_blocks) (void)hipHostFree(net->h_blocks);
  net->h_blocks = nullptr;
  dev_free(net->d_ecols);
  dev_free(net->d_evals);
  dev_free(net->d_erng);
  dev_free(net->d_eent);
  dev_free(net->d_win);
  dev_free(net->d_wint);
  dev_free(net->d_bin);
  dev_free(net->d_wout);
  dev_free(net->d_woutt);
  dev_free(net->d_bout);
  dev_free(net->d_pcols);
  dev_free(net->d_pbeg);
  dev_free(net->d_outs);
  dev_free(net->d_w);
  dev_free(net->d_psame);
  dev_free(net->d_bias);
}

}  // namespace

// ---- plan ------------------------------------------------------------------------------

namespace snicit {

thread_local std::string g_last_error;
thread_local DevArena* g_dev_arena = nullptr;

// ---- static scratch of the allocation-free entry point -------------------------------------
// snicit_infer_noalloc carves every device buffer of its run from this array instead of
// calling hipMalloc.  Each GPU the library runs on holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in, not device memory: the device
// address is resolved once per device with hipGetSymbolAddress (at context creation, with
// that device current) and only that address is used.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) char scratch[kScratchBytes];

struct ScratchSlot {
  char* ptr = nullptr;
  std::mutex run;  // one allocation-free run at a time per device within this process
};
std::mutex g_scratch_mu;
std::map<int, std::unique_ptr<ScratchSlot>> g_scratch;

// Resolves the device address of `scratch` on the current device (once per device).
ScratchSlot* scratch_slot(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto& slot = g_scratch[device];
  if (!slot) {
    auto s = std::make_unique<ScratchSlot>();
    void* ptr = nullptr;
    SNICIT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SNICIT_REQUIRE(ptr != nullptr, "no device address for the static scratch");
    s->ptr = static_cast<char*>(ptr);
    slot = std::move(s);
  }
  return slot.get();
}

bool plain_mode(int32_t mode) { return mode != SNICIT_MODE_SNICIT; }
bool compacting(const snicit_plan* p) {
  return p->net->skip_dead && p->cfg.mode != SNICIT_MODE_PLAIN_DENSE && !p->net->has_io();
}

// Whether a plan's post-threshold layers use the compact layout: exact clustering in one
// batch segment, with every post layer run by the fused panel kernel.
bool compact_post_layout(const snicit_plan* p) {
  if (p->cfg.mode != SNICIT_MODE_SNICIT || p->cfg.cluster != SNICIT_CLUSTER_EXACT) return false;
  if (p->kcap != kMaxCentroidsExact) return false;
  for (int l = p->cfg.threshold; l < p->net->L; ++l)
    if (!post_layer_fused(*p->net, l) || p->net->is_ell(l)) return false;
  return true;
}

// Compact post layout: the delta list is compacted after post layer j (0-based, of n_post):
// after each of the first four, then at doubling gaps up to every kPostRing layers, and
// after the last.  In between, rows that became empty stay in the list as zero rows (their
// deltas stay exactly zero), and every layer reads and writes the same positions.
bool post_compacts_after(int j, int n_post) {
  if (j == n_post - 1 || j < 4) return true;
  if (((j + 1) & j) == 0) return true;  // j + 1 is a power of two
  return (j + 1) % kPostRing == 0;
}

int post_compactions(int n_post) {
  int n = 0;
  for (int j = 0; j < n_post; ++j) n += post_compacts_after(j, n_post) ? 1 : 0;
  return n;
}

// Input of hidden layer 0: the input-layer output when the net has one.
const float* hidden_in(const snicit_plan* p) { return p->net->Nin > 0 ? p->d_h0 : p->d_in; }

LayerArrays layer_arrays(const Net& net, int l, const char* slot) {
  LayerArrays a;
  if (slot) {
    const BlockLayout& b = net.blk[l];
    auto at = [&](int64_t off) { return off < 0 ? nullptr : slot + off; };
    a.pcols = reinterpret_cast<const int32_t*>(at(b.pcols));
    a.pbeg = reinterpret_cast<const int32_t*>(at(b.pbeg));
    a.outs = reinterpret_cast<const int32_t*>(at(b.outs));
    a.w = reinterpret_cast<const float*>(at(b.w));
    a.psame = reinterpret_cast<const uint32_t*>(at(b.psame));
    a.ecols = reinterpret_cast<const int32_t*>(at(b.ecols));
    a.evals = reinterpret_cast<const float*>(at(b.evals));
    a.erng = reinterpret_cast<const int32_t*>(at(b.erng));
    a.eent = reinterpret_cast<const int2*>(at(b.eent));
    return a;
  }
  if (net.is_ell(l)) {
    a.ecols = net.d_ecols + net.ell_off[l];
    a.evals = net.d_evals + net.ell_off[l];
    a.erng = net.d_erng + net.erng_off[l];
    a.eent = net.d_eent + net.eent_off[l];
  } else {
    a.pcols = net.d_pcols + (size_t)net.panel_off[l] * kSlots;
    a.pbeg = net.d_pbeg + net.panel_off[l] + l;
    a.outs = net.d_outs + (size_t)l * net.N;
    a.w = net.d_w + net.w_off[l];
    a.psame = net.d_psame + net.panel_off[l];
  }
  return a;
}

// Layer l's arrays for this plan: resident, or its block in slot l mod stream_buffers.
LayerArrays plan_layer_arrays(const snicit_plan* p, int l) {
  const Net& net = *p->net;
  return layer_arrays(net, l, net.streamed() ? p->d_wslot[l % net.stream_buffers] : nullptr);
}

// Streamed nets, within one stage on the plan's stream.  weights_begin starts the stage's
// bookkeeping.  fetch_layer copies layer l's block into slot l mod slots on the copy stream
// (forking it off the plan's stream at the stage's first copy), after the layer that last
// held the slot in this stage is done with it.  use_layer fetches layer l and the next
// slots-1 layers of the stage, then makes the plan's stream wait for layer l's copy;
// done_layer frees the slot after the layer's kernels.  Every block fetched in a stage is
// used in that stage, so the copy stream joins the plan's stream again before the stage
// ends (graph capture needs that).
void weights_begin(snicit_plan* p) {
  if (!p->net->streamed()) return;
  std::fill(p->slot_layer.begin(), p->slot_layer.end(), -1);
  p->copy_forked = false;
}

void fetch_layer(snicit_plan* p, int l) {
  const Net& net = *p->net;
  const int k = l % net.stream_buffers;
  if (p->slot_layer[k] == l) return;
  if (!p->copy_forked) {
    SNICIT_HIP(hipEventRecord(p->ev_fork, p->stream));
    SNICIT_HIP(hipStreamWaitEvent(p->copy_stream, p->ev_fork, 0));
    p->copy_forked = true;
  }
  if (p->slot_layer[k] >= 0) SNICIT_HIP(hipStreamWaitEvent(p->copy_stream, p->ev_free[k], 0));
  SNICIT_HIP(hipMemcpyAsync(p->d_wslot[k], net.h_blocks + net.blk_off[l], (size_t)net.blk_bytes[l],
                            hipMemcpyHostToDevice, p->copy_stream));
  SNICIT_HIP(hipEventRecord(p->ev_copied[k], p->copy_stream));
  p->slot_layer[k] = l;
}

void use_layer(snicit_plan* p, int l, int l_end) {
  const Net& net = *p->net;
  if (!net.streamed()) return;
  for (int j = l; j < std::min(l + net.stream_buffers, l_end); ++j) fetch_layer(p, j);
  SNICIT_HIP(hipStreamWaitEvent(p->stream, p->ev_copied[l % net.stream_buffers], 0));
}

void done_layer(snicit_plan* p, int l) {
  if (!p->net->streamed()) return;
  SNICIT_HIP(hipEventRecord(p->ev_free[l % p->net->stream_buffers], p->stream));
}

Segs plan_segs(snicit_plan* p) {
  return Segs{p->nseg, p->stride, p->d_seglo, p->d_kseg, p->d_k};
}

// One hidden layer with the plan's conventions; panel or wide-ELL kernel by layer kind.
void launch_layer(snicit_plan* p, const SpmmLaunch& a, hipStream_t s) {
  const Net& net = *p->net;
  if (!net.is_ell(a.layer)) {
    launch_panel_spmm(a, s);
    return;
  }
  EllLaunch e{};
  e.cols = a.la.ecols;
  e.vals = a.la.evals;
  e.K = net.ell_K[a.layer];
  e.rng = a.la.erng;
  e.ent = a.la.eent;
  e.rows = net.N;
  e.in_rows = net.N;
  e.bias_vec = net.bias_vec(a.layer);
  e.bias_s = net.bias_scalar(a.layer);
  e.clip = net.clip;
  e.epilogue = a.epilogue;
  e.yin = a.yin;
  e.ldin = a.ldin;
  e.yout = a.yout;
  e.ldout = a.ldout;
  e.count = a.count;
  e.map = a.map;
  e.flags = a.flags;
  e.zc = a.zc;
  e.kcap = a.kcap;
  e.mcent = a.mcent;
  e.eps_post = a.eps_post;
  e.max_cols = a.max_cols;
  e.shape = a.shape;
  launch_ell_spmm(e, s);
}

// Dense input layer: d_in (Nin rows) -> d_h0, every column (src/beyond/general.hpp:5-33).
void enqueue_input_layer(snicit_plan* p) {
  const Net& net = *p->net;
  if (net.Nin == 0) return;
  EllLaunch e{};
  e.cols = nullptr;
  e.vals = net.d_win;
  e.wt = net.d_wint;
  e.wt_ld = net.win_ld;
  e.K = net.Kin;
  e.rows = net.N;
  e.in_rows = net.Kin;
  e.bias_vec = net.d_bin;
  e.clip = net.in_clip;
  e.epilogue = EPI_PLAIN;
  e.yin = p->d_in;
  e.ldin = p->ld;
  e.yout = p->d_h0;
  e.ldout = p->ld;
  e.count = p->d_counts;
  e.max_cols = p->cap;
  e.shape = p->shape;
  launch_ell_spmm(e, p->stream);
}

// Dense output layer on the hidden activations (all columns, sample order) and argmax
// categories (src/beyond/general.hpp:74-128).
void enqueue_output_layer(snicit_plan* p, const float* y) {
  const Net& net = *p->net;
  hipStream_t s = p->stream;
  EllLaunch e{};
  e.cols = nullptr;
  e.vals = net.d_wout;
  e.wt = net.d_woutt;
  e.wt_ld = net.wout_ld;
  e.K = net.Kout;
  e.rows = net.Nout;
  e.in_rows = net.Kout;
  e.bias_vec = net.d_bout;
  e.clip = 0.f;
  e.epilogue = EPI_LINEAR;
  e.yin = y;
  e.ldin = p->ld;
  e.yout = p->d_logits;
  e.ldout = p->ld;
  e.count = p->d_counts;
  e.max_cols = p->cap;
  e.shape = p->shape;
  launch_ell_spmm(e, s);
  launch_argmax(p->d_logits, p->ld, net.Nout, p->d_counts, p->cap, p->d_cat, s);
}

void prof_mark(snicit_plan* p, int l, int which) {
  if (!p->profiling) return;
  SNICIT_HIP(hipEventRecord(p->prof_ev[(size_t)4 * l + which], p->stream));
}

SpmmLaunch spmm_base(snicit_plan* p, int l, int epi) {
  SpmmLaunch a{};
  a.net = p->net;
  a.layer = l;
  a.la = plan_layer_arrays(p, l);
  a.epilogue = epi;
  a.kcap = p->kcap;
  a.eps_post = tol_cut(p->cfg.lossy_eps_post);
  a.max_cols = p->cap;
  a.shape = p->shape;
  return a;
}

// Stage "pre": layers [0, l_end) of the plain loop.  With compaction every layer reads its
// live columns through d_map and writes them densely, then a compaction builds the next
// list (BF's rowsY, src/SDGC/BF/BF.hpp:183-200).  Leaves the live list for layer l_end in
// d_map / d_counts[l_end] / d_orig[return value].
int enqueue_plain_layers(snicit_plan* p, int l_end, bool final_flags) {
  hipStream_t s = p->stream;
  const bool cmp = compacting(p);
  launch_iota_i32(p->d_orig[0], p->cap, s);
  enqueue_input_layer(p);
  weights_begin(p);
  int cur = 0;
  for (int l = 0; l < l_end; ++l) {
    use_layer(p, l, l_end);
    SpmmLaunch a = spmm_base(p, l, EPI_PLAIN);
    a.yin = l == 0 ? hidden_in(p) : p->d_y[(l - 1) & 1];
    a.ldin = p->ld;
    a.yout = p->d_y[l & 1];
    a.ldout = p->ld;
    if (cmp) {
      a.map = l == 0 ? nullptr : p->d_map;
      a.count = p->d_counts + l;
      a.flags = p->d_flags;
    } else {
      a.map = nullptr;
      a.count = p->d_counts;
      a.flags = (final_flags && l == l_end - 1) ? p->d_flags : nullptr;
    }
    prof_mark(p, l, 0);
    launch_layer(p, a, s);
    prof_mark(p, l, 1);
    done_layer(p, l);
    if (cmp) {
      // the list entering l_end (clustering, or the categories) is always exact
      launch_compact(p->d_flags, p->d_counts + l, p->d_counts + l + 1, p->d_map, nullptr,
                     p->d_orig[cur], p->d_orig[cur ^ 1], nullptr, nullptr, p->cap, p->d_tilecnt, s,
                     l + 1 < l_end ? p->shape.compact_slack : 0, p->d_live + l + 1);
      cur ^= 1;
    }
  }
  return cur;
}

void enqueue_plain_categories(snicit_plan* p, int l_end, int cur) {
  hipStream_t s = p->stream;
  if (p->net->Nout > 0) {
    enqueue_output_layer(p, l_end == 0 ? hidden_in(p) : p->d_y[(l_end - 1) & 1]);
    return;
  }
  launch_fill_i32(p->d_cat, 0, p->cap, s);
  if (compacting(p)) {
    launch_cat_from_list(p->d_orig[cur], p->d_counts + l_end, p->cap, p->d_cat, s);
  } else {
    launch_cat_from_flags(p->d_flags, p->d_counts, p->cap, p->d_cat, s);
  }
}

struct SnicitState {
  const float* Yt = nullptr;  // activations entering layer t
  const int32_t* map = nullptr;
  const int32_t* count = nullptr;
  int orig = 0;
  float* Da = nullptr;
  float* Db = nullptr;
};

SnicitState snicit_state(snicit_plan* p) {
  SnicitState st;
  const int t = p->cfg.threshold;
  st.Yt = t == 0 ? hidden_in(p) : p->d_y[(t - 1) & 1];
  st.Da = t == 0 ? p->d_y[0] : p->d_y[t & 1];
  st.Db = t == 0 ? p->d_y[1] : p->d_y[(t + 1) & 1];
  return st;
}

// Stage "cluster" (src/SDGC/SNICIT/SNICIT.hpp:406-515): live list at t, pool segment sums,
// Alg. 1 pruning, nearest-centroid assignment, non-empty delta list, delta rows.
// Slot of d_orig holding the live list that enters clustering: t compactions starting
// from slot 0, or one compaction of the input when t == 0; slot 0 without compaction.
int cluster_cur(const snicit_plan* p) {
  const int t = p->cfg.threshold;
  if (!compacting(p)) return 0;
  return t == 0 ? 1 : (t & 1);
}

// Stage "cluster", lossy variant (src/beyond/SNICIT/SNICIT.hpp:539-566, kernel.hpp:5-101):
// pool of strided rows, pruning on whole rows, tolerance assignment and pruned deltas.
// Every row is live (no dead-row skipping in this variant), so positions are sample ids.
void enqueue_cluster_lossy(snicit_plan* p) {
  hipStream_t s = p->stream;
  const int t = p->cfg.threshold;
  const int N = p->net->N;
  SnicitState st = snicit_state(p);
  const snicit_run_cfg& c = p->cfg;
  const Segs sg = plan_segs(p);
  launch_segment_bounds(p->d_orig[0], p->d_counts, p->d_segstart, sg, s);
  launch_lossy_pool(st.Yt, p->ld, sg, c.lossy_pool, N, p->d_poolrows, p->d_poolcol, s);
  launch_lossy_select(p->d_poolrows, p->d_poolcol, sg, c.lossy_pool, N, c.lossy_eps_pool,
                      p->d_close, p->d_cent_slot, s);
  launch_gather_centroids(st.Yt, p->ld, nullptr, p->d_cent_slot, sg, N, p->d_C[0], p->kcap, s);
  launch_lossy_assign(st.Yt, p->ld, sg, p->cap, p->d_C[0], p->kcap, p->d_cent_slot, N,
                      c.lossy_eps_assign, p->d_dist, p->d_mcar[0], p->d_centmark, p->d_rec, s);
  launch_lossy_delta(st.Yt, p->ld, p->d_C[0], p->kcap, p->d_mcar[0], p->d_centmark, p->d_counts,
                     p->cap, N, c.lossy_eps_assign, c.lossy_positive_only != 0, st.Da, p->ld,
                     p->d_flags, s);
  // non-empty delta rows: positions in Da, with original id and centroid carried
  launch_compact(p->d_flags, p->d_counts, p->d_nerow + t, p->d_mapD, nullptr, p->d_orig[0],
                 p->d_orig[1], p->d_mcar[0], p->d_mcar[1], p->cap, p->d_tilecnt, s);
}

void enqueue_cluster(snicit_plan* p) {
  if (p->cfg.cluster == SNICIT_CLUSTER_LOSSY) {
    enqueue_cluster_lossy(p);
    return;
  }
  hipStream_t s = p->stream;
  const int t = p->cfg.threshold;
  const int N = p->net->N;
  SnicitState st = snicit_state(p);
  const int cur = cluster_cur(p);
  const int32_t* map = nullptr;
  const int32_t* count = p->d_counts;
  if (compacting(p)) {
    if (t == 0) {
      // Rows of the input that are all zero stay zero (f(b) = 0); drop them before pooling.
      launch_iota_i32(p->d_orig[0], p->cap, s);
      launch_row_any(p->d_in, p->ld, p->d_counts, p->cap, N, p->d_flags, s);
      launch_compact(p->d_flags, p->d_counts, p->d_counts + 1, p->d_map, nullptr, p->d_orig[0],
                     p->d_orig[1], nullptr, nullptr, p->cap, p->d_tilecnt, s, 0, p->d_live + 1);
      count = p->d_counts + 1;
    } else {
      count = p->d_counts + t;
    }
    map = p->d_map;
  } else {
    if (t == 0) launch_iota_i32(p->d_orig[0], p->cap, s);
    count = p->d_counts;
    map = nullptr;
  }
  launch_fill_i32(p->d_rec, -1, p->cap, s);
  const Segs sg = plan_segs(p);
  launch_segment_bounds(p->d_orig[cur], count, p->d_segstart, sg, s);
  launch_seg_sums(st.Yt, p->ld, map, sg, p->cfg.sample_size, p->cfg.segments, N, p->d_F, s);
  launch_select_exact(p->d_F, sg, p->cfg.sample_size, p->cfg.segments, p->cfg.eps,
                      p->d_cent_slot, s);
  launch_gather_centroids(st.Yt, p->ld, map, p->d_cent_slot, sg, N, p->d_C[0], p->kcap, s);
  launch_assign_exact(st.Yt, p->ld, map, sg, p->cap, p->d_C[0], p->kcap, p->d_cent_slot, N,
                      p->d_dist, p->d_flags, p->d_mcar[0], p->d_orig[cur], p->d_rec, s);
  // Non-empty delta rows: list of (column in Yt, centroid, original id).
  const int32_t* ident = nullptr;
  launch_compact(p->d_flags, count, p->d_nerow + t, p->d_mapD, map ? map : ident, p->d_orig[cur],
                 p->d_orig[cur ^ 1], p->d_mcar[0], p->d_mcar[1], p->cap, p->d_tilecnt, s);
  launch_make_delta(st.Yt, p->ld, p->d_mapD, p->d_C[0], p->kcap, p->d_mcar[1], p->d_nerow + t,
                    p->cap, N, st.Da, p->ld, p->compact_post, s);
}

// List positions post layer j of the compact layout processes (device counter).
const int32_t* post_count(const snicit_plan* p, int j) {
  const int t = p->cfg.threshold;
  return j == 0 ? p->d_nerow + t : p->d_nlist + t + j;
}

// Stage "post" in the compact layout: one fused launch per layer, reading and writing the
// delta rows at fixed list positions with stride compact_ld(positions); at the layers
// post_compacts_after names, the per-layer flags since the last compaction are counted
// (non-empty rows after each layer) and the list is compacted.  Every layer's centroid
// rows, delta rows and counts equal those of the per-layer compaction.
// Whether the post stage's delta launches take chain records computed one layer ahead:
// every post layer's panels hold at most two distinct weight rows, the records buffer
// exists, and the weights are resident.
bool post_chains(const snicit_plan* p) {
  if (!p->compact_post || !p->d_zchain || p->net->streamed() || p->shape.post_form < 2)
    return false;
  for (int l = p->cfg.threshold; l < p->net->L; ++l)
    if (!p->net->few_rows[l]) return false;
  return true;
}

// Stage "post" in the compact layout: one fused launch per layer, reading and writing the
// delta rows at fixed list positions with stride compact_ld(positions); at the layers
// post_compacts_after names, the per-layer flags since the last compaction are counted
// (non-empty rows after each layer) and the list is compacted.  Every layer's centroid
// rows, delta rows and counts equal those of the per-layer compaction.
// With chain records (post_chains): the centroid rows depend on the centroid columns only,
// so a post_chain_kernel launch computes layer 0's chain records and layer 1's centroid
// columns, and the delta launch of layer j computes, beside its delta rows, layer j+1's
// records and layer j+2's centroid columns (records in two slots, centroid columns in two
// buffers, alternating).
void enqueue_post_compact(snicit_plan* p) {
  hipStream_t s = p->stream;
  const Net& net = *p->net;
  const int t = p->cfg.threshold, L = net.L, n_post = L - t;
  const int N = net.N;
  SnicitState st = snicit_state(p);
  int orig = cluster_cur(p) ^ 1, mc = 1;
  int j0 = 0;  // first layer since the last compaction
  const bool chains = post_chains(p);
  auto zslot = [&](int j) {
    return p->d_zchain + (size_t)(j & 1) * p->chain_panels * kPostChainRec;
  };
  // centroid columns at the compact stride for the post layers, back at kcap in d_C[1]
  launch_centroid_restride(p->d_C[0], false, p->d_Cp[0], true, p->kcap, p->d_k, N, s);
  if (chains && n_post > 0)
    launch_post_chain(net, t, plan_layer_arrays(p, t), p->d_Cp[0], p->d_Cp[1], p->kcap, p->d_k,
                      true, zslot(0), s);
  weights_begin(p);
  for (int j = 0; j < n_post; ++j) {
    const int l = t + j;
    use_layer(p, l, L);
    const int q = j & 1;
    const bool after_compaction = j > 0 && post_compacts_after(j - 1, n_post);
    const bool compact_next = post_compacts_after(j, n_post);
    int32_t* ring = p->d_ring + (int64_t)(j % kPostRing) * p->ld;
    NextChain nx;
    if (chains && j + 1 < n_post) {
      const LayerArrays la = plan_layer_arrays(p, l + 1);
      nx.pcols = la.pcols;
      nx.pbeg = la.pbeg;
      nx.outs = la.outs;
      nx.w = la.w;
      nx.psame = la.psame;
      nx.bias_vec = net.bias_vec(l + 1);
      nx.bias_s = net.bias_scalar(l + 1);
      nx.cin = p->d_Cp[(j + 1) & 1];
      nx.cout = p->d_Cp[j & 1];
      nx.zchain = zslot(j + 1);
      nx.P = net.panels(l + 1);
    }
    prof_mark(p, l, 0);
    launch_panel_post(net, l, plan_layer_arrays(p, l), false, p->d_Cp[q], p->d_Cp[q ^ 1],
                      p->kcap, p->d_k, q == 0 ? st.Da : st.Db, p->ld,
                      post_count(p, j == 0 ? 0 : j - 1), q == 0 ? st.Db : st.Da, p->ld, true,
                      post_count(p, j), compact_next ? nullptr : p->d_nlist + l + 1,
                      after_compaction ? p->d_mapD : nullptr, p->d_mcar[mc], ring, 0.f, p->cap,
                      p->shape, s, true, chains ? zslot(j) : nullptr, nx);
    prof_mark(p, l, 1);
    prof_mark(p, l, 2);
    prof_mark(p, l, 3);
    done_layer(p, l);
    if (compact_next) {
      launch_ring_count(p->d_ring, p->ld, j0, j - j0 + 1, post_count(p, j),
                        p->d_nerow + t + j0 + 1, s);
      launch_compact(ring, post_count(p, j), p->d_nlist + l + 1, p->d_mapD, nullptr,
                     p->d_orig[orig], p->d_orig[orig ^ 1], p->d_mcar[mc], p->d_mcar[mc ^ 1],
                     p->cap, nullptr, s);
      orig ^= 1;
      mc ^= 1;
      j0 = j + 1;
    }
  }
  launch_centroid_restride(p->d_Cp[n_post & 1], true, p->d_C[1], false, p->kcap, p->d_k, N, s);
}

// Stage "post" (src/SDGC/SNICIT/SNICIT.hpp:526-538,603-614): per layer the centroid rows,
// then the delta rows with the fused residue epilogue, then the device-side recompaction.
void enqueue_post(snicit_plan* p) {
  if (p->compact_post) {
    enqueue_post_compact(p);
    return;
  }
  hipStream_t s = p->stream;
  const int t = p->cfg.threshold, L = p->net->L;
  SnicitState st = snicit_state(p);
  // clustering leaves the delta list's original ids in slot cur^1 and centroids in slot 1
  int orig = cluster_cur(p) ^ 1, mc = 1;
  const bool lossy = p->cfg.cluster == SNICIT_CLUSTER_LOSSY;
  weights_begin(p);
  for (int l = t; l < L; ++l) {
    use_layer(p, l, L);
    const int q = (l - t) & 1;
    const int32_t* map_in = (l == t && !lossy) ? nullptr : p->d_mapD;
    if (post_layer_fused(*p->net, l) && !p->net->is_ell(l) && p->kcap == kMaxCentroidsExact) {
      prof_mark(p, l, 0);
      launch_panel_post(*p->net, l, plan_layer_arrays(p, l), lossy, p->d_C[q], p->d_C[q ^ 1],
                        p->kcap, p->d_k, q == 0 ? st.Da : st.Db, p->ld, nullptr,
                        q == 0 ? st.Db : st.Da, p->ld, false, p->d_nerow + l, nullptr, map_in,
                        p->d_mcar[mc], p->d_flags, tol_cut(p->cfg.lossy_eps_post), p->cap,
                        p->shape, s);
      prof_mark(p, l, 1);
      prof_mark(p, l, 2);
      prof_mark(p, l, 3);
    } else {
      SpmmLaunch c = spmm_base(p, l, EPI_CENTROID);
      c.yin = p->d_C[q];
      c.ldin = p->kcap;
      c.yout = p->d_C[q ^ 1];
      c.ldout = p->kcap;
      c.count = p->d_k;
      c.zc = p->d_zc;
      c.max_cols = p->kcap;
      prof_mark(p, l, 2);
      launch_layer(p, c, s);
      prof_mark(p, l, 3);
      SpmmLaunch d = spmm_base(p, l, lossy ? EPI_DELTA_LOSSY : EPI_DELTA);
      d.yin = q == 0 ? st.Da : st.Db;
      d.ldin = p->ld;
      d.yout = q == 0 ? st.Db : st.Da;
      d.ldout = p->ld;
      d.map = map_in;
      d.count = p->d_nerow + l;
      d.flags = p->d_flags;
      d.zc = p->d_zc;
      d.mcent = p->d_mcar[mc];
      prof_mark(p, l, 0);
      launch_layer(p, d, s);
      prof_mark(p, l, 1);
    }
    done_layer(p, l);
    // delta rows: a short list (one launch, one workgroup)
    launch_compact(p->d_flags, p->d_nerow + l, p->d_nerow + l + 1, p->d_mapD, nullptr,
                   p->d_orig[orig], p->d_orig[orig ^ 1], p->d_mcar[mc], p->d_mcar[mc ^ 1],
                   p->cap, nullptr, s);
    orig ^= 1;
    mc ^= 1;
  }
}

// Delta-list positions each post layer processed in the last run ([L + 2], index = layer):
// the non-empty rows entering it, or in the compact layout the list length, which keeps
// rows that became empty until the next compaction.
std::vector<int32_t> post_list_lengths(const snicit_plan* p, const std::vector<int32_t>& nerow) {
  if (!p->compact_post) return nerow;
  std::vector<int32_t> list(p->net->L + 2);
  SNICIT_HIP(hipMemcpy(list.data(), p->d_nlist, list.size() * 4, hipMemcpyDeviceToHost));
  list[p->cfg.threshold] = nerow[p->cfg.threshold];
  return list;
}

// Positions the last post layer wrote (the stride of last_D in the compact layout), or null.
const int32_t* last_D_count(const snicit_plan* p) {
  return p->compact_post ? post_count(p, p->net->L - p->cfg.threshold - 1) : nullptr;
}

void enqueue_recover(snicit_plan* p) {
  hipStream_t s = p->stream;
  const int L = p->net->L, N = p->net->N;
  SNICIT_REQUIRE(p->last_C != nullptr && p->last_D != nullptr,
                 "internal: recovery enqueued before the final-state bookkeeping");
  if (p->net->Nout > 0) {
    // recovered hidden activations yhat_i + yhat_m(i) for every sample (the reference adds
    // the centroid back in place, src/beyond/SNICIT/kernel.hpp:170-183), then the output
    // layer and argmax
    launch_expand_centroids(p->d_rec, p->d_counts, p->cap, p->last_C, p->kcap, N, p->d_h0, p->ld,
                            s);
    launch_scatter_delta(p->last_D, p->ld, last_D_count(p), p->d_mapD, p->d_mcar[p->last_mcar],
                         p->d_orig[p->last_orig], p->d_nerow + L, p->cap, p->last_C, p->kcap, N,
                         p->d_h0, p->ld, s);
    enqueue_output_layer(p, p->d_h0);
    return;
  }
  launch_centroid_cats(p->last_C, p->kcap, p->d_k, N, p->d_catc, s);
  launch_recover_base(p->d_rec, p->d_counts, p->cap, p->d_catc, p->d_cat, s);
  launch_recover_rows(p->last_D, p->ld, last_D_count(p), p->d_mapD, p->d_mcar[p->last_mcar],
                      p->d_orig[p->last_orig], p->d_nerow + L, p->cap, p->last_C, p->kcap, N,
                      p->d_act, p->d_cat, s);
}

// Runs `body` directly or through a captured graph cached in slot `g`.  Plans on streamed
// nets run directly: a captured graph runs its host-to-device copy nodes as blit kernels on
// the queue it is launched on, in line with the layer kernels (measured: every copy then
// adds its full time), while directly issued copies run on the copy stream's DMA engine
// beside them.
void run_stage(snicit_plan* p, int g, const std::function<void()>& body) {
  hipStream_t s = p->stream;
  if (!p->cfg.use_graph || p->net->streamed()) {
    body();
    return;
  }
  if (!p->graph[g]) {
    hipGraph_t graph = nullptr;
    SNICIT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      body();
    } catch (...) {
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(s, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      throw;
    }
    SNICIT_HIP(hipStreamEndCapture(s, &graph));
    hipError_t e = hipGraphInstantiate(&p->graph[g], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SNICIT_HIP(e);
  }
  SNICIT_HIP(hipGraphLaunch(p->graph[g], s));
}

void drop_graphs(snicit_plan* p) {
  for (auto& g : p->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Where a run leaves its final state; the stage bodies run at capture only, so this follows
// from the configuration alone.  Must run before the stages are enqueued: enqueue_recover
// and the activation readers take these pointers.
void plan_bookkeeping(snicit_plan* p, int32_t stop) {
  const int L = p->net->L;
  if (plain_mode(p->cfg.mode)) {
    const int l_end = stop < 0 ? L : stop;
    p->last_orig = compacting(p) ? (l_end & 1) : 0;
    p->last_Y = l_end == 0 ? p->d_in : p->d_y[(l_end - 1) & 1];
    p->last_list = compacting(p) && l_end > 0;
    p->stop_layer = l_end;
  } else {
    // each compaction of the delta list flips the original-id and centroid slots: one per
    // post layer, or post_compactions() of them in the compact layout
    const int t = p->cfg.threshold;
    SnicitState st = snicit_state(p);
    const int n_post = L - t;
    const int flips = p->compact_post ? post_compactions(n_post) : n_post;
    p->last_orig = (cluster_cur(p) ^ 1) ^ (flips & 1);
    p->last_mcar = 1 ^ (flips & 1);
    p->last_C = p->compact_post ? p->d_C[1] : p->d_C[n_post & 1];
    p->last_D = (((n_post - 1) & 1) == 0) ? st.Db : st.Da;
    p->stop_layer = L;
  }
}

// Every stage of one run on the plan's stream, for capture into a caller's graph.
void enqueue_all(snicit_plan* p) {
  if (plain_mode(p->cfg.mode)) {
    const int c = enqueue_plain_layers(p, p->net->L, true);
    enqueue_plain_categories(p, p->net->L, c);
  } else {
    enqueue_plain_layers(p, p->cfg.threshold, false);
    enqueue_cluster(p);
    enqueue_post(p);
    enqueue_recover(p);
  }
}

void plan_launch(snicit_plan* p) {
  if (!p->have_input) throw Error(SNICIT_E_STATE, "plan has no input");
  if (p->graph_stop != -1) {
    drop_graphs(p);
    p->graph_stop = -1;
  }
  plan_bookkeeping(p, -1);
  if (plain_mode(p->cfg.mode)) {
    run_stage(p, 0, [&] {
      const int c = enqueue_plain_layers(p, p->net->L, true);
      enqueue_plain_categories(p, p->net->L, c);
    });
  } else {
    run_stage(p, 0, [&] { enqueue_plain_layers(p, p->cfg.threshold, false); });
    run_stage(p, 1, [&] { enqueue_cluster(p); });
    run_stage(p, 2, [&] { enqueue_post(p); });
    run_stage(p, 3, [&] { enqueue_recover(p); });
  }
}

void plan_run(snicit_plan* p, int32_t stop, snicit_stats* stats) {
  if (!p->have_input) throw Error(SNICIT_E_STATE, "plan has no input (call set_input first)");
  Net* net = p->net;
  hipStream_t s = p->stream;
  const int L = net->L;
  if (p->graph_stop != stop) {
    drop_graphs(p);
    p->graph_stop = stop;
  }
  // the recovery stage reads where the post stage leaves its state when it is enqueued
  plan_bookkeeping(p, stop);
  const bool timed = p->ev[0] != nullptr;
  auto mark = [&](int i) {
    if (timed) SNICIT_HIP(hipEventRecord(p->ev[i], s));
  };
  mark(0);
  if (plain_mode(p->cfg.mode)) {
    const int l_end = stop < 0 ? L : stop;
    run_stage(p, 0, [&] {
      const int c = enqueue_plain_layers(p, l_end, true);
      enqueue_plain_categories(p, l_end, c);
    });
    for (int i = 1; i < 5; ++i) mark(i);
  } else {
    run_stage(p, 0, [&] { enqueue_plain_layers(p, p->cfg.threshold, false); });
    mark(1);
    run_stage(p, 1, [&] { enqueue_cluster(p); });
    mark(2);
    run_stage(p, 2, [&] { enqueue_post(p); });
    mark(3);
    run_stage(p, 3, [&] { enqueue_recover(p); });
    mark(4);
  }
  if (timed) SNICIT_HIP(hipEventSynchronize(p->ev[4]));
  else SNICIT_HIP(hipStreamSynchronize(s));
  p->have_run = true;
  if (stats) {
    float ms[4];
    for (int i = 0; i < 4; ++i) {
      ms[i] = 0.f;
      if (timed) SNICIT_HIP(hipEventElapsedTime(&ms[i], p->ev[i], p->ev[i + 1]));
    }
    plan_stats(p, stats, ms);
  }
}

// Statistics of the last run from the device counters; stage_ms (nullable) = pre, cluster,
// post, recover milliseconds.
void plan_stats(snicit_plan* p, snicit_stats* stats, const float* stage_ms) {
  Net* net = p->net;
  const int L = net->L;
  const bool plain = plain_mode(p->cfg.mode);
  {
    std::memset(stats, 0, sizeof(*stats));
    if (stage_ms) {
      stats->ms_pre = stage_ms[0];
      stats->ms_cluster = stage_ms[1];
      stats->ms_post = stage_ms[2];
      stats->ms_recover = stage_ms[3];
      stats->ms_total = stage_ms[0] + stage_ms[1] + stage_ms[2] + stage_ms[3];
    }
    stats->batch = p->B;
    stats->layers = p->stop_layer;
    stats->edges = (double)p->B * ((double)net->nnz_hidden * (double)p->stop_layer +
                                   (double)net->N * net->Nin + (double)net->Nout * net->N);
    std::vector<int32_t> counts(L + 4), nerow(L + 2);
    SNICIT_HIP(hipMemcpy(counts.data(), p->d_counts, counts.size() * 4, hipMemcpyDeviceToHost));
    SNICIT_HIP(hipMemcpy(nerow.data(), p->d_nerow, nerow.size() * 4, hipMemcpyDeviceToHost));
    int64_t rows = 0;
    if (plain) {
      for (int l = 0; l < p->stop_layer; ++l) rows += compacting(p) ? counts[l] : counts[0];
    } else {
      const int t = p->cfg.threshold;
      for (int l = 0; l < t; ++l) rows += compacting(p) ? counts[l] : counts[0];
      int32_t k = 0;  // centroid columns the post layers process
      SNICIT_HIP(hipMemcpy(&k, p->d_k, 4, hipMemcpyDeviceToHost));
      std::vector<int32_t> kseg(p->nseg);
      SNICIT_HIP(hipMemcpy(kseg.data(), p->d_kseg, p->nseg * 4, hipMemcpyDeviceToHost));
      int32_t ktot = 0;
      for (int32_t v : kseg) ktot += v;
      stats->centroids = ktot;
      stats->live_at_threshold = compacting(p) ? counts[t == 0 ? 1 : t] : counts[0];
      stats->nerow_first = nerow[t];
      stats->nerow_last = nerow[L];
      const std::vector<int32_t> list = post_list_lengths(p, nerow);
      for (int l = t; l < L; ++l) rows += list[l] + k;
    }
    stats->rows_processed = rows;
    std::vector<int32_t> cat(p->B);
    SNICIT_HIP(hipMemcpy(cat.data(), p->d_cat, p->B * 4, hipMemcpyDeviceToHost));
    int64_t act = 0;
    for (int32_t c : cat) act += c != 0;
    stats->active = act;
  }
}

void plan_free(snicit_plan* p) {
  drop_graphs(p);
  if (p->copy_stream) (void)hipStreamSynchronize(p->copy_stream);
  for (char* q : p->d_wslot) dev_free(q);
  p->d_wslot.clear();
  for (auto* v : {&p->ev_copied, &p->ev_free})
    for (auto& e : *v)
      if (e) (void)hipEventDestroy(e);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->copy_stream) (void)hipStreamDestroy(p->copy_stream);
  p->copy_stream = nullptr;

  for (auto& e : p->prof_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : p->ev)
    if (e) (void)hipEventDestroy(e);
  void* ptrs[] = {p->d_in,     p->d_y[0],    p->d_y[1],     p->d_flags,   p->d_map,
                  p->d_mapD,   p->d_orig[0], p->d_orig[1],  p->d_mcar[0], p->d_mcar[1],
                  p->d_counts, p->d_tilecnt, p->d_nerow,   p->d_k,       p->d_F,
                  p->d_cent_slot,
                  p->d_C[0],   p->d_C[1],    p->d_zc,       p->d_rec,     p->d_catc,
                  p->d_dist,     p->d_h0,      p->d_logits,   p->d_poolrows, p->d_poolcol,
                  p->d_centmark, p->d_close, p->d_segstart, p->d_seglo, p->d_kseg,
                  p->d_act,    p->d_cat,     p->d_img_params, p->d_img_blobs, p->d_img_proto,
                  p->d_stage,  p->d_nonfinite, p->d_nlist, p->d_ring,   p->d_Cp[0],
                  p->d_Cp[1],  p->d_zchain, p->d_live};
  for (void* q : ptrs) dev_free(q);
}

}  // namespace snicit

// ---- C-ABI ------------------------------------------------------------------------------
extern "C" {

const char* snicit_last_error(void) { return g_last_error.c_str(); }

void snicit_run_cfg_default(snicit_run_cfg* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->mode = SNICIT_MODE_SNICIT;
  cfg->threshold = 30;
  cfg->cluster = SNICIT_CLUSTER_EXACT;
  cfg->sample_size = 32;
  cfg->segments = 16;
  cfg->eps = 0.03f;
  cfg->lossy_pool = 128;
  cfg->lossy_eps_pool = 0.03f;
  cfg->lossy_eps_assign = 0.04f;
  cfg->lossy_eps_post = 0.05f;
  cfg->use_graph = 1;
}

int snicit_ctx_create(int device, snicit_ctx** out) {
  return guarded([&] {
    SNICIT_REQUIRE(out != nullptr, "out is null");
    int n = 0;
    SNICIT_HIP(hipGetDeviceCount(&n));
    SNICIT_REQUIRE(device >= 0 && device < n, "no such device " + std::to_string(device));
    SNICIT_HIP(hipSetDevice(device));
    auto c = std::make_unique<snicit_ctx>();
    c->device = device;
    SNICIT_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    SNICIT_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    (void)scratch_slot(device);  // the static scratch's address on this device
    *out = c.release();
  });
}

int snicit_ctx_set_grouping(snicit_ctx* ctx, int32_t grouping) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx != nullptr, "null context");
    SNICIT_REQUIRE(grouping == SNICIT_GROUP_NONE || grouping == SNICIT_GROUP_CONSECUTIVE ||
                       grouping == SNICIT_GROUP_SEARCH,
                   "grouping must be SNICIT_GROUP_NONE, _CONSECUTIVE or _SEARCH");
    ctx->grouping = grouping;
  });
}

int snicit_ctx_destroy(snicit_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int snicit_net_create_ell(snicit_ctx* ctx, int32_t N, int32_t L, int32_t K, const int32_t* cols,
                          const float* vals, int32_t bias_kind, const float* bias, float clip,
                          snicit_net** out) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx && out && cols && vals, "null argument");
    SNICIT_REQUIRE(N > 0 && L > 0 && K > 0 && K <= N, "need N > 0, L > 0 and 0 < K <= N");
    SNICIT_REQUIRE(clip > 0.f, "clip must be positive");
    SNICIT_HIP(hipSetDevice(ctx->device));
    auto net = std::make_unique<snicit_net>();
    net->ctx = ctx;
    net->N = N;
    net->L = L;
    net->K = K;
    net->clip = clip;
    net->grouping = ctx->grouping;
    net_set_bias(net.get(), bias_kind, bias);
    net->panel_off.push_back(0);
    net->nnz_hidden = (int64_t)N * K;
    NetBuilder nb{net.get()};
    HostLayerPanels hp;
    for (int32_t l = 0; l < L; ++l) {
      if (K > kSlots) {
        // rows wider than a panel: wide-ELL layer
        nb.add_ell(cols + (size_t)l * N * K, vals + (size_t)l * N * K, K);
        continue;
      }
      pack_layer_ell(N, K, cols + (size_t)l * N * K, vals + (size_t)l * N * K, hp, ctx->grouping);
      nb.add(hp);
    }
    try {
      upload_net(net.get(), nb, 0.f);
    } catch (...) {
      destroy_net(net.get());
      throw;
    }
    *out = net.release();
  });
}

int snicit_net_create_csr(snicit_ctx* ctx, int32_t N, int32_t L, const int64_t* row_ptr,
                          const int32_t* cols, const float* vals, int32_t rows_are_inputs,
                          int32_t bias_kind, const float* bias, float clip, snicit_net** out) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx && out && row_ptr && (cols || row_ptr[(size_t)L * (N + 1) - 1] == 0) &&
                       (vals || row_ptr[(size_t)L * (N + 1) - 1] == 0),
                   "null argument");
    SNICIT_REQUIRE(N > 0 && L > 0, "need N > 0 and L > 0");
    // per output row, (input, value) in ascending input (stable for repeated inputs)
    std::vector<std::vector<std::vector<std::pair<int32_t, float>>>> rows(L);
    int32_t K = 1;
    int64_t nnz = 0;
    for (int32_t l = 0; l < L; ++l) {
      const int64_t* rp = row_ptr + (size_t)l * (N + 1);
      auto& R = rows[l];
      R.assign(N, {});
      for (int32_t i = 0; i < N; ++i) {
        SNICIT_REQUIRE(rp[i] >= 0 && rp[i] <= rp[i + 1], "row_ptr must be non-decreasing");
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
          const int32_t c = cols[e];
          SNICIT_REQUIRE(c >= 0 && c < N, "CSR column out of range in layer " + std::to_string(l));
          if (rows_are_inputs) R[c].push_back({i, vals[e]});  // input i feeds output c
          else R[i].push_back({c, vals[e]});
        }
      }
      for (auto& r : R) {
        if (!rows_are_inputs)
          std::stable_sort(r.begin(), r.end(),
                           [](const std::pair<int32_t, float>& x,
                              const std::pair<int32_t, float>& y) { return x.first < y.first; });
        K = std::max<int32_t>(K, (int32_t)r.size());
        nnz += (int64_t)r.size();
      }
    }
    SNICIT_REQUIRE(K <= N, "a row has more entries than N");
    // padded per-row lists: value-0 entries repeat the last input (or input 0 for empty rows)
    std::vector<int32_t> ec((size_t)L * N * K);
    std::vector<float> ev((size_t)L * N * K);
    for (int32_t l = 0; l < L; ++l)
      for (int32_t o = 0; o < N; ++o) {
        const auto& r = rows[l][o];
        const size_t at = ((size_t)l * N + o) * K;
        for (int32_t k = 0; k < K; ++k) {
          const bool real = k < (int32_t)r.size();
          ec[at + k] = real ? r[k].first : (r.empty() ? 0 : r.back().first);
          ev[at + k] = real ? r[k].second : 0.f;
        }
      }
    rows.clear();
    const int rc = snicit_net_create_ell(ctx, N, L, K, ec.data(), ev.data(), bias_kind, bias, clip,
                                         out);
    if (rc != SNICIT_OK) throw Error(rc, g_last_error);
    (*out)->nnz_hidden = nnz / L;
  });
}

int snicit_net_generate_sdgc(snicit_ctx* ctx, int32_t N, int32_t L, uint64_t seed, float weight,
                             float bias, float clip, snicit_net** out) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx && out, "null argument");
    SNICIT_REQUIRE(L > 0, "L must be positive");
    SNICIT_REQUIRE(clip > 0.f, "clip must be positive");
    SNICIT_HIP(hipSetDevice(ctx->device));
    auto net = std::make_unique<snicit_net>();
    net->ctx = ctx;
    net->N = N;
    net->L = L;
    net->K = kSlots;
    net->clip = clip;
    net_set_bias(net.get(), SNICIT_BIAS_SCALAR, &bias);
    net->panel_off.push_back(0);
    net->nnz_hidden = (int64_t)N * kSlots;
    NetBuilder nb{net.get()};
    nb.weights_given = false;
    HostLayerPanels hp;
    for (int32_t l = 0; l < L; ++l) {
      sdgc_layer_panels(N, l, seed, hp);
      nb.add(hp);
    }
    try {
      upload_net(net.get(), nb, weight);
    } catch (...) {
      destroy_net(net.get());
      throw;
    }
    *out = net.release();
  });
}

int snicit_net_set_io(snicit_net* net, int32_t Nin, const float* w_in, const float* b_in,
                      float in_clip, int32_t Nout, const float* w_out, const float* b_out) {
  return guarded([&] {
    SNICIT_REQUIRE(net != nullptr, "null net");
    SNICIT_REQUIRE(Nin >= 0 && Nout >= 0, "sizes must be >= 0");
    SNICIT_REQUIRE(Nin == 0 || (w_in && b_in && in_clip > 0.f), "input layer needs weights, bias, clip");
    SNICIT_REQUIRE(Nout == 0 || (w_out && b_out), "output layer needs weights and bias");
    SNICIT_REQUIRE(Nout == 0 || net->N % 16 == 0, "an output layer needs N divisible by 16");
    SNICIT_REQUIRE(net->live_plans == 0,
                   "set the input/output layers before creating plans on the net");
    SNICIT_HIP(hipSetDevice(net->ctx->device));
    const int32_t N = net->N;
    dev_free(net->d_win);
    dev_free(net->d_wint);
    dev_free(net->d_bin);
    dev_free(net->d_wout);
    dev_free(net->d_woutt);
    dev_free(net->d_bout);
    net->d_win = net->d_wint = net->d_bin = net->d_wout = net->d_woutt = net->d_bout = nullptr;
    net->Nin = Nin;
    net->Nout = Nout;
    net->Kin = (Nin + 15) / 16 * 16;
    net->Kout = (N + 15) / 16 * 16;
    net->in_clip = in_clip;
    if (Nin > 0) {
      std::vector<float> w((size_t)N * net->Kin, 0.f);
      for (int32_t o = 0; o < N; ++o)
        std::memcpy(&w[(size_t)o * net->Kin], w_in + (size_t)o * Nin, (size_t)Nin * 4);
      net->win_ld = (N + 63) / 64 * 64;
      const size_t ldw = (size_t)net->win_ld;
      std::vector<float> wt((size_t)net->Kin * ldw, 0.f);
      for (int32_t o = 0; o < N; ++o)
        for (int32_t k = 0; k < Nin; ++k) wt[(size_t)k * ldw + o] = w_in[(size_t)o * Nin + k];
      net->d_win = dev_alloc<float>(w.size(), &net->device_bytes);
      net->d_wint = dev_alloc<float>(wt.size(), &net->device_bytes);
      net->d_bin = dev_alloc<float>(N, &net->device_bytes);
      SNICIT_HIP(hipMemcpy(net->d_win, w.data(), w.size() * 4, hipMemcpyHostToDevice));
      SNICIT_HIP(hipMemcpy(net->d_wint, wt.data(), wt.size() * 4, hipMemcpyHostToDevice));
      SNICIT_HIP(hipMemcpy(net->d_bin, b_in, (size_t)N * 4, hipMemcpyHostToDevice));
    }
    if (Nout > 0) {
      std::vector<float> w((size_t)Nout * net->Kout, 0.f);
      for (int32_t o = 0; o < Nout; ++o)
        std::memcpy(&w[(size_t)o * net->Kout], w_out + (size_t)o * N, (size_t)N * 4);
      net->wout_ld = (Nout + 31) / 32 * 32;
      const size_t ldw = (size_t)net->wout_ld;
      std::vector<float> wt((size_t)net->Kout * ldw, 0.f);
      for (int32_t o = 0; o < Nout; ++o)
        for (int32_t k = 0; k < N; ++k) wt[(size_t)k * ldw + o] = w_out[(size_t)o * N + k];
      net->d_wout = dev_alloc<float>(w.size(), &net->device_bytes);
      net->d_woutt = dev_alloc<float>(wt.size(), &net->device_bytes);
      net->d_bout = dev_alloc<float>(Nout, &net->device_bytes);
      SNICIT_HIP(hipMemcpy(net->d_wout, w.data(), w.size() * 4, hipMemcpyHostToDevice));
      SNICIT_HIP(hipMemcpy(net->d_woutt, wt.data(), wt.size() * 4, hipMemcpyHostToDevice));
      SNICIT_HIP(hipMemcpy(net->d_bout, b_out, (size_t)Nout * 4, hipMemcpyHostToDevice));
    }
  });
}

int snicit_net_stream_weights(snicit_net* net, int32_t buffers) {
  return guarded([&] {
    SNICIT_REQUIRE(net != nullptr, "null net");
    SNICIT_REQUIRE(buffers >= 2 && buffers <= 8, "buffers must be in [2, 8]");
    SNICIT_REQUIRE(!net->streamed(), "the net's weights are already streamed");
    SNICIT_REQUIRE(net->live_plans == 0, "stream the weights before creating plans on the net");
    SNICIT_HIP(hipSetDevice(net->ctx->device));
    const int L = net->L;
    const int64_t N = net->N;
    const int64_t nch = ell_chunks(net->N);
    auto align = [](int64_t b) { return (b + 255) / 256 * 256; };
    std::vector<BlockLayout> blk(L);
    std::vector<int64_t> bytes(L), eent_len(L, 0);
    int64_t freed = 0;  // device bytes of the arrays that move
    for (int l = 0; l < L; ++l) {
      BlockLayout& b = blk[l];
      int64_t at = 0;
      if (net->is_ell(l)) {
        const int64_t K = net->ell_K[l];
        b.ecols = at;
        at = align(at + N * K * 4);
        b.evals = at;
        at = align(at + N * K * 4);
        b.erng = at;
        at = align(at + N * (nch + 1) * 4);
        // the layer's stream length is its last row's end (relative to the layer)
        int32_t end = 0;
        SNICIT_HIP(hipMemcpy(&end, net->d_erng + net->erng_off[l] + N * (nch + 1) - 1, 4,
                             hipMemcpyDeviceToHost));
        eent_len[l] = end;
        b.eent = at;
        at = align(at + ((int64_t)end + kEllGroup) * 8);
        freed += N * K * 8 + N * (nch + 1) * 4 + (int64_t)end * 8;
      } else {
        const int64_t P = net->panels(l);
        const int64_t wf = net->mfma[l] ? P * kSlots * 32 : N * kSlots;
        b.pcols = at;
        at = align(at + P * kSlots * 4);
        b.pbeg = at;
        at = align(at + (P + 1) * 4);
        b.outs = at;
        at = align(at + N * 4);
        b.w = at;
        at = align(at + wf * 4);
        b.psame = at;
        at = align(at + P * 4);
        freed += P * kSlots * 4 + (P + 1) * 4 + N * 4 + wf * 4 + P * 4;
      }
      bytes[l] = at;
    }
    std::vector<int64_t> off(L + 1, 0);
    for (int l = 0; l < L; ++l) off[l + 1] = off[l] + bytes[l];
    char* h = nullptr;
    SNICIT_HIP(hipHostMalloc((void**)&h, (size_t)std::max<int64_t>(off[L], 256), 0));
    try {
      for (int l = 0; l < L; ++l) {
        const LayerArrays a = layer_arrays(*net, l, nullptr);
        const BlockLayout& b = blk[l];
        char* base = h + off[l];
        if (net->is_ell(l)) {
          const int64_t K = net->ell_K[l];
          SNICIT_HIP(hipMemcpy(base + b.ecols, a.ecols, N * K * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.evals, a.evals, N * K * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.erng, a.erng, N * (nch + 1) * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.eent, a.eent, eent_len[l] * 8, hipMemcpyDeviceToHost));
          // one group of zero slack after the stream, as in device memory
          std::memset(base + b.eent + eent_len[l] * 8, 0, kEllGroup * 8);
        } else {
          const int64_t P = net->panels(l);
          const int64_t wf = net->mfma[l] ? P * kSlots * 32 : N * kSlots;
          SNICIT_HIP(hipMemcpy(base + b.pcols, a.pcols, P * kSlots * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.pbeg, a.pbeg, (P + 1) * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.outs, a.outs, N * 4, hipMemcpyDeviceToHost));
          SNICIT_HIP(hipMemcpy(base + b.w, a.w, wf * 4, hipMemcpyDeviceToHost));
          if (P > 0)
            SNICIT_HIP(hipMemcpy(base + b.psame, a.psame, P * 4, hipMemcpyDeviceToHost));
        }
      }
    } catch (...) {
      (void)hipHostFree(h);
      throw;
    }
    for (void* q : {(void*)net->d_pcols, (void*)net->d_pbeg, (void*)net->d_outs,
                    (void*)net->d_w, (void*)net->d_psame, (void*)net->d_ecols,
                    (void*)net->d_evals, (void*)net->d_erng, (void*)net->d_eent})
      dev_free(q);
    net->d_psame = nullptr;
    net->d_pcols = net->d_pbeg = net->d_outs = net->d_ecols = net->d_erng = nullptr;
    net->d_w = net->d_evals = nullptr;
    net->d_eent = nullptr;
    net->device_bytes = std::max<int64_t>(0, net->device_bytes - freed);
    net->h_blocks = h;
    net->blk = std::move(blk);
    net->blk_off.assign(off.begin(), off.end() - 1);
    net->blk_bytes = bytes;
    net->blk_max = std::max<int64_t>(256, *std::max_element(bytes.begin(), bytes.end()));
    net->stream_buffers = buffers;
  });
}

int snicit_net_destroy(snicit_net* net) {
  return guarded([&] {
    if (!net) return;
    destroy_net(net);
    delete net;
  });
}

int snicit_net_info(const snicit_net* net, int32_t* N, int32_t* L, int32_t* K,
                    int64_t* panels_total, int64_t* device_bytes) {
  return guarded([&] {
    SNICIT_REQUIRE(net != nullptr, "null net");
    if (N) *N = net->N;
    if (L) *L = net->L;
    if (K) *K = net->K;
    if (panels_total) *panels_total = net->panel_off.back();
    if (device_bytes) *device_bytes = net->device_bytes;
  });
}

}  // extern "C"

namespace snicit {
// Plan of one net over max_batch columns split into batch segments at segstart (nseg + 1
// entries, ascending; the last one bounds every column).  Runs on `stream`.
snicit_plan* plan_make(snicit_net* net, const snicit_run_cfg* cfg, int64_t max_batch,
                       const std::vector<int32_t>& segstart, hipStream_t stream) {
  {
    SNICIT_REQUIRE(net && cfg, "null argument");
    const int32_t nseg = (int32_t)segstart.size() - 1;
    SNICIT_REQUIRE(nseg >= 1, "a plan needs at least one batch segment");
    for (int32_t i = 0; i < nseg; ++i)
      SNICIT_REQUIRE(segstart[i] <= segstart[i + 1], "batch segments must be ascending");
    SNICIT_REQUIRE(max_batch > 0 && max_batch < (int64_t(1) << 31) - 1024,
                   "max_batch out of range");
    SNICIT_REQUIRE(cfg->mode >= SNICIT_MODE_SNICIT && cfg->mode <= SNICIT_MODE_PLAIN_DENSE,
                   "bad mode");
    if (cfg->mode == SNICIT_MODE_SNICIT) {
      SNICIT_REQUIRE(cfg->threshold >= 0 && cfg->threshold < net->L,
                     "threshold must satisfy 0 <= t < L (got " + std::to_string(cfg->threshold) +
                         ", L = " + std::to_string(net->L) + ")");
      SNICIT_REQUIRE(cfg->cluster == SNICIT_CLUSTER_EXACT || cfg->cluster == SNICIT_CLUSTER_LOSSY,
                     "bad cluster variant");
      if (cfg->cluster == SNICIT_CLUSTER_EXACT) {
        SNICIT_REQUIRE(cfg->sample_size >= 1 && cfg->sample_size <= kMaxCentroidsExact,
                       "sample_size must be in [1, 32]");
        SNICIT_REQUIRE(cfg->segments >= 1 && cfg->segments <= 32 && net->N % cfg->segments == 0,
                       "segments must divide N and be <= 32");
      } else {
        SNICIT_REQUIRE(cfg->lossy_pool >= 1 && cfg->lossy_pool <= kMaxCentroidsLossy,
                       "lossy_pool must be in [1, 128]");
        SNICIT_REQUIRE(!net->skip_dead || net->has_io(),
                       "lossy clustering runs on nets without dead-row skipping (a bias > 0 or "
                       "dense input/output layers)");
      }
    }
    SNICIT_HIP(hipSetDevice(net->ctx->device));
    auto p = std::make_unique<snicit_plan>();
    p->net = net;
    p->cfg = *cfg;
    p->stream = stream;
    p->cap = max_batch;
    p->ld = (max_batch + 63) / 64 * 64;
    const bool lossy = cfg->mode == SNICIT_MODE_SNICIT && cfg->cluster == SNICIT_CLUSTER_LOSSY;
    p->nseg = nseg;
    p->stride = lossy ? kMaxCentroidsLossy : kMaxCentroidsExact;
    p->kcap = p->stride * nseg;
    p->in_rows = net->Nin > 0 ? net->Kin : net->N;
    const int N = net->N, L = net->L;
    try {
      int64_t* acct = &p->device_bytes;
      p->d_in = dev_alloc_big<float>((size_t)p->in_rows * p->ld, acct);
      if (net->Nin > 0)  // rows Nin..Kin-1 meet zero weights and stay zero
        SNICIT_HIP(hipMemset(p->d_in, 0, (size_t)p->in_rows * p->ld * 4));
      if (net->has_io()) p->d_h0 = dev_alloc_big<float>((size_t)N * p->ld, acct);
      if (net->Nout > 0) p->d_logits = dev_alloc<float>((size_t)net->Nout * p->ld, acct);
      if (lossy) {
        p->d_poolrows = dev_alloc<float>((size_t)kMaxCentroidsLossy * nseg * N, acct);
        p->d_poolcol = dev_alloc<int32_t>((size_t)kMaxCentroidsLossy * nseg, acct);
        p->d_centmark = dev_alloc<int32_t>(p->ld, acct);
        p->d_close = dev_alloc<uint64_t>((size_t)kMaxCentroidsLossy * 2 * nseg, acct);
      }
      p->d_segstart = dev_alloc<int32_t>(nseg + 1, acct);
      p->d_seglo = dev_alloc<int32_t>(nseg + 1, acct);
      p->d_kseg = dev_alloc<int32_t>(nseg, acct);
      SNICIT_HIP(hipMemcpy(p->d_segstart, segstart.data(), (nseg + 1) * 4,
                           hipMemcpyHostToDevice));
      SNICIT_HIP(hipMemset(p->d_kseg, 0, nseg * 4));
      p->d_y[0] = dev_alloc_big<float>((size_t)N * p->ld, acct);
      p->d_y[1] = dev_alloc_big<float>((size_t)N * p->ld, acct);
      p->d_flags = dev_alloc<int32_t>(p->ld, acct);
      p->d_map = dev_alloc<int32_t>(p->ld, acct);
      p->d_mapD = dev_alloc<int32_t>(p->ld, acct);
      for (int i = 0; i < 2; ++i) {
        p->d_orig[i] = dev_alloc<int32_t>(p->ld, acct);
        p->d_mcar[i] = dev_alloc<int32_t>(p->ld, acct);
        p->d_C[i] = dev_alloc<float>((size_t)N * p->kcap, acct);
      }
      p->d_counts = dev_alloc<int32_t>(L + 4, acct);
      p->d_live = dev_alloc<int32_t>(L + 4, acct);
      p->d_tilecnt = dev_alloc<int32_t>((size_t)compact_tiles(p->cap), acct);
      p->d_nerow = dev_alloc<int32_t>(L + 2, acct);
      p->d_k = dev_alloc<int32_t>(1, acct);
      p->d_F = dev_alloc<float>((size_t)32 * 32 * nseg, acct);
      p->d_cent_slot = dev_alloc<int32_t>(p->kcap, acct);
      p->d_zc = dev_alloc<float>((size_t)N * p->kcap, acct);
      p->d_rec = dev_alloc<int32_t>(p->ld, acct);
      p->d_dist = dev_alloc<int32_t>((size_t)p->ld * p->stride, acct);
      SNICIT_HIP(hipMemset(p->d_dist, 0, (size_t)p->ld * p->stride * 4));
      p->d_catc = dev_alloc<int32_t>(p->kcap, acct);
      p->d_act = dev_alloc<int32_t>(p->ld, acct);
      p->d_cat = dev_alloc<int32_t>(p->ld, acct);
      SNICIT_HIP(hipMemset(p->d_flags, 0, p->ld * 4));
      SNICIT_HIP(hipMemset(p->d_act, 0, p->ld * 4));
      SNICIT_HIP(hipMemset(p->d_counts, 0, (L + 4) * 4));
      SNICIT_HIP(hipMemset(p->d_live, 0, (L + 4) * 4));
      SNICIT_HIP(hipMemset(p->d_nerow, 0, (L + 2) * 4));
      SNICIT_HIP(hipMemset(p->d_k, 0, 4));
      SNICIT_HIP(hipMemset(p->d_C[0], 0, (size_t)N * p->kcap * 4));
      SNICIT_HIP(hipMemset(p->d_C[1], 0, (size_t)N * p->kcap * 4));
      SNICIT_HIP(hipMemset(p->d_zc, 0, (size_t)N * p->kcap * 4));
      p->compact_post = compact_post_layout(p.get());
      if (p->compact_post) {
        for (int i = 0; i < 2; ++i) {
          p->d_Cp[i] = dev_alloc<float>((size_t)N * p->kcap, acct);
          SNICIT_HIP(hipMemset(p->d_Cp[i], 0, (size_t)N * p->kcap * 4));
        }
        // chain records of two post layers (panels with at most two distinct weight rows)
        bool few = true;
        int64_t maxP = 0;
        for (int l = cfg->threshold; l < L; ++l) {
          few = few && net->few_rows[l];
          maxP = std::max<int64_t>(maxP, net->panels(l));
        }
        if (few && maxP > 0) {
          p->chain_panels = maxP;
          p->d_zchain = dev_alloc<float>((size_t)2 * maxP * kPostChainRec, acct);
        }
        p->d_nlist = dev_alloc<int32_t>(L + 2, acct);
        p->d_ring = dev_alloc<int32_t>((size_t)kPostRing * p->ld, acct);
        SNICIT_HIP(hipMemset(p->d_nlist, 0, (L + 2) * 4));
        SNICIT_HIP(hipMemset(p->d_ring, 0, (size_t)kPostRing * p->ld * 4));
      }
      if (!g_dev_arena)  // the allocation-free entry point runs untimed
        for (auto& e : p->ev) SNICIT_HIP(hipEventCreate(&e));
      if (net->streamed()) {
        const int nb = net->stream_buffers;
        p->slot_layer.assign(nb, -1);
        p->d_wslot.assign(nb, nullptr);
        p->ev_copied.assign(nb, nullptr);
        p->ev_free.assign(nb, nullptr);
        for (int i = 0; i < nb; ++i) {
          p->d_wslot[i] = dev_alloc<char>((size_t)net->blk_max, acct);
          SNICIT_HIP(hipEventCreateWithFlags(&p->ev_copied[i], hipEventDisableTiming));
          SNICIT_HIP(hipEventCreateWithFlags(&p->ev_free[i], hipEventDisableTiming));
        }
        SNICIT_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
        // default priority, like every stream of the library
        SNICIT_HIP(hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking));
      }
    } catch (...) {
      plan_free(p.get());
      throw;
    }
    ++net->live_plans;
    return p.release();
  }
}
}  // namespace snicit

extern "C" {

int snicit_plan_create(snicit_net* net, const snicit_run_cfg* cfg, int64_t max_batch,
                       snicit_plan** out) {
  return guarded([&] {
    SNICIT_REQUIRE(net && cfg && out, "null argument");
    *out = plan_make(net, cfg, max_batch, {0, INT32_MAX}, net->ctx->stream);
  });
}

int snicit_plan_destroy(snicit_plan* plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipStreamSynchronize(plan->stream);
    --plan->net->live_plans;
    plan_free(plan);
    delete plan;
  });
}

}  // extern "C"

namespace snicit {
void ensure_stage(snicit_plan* p) {
  const Net& n = *p->net;
  const int64_t w = std::max<int64_t>(n.N, std::max<int64_t>(n.Nin, n.Nout));
  if (!p->d_stage) p->d_stage = dev_alloc<float>((size_t)p->cap * w, &p->device_bytes);
  if (!p->d_nonfinite) p->d_nonfinite = dev_alloc<int32_t>(1, &p->device_bytes);
}
void load_host_input(snicit_plan* p, const float* y0, int64_t B, int64_t col_off);
void load_images(snicit_plan* p, uint64_t seed, int64_t first_sample, int64_t B,
                 int64_t col_off);
void set_count(snicit_plan* p, int64_t B) {
  int32_t b32 = (int32_t)B;
  hipStream_t s = p->stream;
  SNICIT_HIP(hipMemcpyAsync(p->d_counts, &b32, 4, hipMemcpyHostToDevice, s));
  SNICIT_HIP(hipMemcpyAsync(p->d_live, &b32, 4, hipMemcpyHostToDevice, s));
  SNICIT_HIP(hipStreamSynchronize(s));
  p->B = B;
  p->have_input = true;
  p->have_run = false;
}
}  // namespace snicit

extern "C" {

int snicit_plan_set_input(snicit_plan* p, const float* y0, int64_t B) {
  return guarded([&] {
    SNICIT_REQUIRE(p && (y0 || B == 0), "null argument");
    SNICIT_REQUIRE(B >= 0 && B <= p->cap, "batch exceeds plan capacity");
    SNICIT_REQUIRE(p->nseg == 1, "batch plans take their input per segment");
    load_host_input(p, y0, B, 0);
    set_count(p, B);
  });
}

}  // extern "C"

namespace snicit {
// Renders the blob images of samples first_sample .. first_sample+B-1 on the device into
// columns col_off .. col_off+B-1 of the plan's input.
void load_images(snicit_plan* p, uint64_t seed, int64_t first_sample, int64_t B,
                 int64_t col_off) {
  SNICIT_REQUIRE(p->net->Nin == 0, "image inputs feed nets without an input layer");
  if (B <= 0) return;
  const int N = p->net->N;
  hipStream_t s = p->stream;
  ImageGeom g = image_geometry(N, seed);
  std::vector<int32_t> params((size_t)B * 4);
  for (int64_t b = 0; b < B; ++b) {
    SampleParams sp = image_sample_params(g, seed, first_sample + b);
    params[b * 4 + 0] = sp.cls;
    params[b * 4 + 1] = sp.dx;
    params[b * 4 + 2] = sp.dy;
    params[b * 4 + 3] = sp.keep;
  }
  std::vector<int32_t> blobs(g.blobs.size() * 4), proto(20);
  for (size_t i = 0; i < g.blobs.size(); ++i) {
    blobs[i * 4 + 0] = g.blobs[i].cx;
    blobs[i * 4 + 1] = g.blobs[i].cy;
    blobs[i * 4 + 2] = g.blobs[i].r;
    blobs[i * 4 + 3] = 0;
  }
  for (int c = 0; c < 10; ++c) {
    proto[c] = g.first_blob[c];
    proto[10 + c] = g.nproto_blobs[c];
  }
  SNICIT_REQUIRE(blobs.size() <= 64 * 4, "too many blobs");
  // an earlier render may still read the parameter buffers
  SNICIT_HIP(hipStreamSynchronize(s));
  if (p->img_params_cap < B) {
    dev_free(p->d_img_params);
    p->d_img_params = dev_alloc<int32_t>((size_t)B * 4);
    p->img_params_cap = B;
  }
  if (!p->d_img_blobs) p->d_img_blobs = dev_alloc<int32_t>(64 * 4);
  if (!p->d_img_proto) p->d_img_proto = dev_alloc<int32_t>(20);
  SNICIT_HIP(hipMemcpy(p->d_img_params, params.data(), params.size() * 4, hipMemcpyHostToDevice));
  SNICIT_HIP(hipMemcpy(p->d_img_blobs, blobs.data(), blobs.size() * 4, hipMemcpyHostToDevice));
  SNICIT_HIP(hipMemcpy(p->d_img_proto, proto.data(), proto.size() * 4, hipMemcpyHostToDevice));
  launch_images(g, p->d_img_params, p->d_img_blobs, p->d_img_proto, B, N, p->d_in + col_off,
                p->ld, s);
}

// Copies B host samples (B x width, sample-major) into columns col_off.. of the plan's input.
// Inputs must be finite: rows shorter than a panel's 32 slots and merged short-row panels
// add value-0 steps fma(0, x, acc), which equal acc only for finite x, and the reference
// never reads those inputs for such a row.  A non-finite value is refused here.
void load_host_input(snicit_plan* p, const float* y0, int64_t B, int64_t col_off) {
  if (B <= 0) return;
  const int width = p->net->Nin > 0 ? p->net->Nin : p->net->N;
  hipStream_t s = p->stream;
  ensure_stage(p);
  SNICIT_HIP(hipMemcpyAsync(p->d_stage, y0, (size_t)B * width * 4, hipMemcpyHostToDevice, s));
  launch_fill_i32(p->d_nonfinite, 0, 1, s);
  launch_transpose_in(p->d_stage, B, width, p->d_in + col_off, p->ld, p->d_nonfinite, s);
  int32_t bad = 0;
  SNICIT_HIP(hipMemcpyAsync(&bad, p->d_nonfinite, 4, hipMemcpyDeviceToHost, s));
  SNICIT_HIP(hipStreamSynchronize(s));
  if (bad) {
    p->have_input = false;
    p->have_run = false;
    throw Error(SNICIT_E_INVALID, "input holds a non-finite value (inf or NaN)");
  }
}
}  // namespace snicit

extern "C" {

int snicit_plan_set_input_images(snicit_plan* p, uint64_t seed, int64_t first_sample, int64_t B) {
  return guarded([&] {
    SNICIT_REQUIRE(p != nullptr, "null plan");
    SNICIT_REQUIRE(B >= 0 && B <= p->cap, "batch exceeds plan capacity");
    load_images(p, seed, first_sample, B, 0);
    set_count(p, B);
  });
}

int snicit_plan_run(snicit_plan* p, snicit_stats* stats) {
  return guarded([&] {
    SNICIT_REQUIRE(p != nullptr, "null plan");
    SNICIT_HIP(hipSetDevice(p->net->ctx->device));
    plan_run(p, -1, stats);
  });
}

int snicit_plan_run_layers(snicit_plan* p, int32_t stop_layer, snicit_stats* stats) {
  return guarded([&] {
    SNICIT_REQUIRE(p != nullptr, "null plan");
    SNICIT_REQUIRE(plain_mode(p->cfg.mode), "run_layers needs a plain mode");
    SNICIT_REQUIRE(stop_layer >= 0 && stop_layer <= p->net->L, "stop_layer out of range");
    SNICIT_HIP(hipSetDevice(p->net->ctx->device));
    plan_run(p, stop_layer, stats);
  });
}

int snicit_plan_categories(snicit_plan* p, int32_t* cat) {
  return guarded([&] {
    SNICIT_REQUIRE(p && (cat || p->B == 0), "null argument");
    if (!p->have_run) throw Error(SNICIT_E_STATE, "no completed run");
    if (p->B) SNICIT_HIP(hipMemcpy(cat, p->d_cat, p->B * 4, hipMemcpyDeviceToHost));
  });
}

int snicit_plan_activations(snicit_plan* p, float* y) {
  return guarded([&] {
    SNICIT_REQUIRE(p && (y || p->B == 0), "null argument");
    if (!p->have_run) throw Error(SNICIT_E_STATE, "no completed run");
    if (p->B == 0) return;
    ensure_stage(p);
    Net* net = p->net;
    const int N = net->N;
    hipStream_t s = p->stream;
    // Assemble Y_final[n * ld + b] in original sample order inside d_stage's companion
    // scratch (the first plan buffer not holding final state), then transpose to [B][N].
    float* full = nullptr;
    full = dev_alloc<float>((size_t)N * p->ld);
    try {
      if (plain_mode(p->cfg.mode)) {
        if (p->last_list) {
          launch_fill_f32(full, 0.f, (int64_t)N * p->ld, s);
          launch_scatter_cols(p->last_Y, p->ld, p->d_map, p->d_orig[p->last_orig],
                              p->d_counts + p->stop_layer, p->cap, N, full, p->ld, s);
        } else {
          SNICIT_HIP(hipMemcpyAsync(full, p->last_Y, (size_t)N * p->ld * 4,
                                    hipMemcpyDeviceToDevice, s));
        }
      } else {
        launch_expand_centroids(p->d_rec, p->d_counts, p->cap, p->last_C, p->kcap, N, full, p->ld,
                                s);
        launch_scatter_delta(p->last_D, p->ld, last_D_count(p), p->d_mapD,
                             p->d_mcar[p->last_mcar], p->d_orig[p->last_orig], p->d_nerow + net->L,
                             p->cap, p->last_C, p->kcap, N, full, p->ld, s);
      }
      launch_transpose_out(full, p->ld, p->B, N, p->d_stage, s);
      SNICIT_HIP(hipMemcpyAsync(y, p->d_stage, (size_t)p->B * N * 4, hipMemcpyDeviceToHost, s));
      SNICIT_HIP(hipStreamSynchronize(s));
    } catch (...) {
      dev_free(full);
      throw;
    }
    dev_free(full);
  });
}

int64_t snicit_plan_trace(snicit_plan* p, int32_t what, void* out, int64_t cap) {
  int64_t n = -1;
  int rc = guarded([&] {
    SNICIT_REQUIRE(p != nullptr, "null plan");
    if (!p->have_run) throw Error(SNICIT_E_STATE, "no completed run");
    const int L = p->net->L;
    const int t = p->cfg.threshold;
    const bool snic = !plain_mode(p->cfg.mode);
    // live columns (the lists may carry dead ones between compactions)
    std::vector<int32_t> counts(L + 4), nerow(L + 2);
    SNICIT_HIP(hipMemcpy(counts.data(), p->d_live, counts.size() * 4, hipMemcpyDeviceToHost));
    SNICIT_HIP(hipMemcpy(nerow.data(), p->d_nerow, nerow.size() * 4, hipMemcpyDeviceToHost));
    int32_t k = 0;
    SNICIT_HIP(hipMemcpy(&k, p->d_k, 4, hipMemcpyDeviceToHost));
    const int64_t live_t =
        snic ? (compacting(p) ? counts[t == 0 ? 1 : t] : counts[0]) : 0;
    const int pool = snic ? (int)std::min<int64_t>(p->cfg.sample_size, live_t) : 0;
    switch (what) {
      case 0: {  // live rows after each layer
        n = L;
        if (out) {
          int64_t* o = static_cast<int64_t*>(out);
          for (int l = 0; l < std::min<int64_t>(L, cap); ++l) {
            int64_t v;
            if (!compacting(p)) v = -1;
            else if (snic && l >= t) v = -1;
            else if (l + 1 <= p->stop_layer || snic) v = counts[l + 1];
            else v = -1;
            o[l] = v;
          }
        }
        break;
      }
      case 1: {
        n = L;
        if (out) {
          int64_t* o = static_cast<int64_t*>(out);
          for (int l = 0; l < std::min<int64_t>(L, cap); ++l)
            o[l] = (snic && l >= t) ? nerow[l + 1] : 0;
        }
        break;
      }
      case 2:
      case 4: {
        if (!snic) throw Error(SNICIT_E_STATE, "trace needs a SNICIT run");
        if (what == 2 && p->cfg.cluster == SNICIT_CLUSTER_LOSSY) {
          n = p->B > 0 ? p->cfg.lossy_pool : 0;
          std::vector<int32_t> pc(std::max<int64_t>(n, 1));
          if (n) SNICIT_HIP(hipMemcpy(pc.data(), p->d_poolcol, n * 4, hipMemcpyDeviceToHost));
          if (out) std::memcpy(out, pc.data(), std::min<int64_t>(n, cap) * 4);
          break;
        }
        // pool rows / centroids as original ids: the orig list at t
        std::vector<int32_t> slots;
        if (what == 2) {
          for (int s = 0; s < pool; ++s) slots.push_back(s);
        } else {
          std::vector<int32_t> cs(std::max(k, 1));
          SNICIT_HIP(hipMemcpy(cs.data(), p->d_cent_slot, k * 4, hipMemcpyDeviceToHost));
          slots.assign(cs.begin(), cs.begin() + k);
        }
        // rebuild the orig list at t on the host from the per-sample record: live rows at t
        // in ascending original id are exactly the compacted list.
        std::vector<int32_t> rec(p->B);
        SNICIT_HIP(hipMemcpy(rec.data(), p->d_rec, p->B * 4, hipMemcpyDeviceToHost));
        std::vector<int32_t> live;
        for (int64_t b = 0; b < p->B; ++b)
          if (rec[b] >= 0) live.push_back((int32_t)b);
        n = (int64_t)slots.size();
        if (out) {
          int32_t* o = static_cast<int32_t*>(out);
          for (int64_t i = 0; i < std::min<int64_t>(n, cap); ++i) o[i] = live[slots[i]];
        }
        break;
      }
      case 3: {
        if (!snic) throw Error(SNICIT_E_STATE, "trace needs a SNICIT run");
        n = p->cfg.cluster == SNICIT_CLUSTER_LOSSY ? 0 : (int64_t)pool * p->cfg.segments;
        std::vector<float> F(32 * 32);
        SNICIT_HIP(hipMemcpy(F.data(), p->d_F, F.size() * 4, hipMemcpyDeviceToHost));
        if (out) std::memcpy(out, F.data(), std::min<int64_t>(n, cap) * 4);
        break;
      }
      case 5: {
        if (!snic) throw Error(SNICIT_E_STATE, "trace needs a SNICIT run");
        n = p->B;
        if (out && p->B) {
          std::vector<int32_t> rec(p->B);
          SNICIT_HIP(hipMemcpy(rec.data(), p->d_rec, p->B * 4, hipMemcpyDeviceToHost));
          std::memcpy(out, rec.data(), std::min<int64_t>(n, cap) * 4);
        }
        break;
      }
      case 6: {  // output-layer logits, B x Nout, sample-major
        const Net& net = *p->net;
        if (net.Nout == 0) throw Error(SNICIT_E_STATE, "net has no output layer");
        n = p->B * net.Nout;
        if (out && p->B) {
          ensure_stage(p);
          hipStream_t s = p->stream;
          launch_transpose_out(p->d_logits, p->ld, p->B, net.Nout, p->d_stage, s);
          SNICIT_HIP(hipStreamSynchronize(s));
          std::vector<float> lg((size_t)n);
          SNICIT_HIP(hipMemcpy(lg.data(), p->d_stage, lg.size() * 4, hipMemcpyDeviceToHost));
          std::memcpy(out, lg.data(), std::min<int64_t>(n, cap) * 4);
        }
        break;
      }
      default:
        throw Error(SNICIT_E_INVALID, "unknown trace kind");
    }
  });
  return rc == SNICIT_OK ? n : rc;
}

int snicit_plan_profile(snicit_plan* p, double* ms, int64_t* rows, double* cent_ms,
                        int64_t cap) {
  return guarded([&] {
    SNICIT_REQUIRE(p != nullptr, "null plan");
    SNICIT_HIP(hipSetDevice(p->net->ctx->device));
    const int L = p->net->L;
    if (p->prof_ev.empty()) {
      p->prof_ev.resize((size_t)4 * L);
      for (auto& e : p->prof_ev) SNICIT_HIP(hipEventCreate(&e));
    }
    const int32_t saved_graph = p->cfg.use_graph;
    p->cfg.use_graph = 0;
    p->profiling = true;
    try {
      plan_run(p, -1, nullptr);
    } catch (...) {
      p->profiling = false;
      p->cfg.use_graph = saved_graph;
      throw;
    }
    p->profiling = false;
    p->cfg.use_graph = saved_graph;
    std::vector<int32_t> counts(L + 4), nerow(L + 2);
    SNICIT_HIP(hipMemcpy(counts.data(), p->d_counts, counts.size() * 4, hipMemcpyDeviceToHost));
    SNICIT_HIP(hipMemcpy(nerow.data(), p->d_nerow, nerow.size() * 4, hipMemcpyDeviceToHost));
    const bool snic = !plain_mode(p->cfg.mode);
    const int t = snic ? p->cfg.threshold : L;
    const std::vector<int32_t> list = snic ? post_list_lengths(p, nerow) : nerow;
    for (int l = 0; l < std::min<int64_t>(L, cap); ++l) {
      float a = 0.f, c = 0.f;
      SNICIT_HIP(hipEventElapsedTime(&a, p->prof_ev[4 * l], p->prof_ev[4 * l + 1]));
      if (snic && l >= t)
        SNICIT_HIP(hipEventElapsedTime(&c, p->prof_ev[4 * l + 2], p->prof_ev[4 * l + 3]));
      if (ms) ms[l] = a;
      if (cent_ms) cent_ms[l] = c;
      if (rows) rows[l] = (snic && l >= t) ? list[l] : (compacting(p) ? counts[l] : counts[0]);
    }
  });
}

int snicit_infer(snicit_net* net, const snicit_run_cfg* cfg, const float* y0, int64_t B,
                 int64_t batch, int32_t* cat, snicit_stats* stats) {
  return guarded([&] {
    SNICIT_REQUIRE(net && cfg && (y0 || B == 0) && (cat || B == 0), "null argument");
    SNICIT_REQUIRE(batch > 0, "batch must be positive");
    SNICIT_REQUIRE(B % batch == 0, "batch must divide the number of inputs");
    snicit_plan* p = nullptr;
    int rc = snicit_plan_create(net, cfg, batch, &p);
    if (rc) throw Error(rc, g_last_error);
    std::unique_ptr<snicit_plan, int (*)(snicit_plan*)> hold(p, snicit_plan_destroy);
    snicit_stats acc{};
    const int64_t width = net->Nin > 0 ? net->Nin : net->N;  // samples are Nin or N floats
    for (int64_t off = 0; off < B; off += batch) {
      rc = snicit_plan_set_input(p, y0 + (size_t)off * width, batch);
      if (rc) throw Error(rc, g_last_error);
      snicit_stats st{};
      rc = snicit_plan_run(p, &st);
      if (rc) throw Error(rc, g_last_error);
      rc = snicit_plan_categories(p, cat + off);
      if (rc) throw Error(rc, g_last_error);
      acc.ms_total += st.ms_total;
      acc.ms_pre += st.ms_pre;
      acc.ms_cluster += st.ms_cluster;
      acc.ms_post += st.ms_post;
      acc.ms_recover += st.ms_recover;
      acc.edges += st.edges;
      acc.batch += st.batch;
      acc.active += st.active;
      acc.rows_processed += st.rows_processed;
      acc.layers = st.layers;
      acc.centroids = std::max(acc.centroids, st.centroids);
      acc.live_at_threshold += st.live_at_threshold;
      acc.nerow_first += st.nerow_first;
      acc.nerow_last += st.nerow_last;
    }
    if (stats) *stats = acc;
  });
}

// The one-call form with every device buffer taken from the static scratch: no hipMalloc or
// hipFree, no graph capture, no events (the stage times in stats are 0).
int snicit_infer_noalloc(snicit_net* net, const snicit_run_cfg* cfg, const float* y0, int64_t B,
                         int64_t batch, int32_t* cat, snicit_stats* stats) {
  return guarded([&] {
    SNICIT_REQUIRE(net && cfg && (y0 || B == 0) && (cat || B == 0), "null argument");
    SNICIT_REQUIRE(batch > 0, "batch must be positive");
    SNICIT_REQUIRE(B % batch == 0, "batch must divide the number of inputs");
    SNICIT_REQUIRE(!net->streamed(),
                   "nets with streamed weights need copy streams and events; use snicit_infer");
    SNICIT_HIP(hipSetDevice(net->ctx->device));
    ScratchSlot* slot = scratch_slot(net->ctx->device);
    std::lock_guard<std::mutex> one_run(slot->run);
    DevArena arena;
    arena.base = slot->ptr;
    arena.cap = kScratchBytes;
    snicit_run_cfg c = *cfg;
    c.use_graph = 0;
    struct ArenaScope {
      explicit ArenaScope(DevArena* a) { g_dev_arena = a; }
      ~ArenaScope() { g_dev_arena = nullptr; }
    } scope(&arena);
    snicit_plan* p = nullptr;
    int rc = snicit_plan_create(net, &c, batch, &p);
    if (rc) throw Error(rc, g_last_error);
    // destroyed while the arena is still set: its buffers are arena memory
    std::unique_ptr<snicit_plan, int (*)(snicit_plan*)> hold(p, snicit_plan_destroy);
    snicit_stats acc{};
    const int64_t width = net->Nin > 0 ? net->Nin : net->N;
    for (int64_t off = 0; off < B; off += batch) {
      rc = snicit_plan_set_input(p, y0 + (size_t)off * width, batch);
      if (rc) throw Error(rc, g_last_error);
      snicit_stats st{};
      rc = snicit_plan_run(p, &st);
      if (rc) throw Error(rc, g_last_error);
      rc = snicit_plan_categories(p, cat + off);
      if (rc) throw Error(rc, g_last_error);
      acc.edges += st.edges;
      acc.batch += st.batch;
      acc.active += st.active;
      acc.rows_processed += st.rows_processed;
      acc.layers = st.layers;
      acc.centroids = std::max(acc.centroids, st.centroids);
      acc.live_at_threshold += st.live_at_threshold;
      acc.nerow_first += st.nerow_first;
      acc.nerow_last += st.nerow_last;
    }
    if (stats) *stats = acc;
  });
}

// ---- launch-shape autotuner -------------------------------------------------------------
}  // extern "C"

namespace {
std::string tune_key(const snicit_plan* p, int32_t density, uint32_t flags) {
  const Net& n = *p->net;
  const char* kind = n.is_ell(0)       ? "ell"
                     : n.mfma[0]       ? "panel-mfma"
                     : n.plain_mfma    ? "panel-mx"
                                       : "panel-valu";
  const int32_t K = n.is_ell(0) ? n.ell_K[0] : kSlots;
  const int64_t rows = n.is_ell(0) ? 1 : n.N / std::max<int64_t>(1, n.panels(0));
  int64_t bucket = 1;
  while (bucket < p->B) bucket <<= 1;
  char buf[96];
  std::snprintf(buf, sizeof(buf), "%s/N%d/K%d/R%lld/B%lld/D%d/G%d%s", kind, n.N, K,
                (long long)rows, (long long)bucket, density, n.grouping,
                (flags & SNICIT_TUNE_SAMPLE) ? "/S" : "");
  return buf;
}

// Input density in halving steps: 0 when at least half of the plan's input entries are
// nonzero, d when the nonzero fraction lies in [2^-(d+1), 2^-d), 31 for an all-zero input.
int32_t input_density(snicit_plan* p) {
  unsigned long long* d_n = dev_alloc<unsigned long long>(1);
  unsigned long long nz = 0;
  try {
    SNICIT_HIP(hipMemsetAsync(d_n, 0, sizeof(nz), p->stream));
    launch_count_nonzero(p->d_in, p->in_rows, p->ld, p->B, d_n, p->stream);
    SNICIT_HIP(hipMemcpyAsync(&nz, d_n, sizeof(nz), hipMemcpyDeviceToHost, p->stream));
    SNICIT_HIP(hipStreamSynchronize(p->stream));
  } catch (...) {
    dev_free(d_n);
    throw;
  }
  dev_free(d_n);
  if (nz == 0) return 31;
  const double total = (double)p->in_rows * (double)p->B;
  int32_t d = 0;
  while (d < 30 && (double)nz * std::ldexp(1.0, d + 1) < total) ++d;
  return d;
}

// Cache file: a JSON object, one entry per line:
//   "key": {"grid_target": G, "narrow": W, "ell_lds_kb": E, "ell_chunked": C,
//           "post_grid_target": Q, "post_narrow": V, "post_form": F, "sample_size": S,
//           "compact_slack": C, "plain_tiles": N, "ms": T, "ms_post": U},
struct TuneEntry {
  int64_t grid_target;
  int32_t narrow;
  int32_t ell_lds_kb;
  int32_t ell_chunked;
  int64_t post_grid_target;
  int32_t post_narrow;
  int32_t post_form;
  int32_t sample_size;
  int32_t compact_slack;
  int32_t plain_tiles;
  double ms;
  double ms_post;
};
std::vector<std::pair<std::string, TuneEntry>> tune_cache_read(const char* path) {
  std::vector<std::pair<std::string, TuneEntry>> out;
  if (!path) return out;
  FILE* f = std::fopen(path, "r");
  if (!f) return out;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    char key[128];
    long long g = 0, q = 0;
    int w = 0, e = 0, ch = 0, pn = 0, pf = 0, ss = 0, cs = 0, pt = 0;
    double ms = 0, msp = 0;
    if (std::sscanf(line,
                    " \"%127[^\"]\": {\"grid_target\": %lld, \"narrow\": %d, \"ell_lds_kb\": %d, "
                    "\"ell_chunked\": %d, \"post_grid_target\": %lld, \"post_narrow\": %d, "
                    "\"post_form\": %d, \"sample_size\": %d, \"compact_slack\": %d, "
                    "\"plain_tiles\": %d, \"ms\": %lf, \"ms_post\": %lf",
                    key, &g, &w, &e, &ch, &q, &pn, &pf, &ss, &cs, &pt, &ms, &msp) == 13)
      out.push_back({key, TuneEntry{g, w, e, ch, q, pn, pf, ss, cs, pt, ms, msp}});
  }
  std::fclose(f);
  return out;
}
void tune_cache_write(const char* path,
                      const std::vector<std::pair<std::string, TuneEntry>>& entries) {
  if (!path) return;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  SNICIT_REQUIRE(f != nullptr, std::string("cannot write tuning cache ") + tmp);
  std::fprintf(f, "{\n");
  for (size_t i = 0; i < entries.size(); ++i) {
    const TuneEntry& t = entries[i].second;
    std::fprintf(f,
                 "  \"%s\": {\"grid_target\": %lld, \"narrow\": %d, \"ell_lds_kb\": %d, "
                 "\"ell_chunked\": %d, \"post_grid_target\": %lld, \"post_narrow\": %d, "
                 "\"post_form\": %d, \"sample_size\": %d, \"compact_slack\": %d, "
                 "\"plain_tiles\": %d, \"ms\": %.6f, \"ms_post\": %.6f}%s\n",
                 entries[i].first.c_str(), (long long)t.grid_target, t.narrow, t.ell_lds_kb,
                 t.ell_chunked, (long long)t.post_grid_target, t.post_narrow, t.post_form,
                 t.sample_size, t.compact_slack, t.plain_tiles, t.ms, t.ms_post,
                 i + 1 < entries.size() ? "," : "");
  }
  std::fprintf(f, "}\n");
  std::fclose(f);
  SNICIT_REQUIRE(std::rename(tmp.c_str(), path) == 0,
                 std::string("cannot replace tuning cache ") + path);
}

// Median time of one plain launch of layer 0 over *count columns with the given shape.
double time_layer0(snicit_plan* p, const LaunchShape& shape, const int32_t* count,
                   hipEvent_t e0, hipEvent_t e1) {
  SpmmLaunch a = spmm_base(p, 0, EPI_PLAIN);
  a.yin = hidden_in(p);
  a.ldin = p->ld;
  a.yout = p->d_y[0];
  a.ldout = p->ld;
  a.count = count;
  a.shape = shape;
  hipStream_t s = p->stream;
  const Net& net = *p->net;
  if (net.streamed()) {  // layer 0's block into its slot, after any copy still running
    SNICIT_HIP(hipStreamSynchronize(p->copy_stream));
    SNICIT_HIP(hipMemcpyAsync(p->d_wslot[0], net.h_blocks + net.blk_off[0],
                              (size_t)net.blk_bytes[0], hipMemcpyHostToDevice, s));
  }
  for (int i = 0; i < 2; ++i) launch_layer(p, a, s);
  std::vector<float> ms(7);
  for (float& m : ms) {
    SNICIT_HIP(hipEventRecord(e0, s));
    launch_layer(p, a, s);
    SNICIT_HIP(hipEventRecord(e1, s));
    SNICIT_HIP(hipEventSynchronize(e1));
    SNICIT_HIP(hipEventElapsedTime(&m, e0, e1));
  }
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

// Medians over three whole runs (after one that builds the stage graphs) of the pre-threshold
// stage, the post stage and the whole run; stop < 0 runs every layer.
struct RunTimes {
  double pre, post, total;
};
RunTimes time_runs(snicit_plan* p, int32_t stop) {
  drop_graphs(p);
  snicit_stats st{};
  plan_run(p, stop, &st);
  double pre[3], post[3], tot[3];
  for (int i = 0; i < 3; ++i) {
    plan_run(p, stop, &st);
    pre[i] = st.ms_pre;
    post[i] = st.ms_post;
    tot[i] = st.ms_total;
  }
  std::sort(pre, pre + 3);
  std::sort(post, post + 3);
  std::sort(tot, tot + 3);
  return {pre[1], post[1], tot[1]};
}

std::vector<int32_t> run_categories(snicit_plan* p) {
  std::vector<int32_t> c((size_t)p->B);
  if (p->B) SNICIT_HIP(hipMemcpy(c.data(), p->d_cat, (size_t)p->B * 4, hipMemcpyDeviceToHost));
  return c;
}

// Whole runs on up to tries allocations of the activation ping-pong buffers (the plan's own
// first); keeps the fastest and frees the rest.  A fresh pair is tried only while the device
// has room for it and 2 GiB more.  The runs are issued directly, without stage graphs, so
// the candidates cost no graph captures; the kept buffers' graphs are captured by the next
// run.  Returns the whole-run time on the plan's own buffers.
double tune_placement(snicit_plan* p, int tries, snicit_tune_info& res) {
  struct Direct {
    snicit_plan* p;
    int32_t use_graph;
    ~Direct() { p->cfg.use_graph = use_graph; }
  } direct{p, p->cfg.use_graph};
  p->cfg.use_graph = 0;
  const size_t n = (size_t)p->net->N * p->ld;
  std::vector<std::pair<float*, float*>> cand{{p->d_y[0], p->d_y[1]}};
  const double own = time_runs(p, -1).total;
  double best = own;
  int keep = 0;
  try {
    for (int r = 1; r < tries; ++r) {
      size_t free_b = 0, total_b = 0;
      SNICIT_HIP(hipMemGetInfo(&free_b, &total_b));
      if (free_b < 2 * n * sizeof(float) + ((size_t)2 << 30)) break;
      float* a = dev_alloc_big<float>(n);
      float* b = nullptr;
      try {
        b = dev_alloc_big<float>(n);
      } catch (...) {
        dev_free(a);
        throw;
      }
      cand.push_back({a, b});
      p->d_y[0] = a;
      p->d_y[1] = b;
      const double t = time_runs(p, -1).total;
      if (t < best) {
        best = t;
        keep = r;
      }
    }
  } catch (...) {
    p->d_y[0] = cand[0].first;
    p->d_y[1] = cand[0].second;
    drop_graphs(p);
    for (size_t i = 1; i < cand.size(); ++i) {
      dev_free(cand[i].first);
      dev_free(cand[i].second);
    }
    throw;
  }
  p->d_y[0] = cand[keep].first;
  p->d_y[1] = cand[keep].second;
  drop_graphs(p);
  for (size_t i = 0; i < cand.size(); ++i)
    if ((int)i != keep) {
      dev_free(cand[i].first);
      dev_free(cand[i].second);
    }
  res.placements = (int32_t)cand.size();
  res.placement = keep;
  return own;
}

void plan_autotune(snicit_plan* p, const char* cache_path, uint32_t flags,
                   snicit_tune_info* info) {
  if (!p->have_input) throw Error(SNICIT_E_STATE, "autotuning needs the plan's input");
  SNICIT_REQUIRE(p->B > 0, "autotuning needs a non-empty batch");
  SNICIT_REQUIRE((flags & ~(SNICIT_TUNE_SAMPLE | SNICIT_TUNE_PLACEMENT)) == 0,
                 "unknown autotune flags");
  const bool snic = !plain_mode(p->cfg.mode) && p->cfg.threshold < p->net->L;
  const bool exact = p->cfg.mode == SNICIT_MODE_SNICIT && p->cfg.cluster == SNICIT_CLUSTER_EXACT;
  // whole runs need the plan's own stage events (the allocation-free entry point's plan has
  // none); a batch's plans run alone here, over all their segments, before the batch refines
  // them on whole-batch launches
  const bool whole = p->ev[0] != nullptr;
  const int32_t pre_stop = plain_mode(p->cfg.mode) ? std::min(p->net->L, 64) : -1;
  const std::string key = tune_key(p, input_density(p), flags);
  auto cache = tune_cache_read(cache_path);
  snicit_tune_info res{};
  std::snprintf(res.key, sizeof(res.key), "%s", key.c_str());
  bool found = false;
  for (const auto& e : cache)
    if (e.first == key) {
      p->shape.grid_target = e.second.grid_target;
      p->shape.narrow = e.second.narrow;
      p->shape.ell_lds_kb = e.second.ell_lds_kb;
      p->shape.ell_chunked = e.second.ell_chunked;
      p->shape.post_grid_target = e.second.post_grid_target;
      p->shape.post_narrow = e.second.post_narrow;
      p->shape.post_form = e.second.post_form;
      p->shape.compact_slack = e.second.compact_slack;
      p->shape.plain_tiles = e.second.plain_tiles;
      if ((flags & SNICIT_TUNE_SAMPLE) && exact && e.second.sample_size > 0)
        p->cfg.sample_size = e.second.sample_size;
      res.ms_tuned = e.second.ms;
      res.ms_post_tuned = e.second.ms_post;
      res.from_cache = 1;
      found = true;
    }
  if (!found) {
    hipStream_t s = p->stream;
    SNICIT_HIP(hipStreamSynchronize(s));
    const LaunchShape def{};
    p->shape = def;
    if (whole) {
      const RunTimes r0 = time_runs(p, -1);
      res.ms_run_default = r0.total;
      res.ms_pre_default = pre_stop < 0 ? r0.pre : time_runs(p, pre_stop).pre;
      res.ms_post_default = r0.post;
    }
    if (p->net->Nin > 0) enqueue_input_layer(p);  // layer 0 reads the input layer's output
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int32_t* d_cnt = nullptr;
    try {
      SNICIT_HIP(hipEventCreate(&e0));
      SNICIT_HIP(hipEventCreate(&e1));
      d_cnt = dev_alloc<int32_t>(1);
      // layer 0 at the full batch: wide-row forms, then the grid size
      LaunchShape best = def;
      res.ms_default = time_layer0(p, def, p->d_counts, e0, e1);
      double best_ms = res.ms_default;
      if (p->net->is_ell(0)) {
        const int32_t cand[][2] = {{0, 0}, {64, 0}, {128, 0}, {40, 1}, {144, 1}, {80, 2}};
        for (const auto& kc : cand) {
          LaunchShape sh = def;
          sh.ell_lds_kb = kc[0];
          sh.ell_chunked = kc[1];
          const double m = time_layer0(p, sh, p->d_counts, e0, e1);
          if (m < best_ms) {
            best_ms = m;
            best = sh;
          }
        }
      }
      const LaunchShape base = best;
      for (int64_t g : {2048, 4096, 8192, 32768, 65536, 131072}) {
        LaunchShape sh = base;
        sh.grid_target = g;
        const double m = time_layer0(p, sh, p->d_counts, e0, e1);
        if (m < best_ms) {
          best_ms = m;
          best = sh;
        }
      }
      // narrow path limit: the largest column count at which it beats the wavefront layout
      int32_t narrow = 0;
      for (int32_t c : {32, 64, 128, 256, 512}) {
        if (c > p->cap) break;
        SNICIT_HIP(hipMemcpy(d_cnt, &c, 4, hipMemcpyHostToDevice));
        LaunchShape nar = best, wide = best;
        nar.narrow = c;
        wide.narrow = 0;
        if (time_layer0(p, nar, d_cnt, e0, e1) < time_layer0(p, wide, d_cnt, e0, e1)) narrow = c;
        else break;
      }
      best.narrow = narrow;
      res.ms_tuned = best_ms;
      p->shape = best;
      if (whole) {
        // pre-threshold stage at the plan's own live counts: the grid around layer 0's
        // choice (plain modes: the first min(L, 64) layers)
        double best_pre = time_runs(p, pre_stop).pre;
        int64_t best_g = best.grid_target;
        for (int64_t g : {best.grid_target / 4, best.grid_target / 2, best.grid_target * 2,
                          best.grid_target * 4, def.grid_target}) {
          if (g < 1024 || g > 262144 || g == best.grid_target) continue;
          p->shape.grid_target = g;
          const double m = time_runs(p, pre_stop).pre;
          if (m < best_pre) {
            best_pre = m;
            best_g = g;
          }
        }
        p->shape.grid_target = best_g;
        // compaction slack (lists with dead columns carried, read aligned)
        if (compacting(p)) {
          int32_t best_slack = p->shape.compact_slack;
          for (int32_t sl : {0, 16, 32, 64, 128}) {
            if (sl == best_slack) continue;
            p->shape.compact_slack = sl;
            const double m = time_runs(p, pre_stop).pre;
            if (m < best_pre) {
              best_pre = m;
              best_slack = sl;
            }
          }
          p->shape.compact_slack = best_slack;
        }
        // 64-column tiles per wavefront of the plain vector kernel
        int32_t best_tiles = p->shape.plain_tiles;
        for (int32_t nt : {1, 2, 4}) {
          if (nt == best_tiles) continue;
          p->shape.plain_tiles = nt;
          const double m = time_runs(p, pre_stop).pre;
          if (m < best_pre) {
            best_pre = m;
            best_tiles = nt;
          }
        }
        p->shape.plain_tiles = best_tiles;
        res.ms_pre_tuned = best_pre;
      }
      // post-threshold phase: whole runs at the plan's own delta-row counts, timed by the
      // post stage's events; the grid, then the narrow-path limit, then the post-layer form
      if (snic && whole) {
        auto post_ms = [&](int64_t g, int32_t pn, int32_t form) {
          p->shape.post_grid_target = g;
          p->shape.post_narrow = pn;
          p->shape.post_form = form;
          return time_runs(p, -1).post;
        };
        double best_post = post_ms(def.post_grid_target, def.post_narrow, def.post_form);
        int64_t best_g = def.post_grid_target;
        int32_t best_n = def.post_narrow, best_f = def.post_form;
        for (int64_t g : {512, 1024, 4096, 8192, 16384}) {
          const double m = post_ms(g, best_n, best_f);
          if (m < best_post) {
            best_post = m;
            best_g = g;
          }
        }
        for (int32_t pn : {0, 32, 128, 256, 512}) {
          const double m = post_ms(best_g, pn, best_f);
          if (m < best_post) {
            best_post = m;
            best_n = pn;
          }
        }
        for (int32_t f : {0, 1}) {
          const double m = post_ms(best_g, best_n, f);
          if (m < best_post) {
            best_post = m;
            best_f = f;
          }
        }
        p->shape.post_grid_target = best_g;
        p->shape.post_narrow = best_n;
        p->shape.post_form = best_f;
        res.ms_post_tuned = best_post;
      }
      // clustering sample size (exact SNICIT, on request): whole runs, kept only when the
      // categories equal those of the configured size
      if ((flags & SNICIT_TUNE_SAMPLE) && exact && snic && whole) {
        const int32_t own = p->cfg.sample_size;
        double best_run = time_runs(p, -1).total;
        const std::vector<int32_t> want = run_categories(p);
        int32_t best_ss = own;
        for (int32_t ss : {8, 16, 24, 32}) {
          if (ss == own || ss > kMaxCentroidsExact) continue;
          p->cfg.sample_size = ss;
          const double m = time_runs(p, -1).total;
          if (m < best_run && run_categories(p) == want) {
            best_run = m;
            best_ss = ss;
          }
        }
        p->cfg.sample_size = best_ss;
      }
    } catch (...) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      dev_free(d_cnt);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dev_free(d_cnt);
    if (cache_path) {
      cache.push_back({key, TuneEntry{p->shape.grid_target, p->shape.narrow, p->shape.ell_lds_kb,
                                      p->shape.ell_chunked, p->shape.post_grid_target,
                                      p->shape.post_narrow, p->shape.post_form,
                                      (flags & SNICIT_TUNE_SAMPLE) ? p->cfg.sample_size : 0,
                                      p->shape.compact_slack, p->shape.plain_tiles, res.ms_tuned,
                                      res.ms_post_tuned}});
      tune_cache_write(cache_path, cache);
    }
  }
  // settings from the cache: the run before placement stands for the default
  res.placements = 1;
  if ((flags & SNICIT_TUNE_PLACEMENT) && whole) {
    const double own = tune_placement(p, 5, res);
    if (found) res.ms_run_default = own;
  }
  // the last whole run captured the stage graphs with the final settings: they are kept
  const bool final_run = whole && (!found || (flags & SNICIT_TUNE_PLACEMENT));
  if (final_run) res.ms_run_tuned = time_runs(p, -1).total;
  res.grid_target = p->shape.grid_target;
  res.narrow = p->shape.narrow;
  res.ell_lds_kb = p->shape.ell_lds_kb;
  res.ell_chunked = p->shape.ell_chunked;
  res.post_grid_target = p->shape.post_grid_target;
  res.post_narrow = p->shape.post_narrow;
  res.post_form = p->shape.post_form;
  res.sample_size = p->cfg.sample_size;
  res.compact_slack = p->shape.compact_slack;
  res.plain_tiles = p->shape.plain_tiles;
  if (!final_run) drop_graphs(p);
  p->have_run = false;
  if (info) *info = res;
}
}  // namespace

extern "C" {

int snicit_plan_autotune(snicit_plan* plan, const char* cache_path, snicit_tune_info* info) {
  return snicit_plan_autotune_ex(plan, cache_path, 0, info);
}

int snicit_plan_autotune_ex(snicit_plan* plan, const char* cache_path, uint32_t flags,
                            snicit_tune_info* info) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_HIP(hipSetDevice(plan->net->ctx->device));
    plan_autotune(plan, cache_path, flags, info);
  });
}

int snicit_plan_set_shape(snicit_plan* plan, int64_t grid_target, int32_t narrow,
                          int32_t ell_lds_kb, int32_t ell_chunked) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_REQUIRE(grid_target >= 1 && narrow >= 0 && ell_lds_kb >= 0 && ell_lds_kb <= 160 &&
                       (ell_chunked >= 0 && ell_chunked <= 2),
                   "bad launch shape");
    plan->shape.grid_target = grid_target;
    plan->shape.narrow = narrow;
    plan->shape.ell_lds_kb = ell_lds_kb;
    plan->shape.ell_chunked = ell_chunked;
    drop_graphs(plan);
  });
}

int snicit_plan_set_post_shape(snicit_plan* plan, int64_t post_grid_target,
                               int32_t post_narrow) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_REQUIRE(post_grid_target >= 1 && post_narrow >= 0, "bad post launch shape");
    plan->shape.post_grid_target = post_grid_target;
    plan->shape.post_narrow = post_narrow;
    drop_graphs(plan);
  });
}

int snicit_plan_set_plain_tiles(snicit_plan* plan, int32_t tiles) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_REQUIRE(tiles == 1 || tiles == 2 || tiles == 4, "plain_tiles must be 1, 2 or 4");
    plan->shape.plain_tiles = tiles;
    drop_graphs(plan);
  });
}

int snicit_plan_set_compact_slack(snicit_plan* plan, int32_t slack) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_REQUIRE(slack >= 0 && slack < 1024, "compact_slack must be in [0, 1024)");
    plan->shape.compact_slack = slack;
    drop_graphs(plan);
  });
}

int snicit_plan_set_post_form(snicit_plan* plan, int32_t post_form) {
  return guarded([&] {
    SNICIT_REQUIRE(plan != nullptr, "null plan");
    SNICIT_REQUIRE(post_form >= 0 && post_form <= 2, "post_form must be 0, 1 or 2");
    plan->shape.post_form = post_form;
    drop_graphs(plan);
  });
}

// ---- batches ------------------------------------------------------------------------
struct snicit_batch {
  Ctx* ctx = nullptr;
  std::vector<snicit_plan*> plans;   // one per net
  std::vector<hipStream_t> streams;  // streams[0] is the context's (not owned)
  struct Seg {
    int32_t net;
    int64_t off, samples;
    bool loaded;
  };
  std::vector<Seg> segs;
  hipGraphExec_t graph = nullptr;
  hipEvent_t ev[2] = {};
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  bool have_run = false;
};

}  // extern "C"

namespace {
void batch_free(snicit_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (snicit_plan* p : b->plans) {
    if (!p) continue;
    (void)hipStreamSynchronize(p->stream);
    --p->net->live_plans;
    plan_free(p);
    delete p;
  }
  for (size_t i = 1; i < b->streams.size(); ++i)
    if (b->streams[i]) (void)hipStreamDestroy(b->streams[i]);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (hipEvent_t e : b->join)
    if (e) (void)hipEventDestroy(e);
}

snicit_batch::Seg& batch_seg(snicit_batch* b, int32_t seg) {
  SNICIT_REQUIRE(b != nullptr, "null batch");
  SNICIT_REQUIRE(seg >= 0 && seg < (int32_t)b->segs.size(), "segment index out of range");
  return b->segs[seg];
}

// One graph over every net: a fork from the context's stream to one stream per net, each
// net's whole layer loop, and a join back.
void batch_capture(snicit_batch* b) {
  hipStream_t s0 = b->streams[0];
  hipGraph_t graph = nullptr;
  SNICIT_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  try {
    SNICIT_HIP(hipEventRecord(b->fork, s0));
    for (size_t i = 1; i < b->plans.size(); ++i)
      SNICIT_HIP(hipStreamWaitEvent(b->streams[i], b->fork, 0));
    for (snicit_plan* p : b->plans) enqueue_all(p);
    for (size_t i = 1; i < b->plans.size(); ++i) {
      SNICIT_HIP(hipEventRecord(b->join[i], b->streams[i]));
      SNICIT_HIP(hipStreamWaitEvent(s0, b->join[i], 0));
    }
  } catch (...) {
    hipGraph_t junk = nullptr;
    (void)hipStreamEndCapture(s0, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    throw;
  }
  SNICIT_HIP(hipStreamEndCapture(s0, &graph));
  hipError_t e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  SNICIT_HIP(e);
}
}  // namespace

extern "C" {

int snicit_batch_create(snicit_ctx* ctx, snicit_net* const* nets, const snicit_run_cfg* cfgs,
                        int32_t n_nets, const snicit_segment* segs, int32_t n_segs,
                        snicit_batch** out) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx && nets && cfgs && segs && out, "null argument");
    SNICIT_REQUIRE(n_nets >= 1 && n_segs >= 1, "a batch needs nets and segments");
    SNICIT_HIP(hipSetDevice(ctx->device));
    auto b = std::make_unique<snicit_batch>();
    b->ctx = ctx;
    std::vector<std::vector<int32_t>> starts(n_nets);
    std::vector<int64_t> total(n_nets, 0);
    for (int32_t i = 0; i < n_segs; ++i) {
      const int32_t n = segs[i].net;
      SNICIT_REQUIRE(n >= 0 && n < n_nets, "segment " + std::to_string(i) + " names no net");
      SNICIT_REQUIRE(segs[i].samples > 0, "segment " + std::to_string(i) + " is empty");
      starts[n].push_back((int32_t)total[n]);
      b->segs.push_back({n, total[n], segs[i].samples, false});
      total[n] += segs[i].samples;
      SNICIT_REQUIRE(total[n] < (int64_t(1) << 31) - 1024, "too many samples for one net");
    }
    try {
      b->streams.push_back(ctx->stream);
      b->join.assign(n_nets, nullptr);
      for (int i = 0; i < 2; ++i) SNICIT_HIP(hipEventCreateWithFlags(&b->ev[i], 0));
      SNICIT_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
      for (int32_t n = 0; n < n_nets; ++n) {
        SNICIT_REQUIRE(nets[n] && nets[n]->ctx == ctx, "nets must belong to the batch's context");
        SNICIT_REQUIRE(total[n] > 0, "net " + std::to_string(n) + " has no segment");
        if (n > 0) {
          hipStream_t st = nullptr;
          SNICIT_HIP(hipStreamCreate(&st));
          b->streams.push_back(st);
          SNICIT_HIP(hipEventCreateWithFlags(&b->join[n], hipEventDisableTiming));
        }
        std::vector<int32_t> st = starts[n];
        st.push_back((int32_t)total[n]);
        b->plans.push_back(plan_make(nets[n], &cfgs[n], total[n], st, b->streams[n]));
        set_count(b->plans.back(), total[n]);
        b->plans.back()->have_input = false;
      }
    } catch (...) {
      batch_free(b.get());
      throw;
    }
    *out = b.release();
  });
}

int snicit_batch_destroy(snicit_batch* batch) {
  return guarded([&] {
    if (!batch) return;
    batch_free(batch);
    delete batch;
  });
}

int snicit_batch_set_input(snicit_batch* batch, int32_t seg, const float* y0) {
  return guarded([&] {
    snicit_batch::Seg& sg = batch_seg(batch, seg);
    SNICIT_REQUIRE(y0 != nullptr, "null input");
    SNICIT_HIP(hipSetDevice(batch->ctx->device));
    snicit_plan* p = batch->plans[sg.net];
    load_host_input(p, y0, sg.samples, sg.off);
    SNICIT_HIP(hipStreamSynchronize(p->stream));
    sg.loaded = true;
  });
}

int snicit_batch_set_input_images(snicit_batch* batch, int32_t seg, uint64_t seed,
                                  int64_t first_sample) {
  return guarded([&] {
    snicit_batch::Seg& sg = batch_seg(batch, seg);
    SNICIT_HIP(hipSetDevice(batch->ctx->device));
    snicit_plan* p = batch->plans[sg.net];
    load_images(p, seed, first_sample, sg.samples, sg.off);
    SNICIT_HIP(hipStreamSynchronize(p->stream));
    sg.loaded = true;
  });
}

}  // extern "C"

namespace {
// One launch of the batch graph (captured first when there is none); returns its HIP-event
// time.  Every segment must hold its input.
double batch_run(snicit_batch* batch, snicit_stats* per_net) {
  hipStream_t s0 = batch->streams[0];
  for (snicit_plan* p : batch->plans) SNICIT_HIP(hipStreamSynchronize(p->stream));
  // before the capture: the recovery stage reads where the post stage leaves its state
  for (snicit_plan* p : batch->plans) plan_bookkeeping(p, -1);
  if (!batch->graph) batch_capture(batch);
  SNICIT_HIP(hipEventRecord(batch->ev[0], s0));
  SNICIT_HIP(hipGraphLaunch(batch->graph, s0));
  SNICIT_HIP(hipEventRecord(batch->ev[1], s0));
  SNICIT_HIP(hipEventSynchronize(batch->ev[1]));
  float t = 0.f;
  SNICIT_HIP(hipEventElapsedTime(&t, batch->ev[0], batch->ev[1]));
  for (size_t i = 0; i < batch->plans.size(); ++i) {
    snicit_plan* p = batch->plans[i];
    p->have_input = p->have_run = true;
    if (per_net) plan_stats(p, &per_net[i], nullptr);
  }
  batch->have_run = true;
  return t;
}

void batch_drop_graph(snicit_batch* batch) {
  if (batch->graph) {
    (void)hipGraphExecDestroy(batch->graph);
    batch->graph = nullptr;
  }
}

// Median of three batch launches after one that captures the batch graph.
double batch_time(snicit_batch* batch) {
  batch_drop_graph(batch);
  batch_run(batch, nullptr);
  double v[3];
  for (double& m : v) m = batch_run(batch, nullptr);
  std::sort(v, v + 3);
  return v[1];
}

// Each plan's layer-0 shape (plan_autotune), then, on whole-batch launches, each plan's
// pre-threshold grid and (SNICIT modes) post grid, narrow-path limit and post-layer form by
// coordinate search; the refined shapes replace the plans' cache entries.
void batch_autotune(snicit_batch* batch, const char* cache_path, snicit_tune_info* info) {
  const size_t n = batch->plans.size();
  std::vector<snicit_tune_info> res(n);
  const LaunchShape def{};
  for (snicit_plan* p : batch->plans) {
    p->have_input = true;
    p->shape = def;
  }
  bool cached = true;
  for (size_t i = 0; i < n; ++i) {
    plan_autotune(batch->plans[i], cache_path, 0, &res[i]);
    cached = cached && res[i].from_cache;
  }
  if (!cached) {
    // the batch with default shapes, for the record
    std::vector<LaunchShape> tuned(n);
    for (size_t i = 0; i < n; ++i) {
      tuned[i] = batch->plans[i]->shape;
      batch->plans[i]->shape = def;
    }
    const double ms_default = batch_time(batch);
    for (size_t i = 0; i < n; ++i) batch->plans[i]->shape = tuned[i];
    double best = batch_time(batch);
    auto probe = [&](snicit_plan* p, LaunchShape sh) {
      const LaunchShape keep = p->shape;
      p->shape = sh;
      const double m = batch_time(batch);
      if (m < best) best = m;
      else p->shape = keep;
    };
    for (snicit_plan* p : batch->plans) {
      const int64_t g0 = p->shape.grid_target;
      for (int64_t g : {g0 / 2, g0 * 2, def.grid_target}) {
        if (g < 1024 || g > 262144 || g == p->shape.grid_target) continue;
        LaunchShape sh = p->shape;
        sh.grid_target = g;
        probe(p, sh);
      }
      if (compacting(p))
        for (int32_t sl : {0, 64}) {
          LaunchShape sh = p->shape;
          sh.compact_slack = sl;
          if (sl != p->shape.compact_slack) probe(p, sh);
        }
      for (int32_t nt : {1, 4}) {
        LaunchShape sh = p->shape;
        sh.plain_tiles = nt;
        if (nt != p->shape.plain_tiles) probe(p, sh);
      }
      if (plain_mode(p->cfg.mode) || p->cfg.threshold >= p->net->L) continue;
      for (int64_t g : {512, 1024, 4096, 8192}) {
        LaunchShape sh = p->shape;
        sh.post_grid_target = g;
        probe(p, sh);
      }
      for (int32_t pn : {0, 256, 512}) {
        LaunchShape sh = p->shape;
        sh.post_narrow = pn;
        probe(p, sh);
      }
      for (int32_t f : {0, 1}) {
        LaunchShape sh = p->shape;
        sh.post_form = f;
        probe(p, sh);
      }
    }
    for (size_t i = 0; i < n; ++i) {
      const LaunchShape& sh = batch->plans[i]->shape;
      res[i].grid_target = sh.grid_target;
      res[i].post_grid_target = sh.post_grid_target;
      res[i].post_narrow = sh.post_narrow;
      res[i].post_form = sh.post_form;
      res[i].compact_slack = sh.compact_slack;
      res[i].plain_tiles = sh.plain_tiles;
      res[i].ms_run_default = ms_default;
      res[i].ms_run_tuned = best;
    }
    if (cache_path) {
      auto cache = tune_cache_read(cache_path);
      for (size_t i = 0; i < n; ++i) {
        const LaunchShape& sh = batch->plans[i]->shape;
        for (auto& e : cache)
          if (e.first == res[i].key) {
            e.second.grid_target = sh.grid_target;
            e.second.post_grid_target = sh.post_grid_target;
            e.second.post_narrow = sh.post_narrow;
            e.second.post_form = sh.post_form;
            e.second.compact_slack = sh.compact_slack;
            e.second.plain_tiles = sh.plain_tiles;
          }
      }
      tune_cache_write(cache_path, cache);
    }
  }
  batch_drop_graph(batch);
  batch->have_run = false;
  for (snicit_plan* p : batch->plans) p->have_run = false;
  if (info)
    for (size_t i = 0; i < n; ++i) info[i] = res[i];
}
}  // namespace

extern "C" {

int snicit_batch_run(snicit_batch* batch, double* ms, snicit_stats* per_net) {
  return guarded([&] {
    SNICIT_REQUIRE(batch != nullptr, "null batch");
    for (const auto& sg : batch->segs)
      if (!sg.loaded) throw Error(SNICIT_E_STATE, "every segment needs its input before a run");
    SNICIT_HIP(hipSetDevice(batch->ctx->device));
    const double t = batch_run(batch, per_net);
    if (ms) *ms = t;
  });
}

int snicit_batch_categories(snicit_batch* batch, int32_t seg, int32_t* cat) {
  return guarded([&] {
    snicit_batch::Seg& sg = batch_seg(batch, seg);
    SNICIT_REQUIRE(cat != nullptr, "null output");
    if (!batch->have_run) throw Error(SNICIT_E_STATE, "no completed run");
    SNICIT_HIP(hipSetDevice(batch->ctx->device));
    const snicit_plan* p = batch->plans[sg.net];
    SNICIT_HIP(hipMemcpy(cat, p->d_cat + sg.off, sg.samples * 4, hipMemcpyDeviceToHost));
  });
}

int snicit_batch_activations(snicit_batch* batch, int32_t seg, float* y) {
  return guarded([&] {
    snicit_batch::Seg& sg = batch_seg(batch, seg);
    SNICIT_REQUIRE(y != nullptr, "null output");
    if (!batch->have_run) throw Error(SNICIT_E_STATE, "no completed run");
    snicit_plan* p = batch->plans[sg.net];
    const int64_t N = p->net->N;
    std::vector<float> all((size_t)p->B * N);
    const int rc = snicit_plan_activations(p, all.data());
    if (rc != SNICIT_OK) throw Error(rc, snicit_last_error());
    std::memcpy(y, all.data() + (size_t)sg.off * N, (size_t)sg.samples * N * 4);
  });
}

int snicit_batch_autotune(snicit_batch* batch, const char* cache_path, snicit_tune_info* info) {
  return guarded([&] {
    SNICIT_REQUIRE(batch != nullptr, "null batch");
    for (const auto& sg : batch->segs)
      if (!sg.loaded) throw Error(SNICIT_E_STATE, "every segment needs its input first");
    SNICIT_HIP(hipSetDevice(batch->ctx->device));
    batch_autotune(batch, cache_path, info);
  });
}

int snicit_gen_sdgc_ell(int32_t N, int32_t L, uint64_t seed, int32_t* cols) {
  return guarded([&] {
    SNICIT_REQUIRE(cols != nullptr && L > 0, "bad argument");
    for (int32_t l = 0; l < L; ++l) sdgc_layer_ell(N, l, seed, cols + (size_t)l * N * kSlots);
  });
}

int snicit_gen_unstructured(int32_t Nin, int32_t N, int32_t L, int32_t K, int32_t Nout,
                            uint64_t seed, int32_t* cols, float* vals, float* w_in, float* w_out) {
  return guarded([&] {
    SNICIT_REQUIRE(cols && vals && N > 0 && L > 0 && K > 0 && K <= N, "bad argument");
    SNICIT_REQUIRE(Nin == 0 || w_in, "w_in needed when Nin > 0");
    SNICIT_REQUIRE(Nout == 0 || w_out, "w_out needed when Nout > 0");
    for (int32_t l = 0; l < L; ++l)
      unstructured_layer(N, K, l, seed, cols + (size_t)l * N * K, vals + (size_t)l * N * K);
    if (Nin > 0) dense_layer(N, Nin, seed, 0x1D1D1D1D1ULL, w_in);
    if (Nout > 0) dense_layer(Nout, N, seed, 0x0F0F0F0F0ULL, w_out);
  });
}

int snicit_gen_features(int32_t Nin, uint64_t seed, int64_t first_sample, int64_t B, float* out) {
  return guarded([&] {
    SNICIT_REQUIRE(Nin > 0 && (out || B == 0), "bad argument");
    uniform_features(Nin, seed, first_sample, B, out);
  });
}

int snicit_gen_images(int32_t N, uint64_t seed, int64_t first_sample, int64_t B, float* out) {
  return guarded([&] {
    SNICIT_REQUIRE(out != nullptr || B == 0, "null out");
    ImageGeom g = image_geometry(N, seed);
    for (int64_t b = 0; b < B; ++b)
      image_render(g, image_sample_params(g, seed, first_sample + b), N, out + (size_t)b * N);
  });
}

}  // extern "C"
