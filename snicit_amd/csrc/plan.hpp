// Plans (activation buffers of one batch shape on one net) and the helpers the C-ABI
// entry points, batches and the serving loop share.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "kernels.hpp"

struct snicit_ctx : snicit::Ctx {};
struct snicit_net : snicit::Net {};

struct snicit_plan {
  snicit::Net* net = nullptr;
  snicit_run_cfg cfg{};
  hipStream_t stream = nullptr;  // the context's stream, or a batch's per-net stream
  int64_t cap = 0, ld = 0;
  int64_t B = 0;
  bool have_input = false, have_run = false;
  // batch segments: columns [segstart[s], segstart[s+1]) are clustered on their own; one
  // segment covering every column for a plain plan
  int32_t nseg = 1;
  int32_t stride = snicit::kMaxCentroidsExact;  // centroid columns per segment
  int32_t* d_segstart = nullptr;        // [nseg + 1]
  int32_t* d_seglo = nullptr;           // [nseg + 1] segment bounds in the list at t
  int32_t* d_kseg = nullptr;            // [nseg] centroids per segment
  int32_t kcap = snicit::kMaxCentroidsExact;    // stride * nseg
  snicit::LaunchShape shape;                    // layer-kernel launch shape (autotuned)
  int32_t stop_layer = -1;  // layers actually run by the last run
  // activations
  float* d_in = nullptr;
  float* d_y[2] = {nullptr, nullptr};
  // bookkeeping
  int32_t* d_flags = nullptr;
  int32_t* d_map = nullptr;   // pre-threshold / plain live list
  int32_t* d_mapD = nullptr;  // post-threshold delta list
  int32_t* d_orig[2] = {nullptr, nullptr};
  int32_t* d_mcar[2] = {nullptr, nullptr};
  int32_t* d_counts = nullptr;  // [L + 4]: listed columns entering layer l; [L+1..] scratch
  int32_t* d_live = nullptr;    // [L + 4]: live columns entering layer l (the list may carry
                                // dead ones: compact_slack)
  int32_t* d_tilecnt = nullptr; // [compact_tiles(cap)] per-tile counts of the compaction
  int32_t* d_nerow = nullptr;   // [L + 2]: delta columns entering layer l
  // compact post layout (exact clustering, one segment, every post layer fused): the delta
  // list keeps its positions between compactions, which run at the layers post_compacts_after
  // names; d_nlist[l] = list positions layer l processes (l > t), d_ring = per-layer flags
  bool compact_post = false;
  int32_t* d_nlist = nullptr;   // [L + 2]
  int32_t* d_ring = nullptr;    // [kPostRing][ld]
  int32_t* d_k = nullptr;  // centroid columns in use (k with one segment)
  float* d_F = nullptr;
  int32_t* d_cent_slot = nullptr;
  float* d_C[2] = {nullptr, nullptr};
  // compact post layout: centroid columns of the post layers at the compact stride
  // compact_kst(k, kcap), ping-pong (d_C[1] receives the last layer's at stride kcap)
  float* d_Cp[2] = {nullptr, nullptr};
  // chain records of the post stage (layers whose panels have at most two distinct weight
  // rows): layer j's in slot j mod 2, written by the launch of layer j - 1
  float* d_zchain = nullptr;  // [2][chain_panels][kPostChainRec]
  int64_t chain_panels = 0;   // most panels of a post layer
  float* d_zc = nullptr;
  int32_t* d_rec = nullptr;
  int32_t* d_dist = nullptr;  // [cap][stride] distance accumulators (kept zero between runs)
  int32_t* d_catc = nullptr;
  int32_t* d_act = nullptr;
  int32_t* d_cat = nullptr;
  int32_t* d_img_params = nullptr;
  int64_t img_params_cap = 0;
  int32_t* d_img_blobs = nullptr;
  int32_t* d_img_proto = nullptr;
  float* d_stage = nullptr;  // [cap][max(N, Nin, Nout)] staging for host transfers
  int32_t* d_nonfinite = nullptr;  // [1] set when a host input holds inf or NaN
  // nets with dense input / output layers
  float* d_h0 = nullptr;      // [N][ld] input-layer output (hidden loop input), later the
                              //   recovered hidden activations
  float* d_logits = nullptr;  // [Nout][ld]
  int64_t in_rows = 0;        // rows of d_in: N, or Nin padded to 16
  // lossy clustering
  float* d_poolrows = nullptr;  // [nseg * pool][N]
  int32_t* d_poolcol = nullptr;
  int32_t* d_centmark = nullptr;
  uint64_t* d_close = nullptr;  // [nseg][128][2] pool pruning masks
  // where the last run left its final state
  int last_orig = 0, last_mcar = 0;
  const float* last_D = nullptr;
  const float* last_Y = nullptr;
  const float* last_C = nullptr;
  bool last_list = false;  // final columns selected by d_map (plain) / d_mapD (SNICIT)
  // graphs: stage 0 pre, 1 cluster, 2 post, 3 recover
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  int32_t graph_stop = -2;
  hipEvent_t ev[5] = {};
  int64_t device_bytes = 0;
  // per-launch profiling (eager only): events [4*l .. 4*l+3] = main start/stop, centroid
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;
  // streamed nets: device slots for layer blocks (layer l in slot l mod slots), the copy
  // stream, per-slot events (block copied, slot free) and the fork event of a stage
  std::vector<char*> d_wslot;
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> ev_copied, ev_free;
  hipEvent_t ev_fork = nullptr;
  std::vector<int32_t> slot_layer;  // layer each slot holds in the stage being enqueued
  bool copy_forked = false;         // the copy stream has joined the stage being enqueued
};

namespace snicit {

extern thread_local std::string g_last_error;

template <class F>
inline int guarded(F&& f) {
  try {
    f();
    return SNICIT_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SNICIT_E_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SNICIT_E_INVALID;
  }
}

// Device memory of the allocation-free entry point: while g_dev_arena is set on the calling
// thread, dev_alloc carves buffers from [base, base + cap) (256-byte aligned, never freed
// one by one) and dev_free leaves them alone.
struct DevArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
extern thread_local DevArena* g_dev_arena;

template <class T>
inline T* dev_alloc(size_t n, int64_t* acct = nullptr) {
  void* p = nullptr;
  if (n == 0) n = 1;
  if (DevArena* a = g_dev_arena) {
    const size_t off = (a->used + 255) / 256 * 256, bytes = n * sizeof(T);
    if (off > a->cap || bytes > a->cap - off)
      throw Error(SNICIT_E_NOMEM, "the run needs more device memory than the " +
                                      std::to_string(a->cap >> 20) + " MiB static scratch");
    a->used = off + bytes;
    if (acct) *acct += (int64_t)bytes;
    return reinterpret_cast<T*>(a->base + off);
  }
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? SNICIT_E_NOMEM : SNICIT_E_HIP,
                "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                    hipGetErrorString(e));
  if (acct) *acct += (int64_t)(n * sizeof(T));
  return static_cast<T*>(p);
}

// Activation buffers of 1 GiB and more: physically contiguous when the driver grants it,
// else hipMalloc.  Each wavefront gathers its 32 input rows one buffer row (ld floats, up to
// 240 KB) apart, so the span a layer touches depends on the address translation's reach:
// with ordinary allocations the C3 layer stages varied by up to 11% from one process to the
// next (where the buffers landed); contiguous buffers ran at one speed in every process.
template <class T>
inline T* dev_alloc_big(size_t n, int64_t* acct = nullptr) {
  if (!g_dev_arena && n * sizeof(T) >= ((size_t)1 << 30)) {
    void* p = nullptr;
    if (hipExtMallocWithFlags(&p, n * sizeof(T), hipDeviceMallocContiguous) == hipSuccess && p) {
      if (acct) *acct += (int64_t)(n * sizeof(T));
      return static_cast<T*>(p);
    }
    (void)hipGetLastError();  // refused: fall back to an ordinary allocation
  }
  return dev_alloc<T>(n, acct);
}

inline void dev_free(void* p) {
  if (DevArena* a = g_dev_arena) {
    const char* q = static_cast<const char*>(p);
    if (q >= a->base && q < a->base + a->cap) return;  // arena memory is never freed
  }
  if (p) (void)hipFree(p);
}

// Plan of one net over max_batch columns split into batch segments at segstart (nseg + 1
// entries, ascending; entries past the last used segment may repeat the total or exceed it).
snicit_plan* plan_make(snicit_net* net, const snicit_run_cfg* cfg, int64_t max_batch,
                       const std::vector<int32_t>& segstart, hipStream_t stream);
void plan_free(snicit_plan* p);
// Where a run leaves its final state; before any stage is enqueued.
void plan_bookkeeping(snicit_plan* p, int32_t stop);
// Every stage of one run on the plan's stream (for capture into a caller's graph).
void enqueue_all(snicit_plan* p);
// One full run on the plan's stream, stage graphs captured on first use; does not wait.
void plan_launch(snicit_plan* p);
void plan_stats(snicit_plan* p, snicit_stats* stats, const float* stage_ms);
void drop_graphs(snicit_plan* p);
void ensure_stage(snicit_plan* p);
void set_count(snicit_plan* p, int64_t B);
void load_host_input(snicit_plan* p, const float* y0, int64_t B, int64_t col_off);
void load_images(snicit_plan* p, uint64_t seed, int64_t first_sample, int64_t B,
                 int64_t col_off);

}  // namespace snicit
