// Serving loop: continuous batching of inference requests over resident nets.
//
// The reference serves nothing; its driver runs fixed rounds of one batch size
// (src/SDGC/SNICIT/SNICIT.hpp:87-203).  Here a worker thread forms launch groups from
// per-net request queues and runs them in slots: each slot is a plan sized for
// slot_samples samples in slot_segments batch segments, with its own stream, pinned
// staging buffers and stage graphs captured once.  The segment table of a group is copied
// into the plan's device table before the graphs replay, so launch sizes never depend on the
// group.  Completion is observed with an event per slot; results go to the waiting caller.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "plan.hpp"

using namespace snicit;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t id = 0;
  int32_t net = 0;
  int64_t samples = 0;
  bool images = false;
  uint64_t seed = 0;
  int64_t first = 0;
  std::vector<float> x;    // host input when it is not in the pinned arena
  float* pin = nullptr;    // host input in the server's pinned arena, or null
  int64_t pin_bytes = 0;
  Clock::time_point t_submit, t_launch, t_done;
  std::vector<int32_t> cat;
  int status = 0;  // 0 queued, 1 running, 2 done, < 0 failed (status code)
  std::string error;
  int64_t group_samples = 0;
  int32_t group_requests = 0;
  int32_t slot = -1;
};

struct Slot {
  int32_t net = 0;
  snicit_plan* plan = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  float* h_in = nullptr;     // pinned [slot_samples][width]
  int32_t* h_cat = nullptr;  // pinned [slot_samples]
  int32_t* h_seg = nullptr;  // pinned [nseg + 2]: segment starts, then the sample count
  std::vector<Request*> reqs;
  bool busy = false;
};

}  // namespace

struct snicit_server {
  snicit_ctx* ctx = nullptr;
  std::vector<snicit_net*> nets;
  std::vector<snicit_run_cfg> cfgs;
  snicit_server_cfg cfg{};
  std::vector<std::unique_ptr<Slot>> slots;
  int64_t device_bytes = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::deque<Request*>> queues;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  // pinned arena for request inputs: free blocks by offset, first fit, merged on release
  char* pin_base = nullptr;
  std::map<int64_t, int64_t> pin_free;
  std::mutex pin_mu;
  std::thread worker;
  int64_t requests_done = 0, samples_done = 0, groups = 0;
  // fixed-period grouping (cfg.round_ms > 0): the next period boundary, and per net the
  // last boundary passed while it had requests queued (requests submitted by then are due)
  Clock::time_point next_round;
  std::vector<Clock::time_point> due;
  std::vector<uint8_t> has_due;
};

namespace {

int32_t width_of(const snicit_net* n) { return n->Nin > 0 ? n->Nin : n->N; }

constexpr int64_t kPinAlign = 256;
// Inputs at least this large are copied to the device straight from their pinned block;
// smaller ones are packed into the slot's staging rows first (one copy for all of them).
constexpr int64_t kPinDirectBytes = int64_t(1) << 20;

// A block of `bytes` from the pinned arena, or null when no free block is large enough.
float* pin_alloc(snicit_server* S, int64_t bytes) {
  if (!S->pin_base) return nullptr;
  const int64_t need = (bytes + kPinAlign - 1) / kPinAlign * kPinAlign;
  std::lock_guard<std::mutex> g(S->pin_mu);
  for (auto it = S->pin_free.begin(); it != S->pin_free.end(); ++it) {
    if (it->second < need) continue;
    const int64_t off = it->first, len = it->second;
    S->pin_free.erase(it);
    if (len > need) S->pin_free[off + need] = len - need;
    return reinterpret_cast<float*>(S->pin_base + off);
  }
  return nullptr;
}

void pin_release(snicit_server* S, float* p, int64_t bytes) {
  const int64_t need = (bytes + kPinAlign - 1) / kPinAlign * kPinAlign;
  int64_t off = reinterpret_cast<char*>(p) - S->pin_base, len = need;
  std::lock_guard<std::mutex> g(S->pin_mu);
  auto next = S->pin_free.lower_bound(off);
  if (next != S->pin_free.end() && next->first == off + len) {
    len += next->second;
    next = S->pin_free.erase(next);
  }
  if (next != S->pin_free.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      off = prev->first;
      len += prev->second;
      S->pin_free.erase(prev);
    }
  }
  S->pin_free[off] = len;
}

void slot_free(Slot* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->plan) {
    --s->plan->net->live_plans;
    plan_free(s->plan);
    delete s->plan;
  }
  if (s->done) (void)hipEventDestroy(s->done);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->h_in) (void)hipHostFree(s->h_in);
  if (s->h_cat) (void)hipHostFree(s->h_cat);
  if (s->h_seg) (void)hipHostFree(s->h_seg);
}

// Clustering segments of a group: one per request, except that requests of an exact net
// smaller than cluster_min join the open segment until it holds cluster_min samples.
std::vector<int32_t> group_segments(const snicit_server* S, int32_t net,
                                    const std::vector<Request*>& g) {
  const snicit_run_cfg& c = S->cfgs[net];
  const bool merge = c.mode == SNICIT_MODE_SNICIT && c.cluster == SNICIT_CLUSTER_EXACT &&
                     S->cfg.cluster_min > 1;
  std::vector<int32_t> starts;
  int64_t off = 0, open = -1;  // samples in the open (mergeable) segment
  for (const Request* r : g) {
    if (merge && r->samples < S->cfg.cluster_min && open >= 0 && open < S->cfg.cluster_min) {
      open += r->samples;
    } else {
      starts.push_back((int32_t)off);
      open = (merge && r->samples < S->cfg.cluster_min) ? r->samples : S->cfg.cluster_min;
    }
    off += r->samples;
  }
  return starts;
}

// Copies the group's inputs and segment table to the device and launches its run; the
// slot's event marks completion.  Runs on the worker thread, outside the server lock.
void launch_group(snicit_server* S, Slot* sl) {
  snicit_plan* p = sl->plan;
  const snicit_net* net = S->nets[sl->net];
  const int32_t width = width_of(net);
  hipStream_t s = sl->stream;
  int64_t total = 0;
  for (const Request* r : sl->reqs) total += r->samples;
  // host inputs to the device staging rows of their samples: large pinned inputs straight
  // from their arena blocks, the others packed into the slot's pinned staging rows and
  // copied in runs of consecutive samples
  int64_t off = 0, run = -1;  // first sample of the open packed run
  bool any_host = false;
  auto flush = [&](int64_t end) {
    if (run < 0) return;
    SNICIT_HIP(hipMemcpyAsync(p->d_stage + (size_t)run * width, sl->h_in + (size_t)run * width,
                              (size_t)(end - run) * width * 4, hipMemcpyHostToDevice, s));
    run = -1;
  };
  for (const Request* r : sl->reqs) {
    const size_t bytes = (size_t)r->samples * width * 4;
    if (r->images) {
      flush(off);
    } else if (r->pin && (int64_t)bytes >= kPinDirectBytes) {
      flush(off);
      SNICIT_HIP(hipMemcpyAsync(p->d_stage + (size_t)off * width, r->pin, bytes,
                                hipMemcpyHostToDevice, s));
      any_host = true;
    } else {
      std::memcpy(sl->h_in + (size_t)off * width, r->pin ? r->pin : r->x.data(), bytes);
      if (run < 0) run = off;
      any_host = true;
    }
    off += r->samples;
  }
  flush(off);
  // rows of image requests in the staging buffer are overwritten by their renders below
  if (any_host) launch_transpose_in(p->d_stage, total, width, p->d_in, p->ld, nullptr, s);
  off = 0;
  for (const Request* r : sl->reqs) {
    if (r->images) load_images(p, r->seed, r->first, r->samples, off);
    off += r->samples;
  }
  // segment table (unused segments start at the total: empty) and the sample count
  const std::vector<int32_t> starts = group_segments(S, sl->net, sl->reqs);
  const int32_t nseg = p->nseg;
  for (int32_t i = 0; i <= nseg; ++i)
    sl->h_seg[i] = i < (int32_t)starts.size() ? starts[i] : (int32_t)total;
  sl->h_seg[nseg + 1] = (int32_t)total;
  SNICIT_HIP(hipMemcpyAsync(p->d_segstart, sl->h_seg, (nseg + 1) * 4, hipMemcpyHostToDevice, s));
  SNICIT_HIP(hipMemcpyAsync(p->d_counts, sl->h_seg + nseg + 1, 4, hipMemcpyHostToDevice, s));
  p->B = total;
  p->have_input = true;
  plan_launch(p);
  SNICIT_HIP(hipMemcpyAsync(sl->h_cat, p->d_cat, (size_t)total * 4, hipMemcpyDeviceToHost, s));
  SNICIT_HIP(hipEventRecord(sl->done, s));
}

// Samples at which a group closes.
int64_t group_cap(const snicit_server* S) {
  const int64_t g = S->cfg.group_samples;
  return g > 0 && g < S->cfg.slot_samples ? g : S->cfg.slot_samples;
}

bool rounds(const snicit_server* S) { return S->cfg.round_ms > 0.0; }

// With fixed periods, whether the queue head of `net` was submitted by its last boundary.
bool head_due(const snicit_server* S, int32_t net) {
  const auto& q = S->queues[net];
  return !q.empty() && S->has_due[net] && q.front()->t_submit <= S->due[net];
}

// Moves queued requests of `net` into the free slot while they fit (the first one always
// does: every request fits a slot); with fixed periods, only requests that are due.
void fill_group(snicit_server* S, Slot* sl) {
  auto& q = S->queues[sl->net];
  const int64_t cap = group_cap(S);
  int64_t total = 0;
  sl->reqs.clear();
  while (!q.empty() && (int32_t)sl->reqs.size() < S->cfg.slot_segments &&
         (!rounds(S) || head_due(S, sl->net)) &&
         (sl->reqs.empty() || total + q.front()->samples <= cap)) {
    Request* r = q.front();
    q.pop_front();
    total += r->samples;
    sl->reqs.push_back(r);
  }
}

bool group_ready(snicit_server* S, int32_t net, Clock::time_point now) {
  auto& q = S->queues[net];
  if (q.empty()) return false;
  if (rounds(S)) return head_due(S, net);
  bool net_busy = false;
  for (const auto& sl : S->slots)
    if (sl->net == net && sl->busy) net_busy = true;
  if (!net_busy) return true;
  if (ms_between(q.front()->t_submit, now) >= S->cfg.group_window_ms) return true;
  const int64_t cap = group_cap(S);
  int64_t total = 0;
  int32_t n = 0;
  for (const Request* r : q) {
    if (n == S->cfg.slot_segments || (n > 0 && total + r->samples > cap)) return true;
    total += r->samples;
    ++n;
  }
  return total >= cap;
}

void finish_group(snicit_server* S, Slot* sl, Clock::time_point now, int status,
                  const std::string& err) {
  int64_t off = 0, total = 0;
  for (const Request* r : sl->reqs) total += r->samples;
  for (Request* r : sl->reqs) {
    // the group's copies are done (or failed): the input is no longer needed
    if (r->pin) pin_release(S, r->pin, r->pin_bytes);
    r->pin = nullptr;
    std::vector<float>().swap(r->x);
    if (status == 0) r->cat.assign(sl->h_cat + off, sl->h_cat + off + r->samples);
    r->status = status == 0 ? 2 : status;
    r->error = err;
    r->t_done = now;
    r->group_samples = total;
    r->group_requests = (int32_t)sl->reqs.size();
    off += r->samples;
    if (status == 0) {
      ++S->requests_done;
      S->samples_done += r->samples;
    }
  }
  sl->reqs.clear();
  sl->busy = false;
}

void worker_main(snicit_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::unique_lock<std::mutex> lk(S->mu);
  while (!S->stop) {
    bool progressed = false;
    const Clock::time_point now = Clock::now();
    if (rounds(S) && now >= S->next_round) {
      // a period boundary: every request queued by now is due for the next free slot
      for (size_t n = 0; n < S->queues.size(); ++n)
        if (!S->queues[n].empty()) {
          S->due[n] = now;
          S->has_due[n] = 1;
        }
      const auto period = std::chrono::duration_cast<Clock::duration>(
          std::chrono::duration<double, std::milli>(S->cfg.round_ms));
      while (S->next_round <= now) S->next_round += period;
    }
    for (auto& slp : S->slots) {
      Slot* sl = slp.get();
      if (!sl->busy) continue;
      const hipError_t e = hipEventQuery(sl->done);
      if (e == hipErrorNotReady) continue;
      finish_group(S, sl, Clock::now(), e == hipSuccess ? 0 : SNICIT_E_HIP,
                   e == hipSuccess ? "" : std::string("group failed: ") + hipGetErrorString(e));
      progressed = true;
    }
    if (progressed) S->cv_done.notify_all();
    for (int32_t n = 0; n < (int32_t)S->nets.size(); ++n) {
      for (auto& slp : S->slots) {
        Slot* sl = slp.get();
        if (sl->net != n || sl->busy || !group_ready(S, n, now)) continue;
        fill_group(S, sl);
        if (sl->reqs.empty()) continue;
        const Clock::time_point t = Clock::now();
        for (Request* r : sl->reqs) {
          r->status = 1;
          r->t_launch = t;
          r->slot = (int32_t)(&slp - &S->slots[0]);
        }
        sl->busy = true;
        ++S->groups;
        lk.unlock();
        int rc = SNICIT_OK;
        std::string err;
        try {
          launch_group(S, sl);
        } catch (const Error& e) {
          rc = e.code;
          err = e.what();
        } catch (const std::exception& e) {
          rc = SNICIT_E_INVALID;
          err = e.what();
        }
        if (rc != SNICIT_OK) {
          // copies from the arena blocks and the staging rows may already be queued on the
          // slot's stream: let them end before the blocks and the slot are released
          (void)hipStreamSynchronize(sl->stream);
        }
        lk.lock();
        if (rc != SNICIT_OK) {
          finish_group(S, sl, Clock::now(), rc, err);
          S->cv_done.notify_all();
        }
        progressed = true;
      }
    }
    if (!progressed) S->cv_work.wait_for(lk, std::chrono::microseconds(50));
  }
}

void server_free(snicit_server* S) {
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto& sl : S->slots) slot_free(sl.get());
  S->slots.clear();
  if (S->pin_base) (void)hipHostFree(S->pin_base);
  S->pin_base = nullptr;
  S->pin_free.clear();
}

// One slot of `net`: plan with the segment capacity, stream, event, pinned staging.
std::unique_ptr<Slot> slot_make(snicit_server* S, int32_t n) {
  auto sl = std::make_unique<Slot>();
  sl->net = n;
  const snicit_server_cfg& c = S->cfg;
  try {
    // non-blocking: the worker's synchronous host calls must not wait for other slots
    SNICIT_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    SNICIT_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    std::vector<int32_t> segstart(c.slot_segments + 1, 0);
    sl->plan = plan_make(S->nets[n], &S->cfgs[n], c.slot_samples, segstart, sl->stream);
    ensure_stage(sl->plan);
    if (S->nets[n]->Nin == 0) {
      // image-render parameters sized once, so launches never allocate
      snicit_plan* p = sl->plan;
      p->d_img_params = dev_alloc<int32_t>((size_t)c.slot_samples * 4);
      p->img_params_cap = c.slot_samples;
      p->d_img_blobs = dev_alloc<int32_t>(64 * 4);
      p->d_img_proto = dev_alloc<int32_t>(20);
    }
    const int32_t w = width_of(S->nets[n]);
    SNICIT_HIP(hipHostMalloc((void**)&sl->h_in, (size_t)c.slot_samples * w * 4, 0));
    SNICIT_HIP(hipHostMalloc((void**)&sl->h_cat, (size_t)c.slot_samples * 4, 0));
    SNICIT_HIP(hipHostMalloc((void**)&sl->h_seg, (size_t)(c.slot_segments + 2) * 4, 0));
    if (c.warm) {
      // one run over zero inputs captures the stage graphs
      snicit_plan* p = sl->plan;
      launch_fill_f32(p->d_in, 0.f, (int64_t)p->in_rows * p->ld, sl->stream);
      for (int32_t i = 0; i <= c.slot_segments; ++i) sl->h_seg[i] = (int32_t)c.slot_samples;
      sl->h_seg[0] = 0;
      sl->h_seg[c.slot_segments + 1] = (int32_t)c.slot_samples;
      SNICIT_HIP(hipMemcpyAsync(p->d_segstart, sl->h_seg, (c.slot_segments + 1) * 4,
                                hipMemcpyHostToDevice, sl->stream));
      SNICIT_HIP(hipMemcpyAsync(p->d_counts, sl->h_seg + c.slot_segments + 1, 4,
                                hipMemcpyHostToDevice, sl->stream));
      p->B = c.slot_samples;
      p->have_input = true;
      plan_launch(p);
      SNICIT_HIP(hipStreamSynchronize(sl->stream));
    }
  } catch (...) {
    slot_free(sl.get());
    throw;
  }
  return sl;
}

}  // namespace

extern "C" {

void snicit_server_cfg_default(snicit_server_cfg* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->hbm_budget_bytes = int64_t(32) << 30;
  cfg->slot_samples = 8192;
  cfg->slot_segments = 64;
  cfg->max_slots_per_net = 3;
  cfg->cluster_min = 32;
  cfg->warm = 1;
  cfg->group_window_ms = 2.0;
  cfg->group_samples = 0;
  cfg->pinned_bytes = int64_t(256) << 20;
}

int snicit_server_create(snicit_ctx* ctx, snicit_net* const* nets, const snicit_run_cfg* cfgs,
                         int32_t n_nets, const snicit_server_cfg* cfg, snicit_server** out) {
  return guarded([&] {
    SNICIT_REQUIRE(ctx && nets && cfgs && cfg && out, "null argument");
    SNICIT_REQUIRE(n_nets >= 1, "a server needs nets");
    SNICIT_REQUIRE(cfg->slot_samples >= 1 && cfg->slot_samples < (int64_t(1) << 31) - 1024,
                   "slot_samples out of range");
    SNICIT_REQUIRE(cfg->slot_segments >= 1 && cfg->max_slots_per_net >= 1,
                   "slot_segments and max_slots_per_net must be >= 1");
    SNICIT_REQUIRE(cfg->group_window_ms >= 0.0 && cfg->round_ms >= 0.0 && cfg->cluster_min >= 0 &&
                       cfg->group_samples >= 0,
                   "bad grouping policy");
    SNICIT_REQUIRE(cfg->pinned_bytes >= 0, "pinned_bytes must be >= 0");
    SNICIT_HIP(hipSetDevice(ctx->device));
    auto S = std::make_unique<snicit_server>();
    S->ctx = ctx;
    S->cfg = *cfg;
    for (int32_t n = 0; n < n_nets; ++n) {
      SNICIT_REQUIRE(nets[n] && nets[n]->ctx == ctx, "nets must belong to the server's context");
      S->nets.push_back(nets[n]);
      S->cfgs.push_back(cfgs[n]);
    }
    S->queues.resize(n_nets);
    S->due.resize(n_nets);
    S->has_due.assign(n_nets, 0);
    S->next_round = Clock::now();
    try {
      if (cfg->pinned_bytes > 0) {
        const int64_t bytes = cfg->pinned_bytes / kPinAlign * kPinAlign;
        if (bytes > 0) {
          SNICIT_HIP(hipHostMalloc((void**)&S->pin_base, (size_t)bytes, 0));
          S->pin_free[0] = bytes;
        }
      }
      // slots round-robin over the nets while the budget allows (at least one per net)
      for (int32_t round = 0; round < cfg->max_slots_per_net; ++round) {
        for (int32_t n = 0; n < n_nets; ++n) {
          auto sl = slot_make(S.get(), n);
          const int64_t bytes = sl->plan->device_bytes;
          if (round > 0 && S->device_bytes + bytes > cfg->hbm_budget_bytes) {
            slot_free(sl.get());
            continue;
          }
          if (round == 0 && S->device_bytes + bytes > cfg->hbm_budget_bytes) {
            slot_free(sl.get());
            throw Error(SNICIT_E_NOMEM, "hbm_budget_bytes does not hold one slot per net");
          }
          S->device_bytes += bytes;
          S->slots.push_back(std::move(sl));
        }
      }
      S->worker = std::thread(worker_main, S.get());
    } catch (...) {
      server_free(S.get());
      throw;
    }
    *out = S.release();
  });
}

int snicit_server_destroy(snicit_server* server) {
  return guarded([&] {
    if (!server) return;
    server_free(server);
    delete server;
  });
}

namespace {
int submit(snicit_server* S, std::unique_ptr<Request> r, uint64_t* id) {
  return guarded([&] {
    SNICIT_REQUIRE(S && id, "null argument");
    SNICIT_REQUIRE(r->net >= 0 && r->net < (int32_t)S->nets.size(), "no such net");
    SNICIT_REQUIRE(r->samples >= 1 && r->samples <= S->cfg.slot_samples,
                   "request size must be in [1, slot_samples]");
    r->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> g(S->mu);
      r->id = S->next_id++;
      *id = r->id;
      S->queues[r->net].push_back(r.get());
      S->reqs[r->id] = std::move(r);
    }
    S->cv_work.notify_one();
  });
}
}  // namespace

int snicit_server_submit(snicit_server* server, int32_t net, const float* x, int64_t samples,
                         uint64_t* request_id) {
  auto r = std::make_unique<Request>();
  r->net = net;
  r->samples = samples;
  const int rc = guarded([&] {
    SNICIT_REQUIRE(server && x && request_id, "null argument");
    SNICIT_REQUIRE(net >= 0 && net < (int32_t)server->nets.size(), "no such net");
    SNICIT_REQUIRE(samples >= 1 && samples <= server->cfg.slot_samples,
                   "request size must be in [1, slot_samples]");
    // (the checks above are the ones submit() repeats, so the block cannot leak)
    const size_t n = (size_t)samples * width_of(server->nets[net]);
    r->pin = pin_alloc(server, (int64_t)n * 4);
    if (r->pin) {
      r->pin_bytes = (int64_t)n * 4;
      std::memcpy(r->pin, x, n * 4);
    } else {
      r->x.assign(x, x + n);
    }
  });
  if (rc != SNICIT_OK) return rc;
  return submit(server, std::move(r), request_id);
}

int snicit_server_submit_images(snicit_server* server, int32_t net, uint64_t seed,
                                int64_t first_sample, int64_t samples, uint64_t* request_id) {
  auto r = std::make_unique<Request>();
  r->net = net;
  r->samples = samples;
  r->images = true;
  r->seed = seed;
  r->first = first_sample;
  const int rc = guarded([&] {
    SNICIT_REQUIRE(server != nullptr, "null server");
    SNICIT_REQUIRE(net >= 0 && net < (int32_t)server->nets.size(), "no such net");
    SNICIT_REQUIRE(server->nets[net]->Nin == 0, "image inputs feed nets without an input layer");
  });
  if (rc != SNICIT_OK) return rc;
  return submit(server, std::move(r), request_id);
}

int snicit_server_wait(snicit_server* server, uint64_t request_id, double timeout_ms,
                       int32_t* cat, snicit_request_info* info) {
  return guarded([&] {
    SNICIT_REQUIRE(server != nullptr, "null server");
    std::unique_lock<std::mutex> lk(server->mu);
    auto it = server->reqs.find(request_id);
    SNICIT_REQUIRE(it != server->reqs.end(), "unknown request id");
    Request* r = it->second.get();
    auto done = [&] { return r->status == 2 || r->status < 0; };
    if (timeout_ms < 0) {
      server->cv_done.wait(lk, done);
    } else if (!server->cv_done.wait_for(
                   lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) {
      throw Error(SNICIT_E_TIMEOUT, "request not done yet");
    }
    if (r->status < 0) {
      const int code = r->status;
      const std::string msg = r->error;
      server->reqs.erase(it);
      throw Error(code, msg);
    }
    if (cat) std::memcpy(cat, r->cat.data(), r->cat.size() * 4);
    if (info) {
      info->ms_queue = ms_between(r->t_submit, r->t_launch);
      info->ms_run = ms_between(r->t_launch, r->t_done);
      info->ms_total = ms_between(r->t_submit, r->t_done);
      info->group_samples = r->group_samples;
      info->group_requests = r->group_requests;
      info->slot = r->slot;
    }
    server->reqs.erase(it);
  });
}

int snicit_server_stats_get(snicit_server* server, snicit_server_stats* stats) {
  return guarded([&] {
    SNICIT_REQUIRE(server && stats, "null argument");
    std::lock_guard<std::mutex> g(server->mu);
    std::memset(stats, 0, sizeof(*stats));
    stats->requests_done = server->requests_done;
    stats->samples_done = server->samples_done;
    stats->groups = server->groups;
    stats->device_bytes = server->device_bytes;
    stats->slots = (int32_t)server->slots.size();
    int32_t q = 0;
    for (const auto& d : server->queues) q += (int32_t)d.size();
    stats->queued = q;
  });
}

}  // extern "C"
