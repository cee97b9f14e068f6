// serve.cpp -- the served form (parallax/px_serve.h): a request loop with admission
// against an HBM budget and continuous batching of mixed primitive calls.
//
// One worker thread owns the schedule:
//   1. completions: a slot whose event has fired completes every request of its batch
//      (latency = completion seen - arrival);
//   2. admission, in arrival order: a request is admitted once its output fits in the
//      pool (the budget); the head of the queue blocks the ones behind it, so nothing
//      starves;
//   3. dispatch: every free slot takes the oldest admitted request and all admitted
//      requests of the same (op, dtype, predicate) -- one segmented launch over
//      segments that stay where they are in the pool (no gather copies).
// Slots are streams with their own scratch, so batches of different groups overlap on
// the device while new requests keep arriving.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "parallax/arena.hpp"
#include "parallax/px_device.h"
#include "parallax/px_serve.h"

using parallax::SegSpec;
using parallax::StreamContext;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum class St { PENDING, READY, RUNNING, DONE };

struct Req {
  pxs_request r;
  St st = St::PENDING;
  double t_arr = 0, t_done = 0;
  void* d_out = nullptr;  // server-allocated output (null for sort)
  uint64_t count = 0;
  size_t es = 0;
};

struct Slot {
  std::unique_ptr<StreamContext> ctx;
  hipEvent_t ev = nullptr;
  bool busy = false;
  std::vector<int64_t> ids;
  void* counts = nullptr;       // device counts of a copy_if batch (pool)
  uint64_t* pin_counts = nullptr;
  size_t pin_cap = 0;
};

bool same_group(const Req& a, const Req& b) {
  if (a.r.op != b.r.op || a.r.dtype != b.r.dtype) return false;
  if (a.r.op != PX_OP_FLAGS) return true;
  return a.r.flags.fn == b.r.flags.fn && a.r.flags.flags == b.r.flags.flags &&
         a.r.flags.c0 == b.r.flags.c0 && a.r.flags.c1 == b.r.flags.c1;
}

}  // namespace

struct pxs_server {
  pxs_config cfg{};
  parallax::HipBackend* be = nullptr;
  parallax::UnifiedArena pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending, ready;
  std::unordered_map<int64_t, Req> reqs;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  int64_t next_id = 0;
  uint64_t submitted = 0, completed = 0, batches = 0, rejected = 0;
  std::vector<double> lat_ms;
  double first_arr = 0, last_done = 0;
  size_t peak = 0;

  void loop();
  bool launch(Slot& s, const std::vector<int64_t>& ids);
};

bool pxs_server::launch(Slot& s, const std::vector<int64_t>& ids) {
  const Req& r0 = reqs[ids.front()];
  const size_t es = r0.es;
  const char* base = static_cast<const char*>(pool.device_base());
  const uint32_t nseg = static_cast<uint32_t>(ids.size());
  std::vector<uint64_t> lo(nseg), hi(nseg), out(nseg);
  for (uint32_t i = 0; i < nseg; ++i) {
    const Req& q = reqs[ids[i]];
    lo[i] = static_cast<uint64_t>(static_cast<const char*>(q.r.d_in) - base) / es;
    hi[i] = lo[i] + q.r.n;
    out[i] = q.d_out ? static_cast<uint64_t>(static_cast<const char*>(q.d_out) - base) / es : lo[i];
  }
  SegSpec sp;
  sp.lo = lo.data();
  sp.hi = hi.data();
  sp.out = out.data();
  sp.nseg = nseg;
  void* b = pool.device_base();
  int rc = PX_EOK;
  s.counts = nullptr;
  switch (r0.r.op) {
    case PX_OP_REDUCE: rc = parallax::run_seg_reduce(*s.ctx, r0.r.dtype, b, sp, b); break;
    case PX_OP_SCAN: rc = parallax::run_seg_scan(*s.ctx, r0.r.dtype, b, b, sp); break;
    case PX_OP_SORT: sp.out = nullptr; rc = parallax::run_seg_sort(*s.ctx, r0.r.dtype, b, sp); break;
    case PX_OP_FLAGS: {
      s.counts = pool.allocate(nseg * sizeof(uint64_t));
      if (!s.counts) return false;
      if (s.pin_cap < nseg) {
        if (s.pin_counts) (void)hipHostFree(s.pin_counts);
        s.pin_counts = nullptr;
        s.pin_cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&s.pin_counts), nseg * sizeof(uint64_t),
                          hipHostMallocDefault) != hipSuccess)
          return false;
        s.pin_cap = nseg;
      }
      rc = parallax::run_seg_compact(*s.ctx, r0.r.flags, b, b, sp, static_cast<uint64_t*>(s.counts));
      if (rc == PX_EOK &&
          hipMemcpyAsync(s.pin_counts, s.counts, nseg * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         s.ctx->stream()) != hipSuccess)
        rc = PX_EDEVICE;
      break;
    }
    default: rc = PX_EARG;
  }
  if (rc != PX_EOK) return false;
  if (hipEventRecord(s.ev, s.ctx->stream()) != hipSuccess) return false;
  s.ids = ids;
  s.busy = true;
  ++batches;
  for (int64_t id : ids) reqs[id].st = St::RUNNING;
  return true;
}

void pxs_server::loop() {
  (void)hipSetDevice(be->device());
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    bool progress = false;
    // 1. completions
    for (auto& s : slots) {
      if (!s.busy || hipEventQuery(s.ev) != hipSuccess) continue;
      const double t = now_s();
      for (size_t i = 0; i < s.ids.size(); ++i) {
        Req& q = reqs[s.ids[i]];
        q.st = St::DONE;
        q.t_done = t;
        if (s.counts) q.count = s.pin_counts[i];
        lat_ms.push_back((t - q.t_arr) * 1e3);
        ++completed;
      }
      last_done = t;
      if (s.counts) pool.deallocate(s.counts);
      s.counts = nullptr;
      s.ids.clear();
      s.busy = false;
      progress = true;
    }
    if (progress) cv_done.notify_all();
    // 2. admission in arrival order
    while (!pending.empty()) {
      Req& q = reqs[pending.front()];
      size_t need = 0;
      if (q.r.op == PX_OP_REDUCE) need = q.es;
      else if (q.r.op == PX_OP_SCAN || q.r.op == PX_OP_FLAGS) need = q.r.n * q.es;
      if (need) {
        q.d_out = pool.allocate(need);
        if (!q.d_out) break;
      }
      q.st = St::READY;
      ready.push_back(pending.front());
      pending.pop_front();
      progress = true;
    }
    peak = std::max(peak, pool.used());
    // 3. dispatch to free slots
    for (auto& s : slots) {
      if (s.busy || ready.empty()) continue;
      const Req& head = reqs[ready.front()];
      if (cfg.max_wait_us > 0 && (now_s() - head.t_arr) * 1e6 < cfg.max_wait_us) {
        size_t same = 0;
        for (int64_t id : ready) same += same_group(head, reqs[id]) ? 1 : 0;
        if (static_cast<int>(same) < cfg.max_batch) continue;
      }
      std::vector<int64_t> ids;
      std::deque<int64_t> rest;
      for (int64_t id : ready) {
        if (static_cast<int>(ids.size()) < cfg.max_batch && same_group(head, reqs[id]))
          ids.push_back(id);
        else
          rest.push_back(id);
      }
      if (!launch(s, ids)) {
        // a failed launch completes its requests with no result (counted as rejected)
        for (int64_t id : ids) {
          Req& q = reqs[id];
          q.st = St::DONE;
          q.t_done = now_s();
          ++rejected;
        }
        cv_done.notify_all();
      }
      ready.swap(rest);
      progress = true;
    }
    if (!progress) cv_work.wait_for(lk, std::chrono::microseconds(20));
  }
}

extern "C" {

void pxs_config_default(pxs_config* cfg) {
  if (!cfg) return;
  cfg->slots = 4;
  cfg->hbm_budget = 64ull << 30;
  cfg->max_batch = 4096;
  cfg->max_wait_us = 0.0;
}

pxs_server_t pxs_create(const pxs_config* cfg) {
  parallax::HipBackend* be = parallax::global_backend();
  if (!be) {
    parallax::set_last_error(PX_ENODEV, "no GPU");
    return nullptr;
  }
  auto* s = new pxs_server;
  pxs_config_default(&s->cfg);
  if (cfg) s->cfg = *cfg;
  if (s->cfg.slots < 1) s->cfg.slots = 1;
  if (s->cfg.max_batch < 1) s->cfg.max_batch = 1;
  s->be = be;
  if (!s->pool.initialize(be, s->cfg.hbm_budget, false)) {
    parallax::set_last_error(PX_ENOMEM, "server pool allocation failed");
    delete s;
    return nullptr;
  }
  s->slots.resize(s->cfg.slots);
  for (auto& sl : s->slots) {
    sl.ctx = std::make_unique<StreamContext>(be, nullptr);
    if (!sl.ctx->ok() || hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
      parallax::set_last_error(PX_ENOMEM, "server slot creation failed");
      pxs_destroy(s);
      return nullptr;
    }
  }
  s->worker = std::thread([s] { s->loop(); });
  return s;
}

void pxs_destroy(pxs_server_t s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (auto& sl : s->slots) {
    if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->stream());
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.pin_counts) (void)hipHostFree(sl.pin_counts);
    sl.ctx.reset();
  }
  delete s;
}

void* pxs_alloc(pxs_server_t s, size_t bytes) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  void* p = s->pool.allocate(bytes);
  s->peak = std::max(s->peak, s->pool.used());
  return p;
}

void pxs_free(pxs_server_t s, void* d_ptr) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  s->pool.deallocate(d_ptr);
  s->cv_work.notify_all();
}

int64_t pxs_submit(pxs_server_t s, const pxs_request* req) {
  if (!s || !req) return PX_EARG;
  if (req->dtype < 0 || req->dtype >= PX_NUM_DTYPES) return PX_EARG;
  if (req->op != PX_OP_REDUCE && req->op != PX_OP_SCAN && req->op != PX_OP_SORT &&
      req->op != PX_OP_FLAGS)
    return PX_EARG;
  if (req->op == PX_OP_FLAGS && (req->flags.dtype != static_cast<uint32_t>(req->dtype) ||
                                 req->flags.fn >= PX_NUM_PREDS))
    return PX_EARG;
  const size_t es = pxk::dtype_size(req->dtype);
  std::lock_guard<std::mutex> lk(s->mu);
  if (req->n && (!s->pool.contains(req->d_in) ||
                 !s->pool.contains(static_cast<char*>(req->d_in) + req->n * es - 1)))
    return PX_EARG;
  if (req->n * es > s->pool.capacity()) {
    ++s->rejected;
    return PX_ENOMEM;
  }
  const int64_t id = s->next_id++;
  Req q;
  q.r = *req;
  q.es = es;
  q.t_arr = now_s();
  if (s->first_arr == 0) s->first_arr = q.t_arr;  // first arrival since the last reset
  s->reqs.emplace(id, q);
  s->pending.push_back(id);
  ++s->submitted;
  s->cv_work.notify_all();
  return id;
}

int pxs_wait(pxs_server_t s, int64_t id, double timeout_ms) {
  if (!s) return PX_EARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto done = [&] {
    auto it = s->reqs.find(id);
    return it == s->reqs.end() || it->second.st == St::DONE;
  };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
    return 0;
  }
  return s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done) ? 0 : 1;
}

int pxs_result(pxs_server_t s, int64_t id, void** d_out, uint64_t* count, double* latency_ms) {
  if (!s) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second.st != St::DONE) return PX_EARG;
  const Req& q = it->second;
  if (d_out) *d_out = q.d_out ? q.d_out : q.r.d_in;
  if (count) *count = q.count;
  if (latency_ms) *latency_ms = (q.t_done - q.t_arr) * 1e3;
  return PX_EOK;
}

int pxs_release(pxs_server_t s, int64_t id) {
  if (!s) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second.st != St::DONE) return PX_EARG;
  if (it->second.d_out) s->pool.deallocate(it->second.d_out);
  s->reqs.erase(it);
  s->cv_work.notify_all();
  return PX_EOK;
}

int pxs_drain(pxs_server_t s) {
  if (!s) return PX_EARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->completed + s->rejected >= s->submitted; });
  return PX_EOK;
}

int pxs_get_stats(pxs_server_t s, pxs_stats* out) {
  if (!s || !out) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->batches = s->batches;
  out->rejected = s->rejected;
  std::vector<double> v = s->lat_ms;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      const double idx = p * (v.size() - 1);
      const size_t i = static_cast<size_t>(idx);
      const double f = idx - i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_ms = sum / v.size();
    out->max_ms = v.back();
  }
  out->first_arrival_s = s->first_arr;
  out->last_completion_s = s->last_done;
  const double span = s->last_done - s->first_arr;
  out->req_per_s = span > 0 ? static_cast<double>(v.size()) / span : 0.0;
  out->pool_bytes = s->pool.used();
  out->pool_peak_bytes = s->peak;
  return PX_EOK;
}

int pxs_reset_stats(pxs_server_t s) {
  if (!s) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_ms.clear();
  s->first_arr = 0;
  s->last_done = 0;
  s->batches = 0;
  return PX_EOK;
}

}  // extern "C"
