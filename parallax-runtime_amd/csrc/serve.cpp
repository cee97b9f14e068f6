// serve.cpp -- the served form (parallax/px_serve.h): a request loop with admission
// against an HBM budget and continuous batching of mixed primitive calls.
//
// One worker thread owns the schedule:
//   1. completions: a slot whose event has fired completes every request of its batch
//      with the batch's status (latency = completion seen - arrival);
//   2. admission, in arrival order: a request is admitted once its output fits in the
//      pool (the budget); the head of the queue blocks the ones behind it, so nothing
//      starves.  A request whose output could never fit beside the client's resident
//      inputs fails with PX_ENOMEM instead of blocking the queue;
//   3. dispatch: every free slot takes the oldest admitted request and all admitted
//      requests of the same (op, dtype, predicate) -- one segmented launch over
//      segments that stay where they are in the pool (no gather copies).  Scratch a batch
//      needs (copy_if counts, the segmented radix sort's buffers) comes from the same
//      pool, so it counts against the budget too.
// A slot is one batch in flight with its own scratch (stream context, tables, event).
// Slots share `lanes` streams (default two): a free slot queues its batch behind the
// batches already on its lane, so the device runs them back to back while the worker
// admits and packs the next ones.  The scan, compaction and radix kernels are
// persistent look-back kernels sized to the whole device, and two of them on different
// streams split the CUs: each ran 15x slower than alone (served pass, 8 slots on 8
// streams: 7.8 ms; on one stream, one batch at a time: 5.9 ms; on one stream with
// batches queued: 5.2 ms).  So they all take lane 0, and what has no waits between
// blocks (a static grid, which shares CUs safely) takes the last lane: reduce batches,
// and a sort batch's table upload, short-segment sort and counting kernels, after which
// an event sends its digit passes to lane 0.  Those overlap the look-back kernels of the
// batches ahead on lane 0 (reduces: 4.76 -> 4.65 ms, p50 4.0 -> 2.9 ms).
//
// When nothing runs and the queue head waits for memory only the client can free
// (outputs it has not released), the server is stalled: pxs_wait on a request that is
// not running and pxs_drain return PX_ENOMEM instead of blocking forever; a release or
// free clears the state and the worker tries again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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

size_t round_pool(size_t b) {
  const size_t a = parallax::UnifiedArena::kDefaultAlignment;
  return (b + a - 1) / a * a;
}

enum class St { PENDING, READY, RUNNING, DONE };

struct Req {
  pxs_request r;
  St st = St::PENDING;
  int status = PX_EOK;  // of a DONE request: PX_EOK or the failure's PX_E* code
  double t_arr = 0, t_done = 0;
  void* d_out = nullptr;  // server-allocated output (null for sort)
  uint64_t count = 0;
  size_t es = 0;
};

struct Slot {
  std::unique_ptr<StreamContext> ctx;
  hipEvent_t ev = nullptr;       // the batch is complete
  hipEvent_t ev_prep = nullptr;  // a sort's counting kernels are done (its passes wait)
  hipStream_t done_stream = nullptr;  // where `ev` is recorded (the lane of the last launch)
  bool busy = false;
  int rc = PX_EOK;               // status the batch completes with
  std::vector<int64_t> ids;
  void* counts = nullptr;        // device counts of a copy_if batch (pool)
  void* scratch = nullptr;       // segmented radix buffers of a sort batch (pool)
  uint64_t* pin_counts = nullptr;
  size_t pin_cap = 0;
};

bool same_group(const Req& a, const Req& b) {
  if (a.r.op != b.r.op || a.r.dtype != b.r.dtype) return false;
  if (a.r.op != PX_OP_FLAGS) return true;
  return a.r.flags.fn == b.r.flags.fn && a.r.flags.flags == b.r.flags.flags &&
         a.r.flags.c0 == b.r.flags.c0 && a.r.flags.c1 == b.r.flags.c1;
}

size_t out_bytes(const pxs_request& r, size_t es) {
  if (r.op == PX_OP_REDUCE) return es;
  if (r.op == PX_OP_SCAN || r.op == PX_OP_FLAGS) return r.n * es;
  return 0;
}

// Scratch bytes a sort of [lo, lo + n) takes from the pool (segmented radix buffers).
size_t sort_scratch(int dtype, uint64_t lo, uint64_t n) {
  const uint64_t hi = lo + n;
  SegSpec sp;
  sp.lo = &lo;
  sp.hi = &hi;
  sp.nseg = 1;
  return parallax::seg_sort_scratch_bytes(dtype, sp);
}

enum LaunchResult { kLaunched, kNoMemory, kFailed };

// Served reduces: 256 threads x 16 vectors per tile (a request up to one tile long is
// summed by one block; longer ones leave one partial per 16 KiB of input).
constexpr pxk::TileCfg kServedReduceTile{256, 16};

}  // namespace

struct pxs_server {
  pxs_config cfg{};
  parallax::HipBackend* be = nullptr;
  parallax::UnifiedArena pool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending, ready;
  std::unordered_map<int64_t, Req> reqs;
  std::unordered_map<void*, size_t> client_allocs;  // pxs_alloc'd inputs
  size_t client_bytes = 0;
  std::vector<Slot> slots;
  std::vector<hipStream_t> lanes;  // streams the slots share (slot i on lanes[i % size])
  std::thread worker;
  bool stop = false;
  bool stalled = false;
  std::atomic<bool> wake{false};  // new work or freed memory: re-evaluate the schedule
  int64_t next_id = 0;
  uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, rejected = 0;
  std::vector<double> lat_ms;
  double first_arr = 0, last_done = 0;
  size_t peak = 0;

  void loop();
  int launch(Slot& s, const std::vector<int64_t>& ids);
  void finish_slot(Slot& s, double t);
  void fail(int64_t id, int code, double t);
  // Bytes the pool could give once every server-held and releasable block is free.
  size_t attainable() const {
    return pool.capacity() > client_bytes ? pool.capacity() - client_bytes : 0;
  }
  bool any_busy() const {
    for (const auto& s : slots)
      if (s.busy) return true;
    return false;
  }
};

void pxs_server::fail(int64_t id, int code, double t) {
  Req& q = reqs[id];
  q.st = St::DONE;
  q.status = code;
  q.t_done = t;
  ++failed;
}

void pxs_server::finish_slot(Slot& s, double t) {
  for (size_t i = 0; i < s.ids.size(); ++i) {
    Req& q = reqs[s.ids[i]];
    q.st = St::DONE;
    q.status = s.rc;
    q.t_done = t;
    if (s.rc == PX_EOK) {
      if (s.counts) q.count = s.pin_counts[i];
      lat_ms.push_back((t - q.t_arr) * 1e3);
      ++completed;
    } else {
      ++failed;
    }
  }
  if (s.rc == PX_EOK) last_done = t;
  if (s.counts) pool.deallocate(s.counts);
  if (s.scratch) pool.deallocate(s.scratch);
  s.counts = s.scratch = nullptr;
  s.ids.clear();
  s.busy = false;
  s.rc = PX_EOK;
}

// Launches one batch on a free slot.  kNoMemory: the pool has no room for the batch's
// scratch now (nothing was queued).  kFailed: the batch failed and its requests are
// complete (failed) already, or complete with the failure once work queued before the
// failure has drained (the slot is then busy with rc set).
int pxs_server::launch(Slot& s, const std::vector<int64_t>& ids) {
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
  s.counts = s.scratch = nullptr;
  size_t scratch_bytes = 0;
  if (r0.r.op == PX_OP_SORT) {
    sp.out = nullptr;
    scratch_bytes = parallax::seg_sort_scratch_bytes(r0.r.dtype, sp);
    if (scratch_bytes && !(s.scratch = pool.allocate(scratch_bytes))) return kNoMemory;
  } else if (r0.r.op == PX_OP_FLAGS) {
    if (!(s.counts = pool.allocate(nseg * sizeof(uint64_t)))) return kNoMemory;
  }
  const double t = now_s();
  auto fail_now = [&](int code) {
    if (s.counts) pool.deallocate(s.counts);
    if (s.scratch) pool.deallocate(s.scratch);
    s.counts = s.scratch = nullptr;
    for (int64_t id : ids) fail(id, code, t);
    return kFailed;
  };
  if (r0.r.op == PX_OP_FLAGS && s.pin_cap < nseg) {
    if (s.pin_counts) (void)hipHostFree(s.pin_counts);
    s.pin_counts = nullptr;
    s.pin_cap = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&s.pin_counts), nseg * sizeof(uint64_t),
                      hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail_now(PX_ENOMEM);
    }
    s.pin_cap = nseg;
  }
  // Tile shapes fixed per (op, dtype): a request's float arithmetic (the order its
  // elements are summed in) must not depend on what it is batched with, so its result is
  // the same alone and in any mix.
  int rc = PX_EOK;
  switch (r0.r.op) {
    case PX_OP_REDUCE: rc = parallax::run_seg_reduce(*s.ctx, r0.r.dtype, b, sp, b, kServedReduceTile); break;
    case PX_OP_SCAN: rc = parallax::run_seg_scan(*s.ctx, r0.r.dtype, b, b, sp, pxk::tune().scan[r0.r.dtype]); break;
    case PX_OP_SORT:
      // with two lanes the slot is on the last: its counting kernels run there and its
      // digit passes (look-back kernels) on lane 0
      rc = parallax::run_seg_sort(*s.ctx, r0.r.dtype, b, sp, s.scratch, scratch_bytes,
                                  lanes.size() > 1 ? lanes[0] : nullptr, lanes.size() > 1 ? s.ev_prep : nullptr);
      break;
    case PX_OP_FLAGS:
      rc = parallax::run_seg_compact(*s.ctx, r0.r.flags, b, b, sp, static_cast<uint64_t*>(s.counts),
                                     pxk::tune().compact[r0.r.flags.dtype]);
      if (rc == PX_EOK &&
          hipMemcpyAsync(s.pin_counts, s.counts, nseg * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         s.ctx->stream()) != hipSuccess)
        rc = PX_EDEVICE;
      break;
    default: rc = PX_EARG;
  }
  // Kernels queued before a failure may still run on the request buffers: the batch
  // completes (with rc) when the slot's event fires, never before.  A sort on two lanes
  // ends on lane 0 (everything it queued on its own lane is ordered before that).
  s.done_stream = (r0.r.op == PX_OP_SORT && lanes.size() > 1) ? lanes[0] : s.ctx->stream();
  if (hipEventRecord(s.ev, s.done_stream) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(s.ctx->stream());
    (void)hipStreamSynchronize(s.done_stream);
    return fail_now(rc != PX_EOK ? rc : PX_EDEVICE);
  }
  s.ids = ids;
  s.busy = true;
  s.rc = rc;
  for (int64_t id : ids) reqs[id].st = St::RUNNING;
  if (rc != PX_EOK) return kFailed;
  ++batches;
  return kLaunched;
}

void pxs_server::loop() {
  (void)hipSetDevice(be->device());
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    bool progress = false;
    wake = false;
    // 1. completions
    for (auto& s : slots) {
      if (!s.busy || hipEventQuery(s.ev) == hipErrorNotReady) continue;
      finish_slot(s, now_s());
      progress = true;
    }
    // 2. admission in arrival order
    bool head_waits = false;
    while (!pending.empty()) {
      const int64_t id = pending.front();
      Req& q = reqs[id];
      const size_t need = out_bytes(q.r, q.es);
      if (need) {
        q.d_out = pool.allocate(need);
        if (!q.d_out) {
          if (round_pool(need) > attainable()) {
            // never fits beside the client's inputs: fail it rather than block the queue
            fail(id, PX_ENOMEM, now_s());
            pending.pop_front();
            progress = true;
            continue;
          }
          head_waits = true;
          break;
        }
      }
      q.st = St::READY;
      ready.push_back(id);
      pending.pop_front();
      progress = true;
    }
    peak = std::max(peak, pool.used());
    // 3. dispatch to free slots (slot i on lane i % lanes; with two or more lanes, reduce
    // and sort batches take the last lane -- a sort's digit passes then run on lane 0 --
    // and the others lane 0)
    bool dispatch_waits = false;
    const size_t nl = lanes.size();
    // lane 0's slots first: its own batches then queue ahead of the digit passes that
    // sorts dispatched on the last lane add to it (those wait for their counting kernels)
    for (size_t q = 0; q < slots.size(); ++q) {
      const size_t per = (slots.size() + nl - 1) / nl;
      const size_t si = (q % per) * nl + q / per;  // lane-major order over slot indices
      if (si >= slots.size()) continue;
      Slot& s = slots[si];
      if (s.busy || ready.empty()) continue;
      const size_t lane = si % nl;
      auto lane_of = [&](const Req& q) {
        return (nl > 1 && (q.r.op == PX_OP_REDUCE || q.r.op == PX_OP_SORT)) ? nl - 1 : size_t(0);
      };
      // the oldest request of this lane; on the last lane of several, the oldest sort
      // first (its digit passes wait on lane 0 for its counting kernels)
      int64_t head_id = -1;
      for (int64_t id : ready)
        if (lane_of(reqs[id]) == lane && (head_id < 0 || (nl > 1 && lane == nl - 1 &&
                                                           reqs[id].r.op == PX_OP_SORT &&
                                                           reqs[head_id].r.op != PX_OP_SORT))) {
          head_id = id;
          if (nl == 1 || lane != nl - 1 || reqs[id].r.op == PX_OP_SORT) break;
        }
      if (head_id < 0) continue;
      const Req& head = reqs[head_id];
      if (cfg.max_wait_us > 0 && (now_s() - head.t_arr) * 1e6 < cfg.max_wait_us) {
        size_t same = 0;
        for (int64_t id : ready) same += same_group(head, reqs[id]) ? 1 : 0;
        if (static_cast<int>(same) < cfg.max_batch) continue;
      }
      std::vector<int64_t> ids;
      for (int64_t id : ready)
        if (static_cast<int>(ids.size()) < cfg.max_batch && same_group(head, reqs[id])) ids.push_back(id);
      int lr = launch(s, ids);
      if (lr == kNoMemory && !any_busy() && ids.size() > 1) {
        ids.resize(1);  // the head alone
        lr = launch(s, ids);
      }
      if (lr == kNoMemory) {
        if (any_busy()) break;  // running batches free their scratch when they finish
        const Req& q = reqs[ids.front()];
        const char* base = static_cast<const char*>(pool.device_base());
        const uint64_t lo = static_cast<uint64_t>(static_cast<const char*>(q.r.d_in) - base) / q.es;
        const size_t sc = q.r.op == PX_OP_SORT ? sort_scratch(q.r.dtype, lo, q.r.n) : 64;
        if (round_pool(sc) > attainable()) {
          fail(ids.front(), PX_ENOMEM, now_s());
          ready.erase(std::find(ready.begin(), ready.end(), ids.front()));
          progress = true;
          continue;
        }
        dispatch_waits = true;
        break;
      }
      // launched, or failed: take the batch out of the ready queue
      std::deque<int64_t> rest;
      size_t k = 0;
      for (int64_t id : ready) {
        if (k < ids.size() && ids[k] == id) ++k;
        else rest.push_back(id);
      }
      ready.swap(rest);
      progress = true;
    }
    const bool now_stalled = (head_waits || dispatch_waits) && !any_busy();
    if (now_stalled != stalled) progress = true;
    stalled = now_stalled;
    if (progress) cv_done.notify_all();
    if (progress || wake.load()) continue;
    if (any_busy()) {
      // batches are on the device: poll their events with the lock released (only this
      // thread touches the slots), until one fires, a client wakes us, or 100 us pass
      std::vector<hipEvent_t> evs;
      for (const auto& s : slots)
        if (s.busy) evs.push_back(s.ev);
      lk.unlock();
      const auto t0 = std::chrono::steady_clock::now();
      bool fired = false;
      while (!fired && !wake.load() &&
             std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(100)) {
        for (hipEvent_t e : evs)
          if (hipEventQuery(e) != hipErrorNotReady) fired = true;
        if (!fired) std::this_thread::yield();
      }
      lk.lock();
    } else {
      cv_work.wait(lk, [&] { return stop || wake.load(); });
    }
  }
}

extern "C" {

void pxs_config_default(pxs_config* cfg) {
  if (!cfg) return;
  cfg->slots = 8;
  cfg->hbm_budget = 64ull << 30;
  cfg->max_batch = 4096;
  cfg->max_wait_us = 0.0;
  cfg->lanes = 2;
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
  s->cfg.lanes = std::min(std::max(s->cfg.lanes, 1), std::min(s->cfg.slots, 2));  // a third lane gets no work
  s->be = be;
  if (!s->pool.initialize(be, s->cfg.hbm_budget, false)) {
    parallax::set_last_error(PX_ENOMEM, "server pool allocation failed");
    delete s;
    return nullptr;
  }
  if (hipSetDevice(be->device()) != hipSuccess) {
    parallax::set_last_error(PX_ENODEV, "pxs_create: hipSetDevice");
    delete s;
    return nullptr;
  }
  s->lanes.assign(s->cfg.lanes, nullptr);
  for (auto& l : s->lanes) {
    if (hipStreamCreateWithFlags(&l, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      l = nullptr;
      parallax::set_last_error(PX_ENOMEM, "server stream creation failed");
      pxs_destroy(s);
      return nullptr;
    }
  }
  s->slots.resize(s->cfg.slots);
  for (size_t i = 0; i < s->slots.size(); ++i) {
    Slot& sl = s->slots[i];
    sl.ctx = std::make_unique<StreamContext>(be, s->lanes[i % s->lanes.size()]);
    if (!sl.ctx->ok() || hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev_prep, hipEventDisableTiming) != hipSuccess) {
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
    if (sl.done_stream) (void)hipStreamSynchronize(sl.done_stream);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.ev_prep) (void)hipEventDestroy(sl.ev_prep);
    if (sl.pin_counts) (void)hipHostFree(sl.pin_counts);
    sl.ctx.reset();
  }
  for (hipStream_t l : s->lanes)
    if (l) (void)hipStreamDestroy(l);
  delete s;
}

void* pxs_alloc(pxs_server_t s, size_t bytes) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  void* p = s->pool.allocate(bytes);
  if (p) {
    s->client_allocs[p] = round_pool(bytes);
    s->client_bytes += round_pool(bytes);
  }
  s->peak = std::max(s->peak, s->pool.used());
  return p;
}

void pxs_free(pxs_server_t s, void* d_ptr) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->client_allocs.find(d_ptr);
  if (it == s->client_allocs.end()) return;  // not a pxs_alloc pointer: ignored
  s->client_bytes -= it->second;
  s->client_allocs.erase(it);
  s->pool.deallocate(d_ptr);
  s->stalled = false;
  s->wake = true;
  s->cv_work.notify_all();
}

namespace {
// pxs_submit with the server's lock held (no wake-up of the worker).
int64_t submit_locked(pxs_server_t s, const pxs_request* req);
}  // namespace

int64_t pxs_submit(pxs_server_t s, const pxs_request* req) {
  if (!s || !req) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t id = submit_locked(s, req);
  if (id >= 0) {
    s->wake = true;
    s->cv_work.notify_all();
  }
  return id;
}

int pxs_submit_batch(pxs_server_t s, const pxs_request* reqs, size_t n, int64_t* ids) {
  if (!s || (n && (!reqs || !ids))) return PX_EARG;
  int bad = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < n; ++i) {
      ids[i] = submit_locked(s, &reqs[i]);
      if (ids[i] < 0) ++bad;
    }
    s->wake = true;
  }
  s->cv_work.notify_all();
  return bad;
}

}  // extern "C"

namespace {
int64_t submit_locked(pxs_server_t s, const pxs_request* req) {
  if (req->dtype < 0 || req->dtype >= PX_NUM_DTYPES) return PX_EARG;
  if (req->op != PX_OP_REDUCE && req->op != PX_OP_SCAN && req->op != PX_OP_SORT &&
      req->op != PX_OP_FLAGS)
    return PX_EARG;
  if (req->op == PX_OP_FLAGS && (req->flags.dtype != static_cast<uint32_t>(req->dtype) ||
                                 req->flags.fn >= PX_NUM_PREDS))
    return PX_EARG;
  const size_t es = pxk::dtype_size(req->dtype);
  if (req->n && (!s->pool.contains(req->d_in) ||
                 !s->pool.contains(static_cast<char*>(req->d_in) + req->n * es - 1))) {
    parallax::set_last_error(PX_EARG, "pxs_submit: input outside the server's pool");
    return PX_EARG;
  }
  const char* base = static_cast<const char*>(s->pool.device_base());
  const uint64_t off = static_cast<uint64_t>(static_cast<const char*>(req->d_in) - base);
  if (req->n && off % es != 0) {
    parallax::set_last_error(PX_EARG, "pxs_submit: input not aligned to its element size");
    return PX_EARG;
  }
  // An in-place sort must not share bytes with another request that has not completed
  // (they could run concurrently on different slots).
  if (req->n) {
    const uint64_t a0 = off, a1 = off + req->n * es;
    for (const auto& kv : s->reqs) {
      const Req& q = kv.second;
      if (q.st == St::DONE || q.r.n == 0) continue;
      if (req->op != PX_OP_SORT && q.r.op != PX_OP_SORT) continue;
      const uint64_t b0 = static_cast<uint64_t>(static_cast<const char*>(q.r.d_in) - base);
      const uint64_t b1 = b0 + q.r.n * q.es;
      if (a0 < b1 && b0 < a1) {
        parallax::set_last_error(PX_EARG, "pxs_submit: input overlaps a pending in-place sort");
        return PX_EARG;
      }
    }
  }
  // outputs (and a sort's scratch) that could never fit beside the client's inputs
  const size_t need = req->op == PX_OP_SORT ? sort_scratch(req->dtype, off / es, req->n)
                                             : out_bytes(*req, es);
  if (round_pool(need) > s->attainable()) {
    ++s->rejected;
    parallax::set_last_error(PX_ENOMEM, "pxs_submit: the request's output cannot fit in the pool");
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
  return id;
}
}  // namespace

extern "C" {

int pxs_wait(pxs_server_t s, int64_t id, double timeout_ms) {
  if (!s) return PX_EARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it0 = s->reqs.find(id);
  if (it0 == s->reqs.end()) return PX_EARG;
  auto settled = [&] {
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.st == St::DONE) return true;
    return s->stalled && it->second.st != St::RUNNING;
  };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, settled);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), settled)) {
    return 1;
  }
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return PX_EARG;
  if (it->second.st != St::DONE) {
    parallax::set_last_error(PX_ENOMEM, "pxs_wait: the request waits for pool memory the client holds");
    return PX_ENOMEM;
  }
  return it->second.status;
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
  return q.status;
}

int pxs_release(pxs_server_t s, int64_t id) {
  if (!s) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second.st != St::DONE) return PX_EARG;
  if (it->second.d_out) s->pool.deallocate(it->second.d_out);
  s->reqs.erase(it);
  s->stalled = false;
  s->wake = true;
  s->cv_work.notify_all();
  return PX_EOK;
}

int pxs_drain(pxs_server_t s) {
  if (!s) return PX_EARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->completed + s->failed >= s->submitted || s->stalled; });
  if (s->completed + s->failed >= s->submitted) return PX_EOK;
  parallax::set_last_error(PX_ENOMEM, "pxs_drain: requests wait for pool memory the client holds");
  return PX_ENOMEM;
}

int pxs_get_stats(pxs_server_t s, pxs_stats* out) {
  if (!s || !out) return PX_EARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->batches = s->batches;
  out->rejected = s->rejected;
  out->failed = s->failed;
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
