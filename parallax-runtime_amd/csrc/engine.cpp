// engine.cpp -- stream contexts, tile tables and the primitive runners.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>

#include "dev_tiles.hpp"

namespace parallax {

namespace {
thread_local int t_err_code = 0;
thread_local std::string t_err_text;
EngineStats g_stats;
}  // namespace

void set_last_error(int code, const std::string& what) {
  t_err_code = code;
  t_err_text = what;
}
int last_error_code() { return t_err_code; }
const char* last_error_text() { return t_err_text.c_str(); }
EngineStats& stats() { return g_stats; }

HipBackend* global_backend() {
  static std::mutex mu;
  static std::unique_ptr<HipBackend> be;
  static bool tried = false;
  std::lock_guard<std::mutex> lk(mu);
  if (!tried) {
    tried = true;
    auto b = std::make_unique<HipBackend>();
    if (b->initialize()) be = std::move(b);
  }
  return be.get();
}

// ------------------------------------------------------------------ StreamContext
StreamContext::StreamContext(HipBackend* be, hipStream_t s) : be_(be) {
  if (!be_ || !be_->valid()) return;
  if (hipSetDevice(be_->device()) != hipSuccess) return;
  if (s) {
    stream_ = s;
  } else {
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) return;
    own_stream_ = true;
  }
  cus_ = std::max(1, be_->capabilities().compute_units);
  if (hipMalloc(reinterpret_cast<void**>(&ctl_), sizeof(pxk::Ctl)) != hipSuccess) return;
  pxk::Ctl init;
  std::memset(&init, 0, sizeof(init));
  init.epoch = 1;
  if (hipMemcpy(ctl_, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) return;
  if (hipEventCreateWithFlags(&pin_ev_, hipEventDisableTiming) != hipSuccess) return;
  ok_ = true;
  if (!reserve(1u << 16, 64u << 10)) ok_ = false;
}

StreamContext::~StreamContext() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (ctl_) (void)hipFree(ctl_);
  if (status_) (void)hipFree(status_);
  if (partials_) (void)hipFree(partials_);
  if (work_) (void)hipFree(work_);
  if (pin_) (void)hipHostFree(pin_);
  if (pin_ev_) (void)hipEventDestroy(pin_ev_);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

hipError_t wait_stream(hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  while (true) {
    const hipError_t e = hipStreamQuery(s);
    if (e != hipErrorNotReady) return e;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) break;
  }
  return hipStreamSynchronize(s);
}

bool StreamContext::grow(void** p, size_t* cap, size_t need, bool zero) {
  if (need <= *cap) return true;
  if (frozen_) return false;  // no reallocation inside a graph capture
  size_t ncap = std::max<size_t>(need, *cap * 2);
  ncap = (ncap + 4095) & ~size_t(4095);
  if (hipStreamSynchronize(stream_) != hipSuccess) return false;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, ncap) != hipSuccess) {
    // the doubled size did not fit: try the size asked for
    (void)hipGetLastError();
    *p = nullptr;
    ncap = (need + 4095) & ~size_t(4095);
    if (hipMalloc(p, ncap) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      set_last_error(PX_ERR_NOMEM, "scratch allocation of " + std::to_string(ncap) + " bytes failed");
      return false;
    }
  }
  // cleared in stream order, before any later launch on this stream reads it (never
  // inside a graph capture: a frozen context does not grow)
  if (zero && hipMemsetAsync(*p, 0, ncap, stream_) != hipSuccess) return false;
  *cap = ncap;
  return true;
}

bool StreamContext::reserve(size_t status_words, size_t partial_bytes) {
  if (!grow(reinterpret_cast<void**>(&status_), &status_cap_, status_words * 8, true)) return false;
  if (!grow(&partials_, &partial_cap_, partial_bytes, false)) return false;
  return true;
}

pxk::Scratch StreamContext::scratch() const {
  pxk::Scratch s;
  s.ctl = ctl_;
  s.status = status_;
  s.status_words = status_cap_ / 8;
  s.partials = partials_;
  s.partial_bytes = partial_cap_;
  s.max_blocks = static_cast<uint32_t>(cus_) * 8u;
  return s;
}

bool StreamContext::upload(const void* host, size_t bytes, void** dev_out) {
  if (frozen_) {
    // replaying a warmed-up call: the tables are already on the device
    if (bytes != last_upload_ || !work_) return false;
    *dev_out = work_;
    return true;
  }
  if (!grow(&work_, &work_cap_, bytes, false)) return false;
  last_upload_ = bytes;
  if (pin_pending_) {
    if (hipEventSynchronize(pin_ev_) != hipSuccess) return false;
    pin_pending_ = false;
  }
  if (bytes > pin_cap_) {
    if (pin_) (void)hipHostFree(pin_);
    pin_ = nullptr;
    pin_cap_ = 0;
    const size_t ncap = (std::max(bytes, size_t(1) << 16) + 4095) & ~size_t(4095);
    if (hipHostMalloc(&pin_, ncap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      pin_ = nullptr;
      return false;
    }
    pin_cap_ = ncap;
  }
  std::memcpy(pin_, host, bytes);
  if (hipMemcpyAsync(work_, pin_, bytes, hipMemcpyHostToDevice, stream_) != hipSuccess) return false;
  if (hipEventRecord(pin_ev_, stream_) != hipSuccess) return false;
  pin_pending_ = true;
  *dev_out = work_;
  return true;
}

bool StreamContext::note_status_launch() {
  if (++status_launches_ < pxk::kEpochMax - 2 || frozen_) return true;
  // About to wrap the 30-bit epoch: clear every status word and restart at epoch 1.
  if (hipStreamSynchronize(stream_) != hipSuccess) return false;
  if (status_ && hipMemset(status_, 0, status_cap_) != hipSuccess) return false;
  pxk::Ctl init;
  std::memset(&init, 0, sizeof(init));
  init.epoch = 1;
  if (hipMemcpy(ctl_, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) return false;
  status_launches_ = 1;
  return true;
}

// ------------------------------------------------------------------ tile tables
void build_tile_tables(const SegSpec& sp, uint32_t tile, uint32_t vec, int64_t shift,
                       TileTables& t) {
  t.tile_off.resize(sp.nseg + 1);
  t.tile_seg.clear();
  uint32_t acc = 0;
  for (uint32_t s = 0; s < sp.nseg; ++s) {
    t.tile_off[s] = acc;
    const int64_t lo = shift + static_cast<int64_t>(sp.lo[s]);
    const int64_t hi = shift + static_cast<int64_t>(sp.hi[s]);
    const int64_t w = lo & ~static_cast<int64_t>(vec - 1);
    uint64_t nt = hi > w ? static_cast<uint64_t>((hi - w + tile - 1) / tile) : 1;
    if (nt == 0) nt = 1;
    for (uint64_t k = 0; k < nt; ++k) t.tile_seg.push_back(s);
    acc += static_cast<uint32_t>(nt);
  }
  t.tile_off[sp.nseg] = acc;
  t.ntiles = acc;
}

namespace {

bool valid_dtype(int d) { return d >= 0 && d < PX_NUM_DTYPES; }

int dev_err(hipError_t e, const char* what) {
  if (e == hipSuccess) return PX_OK;
  set_last_error(PX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return PX_ERR_DEVICE;
}

// Zero `bytes` (multiple of 4) of device memory with a kernel (graph-safe).
int zero_dev(StreamContext& c, void* p, size_t bytes) {
  return dev_err(pxk::fill_u32(static_cast<uint32_t*>(p), bytes / 4, 0u, c.stream()), "fill");
}

// Tile of a plain scan / compaction: the forced shape, else the per-type default.
pxk::TileCfg plain_tile(int op, int dtype) {
  const pxk::TileCfg forced = pxk::tile_override(op, dtype);
  if (forced.block) return forced;
  return op == PX_OP_SCAN ? pxk::tune().scan[dtype] : pxk::tune().compact[dtype];
}

pxk::SegGeom plain_geom(uint64_t n, int dtype, pxk::TileCfg k, const void* in) {
  pxk::SegGeom g{};
  const uint32_t tile = pxk::tile_elems(dtype, k);
  const int64_t shift = pxk::elem_shift(in, pxk::dtype_size(dtype));
  uint64_t nt = (static_cast<uint64_t>(shift) + n + tile - 1) / tile;
  if (nt == 0) nt = 1;
  g.n = n;
  g.ntiles = static_cast<uint32_t>(nt);
  g.nseg = 1;
  g.tile_block = k.block;
  g.tile_k = k.k;
  return g;
}

bool valid_spec(const SegSpec& sp) {
  if (!sp.lo || !sp.hi) return false;
  for (uint32_t s = 0; s < sp.nseg; ++s)
    if (sp.hi[s] < sp.lo[s]) return false;
  return true;
}

// Upload segment bounds + tile tables as one blob and fill a segmented geometry.
// `extra` (optional) is appended to the same upload, its device copy returned in
// *extra_dev (one upload per call: a second one would reuse the same device buffer).
int seg_geom(StreamContext& c, const SegSpec& sp, int dtype, pxk::TileCfg k, const void* in,
             pxk::SegGeom& g, TileTables& tt, const std::vector<uint64_t>* extra = nullptr,
             const uint64_t** extra_dev = nullptr) {
  const uint32_t nseg = sp.nseg;
  const uint32_t tile = pxk::tile_elems(dtype, k);
  const uint32_t vec = static_cast<uint32_t>(16 / pxk::dtype_size(dtype));
  const int64_t shift = pxk::elem_shift(in, pxk::dtype_size(dtype));
  build_tile_tables(sp, tile, vec, shift, tt);
  const size_t b64 = static_cast<size_t>(nseg) * sizeof(uint64_t);
  const size_t o_hi = b64, o_out = 2 * b64;
  const size_t o_toff = (sp.out ? 3 : 2) * b64;
  const size_t b_toff = (nseg + 1) * sizeof(uint32_t);
  const size_t o_tseg = (o_toff + b_toff + 15) & ~size_t(15);
  const size_t b_tseg = static_cast<size_t>(tt.ntiles) * sizeof(uint32_t);
  const size_t o_extra = (o_tseg + b_tseg + 15) & ~size_t(15);
  const size_t b_extra = extra ? extra->size() * sizeof(uint64_t) : 0;
  std::vector<unsigned char> blob(o_extra + b_extra);
  if (b64) std::memcpy(blob.data(), sp.lo, b64);
  if (b64) std::memcpy(blob.data() + o_hi, sp.hi, b64);
  if (sp.out && b64) std::memcpy(blob.data() + o_out, sp.out, b64);
  if (b_extra) std::memcpy(blob.data() + o_extra, extra->data(), b_extra);
  std::memcpy(blob.data() + o_toff, tt.tile_off.data(), b_toff);
  if (b_tseg) std::memcpy(blob.data() + o_tseg, tt.tile_seg.data(), b_tseg);
  void* dev = nullptr;
  if (!c.upload(blob.data(), blob.size(), &dev)) {
    set_last_error(PX_ERR_NOMEM, "table upload");
    return PX_ERR_NOMEM;
  }
  unsigned char* d = static_cast<unsigned char*>(dev);
  g.seg_lo = reinterpret_cast<const uint64_t*>(d);
  g.seg_hi = reinterpret_cast<const uint64_t*>(d + o_hi);
  g.seg_out = sp.out ? reinterpret_cast<const uint64_t*>(d + o_out) : nullptr;
  g.tile_off = reinterpret_cast<const uint32_t*>(d + o_toff);
  g.tile_seg = reinterpret_cast<const uint32_t*>(d + o_tseg);
  if (extra_dev) *extra_dev = reinterpret_cast<const uint64_t*>(d + o_extra);
  g.n = 0;
  g.ntiles = tt.ntiles;
  g.nseg = nseg;
  g.tile_block = k.block;
  g.tile_k = k.k;
  return PX_OK;
}

}  // namespace

// ------------------------------------------------------------------ runners
int run_reduce(StreamContext& c, int dtype, const void* in, uint64_t n, void* out_dev) {
  if (!valid_dtype(dtype) || !out_dev || (n && !in)) return PX_ERR_ARG;
  ++g_stats.reduce;
  ++g_stats.launches;
  if (n == 0) return zero_dev(c, out_dev, pxk::dtype_size(dtype));
  if (!c.reserve(0, static_cast<size_t>(c.cu_count()) * 8 * 16)) return PX_ERR_NOMEM;
  return dev_err(pxk::reduce(dtype, in, n, out_dev, c.scratch(), c.stream()), "reduce");
}

int run_map_reduce(StreamContext& c, const px_desc& map, const void* in, uint64_t n, void* out_dev) {
  if (!valid_dtype(map.dtype) || !valid_dtype(map.aux) || !out_dev || (n && !in)) return PX_ERR_ARG;
  if (!pxk::map_reduce_supported(map)) {
    set_last_error(PX_ERR_ARG, "map_reduce: unsupported map for the fused reduce");
    return PX_ERR_ARG;
  }
  ++g_stats.reduce;
  ++g_stats.launches;
  if (n == 0) return zero_dev(c, out_dev, pxk::dtype_size(map.aux));
  if (!c.reserve(0, static_cast<size_t>(c.cu_count()) * 8 * 16)) return PX_ERR_NOMEM;
  return dev_err(pxk::map_reduce(map, in, n, out_dev, c.scratch(), c.stream()), "map_reduce");
}

int run_scan(StreamContext& c, int dtype, const void* in, void* out, uint64_t n, bool exclusive,
             const void* init_host) {
  if (!valid_dtype(dtype) || (n && (!in || !out))) return PX_ERR_ARG;
  if (n == 0) return PX_OK;
  ++g_stats.scan;
  ++g_stats.launches;
  const pxk::SegGeom g = plain_geom(n, dtype, plain_tile(PX_OP_SCAN, dtype), in);
  if (!c.reserve(static_cast<size_t>(g.ntiles) * (pxk::carry_bytes(dtype) / 4), 0))
    return PX_ERR_NOMEM;
  if (!c.note_status_launch()) return PX_ERR_DEVICE;
  return dev_err(pxk::scan(dtype, in, out, g, exclusive, init_host, c.scratch(), c.stream()), "scan");
}

pxk::SegGeom scan_geometry(int dtype, const void* in, uint64_t n) {
  return plain_geom(n, dtype, plain_tile(PX_OP_SCAN, dtype), in);
}

int run_sort(StreamContext& c, int dtype, void* data, uint64_t n, const pxk::LaunchHook& hook) {
  if (!valid_dtype(dtype) || (n && !data)) return PX_ERR_ARG;
  if (n <= 1) return PX_OK;
  ++g_stats.sort;
  ++g_stats.launches;
  if (pxk::radix_sort_ok(dtype, n)) {
    if (!c.reserve(pxk::radix_status_words(dtype, n), n * pxk::dtype_size(dtype))) return PX_ERR_NOMEM;
    for (int i = 0; i < pxk::radix_passes(dtype); ++i)
      if (!c.note_status_launch()) return PX_ERR_DEVICE;
    return dev_err(pxk::radix_sort(dtype, data, n, c.scratch(), c.stream(), hook), "sort");
  }
  return dev_err(pxk::sort(dtype, data, n, c.stream(), hook), "sort");
}

int run_compact(StreamContext& c, const px_desc& flags, const void* in, void* out, uint64_t n,
                uint64_t* count_dev) {
  if (!valid_dtype(flags.dtype) || !count_dev || (n && (!in || !out))) return PX_ERR_ARG;
  ++g_stats.compact;
  ++g_stats.launches;
  if (n == 0) return zero_dev(c, count_dev, sizeof(uint64_t));
  const pxk::SegGeom g = plain_geom(n, flags.dtype, plain_tile(PX_OP_FLAGS, flags.dtype), in);
  const size_t words = n >= (1ull << 32) ? 2 : 1;
  if (!c.reserve(static_cast<size_t>(g.ntiles) * words, 0)) return PX_ERR_NOMEM;
  if (!c.note_status_launch()) return PX_ERR_DEVICE;
  return dev_err(pxk::compact(flags, pxk::kModeKept, in, out, g, count_dev, nullptr,
                              c.scratch(), c.stream()),
                 "compact");
}

// Partition in std::partition's order (stable == false) or the stable order.
//   std order: count -> num_true; (out != in: copy in -> out); compaction of the indices
//   of the misplaced elements into the partials; swap pass on out.  Any in / out overlap
//   is fine only for in == out.
//   stable order: count, then one compaction pass writing every element; in place it
//   compacts into the partials and copies back (rejected elements land ahead of unread
//   input otherwise).
int run_partition(StreamContext& c, const px_desc& flags, const void* in, void* out, uint64_t n,
                  uint64_t* num_true_dev, bool stable) {
  if (!valid_dtype(flags.dtype) || !num_true_dev || (n && (!in || !out))) return PX_ERR_ARG;
  if (flags.fn == PX_PRED_UNIQUE) {
    set_last_error(PX_ERR_ARG, "partition by adjacent-unique flags is undefined");
    return PX_ERR_ARG;
  }
  ++g_stats.compact;
  ++g_stats.launches;
  if (n == 0) return zero_dev(c, num_true_dev, sizeof(uint64_t));
  const size_t es = pxk::dtype_size(flags.dtype);
  const bool wide = n >= (1ull << 32);
  const size_t idx_bytes = wide ? 8 : 4;
  // partials: [0, 256) count word of the index pass, then the index list (or the stable
  // pass's in-place staging copy); the counting pass uses the front before either
  const size_t body = stable ? ((in == out) ? n * es : 0) : n * idx_bytes;
  const size_t need = std::max<size_t>(static_cast<size_t>(c.cu_count()) * 8 * 16, 256 + body);
  const pxk::TileCfg tile = stable ? plain_tile(PX_OP_FLAGS, flags.dtype) : pxk::misplaced_tile(flags.dtype, n);
  const pxk::SegGeom g = plain_geom(n, flags.dtype, tile, in);
  const size_t words = wide ? 2 : 1;
  if (!c.reserve(static_cast<size_t>(g.ntiles) * words, need)) return PX_ERR_NOMEM;
  int rc = dev_err(pxk::count_if(flags, in, n, num_true_dev, c.scratch(), c.stream()), "count");
  if (rc) return rc;
  if (!c.note_status_launch()) return PX_ERR_DEVICE;
  const pxk::Scratch sc = c.scratch();
  unsigned char* part = static_cast<unsigned char*>(sc.partials);
  if (stable) {
    void* dst = (in == out) ? static_cast<void*>(part + 256) : out;
    rc = dev_err(pxk::compact(flags, pxk::kModeStable, in, dst, g, nullptr, num_true_dev, sc, c.stream()),
                 "partition");
    if (rc || in != out) return rc;
    return dev_err(hipMemcpyAsync(out, part + 256, n * es, hipMemcpyDeviceToDevice, c.stream()), "partition copy");
  }
  if (in != out) {
    rc = dev_err(hipMemcpyAsync(out, in, n * es, hipMemcpyDeviceToDevice, c.stream()), "partition copy");
    if (rc) return rc;
  }
  uint64_t* count2 = reinterpret_cast<uint64_t*>(part);
  rc = dev_err(pxk::compact(flags, pxk::kModeMisplaced, in, part + 256, g, count2, num_true_dev, sc,
                            c.stream()),
               "partition");
  if (rc) return rc;
  return dev_err(pxk::swap_pairs(flags.dtype, out, part + 256, wide, count2, n, sc, c.stream()), "partition swap");
}

int run_transform(StreamContext& c, const px_desc& d, const void* in, void* out, uint64_t n,
                  const void* captures, size_t capture_size, uint64_t host_base,
                  uint64_t dev_base, uint64_t arena_bytes) {
  if (!valid_dtype(d.dtype) || !valid_dtype(d.aux) || d.fn >= PX_NUM_MAPOPS) return PX_ERR_ARG;
  if (n && (!in || !out)) return PX_ERR_ARG;
  ++g_stats.transform;
  ++g_stats.launches;
  pxk::MapArgs a;
  std::memset(&a, 0, sizeof(a));
  a.count_lo = static_cast<uint32_t>(n);
  a.host_base = host_base;
  a.dev_base = dev_base;
  a.arena_bytes = arena_bytes;
  const size_t cs = std::min(capture_size, sizeof(a.cap));
  if (captures && cs) std::memcpy(a.cap, captures, cs);
  a.cap_size = captures ? static_cast<uint32_t>(cs) : 0u;
  return dev_err(pxk::transform(d, in, out, n, a, c.stream()), "transform");
}

int run_find(StreamContext& c, const px_desc& pred, const void* in, uint64_t n, uint64_t* out_dev) {
  if (!valid_dtype(pred.dtype) || !out_dev) return PX_ERR_ARG;
  ++g_stats.launches;
  return dev_err(pxk::find_first(pred, in, n, out_dev, c.scratch(), c.stream()), "find");
}

int run_argminmax(StreamContext& c, int dtype, const void* in, uint64_t n, bool want_max,
                  bool want_last, uint64_t* out_dev) {
  if (!valid_dtype(dtype) || !out_dev) return PX_ERR_ARG;
  ++g_stats.launches;
  if (!c.reserve(0, static_cast<size_t>(c.cu_count()) * 8 * 16)) return PX_ERR_NOMEM;
  return dev_err(pxk::argminmax(dtype, in, n, want_max, want_last, out_dev, c.scratch(), c.stream()),
                 "argminmax");
}

int run_mismatch(StreamContext& c, int dtype, const void* a, const void* b, uint64_t n,
                 uint64_t* out_dev) {
  if (!valid_dtype(dtype) || !out_dev) return PX_ERR_ARG;
  ++g_stats.launches;
  return dev_err(pxk::mismatch(dtype, a, b, n, out_dev, c.scratch(), c.stream()), "mismatch");
}

std::vector<pxk::TileCfg> seg_tile_candidates(int op, int dtype) {
  std::vector<pxk::TileCfg> v;
  if (op == PX_OP_REDUCE) {
    for (int k : {4, 8, 16}) v.push_back({256, k});
  } else {
    for (pxk::TileCfg t : {pxk::TileCfg{256, 4}, pxk::TileCfg{256, 8}, pxk::TileCfg{512, 8}, pxk::TileCfg{1024, 4},
                           pxk::TileCfg{1024, 8}})
      if (pxk::tile_cfg_supported(t, op)) v.push_back(t);
  }
  return v;
}

pxk::TileCfg seg_tile_default(int op, int dtype, const SegSpec& sp) {
  const pxk::TileCfg forced = pxk::tile_override(op, dtype);
  if (forced.block) return forced;
  uint64_t total = 0;
  for (uint32_t s = 0; s < sp.nseg; ++s) total += sp.hi[s] - sp.lo[s];
  const uint64_t mean = sp.nseg ? total / sp.nseg : 0;
  const pxk::TileCfg big = op == PX_OP_REDUCE ? pxk::tune().seg_reduce
                          : op == PX_OP_SCAN  ? pxk::tune().scan[dtype]
                                              : pxk::tune().compact[dtype];
  pxk::TileCfg best{0, 0};
  for (const pxk::TileCfg& t : seg_tile_candidates(op, dtype)) {
    if (best.block == 0) best = t;  // smallest candidate first
    if (op == PX_OP_REDUCE) {
      // reduce tiles sync once per tile and leave one partial each: long segments take
      // the largest tile of at most a quarter of the mean length (served f64 reduces of
      // ~2M elements: 256 x 16)
      if (pxk::tile_elems(dtype, t) <= std::max<uint64_t>(mean / 4, pxk::tile_elems(dtype, big))) best = t;
    } else if (pxk::tile_elems(dtype, t) <= mean && pxk::tile_elems(dtype, t) <= pxk::tile_elems(dtype, big)) {
      best = t;
    }
  }
  return best;
}

int run_seg_reduce(StreamContext& c, int dtype, const void* in, const SegSpec& sp, void* out_dev,
                   pxk::TileCfg tile) {
  if (!valid_dtype(dtype) || !out_dev || !in || !valid_spec(sp)) return PX_ERR_ARG;
  if (sp.nseg == 0) return PX_OK;
  ++g_stats.seg;
  ++g_stats.launches;
  pxk::SegGeom g;
  TileTables tt;
  if (tile.block == 0) tile = seg_tile_default(PX_OP_REDUCE, dtype, sp);
  int rc = seg_geom(c, sp, dtype, tile, in, g, tt);
  if (rc) return rc;
  if (!c.reserve(0, static_cast<size_t>(g.ntiles) * 16)) return PX_ERR_NOMEM;
  return dev_err(pxk::seg_reduce(dtype, in, g, out_dev, c.scratch(), c.stream()), "seg_reduce");
}

int run_seg_scan(StreamContext& c, int dtype, const void* in, void* out, const SegSpec& sp,
                 pxk::TileCfg tile) {
  if (!valid_dtype(dtype) || !in || !out || !valid_spec(sp)) return PX_ERR_ARG;
  if (sp.nseg == 0) return PX_OK;
  ++g_stats.seg;
  ++g_stats.launches;
  // Segments of at most kShortSeg elements take one wave each (seg_scan_short); a
  // look-back tile would spend a block's fixed work on them (B1, 16,384 f32 segments:
  // the 9,053 of at most 1,024 elements cost 86 us of a 706 us call as tiles).  A wave
  // walks its segment 64 elements at a time, so longer ones lose: B1 at 1,024 / 2,048 /
  // 4,096 0.589 / 0.600 / 0.623 ms (512 x 8 tiles, flushed).
  constexpr uint64_t kShortSeg = 1024;
  std::vector<uint64_t> llo, lhi, lout, shrt;  // shrt: lo[], hi[], out[] of the short ones
  std::vector<uint32_t> sidx;
  for (uint32_t s = 0; s < sp.nseg; ++s) {
    if (sp.hi[s] - sp.lo[s] <= kShortSeg) {
      sidx.push_back(s);
    } else {
      llo.push_back(sp.lo[s]);
      lhi.push_back(sp.hi[s]);
      if (sp.out) lout.push_back(sp.out[s]);
    }
  }
  const uint32_t nshort = static_cast<uint32_t>(sidx.size());
  shrt.resize(static_cast<size_t>(nshort) * (sp.out ? 3 : 2));
  for (uint32_t j = 0; j < nshort; ++j) {
    shrt[j] = sp.lo[sidx[j]];
    shrt[nshort + j] = sp.hi[sidx[j]];
    if (sp.out) shrt[2 * static_cast<size_t>(nshort) + j] = sp.out[sidx[j]];
  }
  SegSpec lsp;
  lsp.lo = llo.data();
  lsp.hi = lhi.data();
  lsp.out = sp.out ? lout.data() : nullptr;
  lsp.nseg = static_cast<uint32_t>(llo.size());
  pxk::SegGeom g;
  TileTables tt;
  if (tile.block == 0) tile = seg_tile_default(PX_OP_SCAN, dtype, lsp.nseg ? lsp : sp);
  const uint64_t* sdev = nullptr;
  int rc = seg_geom(c, lsp, dtype, tile, in, g, tt, &shrt, &sdev);
  if (rc) return rc;
  if (g.ntiles) {
    if (!c.reserve(static_cast<size_t>(g.ntiles) * (pxk::carry_bytes(dtype) / 4), 0))
      return PX_ERR_NOMEM;
    if (!c.note_status_launch()) return PX_ERR_DEVICE;
    rc = dev_err(pxk::scan(dtype, in, out, g, false, nullptr, c.scratch(), c.stream()), "seg_scan");
    if (rc) return rc;
  }
  return dev_err(pxk::seg_scan_short(dtype, in, out, sdev, sdev + nshort, sp.out ? sdev + 2 * static_cast<size_t>(nshort) : nullptr,
                                     nshort, c.scratch(), c.stream()),
                 "seg_scan");
}

int run_seg_compact(StreamContext& c, const px_desc& flags, const void* in, void* out,
                    const SegSpec& sp, uint64_t* counts_dev, pxk::TileCfg tile) {
  if (!valid_dtype(flags.dtype) || !in || !out || !counts_dev || !valid_spec(sp)) return PX_ERR_ARG;
  if (sp.nseg == 0) return PX_OK;
  ++g_stats.seg;
  ++g_stats.launches;
  pxk::SegGeom g;
  TileTables tt;
  if (tile.block == 0) tile = seg_tile_default(PX_OP_FLAGS, flags.dtype, sp);
  int rc = seg_geom(c, sp, flags.dtype, tile, in, g, tt);
  if (rc) return rc;
  if (!c.reserve(static_cast<size_t>(g.ntiles), 0)) return PX_ERR_NOMEM;
  if (!c.note_status_launch()) return PX_ERR_DEVICE;
  return dev_err(pxk::compact(flags, pxk::kModeKept, in, out, g, counts_dev, nullptr,
                              c.scratch(), c.stream()),
                 "seg_compact");
}

namespace {
// Which segments of a batched sort take the segmented radix sort, and where each sits in
// the packed scratch buffers (segments longer than seg_radix_min_len -- 1,024 4-byte /
// 512 8-byte keys -- and shorter than 2^31 keys, when the radix path is enabled).
struct LongSegs {
  std::vector<uint32_t> idx;  // segment index in the batch
  std::vector<uint64_t> lo, hi, pack;
  uint64_t packed = 0;  // elements of one scratch buffer
};
LongSegs long_segments(int dtype, const SegSpec& sp) {
  LongSegs L;
  if (pxk::tune().sort_radix_min == 0) return L;
  const uint64_t long_min = pxk::seg_radix_min_len(dtype);
  const uint64_t vec = 16 / pxk::dtype_size(dtype);
  for (uint32_t s = 0; s < sp.nseg; ++s) {
    const uint64_t len = sp.hi[s] - sp.lo[s];
    if (len <= long_min || len >= (1ull << 31)) continue;
    // same position within a 16-byte line as in the array
    const uint64_t at = (L.packed + vec - 1) / vec * vec + sp.lo[s] % vec;
    L.idx.push_back(s);
    L.lo.push_back(sp.lo[s]);
    L.hi.push_back(sp.hi[s]);
    L.pack.push_back(at);
    L.packed = at + len;
  }
  return L;
}
size_t round256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

size_t seg_sort_scratch_bytes(int dtype, const SegSpec& sp) {
  if (!valid_dtype(dtype) || !valid_spec(sp)) return 0;
  const LongSegs L = long_segments(dtype, sp);
  if (L.lo.empty()) return 0;
  return 2 * round256(L.packed * pxk::dtype_size(dtype)) +
         pxk::seg_radix_cnt_bytes(dtype, static_cast<uint32_t>(L.lo.size()));
}

int run_seg_sort(StreamContext& c, int dtype, void* data, const SegSpec& sp, void* ext_scratch,
                 size_t ext_bytes, hipStream_t pass_stream, hipEvent_t prep_done) {
  // with a pass stream: whatever this call queued on c.stream() precedes later work there
  auto join = [&](int rc) {
    if (!pass_stream || !prep_done) return rc;
    if (hipEventRecord(prep_done, c.stream()) != hipSuccess ||
        hipStreamWaitEvent(pass_stream, prep_done, 0) != hipSuccess) {
      (void)hipGetLastError();
      if (rc == PX_OK) rc = dev_err(hipErrorUnknown, "seg_sort: stream join");
    }
    return rc;
  };
  if (!valid_dtype(dtype) || !data || !valid_spec(sp)) return PX_ERR_ARG;
  if (sp.nseg == 0) return PX_OK;
  ++g_stats.seg;
  ++g_stats.launches;
  // Long segments go to the segmented radix sort (long_segments); the rest stay on the
  // bitonic path, which sorts each in one launch.
  const LongSegs L = long_segments(dtype, sp);
  std::vector<uint64_t> blo(sp.nseg), bhi(sp.nseg);
  std::vector<uint32_t> rfirst, rtseg;
  for (uint32_t s = 0; s < sp.nseg; ++s) {
    blo[s] = sp.lo[s];
    bhi[s] = sp.hi[s];
  }
  for (uint32_t s : L.idx) bhi[s] = blo[s];  // not on the bitonic path
  const uint32_t nl = static_cast<uint32_t>(L.lo.size());
  const uint32_t rtile = pxk::seg_radix_tile(dtype, L.lo.data(), L.hi.data(), nl);
  for (uint32_t ls = 0; ls < nl; ++ls) {
    rfirst.push_back(static_cast<uint32_t>(rtseg.size()));
    for (uint64_t k = 0; k < (L.hi[ls] - L.lo[ls] + rtile - 1) / rtile; ++k) rtseg.push_back(ls);
  }
  pxk::SegSortPlan p;
  if (!pxk::seg_sort_plan(dtype, blo.data(), bhi.data(), sp.nseg, p)) return PX_ERR_ARG;
  // blob: interleaved bitonic bounds (lo, hi) as u64, the bitonic block tables (u32),
  // then the radix tables: seg_lo, seg_hi, seg_pack (u64), seg_first, tile_seg (u32)
  const size_t b_b = static_cast<size_t>(sp.nseg) * 2 * sizeof(uint64_t);
  const size_t o_r = (b_b + p.tables.size() * sizeof(uint32_t) + 15) & ~size_t(15);
  const size_t o_rhi = o_r + nl * sizeof(uint64_t);
  const size_t o_rpack = o_rhi + nl * sizeof(uint64_t);
  const size_t o_rfirst = o_rpack + nl * sizeof(uint64_t);
  const size_t o_rtseg = o_rfirst + nl * sizeof(uint32_t);
  std::vector<unsigned char> blob(o_rtseg + rtseg.size() * sizeof(uint32_t));
  uint64_t* bounds = reinterpret_cast<uint64_t*>(blob.data());
  for (uint32_t s = 0; s < sp.nseg; ++s) {
    bounds[2 * s] = blo[s];
    bounds[2 * s + 1] = bhi[s];
  }
  if (!p.tables.empty())
    std::memcpy(blob.data() + b_b, p.tables.data(), p.tables.size() * sizeof(uint32_t));
  if (nl) {
    std::memcpy(blob.data() + o_r, L.lo.data(), nl * sizeof(uint64_t));
    std::memcpy(blob.data() + o_rhi, L.hi.data(), nl * sizeof(uint64_t));
    std::memcpy(blob.data() + o_rpack, L.pack.data(), nl * sizeof(uint64_t));
    std::memcpy(blob.data() + o_rfirst, rfirst.data(), nl * sizeof(uint32_t));
    std::memcpy(blob.data() + o_rtseg, rtseg.data(), rtseg.size() * sizeof(uint32_t));
  }
  // scratch: second buffer, third buffer (the long segments packed), then the counts;
  // from the caller (the server's pool, counted against its budget) or the context
  const size_t buf_bytes = round256(L.packed * pxk::dtype_size(dtype));
  const size_t scratch_bytes = nl ? 2 * buf_bytes + pxk::seg_radix_cnt_bytes(dtype, nl) : 0;
  if (nl && !c.reserve(static_cast<size_t>(rtseg.size()) * 256, ext_scratch ? 0 : scratch_bytes))
    return PX_ERR_NOMEM;
  if (nl && ext_scratch && ext_bytes < scratch_bytes) {
    set_last_error(PX_ERR_ARG, "seg_sort: external scratch too small");
    return PX_ERR_ARG;
  }
  void* dev = nullptr;
  if (!c.upload(blob.data(), blob.size(), &dev)) return PX_ERR_NOMEM;
  unsigned char* d = static_cast<unsigned char*>(dev);
  const uint64_t* bd = reinterpret_cast<const uint64_t*>(d);
  const uint32_t* tab_d = reinterpret_cast<const uint32_t*>(d + b_b);
  int rc = dev_err(pxk::seg_sort_run(p, data, bd, tab_d, c.stream()), "seg_sort");
  if (rc != PX_OK || nl == 0) return join(rc);
  for (size_t i = 0; i < pxk::dtype_size(dtype); ++i)
    if (!c.note_status_launch()) return join(PX_ERR_DEVICE);
  const pxk::Scratch sc = c.scratch();
  unsigned char* base = static_cast<unsigned char*>(ext_scratch ? ext_scratch : sc.partials);
  pxk::RadixSegs g;
  g.tile = rtile;
  g.ntiles = static_cast<uint32_t>(rtseg.size());
  g.nseg = nl;
  g.seg_lo = reinterpret_cast<const uint64_t*>(d + o_r);
  g.seg_hi = reinterpret_cast<const uint64_t*>(d + o_rhi);
  g.seg_pack = reinterpret_cast<const uint64_t*>(d + o_rpack);
  g.seg_first = reinterpret_cast<const uint32_t*>(d + o_rfirst);
  g.tile_seg = reinterpret_cast<const uint32_t*>(d + o_rtseg);
  g.tmp_b = base;
  g.tmp2 = base + buf_bytes;
  g.seg_cnt = reinterpret_cast<uint32_t*>(base + 2 * buf_bytes);
  rc = dev_err(pxk::seg_radix_sort(dtype, data, g, sc, c.stream(), pass_stream, prep_done), "seg_radix_sort");
  // seg_radix_sort joins the streams itself once its counting kernels are queued; a
  // failure before that still needs the join
  return rc == PX_OK ? rc : join(rc);
}

}  // namespace parallax
