// engine.hpp -- stream contexts and the primitive runners every front end shares
// (the C ABI's KernelLauncher, the device-resident pxd_* API, batched plans, the
// server).  Internal to libparallax_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "dev_tiles.hpp"
#include "kernels.hpp"
#include "parallax/descriptor.h"
#include "parallax/hip_backend.hpp"

namespace parallax {

// Waits for the stream's work: polls for up to ~200 us (a short call completes and is
// seen within a microsecond or two), then blocks in the runtime for long ones.
hipError_t wait_stream(hipStream_t s);

// Error codes of the pxd_* / pxb_* / pxs_* entry points (0 = success).
enum PxStatus {
  PX_OK = 0,
  PX_ERR_ARG = -1,       // bad argument (dtype, size, null pointer, descriptor)
  PX_ERR_DEVICE = -2,    // HIP error
  PX_ERR_NODEV = -3,     // no GPU / runtime not initialised
  PX_ERR_NOMEM = -4,     // allocation failed
  PX_ERR_UNSUPPORTED = -5
};

// Records the last error of this thread (pxr_last_error()).
void set_last_error(int code, const std::string& what);
int last_error_code();
const char* last_error_text();

// One stream plus the scratch its kernels need.  Kernels on one context run in stream
// order; two contexts never share scratch, so they may run concurrently.
class StreamContext {
 public:
  StreamContext(HipBackend* be, hipStream_t stream);  // stream == nullptr: own stream
  ~StreamContext();
  StreamContext(const StreamContext&) = delete;
  StreamContext& operator=(const StreamContext&) = delete;

  bool ok() const { return ok_; }
  hipStream_t stream() const { return stream_; }
  HipBackend* backend() const { return be_; }
  int cu_count() const { return cus_; }

  // Grow scratch (synchronises the stream before freeing anything in use).
  bool reserve(size_t status_words, size_t partial_bytes);
  pxk::Scratch scratch() const;
  // Device temp area for per-call tables; contents are valid in stream order until
  // the next call that uses it.  Uploads go through a pinned staging buffer.
  bool upload(const void* host, size_t bytes, void** dev_out);
  pxk::Ctl* ctl() const { return ctl_; }
  // Account a launch that uses look-back status; clears status before the epoch wraps.
  bool note_status_launch();
  // While frozen (graph capture of a warmed-up call), upload() returns the previous
  // upload's device copy without copying and reserve() must not need to grow.
  void freeze(bool f) { frozen_ = f; }
  bool frozen() const { return frozen_; }

  std::mutex mu;  // serialises API calls on this context

 private:
  bool grow(void** p, size_t* cap, size_t need, bool zero);
  HipBackend* be_ = nullptr;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  bool ok_ = false;
  int cus_ = 256;
  pxk::Ctl* ctl_ = nullptr;
  uint64_t* status_ = nullptr;
  size_t status_cap_ = 0;  // bytes
  void* partials_ = nullptr;
  size_t partial_cap_ = 0;
  void* work_ = nullptr;
  size_t work_cap_ = 0;
  void* pin_ = nullptr;
  size_t pin_cap_ = 0;
  hipEvent_t pin_ev_ = nullptr;
  bool pin_pending_ = false;
  uint64_t status_launches_ = 0;
  bool frozen_ = false;
  size_t last_upload_ = 0;
};

// Segments of a batched call (host arrays of nseg entries): segment s is input
// elements [lo[s], hi[s]) and writes its outputs from out[s] (out == nullptr: lo[s]).
// The contiguous form (offsets o[0..nseg]) is lo = o, hi = o + 1.
struct SegSpec {
  const uint64_t* lo = nullptr;
  const uint64_t* hi = nullptr;
  const uint64_t* out = nullptr;
  uint32_t nseg = 0;
};

// Tile tables of a segmented call (host side).
struct TileTables {
  std::vector<uint32_t> tile_off;  // nseg + 1
  std::vector<uint32_t> tile_seg;  // ntiles
  uint32_t ntiles = 0;
};
void build_tile_tables(const SegSpec& sp, uint32_t tile, uint32_t vec, int64_t shift,
                       TileTables& t);

// ---- primitive runners on a context (device pointers, stream-ordered, no host sync)
int run_reduce(StreamContext& c, int dtype, const void* in, uint64_t n, void* out_dev);
// Fused Σ map(x) (pxk::map_reduce); the result (map.aux type) goes to out_dev.
int run_map_reduce(StreamContext& c, const px_desc& map, const void* in, uint64_t n, void* out_dev);
int run_scan(StreamContext& c, int dtype, const void* in, void* out, uint64_t n, bool exclusive,
             const void* init_host);
// Tile geometry of a plain scan of n elements at `in` (the shape run_scan uses).
pxk::SegGeom scan_geometry(int dtype, const void* in, uint64_t n);
int run_sort(StreamContext& c, int dtype, void* data, uint64_t n, const pxk::LaunchHook& hook = {});
// copy_if / remove_if / unique: count to *count_dev (device u64)
int run_compact(StreamContext& c, const px_desc& flags, const void* in, void* out, uint64_t n,
                uint64_t* count_dev);
// Partition of in into out (in == out: in place): std::partition's order, or with
// stable = true the stable order (kept first, the rest after, both in input order).
int run_partition(StreamContext& c, const px_desc& flags, const void* in, void* out, uint64_t n,
                  uint64_t* num_true_dev, bool stable = false);
int run_transform(StreamContext& c, const px_desc& d, const void* in, void* out, uint64_t n,
                  const void* captures, size_t capture_size, uint64_t host_base,
                  uint64_t dev_base, uint64_t arena_bytes);
int run_find(StreamContext& c, const px_desc& pred, const void* in, uint64_t n, uint64_t* out_dev);
int run_argminmax(StreamContext& c, int dtype, const void* in, uint64_t n, bool want_max,
                  bool want_last, uint64_t* out_dev);
int run_mismatch(StreamContext& c, int dtype, const void* a, const void* b, uint64_t n,
                 uint64_t* out_dev);

// segmented (batched) runners
// Batched runners.  `tile` picks the tile shape ({0, 0}: seg_tile_default()).
int run_seg_reduce(StreamContext& c, int dtype, const void* in, const SegSpec& sp, void* out_dev,
                   pxk::TileCfg tile = {0, 0});
int run_seg_scan(StreamContext& c, int dtype, const void* in, void* out, const SegSpec& sp,
                 pxk::TileCfg tile = {0, 0});
// Batched sort.  The segmented radix sort of the long segments needs
// seg_sort_scratch_bytes() of scratch: from ext_scratch (device memory of the caller,
// e.g. the server's pool) when given, else from the context.
size_t seg_sort_scratch_bytes(int dtype, const SegSpec& sp);
// With `pass_stream` and `prep_done`, the digit passes (persistent look-back kernels)
// run on pass_stream and everything else on c.stream(); every launch on c.stream() is
// ordered before later work on pass_stream (prep_done recorded there and waited for),
// whether or not any digit pass runs.
int run_seg_sort(StreamContext& c, int dtype, void* data, const SegSpec& sp,
                 void* ext_scratch = nullptr, size_t ext_bytes = 0, hipStream_t pass_stream = nullptr,
                 hipEvent_t prep_done = nullptr);
int run_seg_compact(StreamContext& c, const px_desc& flags, const void* in, void* out,
                    const SegSpec& sp, uint64_t* counts_dev, pxk::TileCfg tile = {0, 0});
// Default tile of a batched call: the largest candidate no longer than the mean
// segment (small segments waste less of a small tile).
pxk::TileCfg seg_tile_default(int op, int dtype, const SegSpec& sp);
// Tile shapes the autotuner tries for `op` (REDUCE, SCAN, FLAGS).
std::vector<pxk::TileCfg> seg_tile_candidates(int op, int dtype);

// Counters of device work actually launched (tests check the GPU path ran).
struct EngineStats {
  uint64_t launches = 0;
  uint64_t reduce = 0, scan = 0, sort = 0, compact = 0, transform = 0, seg = 0;
};
EngineStats& stats();

// The process-wide device (lazy; nullptr without a GPU or with PARALLAX_FORCE_CPU).
HipBackend* global_backend();

}  // namespace parallax
