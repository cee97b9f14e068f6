// staging.hpp -- host <-> device transfers of ordinary (pageable) host memory for the
// C ABI's host operands.  Internal to libparallax_amd.so.
//
// Pageable memory cannot be DMA'd directly (and registering it is not available), so it
// goes through a ring of pinned chunks: host threads copy chunk i into a pinned buffer
// while the DMA engine moves chunk i-1 to the device (and the reverse for results).
// Device-side staging buffers are kept between calls (grow-only), so a call with host
// operands does no hipMalloc / hipFree once sizes have been seen.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

namespace parallax {

class HostStager {
 public:
  HostStager() = default;
  ~HostStager();
  HostStager(const HostStager&) = delete;
  HostStager& operator=(const HostStager&) = delete;

  // Device buffer `slot` (0..kSlots-1) of at least `bytes`, kept for later calls.  The
  // caller guarantees that no work using the previous buffer of the slot is in flight.
  void* device_buffer(int slot, size_t bytes);
  // host -> device on `stream`: returns once the host memory may be reused (the last
  // chunk has been copied into pinned memory); the device copy completes in stream order.
  bool h2d(void* dev_dst, const void* host_src, size_t bytes, hipStream_t stream);
  // device -> host on `stream` after the work already queued there; returns when the
  // host memory holds the bytes.
  bool d2h(void* host_dst, const void* dev_src, size_t bytes, hipStream_t stream);

  static constexpr int kSlots = 3;

 private:
  bool ensure_ring();
  void parallel_copy(void* dst, const void* src, size_t bytes);
  void worker(int id);

  static constexpr size_t kChunk = size_t(16) << 20;  // bytes per pinned chunk
  static constexpr int kRing = 4;                     // pinned chunks in flight

  void* dev_[kSlots] = {};
  size_t dev_cap_[kSlots] = {};
  void* pin_[kRing] = {};
  hipEvent_t ev_[kRing] = {};
  bool ring_ok_ = false;

  // copy threads: every worker copies its share of the current job, the caller the rest
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_job_, cv_done_;
  uint64_t job_gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
  char* job_dst_ = nullptr;
  const char* job_src_ = nullptr;
  size_t job_bytes_ = 0;
  int job_parts_ = 1;
};

}  // namespace parallax
