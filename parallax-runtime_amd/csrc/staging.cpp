// staging.cpp -- pinned-chunk pipeline for pageable host operands (staging.hpp).
#include "staging.hpp"

#include <algorithm>
#include <cstring>

namespace parallax {

HostStager::~HostStager() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_job_.notify_all();
  for (auto& t : threads_) t.join();
  for (int b = 0; b < kRing; ++b) {
    if (ev_[b]) (void)hipEventSynchronize(ev_[b]);
    if (ev_[b]) (void)hipEventDestroy(ev_[b]);
    if (pin_[b]) (void)hipHostFree(pin_[b]);
  }
  for (int s = 0; s < kSlots; ++s)
    if (dev_[s]) (void)hipFree(dev_[s]);
}

void* HostStager::device_buffer(int slot, size_t bytes) {
  if (slot < 0 || slot >= kSlots) return nullptr;
  if (bytes <= dev_cap_[slot] && dev_[slot]) return dev_[slot];
  if (dev_[slot]) (void)hipFree(dev_[slot]);
  dev_[slot] = nullptr;
  dev_cap_[slot] = 0;
  const size_t cap = (std::max<size_t>(bytes, 1) + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
  if (hipMalloc(&dev_[slot], cap) != hipSuccess) {
    (void)hipGetLastError();
    dev_[slot] = nullptr;
    return nullptr;
  }
  dev_cap_[slot] = cap;
  return dev_[slot];
}

bool HostStager::ensure_ring() {
  if (ring_ok_) return true;
  for (int b = 0; b < kRing; ++b) {
    if (!pin_[b] && hipHostMalloc(&pin_[b], kChunk, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      pin_[b] = nullptr;
      return false;
    }
    if (!ev_[b] && hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      ev_[b] = nullptr;
      return false;
    }
  }
  // copy threads: half the host's cores, at most 8 (memcpy bandwidth saturates there)
  const unsigned hw = std::thread::hardware_concurrency();
  const int n = static_cast<int>(std::clamp<unsigned>(hw / 2, 1u, 8u)) - 1;
  for (int i = 0; i < n; ++i) threads_.emplace_back([this, i] { worker(i); });
  ring_ok_ = true;
  return true;
}

namespace {
// Byte range of part p of `parts` over `bytes` (64-byte aligned boundaries).
void part_range(size_t bytes, int parts, int p, size_t* lo, size_t* hi) {
  const size_t per = ((bytes + parts - 1) / parts + 63) & ~size_t(63);
  *lo = std::min(bytes, per * static_cast<size_t>(p));
  *hi = std::min(bytes, *lo + per);
}
}  // namespace

void HostStager::worker(int id) {
  uint64_t seen = 0;
  while (true) {
    char* dst;
    const char* src;
    size_t bytes;
    int parts;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_job_.wait(lk, [&] { return stop_ || job_gen_ != seen; });
      if (stop_) return;
      seen = job_gen_;
      dst = job_dst_;
      src = job_src_;
      bytes = job_bytes_;
      parts = job_parts_;
    }
    size_t lo, hi;
    part_range(bytes, parts, id, &lo, &hi);
    if (hi > lo) std::memcpy(dst + lo, src + lo, hi - lo);
    {
      std::lock_guard<std::mutex> lk(mu_);
      --pending_;
    }
    cv_done_.notify_one();
  }
}

void HostStager::parallel_copy(void* dst, const void* src, size_t bytes) {
  const int nw = static_cast<int>(threads_.size());
  if (nw == 0 || bytes < (size_t(1) << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const int parts = nw + 1;
  {
    std::lock_guard<std::mutex> lk(mu_);
    job_dst_ = static_cast<char*>(dst);
    job_src_ = static_cast<const char*>(src);
    job_bytes_ = bytes;
    job_parts_ = parts;
    pending_ = nw;
    ++job_gen_;
  }
  cv_job_.notify_all();
  size_t lo, hi;
  part_range(bytes, parts, nw, &lo, &hi);  // the caller's share: the last part
  if (hi > lo) std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, hi - lo);
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return pending_ == 0; });
}

bool HostStager::h2d(void* dev_dst, const void* host_src, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return true;
  if (!ensure_ring()) return false;
  const char* src = static_cast<const char*>(host_src);
  char* dst = static_cast<char*>(dev_dst);
  for (size_t off = 0, i = 0; off < bytes; off += kChunk, ++i) {
    const size_t len = std::min(kChunk, bytes - off);
    const int b = static_cast<int>(i % kRing);
    // the chunk's pinned buffer is free once its previous DMA has completed
    if (hipEventSynchronize(ev_[b]) != hipSuccess) return false;
    parallel_copy(pin_[b], src + off, len);
    if (hipMemcpyAsync(dst + off, pin_[b], len, hipMemcpyHostToDevice, stream) != hipSuccess) return false;
    if (hipEventRecord(ev_[b], stream) != hipSuccess) return false;
  }
  return true;
}

bool HostStager::d2h(void* host_dst, const void* dev_src, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return true;
  if (!ensure_ring()) return false;
  const char* src = static_cast<const char*>(dev_src);
  char* dst = static_cast<char*>(host_dst);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t i) {
    const int b = static_cast<int>(i % kRing);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    if (hipEventSynchronize(ev_[b]) != hipSuccess) return false;  // an earlier DMA of the buffer
    if (hipMemcpyAsync(pin_[b], src + off, len, hipMemcpyDeviceToHost, stream) != hipSuccess) return false;
    return hipEventRecord(ev_[b], stream) == hipSuccess;
  };
  for (size_t i = 0; i < std::min<size_t>(n, kRing); ++i)
    if (!issue(i)) return false;
  for (size_t i = 0; i < n; ++i) {
    const int b = static_cast<int>(i % kRing);
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    if (hipEventSynchronize(ev_[b]) != hipSuccess) return false;
    parallel_copy(dst + off, pin_[b], len);
    if (i + kRing < n && !issue(i + kRing)) return false;
  }
  return true;
}

}  // namespace parallax
