// backend_arena.cpp -- HIP device layer and the device arena with its pinned mirror.
// (Reference counterparts: src/backend/vulkan/device.cpp, src/memory/arena.cpp.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>

#include "parallax/arena.hpp"
#include "parallax/hip_backend.hpp"

namespace parallax {

namespace {
inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
bool env_on(const char* name) {
  const char* v = std::getenv(name);
  return v != nullptr && v[0] != '\0' && std::strcmp(v, "0") != 0;
}
}  // namespace

// ------------------------------------------------------------------ HipBackend
HipBackend::~HipBackend() { destroy(); }

bool HipBackend::initialize(int device) {
  if (ok_) return true;
  if (env_on("PARALLAX_FORCE_CPU")) return false;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return false;
  }
  device_ = device < 0 ? 0 : device;
  if (device_ >= count) return false;
  if (hipSetDevice(device_) != hipSuccess) return false;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device_) != hipSuccess) return false;
  caps_.name = p.name;
  caps_.arch = p.gcnArchName;
  caps_.compute_units = p.multiProcessorCount;
  caps_.wave_size = p.warpSize;
  caps_.hbm_bytes = p.totalGlobalMem;
  caps_.lds_per_block = p.sharedMemPerBlock;
  caps_.l2_bytes = p.l2CacheSize;
  caps_.clock_khz = p.clockRate;
  caps_.mem_clock_khz = p.memoryClockRate;
  caps_.mem_bus_bits = p.memoryBusWidth;
  // default priority on both streams
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) return false;
  if (hipStreamCreateWithFlags(&side_, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    return false;
  }
  ok_ = true;
  return true;
}

void HipBackend::destroy() {
  if (!ok_) return;
  (void)hipStreamSynchronize(stream_);
  (void)hipStreamSynchronize(side_);
  (void)hipStreamDestroy(stream_);
  (void)hipStreamDestroy(side_);
  stream_ = side_ = nullptr;
  ok_ = false;
}

bool HipBackend::sync() { return ok_ && hipStreamSynchronize(stream_) == hipSuccess; }

// ------------------------------------------------------------------ SubAllocator
void SubAllocator::reset(size_t capacity, size_t min_align) {
  capacity_ = capacity;
  min_align_ = min_align;
  used_ = high_water_ = 0;
  free_.clear();
  live_.clear();
  if (capacity) free_[0] = capacity;
}

size_t SubAllocator::allocate(size_t size, size_t align) {
  if (size == 0) return SIZE_MAX;
  size_t a = std::max(align, min_align_);
  if (a & (a - 1)) a = min_align_;  // non power of two: fall back to the arena alignment
  const size_t need = round_up(size, min_align_);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    const size_t off = round_up(it->first, a);
    const size_t pad = off - it->first;
    if (it->second < pad + need) continue;
    const size_t blk_off = it->first, blk_size = it->second;
    free_.erase(it);
    if (pad) free_[blk_off] = pad;
    const size_t rem = blk_size - pad - need;
    if (rem) free_[off + need] = rem;
    live_[off] = need;
    used_ += need;
    high_water_ = std::max(high_water_, off + need);
    return off;
  }
  return SIZE_MAX;
}

size_t SubAllocator::deallocate(size_t offset) {
  auto it = live_.find(offset);
  if (it == live_.end()) return 0;
  const size_t size = it->second;
  live_.erase(it);
  used_ -= size;
  auto ins = free_.emplace(offset, size).first;
  auto next = std::next(ins);
  if (next != free_.end() && ins->first + ins->second == next->first) {
    ins->second += next->second;
    free_.erase(next);
  }
  if (ins != free_.begin()) {
    auto prev = std::prev(ins);
    if (prev->first + prev->second == ins->first) {
      prev->second += ins->second;
      free_.erase(ins);
    }
  }
  return size;
}

size_t SubAllocator::largest_free() const {
  size_t m = 0;
  for (const auto& kv : free_) m = std::max(m, kv.second);
  return m;
}

// ------------------------------------------------------------------ UnifiedArena
UnifiedArena::~UnifiedArena() { destroy(); }

bool UnifiedArena::initialize(HipBackend* backend, size_t capacity, bool host_mirror) {
  std::lock_guard<std::mutex> lk(mutex_);
  if (dev_base_) return true;
  backend_ = backend;
  if (!backend_ || !backend_->valid()) return false;
  if (host_mirror) {
    if (const char* env = std::getenv("PARALLAX_ARENA_SIZE")) {
      const unsigned long long mb = std::strtoull(env, nullptr, 10);
      if (mb > 0) capacity = static_cast<size_t>(mb) * 1024 * 1024;
    }
  }
  capacity_ = round_up(capacity, kDefaultAlignment);
  if (capacity_ == 0) return false;
  if (hipSetDevice(backend_->device()) != hipSuccess) return false;
  if (hipMalloc(&dev_base_, capacity_) != hipSuccess) {
    (void)hipGetLastError();
    dev_base_ = nullptr;
    return false;
  }
  if (host_mirror) {
    if (hipHostMalloc(&host_base_, capacity_, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(dev_base_);
      dev_base_ = host_base_ = nullptr;
      return false;
    }
  }
  if (hipEventCreateWithFlags(&ev_, hipEventDisableTiming) != hipSuccess) ev_ = nullptr;
  alloc_.reset(capacity_, kDefaultAlignment);
  return true;
}

void UnifiedArena::destroy() {
  std::lock_guard<std::mutex> lk(mutex_);
  if (!dev_base_) return;
  if (backend_ && backend_->valid()) {
    (void)hipStreamSynchronize(backend_->stream());
    (void)hipStreamSynchronize(backend_->side_stream());
  }
  if (ev_) (void)hipEventDestroy(ev_);
  if (host_base_) (void)hipHostFree(host_base_);
  (void)hipFree(dev_base_);
  ev_ = nullptr;
  host_base_ = dev_base_ = nullptr;
  capacity_ = 0;
  alloc_.reset(0, kDefaultAlignment);
}

void* UnifiedArena::allocate(size_t size, size_t align) {
  std::lock_guard<std::mutex> lk(mutex_);
  if (!dev_base_ || size == 0) return nullptr;
  const size_t off = alloc_.allocate(size, align);
  if (off == SIZE_MAX) return nullptr;
  char* base = static_cast<char*>(host_base_ ? host_base_ : dev_base_);
  return base + off;
}

void UnifiedArena::deallocate(void* ptr) {
  std::lock_guard<std::mutex> lk(mutex_);
  if (!ptr || !dev_base_) return;
  const char* base = static_cast<const char*>(host_base_ ? host_base_ : dev_base_);
  const char* p = static_cast<const char*>(ptr);
  if (p < base || p >= base + capacity_) return;  // not ours: ignored, as the reference
  alloc_.deallocate(static_cast<size_t>(p - base));
}

bool UnifiedArena::contains(const void* ptr) const {
  const char* base = static_cast<const char*>(host_base_ ? host_base_ : dev_base_);
  if (!base || !ptr) return false;
  const char* p = static_cast<const char*>(ptr);
  return p >= base && p < base + capacity_;
}

size_t UnifiedArena::offset_of(const void* ptr) const {
  const char* base = static_cast<const char*>(host_base_ ? host_base_ : dev_base_);
  return static_cast<size_t>(static_cast<const char*>(ptr) - base);
}

void* UnifiedArena::to_device(const void* ptr) const {
  if (!dev_base_) return nullptr;
  const char* p = static_cast<const char*>(ptr);
  const char* d = static_cast<const char*>(dev_base_);
  if (p >= d && p < d + capacity_) return const_cast<char*>(p);
  if (!host_base_) return nullptr;
  const char* h = static_cast<const char*>(host_base_);
  if (p < h || p >= h + capacity_) return nullptr;
  return static_cast<char*>(dev_base_) + (p - h);  // gpu = dev_base + (host - host_base)
}

size_t UnifiedArena::used() const {
  std::lock_guard<std::mutex> lk(mutex_);
  return alloc_.used();
}
size_t UnifiedArena::high_water() const {
  std::lock_guard<std::mutex> lk(mutex_);
  return alloc_.high_water();
}

void UnifiedArena::flush_to_device() {
  const size_t hw = high_water();
  if (host_base_ && hw) flush_range(host_base_, hw);
}

void UnifiedArena::invalidate_from_device() {
  const size_t hw = high_water();
  if (host_base_ && hw) invalidate_range(host_base_, hw);
}

bool UnifiedArena::flush_range(const void* ptr, size_t bytes) {
  if (!host_base_ || bytes == 0) return true;
  if (!contains(ptr)) return false;
  void* dst = to_device(ptr);
  hipStream_t side = backend_->side_stream();
  if (hipMemcpyAsync(dst, ptr, bytes, hipMemcpyHostToDevice, side) != hipSuccess) return false;
  if (hipEventRecord(ev_, side) != hipSuccess) return false;
  return hipStreamWaitEvent(backend_->stream(), ev_, 0) == hipSuccess;
}

bool UnifiedArena::invalidate_range(const void* ptr, size_t bytes) {
  if (!host_base_ || bytes == 0) return true;
  if (!contains(ptr)) return false;
  const void* src = to_device(ptr);
  hipStream_t side = backend_->side_stream();
  if (hipEventRecord(ev_, backend_->stream()) != hipSuccess) return false;
  if (hipStreamWaitEvent(side, ev_, 0) != hipSuccess) return false;
  if (hipMemcpyAsync(const_cast<void*>(ptr), src, bytes, hipMemcpyDeviceToHost, side) != hipSuccess)
    return false;
  return hipStreamSynchronize(side) == hipSuccess;
}

}  // namespace parallax
