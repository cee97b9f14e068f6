// parallax/arena.hpp -- UnifiedArena for the MI355X engine.
//
// Same role and method names as the reference's UnifiedArena (include/parallax/arena.hpp:
// 29-114, src/memory/arena.cpp:21-406): one big region sub-allocated first-fit, 256-byte
// aligned, with host pointers that the funnels memcpy into.  The MI355X version is a
// hipMalloc'd device-local region plus a hipHostMalloc'd (pinned) mirror of the same
// size; an allocation lives at the same offset in both.  A kernel runs on the device
// copy: host_ptr -> dev_base + (host_ptr - host_base).  Operand ranges (not the whole
// used range, as in the reference's flush/invalidate) migrate with hipMemcpyAsync on the
// backend's side stream before and after each call.
//
// Freed blocks coalesce with free neighbours (the reference never coalesces), while the
// first-fit property the reference tests (test_arena.cpp:53-58: a freed block is reused
// at the same offset) still holds.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>

#include "parallax/hip_backend.hpp"

namespace parallax {

// First-fit sub-allocator over [0, capacity) with coalescing free list (no device
// calls; usable on its own).
class SubAllocator {
 public:
  void reset(size_t capacity, size_t min_align);
  // Offset of a block of `size` bytes aligned to max(align, min_align), or SIZE_MAX.
  size_t allocate(size_t size, size_t align);
  // Returns the block's size, or 0 for an unknown offset (ignored, like the reference).
  size_t deallocate(size_t offset);
  bool live(size_t offset) const { return live_.count(offset) != 0; }
  size_t capacity() const { return capacity_; }
  size_t used() const { return used_; }
  size_t high_water() const { return high_water_; }
  size_t largest_free() const;

 private:
  size_t capacity_ = 0;
  size_t min_align_ = 256;
  size_t used_ = 0;
  size_t high_water_ = 0;
  std::map<size_t, size_t> free_;             // offset -> size, sorted, coalesced
  std::unordered_map<size_t, size_t> live_;   // offset -> size
};

class UnifiedArena {
 public:
  static constexpr size_t kDefaultCapacity = 1024ull * 1024 * 1024;  // 1 GiB (env: PARALLAX_ARENA_SIZE MiB)
  static constexpr size_t kDefaultAlignment = 256;

  UnifiedArena() = default;
  ~UnifiedArena();
  UnifiedArena(const UnifiedArena&) = delete;
  UnifiedArena& operator=(const UnifiedArena&) = delete;

  // host_mirror == false gives a device-only pool: allocate() then returns device
  // pointers (used for HBM-resident served buffers).
  bool initialize(HipBackend* backend, size_t capacity = kDefaultCapacity, bool host_mirror = true);
  void destroy();

  void* allocate(size_t size, size_t align = kDefaultAlignment);
  void deallocate(void* ptr);

  bool contains(const void* ptr) const;
  size_t offset_of(const void* ptr) const;  // precondition: contains(ptr)

  // Relocation: the device address of an arena pointer (host mirror or device).
  void* to_device(const void* ptr) const;

  // Whole-used-range migration (the reference's semantics) ...
  void flush_to_device();
  void invalidate_from_device();
  // ... and operand-range migration on the side stream.  flush: host -> device, then
  // the compute stream waits for it.  invalidate: after the compute stream's work,
  // device -> host, and the host waits.
  bool flush_range(const void* ptr, size_t bytes);
  bool invalidate_range(const void* ptr, size_t bytes);

  bool uma() const { return false; }
  bool host_mirrored() const { return host_base_ != nullptr; }
  void* host_base() const { return host_base_; }
  void* device_base() const { return dev_base_; }
  uint64_t device_address() const { return reinterpret_cast<uint64_t>(dev_base_); }
  size_t capacity() const { return capacity_; }
  size_t used() const;
  size_t high_water() const;
  const DeviceCapabilities& capabilities() const { return backend_->capabilities(); }
  bool valid() const { return dev_base_ != nullptr; }
  HipBackend* backend() const { return backend_; }

 private:
  HipBackend* backend_ = nullptr;
  void* dev_base_ = nullptr;
  void* host_base_ = nullptr;
  size_t capacity_ = 0;
  SubAllocator alloc_;
  hipEvent_t ev_ = nullptr;
  mutable std::mutex mutex_;
};

}  // namespace parallax
