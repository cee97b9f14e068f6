// parallax/hip_backend.hpp -- the device layer of the MI355X engine.
//
// Replaces the reference's VulkanBackend (include/parallax/vulkan_backend.hpp,
// src/backend/vulkan/device.cpp:14-345: instance, first compute-capable device, ONE queue
// and a fence).  Here: one HIP device, a compute stream and a migration side stream
// (both default priority), capabilities read from hipDeviceProp_t.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <string>

namespace parallax {

struct DeviceCapabilities {
  // the reference's probe flags (vulkan_backend.hpp); always true on gfx950
  bool shader_int64 = true;
  bool shader_float64 = true;
  bool buffer_device_address = true;
  bool storage_8bit = true;
  bool storage_16bit = true;
  bool external_memory_host = false;
  // gfx950 facts
  std::string name;
  std::string arch;        // gcnArchName, e.g. "gfx950:sramecc+:xnack-"
  int compute_units = 0;
  int wave_size = 64;
  size_t hbm_bytes = 0;
  size_t lds_per_block = 0;
  int l2_bytes = 0;
  int clock_khz = 0;
  int mem_clock_khz = 0;
  int mem_bus_bits = 0;
};

class HipBackend {
 public:
  HipBackend() = default;
  ~HipBackend();
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  // Select `device` (-1: current / 0) and create the streams.  Returns false when no
  // GPU is present or PARALLAX_FORCE_CPU is set (the caller then runs on the CPU).
  bool initialize(int device = -1);
  void destroy();

  bool valid() const { return ok_; }
  int device() const { return device_; }
  hipStream_t stream() const { return stream_; }       // compute
  hipStream_t side_stream() const { return side_; }    // migrations
  const DeviceCapabilities& capabilities() const { return caps_; }
  // Blocks the host until the compute stream is idle (reference KernelLauncher::sync).
  bool sync();

 private:
  bool ok_ = false;
  int device_ = -1;
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;
  DeviceCapabilities caps_;
};

}  // namespace parallax
