// parallax/kernel_launcher.hpp -- KernelLauncher on the MI355X engine.
//
// Same entry points, argument meaning and boolean result as the reference's
// KernelLauncher (include/parallax/kernel_launcher.hpp:19-115); the Vk* types are gone.
// A kernel "name" is the string a descriptor was loaded under (parallax_kernel_load
// assigns "kernel_<n>").  Every launch is synchronous: outputs are visible to the host
// (arena / host pointers) or to the device (device pointers) when it returns.
//
// Operands are resolved per pointer:
//   arena pointer   -> relocated to the device copy; inputs flushed host->device and
//                      outputs invalidated device->host on the side stream;
//   device pointer  -> used in place (no copies);
//   other host ptr  -> staged through a device buffer, results copied back.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "parallax/arena.hpp"
#include "parallax/descriptor.h"
#include "parallax/hip_backend.hpp"

namespace parallax {

class HostStager;

class StreamContext;

class KernelLauncher {
 public:
  KernelLauncher(HipBackend* backend, UnifiedArena* arena);
  ~KernelLauncher();
  KernelLauncher(const KernelLauncher&) = delete;
  KernelLauncher& operator=(const KernelLauncher&) = delete;

  // Registers a primitive descriptor (parallax/descriptor.h) under `name`.
  bool load_kernel(const std::string& name, const uint32_t* words, size_t nwords);
  const px_desc* descriptor(const std::string& name) const;

  // vector_multiply shape: data[i] *= multiplier (a PX_MAP_MUL descriptor).
  bool launch(const std::string& kernel_name, void* buffer, size_t count, float multiplier,
              size_t elem_size = sizeof(float));
  bool launch(const std::string& kernel_name, void* buffer, size_t count,
              size_t elem_size = sizeof(float));
  bool launch_transform(const std::string& kernel_name, void* in_buffer, void* out_buffer,
                        size_t count, size_t elem_size = sizeof(float), size_t out_elem_size = 0,
                        void* captures = nullptr, size_t capture_size = 0);
  bool launch_with_captures(const std::string& kernel_name, void* buffer, size_t count,
                            void* captures, size_t capture_size, size_t elem_size = sizeof(float));
  bool launch_reduce(const std::string& kernel_name, void* data, size_t count, size_t elem_size,
                     void* out_result);
  // Σ map(x) in one pass (transform_reduce / count_if funnels): `map_kernel` is a
  // TRANSFORM descriptor (in dtype, out dtype in aux), `reduce_kernel` a REDUCE of the
  // out dtype.  Maps the fused kernel does not cover (gather) run as transform + reduce
  // on the device.  No counterpart in the reference, which always runs the two passes.
  bool launch_transform_reduce(const std::string& map_kernel, const std::string& reduce_kernel,
                               void* in, size_t count, size_t in_elem_size, size_t out_elem_size,
                               void* out_result);
  size_t launch_argminmax(const std::string& kernel_name, void* data, size_t count,
                          size_t elem_size, bool is_float, bool want_max, bool want_last);
  size_t launch_find(const std::string& kernel_name, void* data, size_t count, size_t elem_size,
                     bool negate, const void* value = nullptr);
  size_t launch_mismatch(const std::string& kernel_name, void* a, void* b, size_t count,
                         size_t elem_size);
  bool launch_scan(const std::string& scan_kernel, const std::string& add_kernel, void* data,
                   size_t count, size_t elem_size);
  bool launch_exclusive_scan(const std::string& scan_kernel, const std::string& add_kernel,
                             const std::string& shift_kernel, void* input, void* output,
                             size_t count, size_t elem_size, const void* init);
  bool launch_sort(const std::string& kernel_name, void* data, size_t count, size_t elem_size);
  bool launch_compact(const std::string& flags_kernel, const std::string& scan_kernel,
                      const std::string& add_kernel, const std::string& scatter_kernel,
                      void* input, void* output, size_t count, size_t elem_size,
                      bool elem_is_float, size_t* out_kept);
  // The reference's relocation probe (test_physptr.cpp:95-106): `arena_ptr` is relocated
  // on the device (gpu = dev_base + (host_ptr - host_base)) and r[i] = r[i]*2+1.
  bool launch_relocate_probe(void* arena_ptr, uint32_t count);

  void sync();
  StreamContext* context() const { return ctx_.get(); }

  // Event timing of the device work of each call, on the launcher's stream, keyed
  // "<op>:<dtype>" (e.g. "scan:u32"); the kernels of a sort are also timed one launch
  // at a time ("sort_tile:u32", "sort_global:u32", "sort_merge:u32").  Off by default.
  void set_profiling(bool on);
  bool profile(const std::string& key, double* total_ms, uint64_t* calls) const;

 private:
  struct Operand;
  // slot: which staging buffer a host operand uses (0 first operand, 1 second)
  bool resolve(void* p, size_t bytes, bool read, Operand& op, int slot = 0);
  bool finish(Operand& op, bool write);
  bool copy_out(void* dst, const void* dev_src, size_t bytes);
  void release(Operand& op);
  const px_desc* find(const std::string& name, const char* who) const;
  void prof_begin();
  void prof_end(const char* op, uint32_t dtype);
  static void prof_hook(void* self, const char* kernel, int phase);

  HipBackend* backend_;
  UnifiedArena* arena_;
  std::unique_ptr<StreamContext> ctx_;
  uint64_t* dev_word_ = nullptr;  // device scratch for counts / indices
  void* dev_value_ = nullptr;     // device scratch for a reduce result
  std::unordered_map<std::string, px_desc> kernels_;
  struct CodeKernel {
    hipModule_t module = nullptr;
    hipFunction_t function = nullptr;
  };
  std::unordered_map<std::string, CodeKernel> code_;  // PX_OP_CODE payloads, by kernel name
  void* dev_caps_ = nullptr;  // device copy of the captures of a code-object launch
  size_t dev_caps_cap_ = 0;
  bool load_code(const std::string& name, const px_desc& d, const uint32_t* words, size_t nwords);
  bool run_code(const std::string& name, const void* in, void* out, uint64_t count,
                const void* captures, size_t capture_size, bool transform);
  std::unique_ptr<HostStager> stager_;  // pageable host operands: pinned chunks, kept buffers
  bool prof_on_ = false;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::unordered_map<std::string, std::pair<double, uint64_t>> prof_;
  struct KernelSpan {
    const char* kernel;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> ev_pool_;   // per-launch events, reused every call
  size_t ev_used_ = 0;
  std::vector<KernelSpan> spans_;     // launches of the current call, resolved at prof_end
};

}  // namespace parallax
