// capi.cpp -- the C ABI of parallax/runtime.h and the funnel registry.
//
// Reference: src/kernel/cache.cpp (registry, handles, launch shims) and
// src/memory/unified_allocator.cpp (lazy globals, arena accessors).  Differences:
//   * handles carry a primitive descriptor, not a pipeline name; loading never compiles;
//   * the registry is mutex-protected and every entry point is serialised (the
//     reference is not thread-safe);
//   * nothing is printed on success; failures always print one "[parallax]" line to
//     stderr (PARALLAX_DEBUG adds registration / miss / load lines);
//   * a lookup miss may be answered by the built-in funnel table: keys of the fixed
//     primitives (device_reduce, device_sort, device_scan, device_exclusive_scan,
//     device_unique) name their op and element type in __PRETTY_FUNCTION__, so they
//     resolve without a compiler-emitted registrar (PARALLAX_BUILTIN_FUNNELS=0 turns
//     this off).  Funnels that carry a user functor (copy_if, partition, transform, ...)
//     still need a registered descriptor.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "parallax/descriptor.h"
#include "parallax/px_device.h"
#include "parallax/runtime.h"
#include "parallax/runtime.hpp"

struct parallax_kernel {
  std::string name;
  px_desc desc;
};

namespace {

std::recursive_mutex& abi_mutex() {
  static std::recursive_mutex m;
  return m;
}

bool env_on(const char* n) {
  const char* v = std::getenv(n);
  return v && v[0] && std::strcmp(v, "0") != 0;
}
bool debug_on() { return env_on("PARALLAX_DEBUG"); }
bool force_cpu() { return env_on("PARALLAX_FORCE_CPU"); }
bool builtin_funnels() {
  const char* v = std::getenv("PARALLAX_BUILTIN_FUNNELS");
  return !(v && std::strcmp(v, "0") == 0);
}

void fail(const char* who, const char* what) {
  std::fprintf(stderr, "[parallax] %s: %s\n", who, what);
}

std::unique_ptr<parallax::UnifiedArena>& arena_slot() {
  static std::unique_ptr<parallax::UnifiedArena> a;
  return a;
}
std::unique_ptr<parallax::KernelLauncher>& launcher_slot() {
  static std::unique_ptr<parallax::KernelLauncher> l;
  return l;
}

struct FunnelEntry {
  const unsigned int* words = nullptr;
  size_t nwords = 0;
  std::vector<unsigned int> owned;  // built-in descriptors
  parallax_kernel_t handle = nullptr;
  bool loaded = false;
};
std::unordered_map<std::string, FunnelEntry>& registry() {
  static std::unordered_map<std::string, FunnelEntry> r;
  return r;
}
std::atomic<uint64_t> g_kernel_counter{0};

// registered external host ranges (parallax_register_buffer)
std::vector<std::pair<const char*, size_t>>& registered() {
  static std::vector<std::pair<const char*, size_t>> r;
  return r;
}

// ---- built-in funnel keys
int dtype_from_name(std::string t) {
  while (!t.empty() && t.back() == ' ') t.pop_back();
  while (!t.empty() && t.front() == ' ') t.erase(t.begin());
  static const std::pair<const char*, int> names[] = {
      {"int", PX_I32},
      {"unsigned int", PX_U32},
      {"long int", PX_I64},
      {"long", PX_I64},
      {"long long int", PX_I64},
      {"long long", PX_I64},
      {"long unsigned int", PX_U64},
      {"unsigned long", PX_U64},
      {"long long unsigned int", PX_U64},
      {"unsigned long long", PX_U64},
      {"float", PX_F32},
      {"double", PX_F64},
  };
  for (const auto& n : names)
    if (t == n.first) return n.second;
  return -1;
}

// The template argument T of a funnel key: GCC "[with T = float; ...]", Clang "[T = float]".
int key_dtype(const std::string& key) {
  size_t b = key.rfind('[');
  if (b == std::string::npos) return -1;
  for (size_t p = key.find("T = ", b); p != std::string::npos; p = key.find("T = ", p + 1)) {
    const char prev = key[p - 1];
    if (prev != ' ' && prev != '[') continue;
    size_t e = key.find_first_of(";]", p);
    if (e == std::string::npos) return -1;
    return dtype_from_name(key.substr(p + 4, e - (p + 4)));
  }
  return -1;
}

bool builtin_descriptor(const std::string& key, px_desc& d) {
  std::memset(&d, 0, sizeof(d));
  d.magic = PX_DESC_MAGIC;
  d.version = PX_DESC_VERSION;
  const int dt = key_dtype(key);
  if (dt < 0) return false;
  d.dtype = static_cast<uint32_t>(dt);
  std::string suffix;
  const size_t br = key.rfind(']');
  if (br != std::string::npos && br + 1 < key.size()) suffix = key.substr(br + 1);
  auto is = [&](const char* fn) { return key.find(std::string("parallax::detail::") + fn + "(") != std::string::npos; };
  if (is("device_reduce") && suffix.empty()) {
    d.op = PX_OP_REDUCE;
  } else if (is("device_sort") && suffix.empty()) {
    d.op = PX_OP_SORT;
  } else if (is("device_scan") || is("device_exclusive_scan")) {
    if (suffix == ":scan") d.op = PX_OP_SCAN;
    else if (suffix == ":add") d.op = PX_OP_SCAN_ADD;
    else if (suffix == ":shift" && is("device_exclusive_scan")) d.op = PX_OP_SCAN_SHIFT;
    else return false;
  } else if (is("device_unique")) {
    if (suffix == ":flags") {
      d.op = PX_OP_FLAGS;
      d.fn = PX_PRED_UNIQUE;
    } else if (suffix == ":scan") {
      d.op = PX_OP_SCAN;
    } else if (suffix == ":add") {
      d.op = PX_OP_SCAN_ADD;
    } else if (suffix == ":scatter") {
      d.op = PX_OP_SCATTER;
      d.aux = PX_SCATTER_COMPACT;
    } else {
      return false;
    }
  } else {
    return false;
  }
  return true;
}

parallax::KernelLauncher* launcher_or_null(const char* who) {
  parallax::KernelLauncher* l = parallax::get_global_launcher();
  if (!l) fail(who, "no GPU runtime (the funnel should have taken its CPU loop)");
  return l;
}

parallax_kernel* handle_of(parallax_kernel_t k) { return reinterpret_cast<parallax_kernel*>(k); }

}  // namespace

// ------------------------------------------------------------------ C++ globals
namespace parallax {

HipBackend* get_global_backend() {
  if (force_cpu()) return nullptr;
  return global_backend();
}

UnifiedArena* get_global_arena() {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  auto& a = arena_slot();
  if (a) return a.get();
  HipBackend* be = get_global_backend();
  if (!be) return nullptr;
  auto arena = std::make_unique<UnifiedArena>();
  if (!arena->initialize(be, UnifiedArena::kDefaultCapacity, true)) {
    fail("arena", "initialisation failed");
    return nullptr;
  }
  a = std::move(arena);
  return a.get();
}

KernelLauncher* get_global_launcher() {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  auto& l = launcher_slot();
  if (l) return l.get();
  HipBackend* be = get_global_backend();
  if (!be) return nullptr;
  l = std::make_unique<KernelLauncher>(be, get_global_arena());
  if (!l->context()) {
    l.reset();
    return nullptr;
  }
  return l.get();
}

}  // namespace parallax

// ------------------------------------------------------------------ C ABI
extern "C" {

void* parallax_umalloc(size_t size, unsigned flags) {
  (void)flags;
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = parallax::get_global_arena();
  void* p = a ? a->allocate(size, 256) : nullptr;
  return p ? p : std::malloc(size);
}

void parallax_ufree(void* ptr) {
  if (!ptr) return;
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = arena_slot().get();
  if (a && a->contains(ptr))
    a->deallocate(ptr);
  else
    std::free(ptr);
}

void* parallax_arena_alloc(size_t size, size_t align) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = parallax::get_global_arena();
  void* p = a ? a->allocate(size, align ? align : 16) : nullptr;
  if (debug_on()) std::fprintf(stderr, "[parallax] arena_alloc %zu bytes -> %p\n", size, p);
  return p;
}

void parallax_arena_free(void* ptr) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = arena_slot().get();
  if (a) a->deallocate(ptr);
}

int parallax_arena_contains(const void* ptr) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = parallax::get_global_arena();
  return (a && a->contains(ptr)) ? 1 : 0;
}

int parallax_heap_contains(const void* ptr) {
  (void)ptr;
  return 0;
}

parallax_kernel_t parallax_kernel_load(const unsigned int* spirv, size_t words) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!spirv || words < PX_DESC_WORDS || spirv[0] != PX_DESC_MAGIC) {
    fail("parallax_kernel_load", "payload is not a primitive descriptor");
    return nullptr;
  }
  parallax::KernelLauncher* l = parallax::get_global_launcher();
  if (!l) {
    if (debug_on()) fail("parallax_kernel_load", "no GPU runtime");
    return nullptr;
  }
  const std::string name = "kernel_" + std::to_string(g_kernel_counter.fetch_add(1));
  if (!l->load_kernel(name, spirv, words)) {
    fail("parallax_kernel_load", "invalid descriptor");
    return nullptr;
  }
  auto* h = new parallax_kernel;  // handles live for the process, as in the reference
  h->name = name;
  std::memcpy(&h->desc, spirv, sizeof(px_desc));
  return reinterpret_cast<parallax_kernel_t>(h);
}

void parallax_kernel_register(const char* key, const unsigned int* spirv, size_t words) {
  if (!key) return;
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  FunnelEntry e;
  e.words = spirv;
  e.nwords = words;
  registry()[key] = std::move(e);
  if (debug_on()) std::fprintf(stderr, "[parallax] register %s (%zu words)\n", key, words);
}

parallax_kernel_t parallax_kernel_lookup(const char* key) {
  if (!key || force_cpu()) return nullptr;
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  auto& reg = registry();
  auto it = reg.find(key);
  if (it == reg.end()) {
    px_desc d;
    if (!builtin_funnels() || !builtin_descriptor(key, d)) {
      if (debug_on()) std::fprintf(stderr, "[parallax] lookup MISS: %s\n", key);
      return nullptr;
    }
    FunnelEntry e;
    e.owned.resize(PX_DESC_WORDS);
    std::memcpy(e.owned.data(), &d, sizeof(d));
    e.words = e.owned.data();
    e.nwords = e.owned.size();
    it = reg.emplace(key, std::move(e)).first;
    it->second.words = it->second.owned.data();
    if (debug_on()) std::fprintf(stderr, "[parallax] lookup built-in: %s\n", key);
  }
  if (!it->second.loaded) {
    it->second.handle = parallax_kernel_load(it->second.words, it->second.nwords);
    it->second.loaded = it->second.handle != nullptr;  // retry later if no device yet
  }
  return it->second.handle;
}

void parallax_kernel_launch(parallax_kernel_t kernel, ...) {
  va_list args;
  va_start(args, kernel);
  void* buffer = va_arg(args, void*);
  size_t count = va_arg(args, size_t);
  size_t elem_size = va_arg(args, size_t);
  va_end(args);
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_kernel_launch", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_kernel_launch");
  if (l) l->launch(handle_of(kernel)->name, buffer, count, elem_size);
}

void parallax_kernel_launch_transform(parallax_kernel_t kernel, ...) {
  va_list args;
  va_start(args, kernel);
  void* in = va_arg(args, void*);
  void* out = va_arg(args, void*);
  size_t count = va_arg(args, size_t);
  size_t elem_size = va_arg(args, size_t);
  va_end(args);
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_kernel_launch_transform", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_kernel_launch_transform");
  if (l) l->launch_transform(handle_of(kernel)->name, in, out, count, elem_size, elem_size);
}

void parallax_kernel_launch_transform2(parallax_kernel_t kernel, void* in_buffer, void* out_buffer,
                                       size_t count, size_t in_elem_size, size_t out_elem_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_kernel_launch_transform2", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_kernel_launch_transform2");
  if (l)
    l->launch_transform(handle_of(kernel)->name, in_buffer, out_buffer, count, in_elem_size,
                        out_elem_size);
}

void parallax_kernel_launch_transform2_captures(parallax_kernel_t kernel, void* in_buffer,
                                                void* out_buffer, size_t count,
                                                size_t in_elem_size, size_t out_elem_size,
                                                void* captures, size_t capture_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_kernel_launch_transform2_captures", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_kernel_launch_transform2_captures");
  if (l)
    l->launch_transform(handle_of(kernel)->name, in_buffer, out_buffer, count, in_elem_size,
                        out_elem_size, captures, capture_size);
}

void parallax_kernel_launch_with_captures(parallax_kernel_t kernel, void* buffer, size_t count,
                                          void* captures, size_t capture_size, size_t elem_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_kernel_launch_with_captures", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_kernel_launch_with_captures");
  if (l) l->launch_with_captures(handle_of(kernel)->name, buffer, count, captures, capture_size, elem_size);
}

void parallax_reduce(parallax_kernel_t kernel, void* data, size_t count, size_t elem_size,
                     void* result) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_reduce", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_reduce");
  if (l) l->launch_reduce(handle_of(kernel)->name, data, count, elem_size, result);
}

void parallax_transform_reduce(parallax_kernel_t map_kernel, parallax_kernel_t reduce_kernel,
                               void* input, size_t count, size_t in_elem_size,
                               size_t out_elem_size, void* result) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!map_kernel || !reduce_kernel) return fail("parallax_transform_reduce", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_transform_reduce");
  if (l)
    l->launch_transform_reduce(handle_of(map_kernel)->name, handle_of(reduce_kernel)->name, input,
                               count, in_elem_size, out_elem_size, result);
}

size_t parallax_argminmax(parallax_kernel_t kernel, void* data, size_t count, size_t elem_size,
                          int is_float, int want_max, int want_last) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return count;
  parallax::KernelLauncher* l = launcher_or_null("parallax_argminmax");
  if (!l) return count;
  return l->launch_argminmax(handle_of(kernel)->name, data, count, elem_size, is_float != 0,
                             want_max != 0, want_last != 0);
}

size_t parallax_find(parallax_kernel_t kernel, void* data, size_t count, size_t elem_size,
                     int negate, const void* value) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return count;
  parallax::KernelLauncher* l = launcher_or_null("parallax_find");
  if (!l) return count;
  return l->launch_find(handle_of(kernel)->name, data, count, elem_size, negate != 0, value);
}

size_t parallax_mismatch(parallax_kernel_t kernel, void* a, void* b, size_t count, size_t elem_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return count;
  parallax::KernelLauncher* l = launcher_or_null("parallax_mismatch");
  if (!l) return count;
  return l->launch_mismatch(handle_of(kernel)->name, a, b, count, elem_size);
}

void parallax_scan(parallax_kernel_t scan_kernel, parallax_kernel_t add_kernel, void* data,
                   size_t count, size_t elem_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!scan_kernel || !add_kernel) return fail("parallax_scan", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_scan");
  if (l) l->launch_scan(handle_of(scan_kernel)->name, handle_of(add_kernel)->name, data, count, elem_size);
}

void parallax_exclusive_scan(parallax_kernel_t scan_kernel, parallax_kernel_t add_kernel,
                             parallax_kernel_t shift_kernel, void* input, void* output,
                             size_t count, size_t elem_size, const void* init) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!scan_kernel || !add_kernel || !shift_kernel) return fail("parallax_exclusive_scan", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_exclusive_scan");
  if (l)
    l->launch_exclusive_scan(handle_of(scan_kernel)->name, handle_of(add_kernel)->name,
                             handle_of(shift_kernel)->name, input, output, count, elem_size, init);
}

void parallax_sort(parallax_kernel_t kernel, void* data, size_t count, size_t elem_size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!kernel) return fail("parallax_sort", "null kernel");
  parallax::KernelLauncher* l = launcher_or_null("parallax_sort");
  if (l) l->launch_sort(handle_of(kernel)->name, data, count, elem_size);
}

size_t parallax_copy_if(parallax_kernel_t flags_kernel, parallax_kernel_t scan_kernel,
                        parallax_kernel_t add_kernel, parallax_kernel_t scatter_kernel, void* input,
                        void* output, size_t count, size_t elem_size, int elem_is_float) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!flags_kernel || !scan_kernel || !add_kernel || !scatter_kernel) {
    fail("parallax_copy_if", "null kernel");
    return 0;
  }
  parallax::KernelLauncher* l = launcher_or_null("parallax_copy_if");
  if (!l) return 0;
  size_t kept = 0;
  if (!l->launch_compact(handle_of(flags_kernel)->name, handle_of(scan_kernel)->name,
                         handle_of(add_kernel)->name, handle_of(scatter_kernel)->name, input,
                         output, count, elem_size, elem_is_float != 0, &kept))
    return 0;
  return kept;
}

}  // extern "C"

bool parallax_register_buffer(void* ptr, size_t size) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  if (!parallax::get_global_launcher()) {
    fail("parallax_register_buffer", "no GPU runtime");
    return false;
  }
  if (!ptr) return false;
  registered().emplace_back(static_cast<const char*>(ptr), size);
  return true;
}

// ---- runtime extras declared in parallax/px_device.h
extern "C" {

// Relocation probe of the reference's test_physptr.cpp on an arena pointer.
int pxr_relocate_probe(void* arena_ptr, unsigned count) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::KernelLauncher* l = parallax::get_global_launcher();
  if (!l) return PX_ENODEV;
  return l->launch_relocate_probe(arena_ptr, count) ? PX_EOK : PX_EDEVICE;
}

// Arena facts: host base, device base, capacity, used (any pointer may be NULL).
int pxr_arena_info(void** host_base, void** dev_base, size_t* capacity, size_t* used) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::UnifiedArena* a = parallax::get_global_arena();
  if (!a) return PX_ENODEV;
  if (host_base) *host_base = a->host_base();
  if (dev_base) *dev_base = a->device_base();
  if (capacity) *capacity = a->capacity();
  if (used) *used = a->used();
  return PX_EOK;
}

int pxr_prof_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::KernelLauncher* l = parallax::get_global_launcher();
  if (!l) return PX_ENODEV;
  l->set_profiling(on != 0);
  return PX_EOK;
}

int pxr_prof_get(const char* key, double* total_ms, uint64_t* calls) {
  std::lock_guard<std::recursive_mutex> lk(abi_mutex());
  parallax::KernelLauncher* l = launcher_slot().get();
  if (!l || !key) return PX_EARG;
  return l->profile(key, total_ms, calls) ? PX_EOK : PX_EARG;
}

// The descriptor a funnel key resolves to without registration (0 = resolved).
int pxr_builtin_descriptor(const char* key, unsigned* words12) {
  if (!key || !words12) return PX_EARG;
  px_desc d;
  if (!builtin_descriptor(key, d)) return PX_EUNSUPPORTED;
  std::memcpy(words12, &d, sizeof(d));
  return PX_EOK;
}

}  // extern "C"
