// launcher.cpp -- KernelLauncher on HIP (reference: src/kernel/kernel_launcher.cpp).
//
// Each reference launch_* expanded into a schedule of synchronous dispatches with a
// fence wait after each; here each is one or two stream-ordered launches of the gfx950
// kernels (engine.cpp) followed by a single wait, with operand migration limited to the
// operand ranges.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "engine.hpp"
#include "parallax/kernel_launcher.hpp"
#include "staging.hpp"

namespace parallax {

struct KernelLauncher::Operand {
  enum Kind { NONE, ARENA, DEVICE, STAGED };
  Kind kind = NONE;
  void* user = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
};

namespace {

bool debug_on() {
  static const bool on = [] {
    const char* v = std::getenv("PARALLAX_DEBUG");
    return v && v[0] && std::strcmp(v, "0") != 0;
  }();
  return on;
}

void report(const char* who, const char* what) {
  std::fprintf(stderr, "[parallax] %s: %s\n", who, what);
  set_last_error(PX_ERR_ARG, std::string(who) + ": " + what);
}

bool size_ok(uint32_t dtype, size_t elem_size) {
  return dtype < PX_NUM_DTYPES && pxk::dtype_size(static_cast<int>(dtype)) == elem_size;
}

bool is_device_pointer(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

KernelLauncher::KernelLauncher(HipBackend* backend, UnifiedArena* arena)
    : backend_(backend), arena_(arena) {
  if (!backend_ || !backend_->valid()) return;
  ctx_ = std::make_unique<StreamContext>(backend_, backend_->stream());
  stager_ = std::make_unique<HostStager>();
  if (hipMalloc(reinterpret_cast<void**>(&dev_word_), 64) != hipSuccess) dev_word_ = nullptr;
  if (hipMalloc(&dev_value_, 64) != hipSuccess) dev_value_ = nullptr;
}

KernelLauncher::~KernelLauncher() {
  if (ctx_) (void)hipStreamSynchronize(ctx_->stream());
  stager_.reset();
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  if (dev_word_) (void)hipFree(dev_word_);
  if (dev_value_) (void)hipFree(dev_value_);
  if (dev_caps_) (void)hipFree(dev_caps_);
  for (auto& kv : code_)
    if (kv.second.module) (void)hipModuleUnload(kv.second.module);
}

namespace {
uint64_t arena_host_base(UnifiedArena* a) {
  return (a && a->host_mirrored()) ? reinterpret_cast<uint64_t>(a->host_base()) : 0;
}
uint64_t arena_dev_base(UnifiedArena* a) { return a ? a->device_address() : 0; }
uint64_t arena_bytes(UnifiedArena* a) { return (a && a->host_mirrored()) ? a->capacity() : 0; }
}  // namespace

bool KernelLauncher::load_kernel(const std::string& name, const uint32_t* words, size_t nwords) {
  if (!words || nwords < PX_DESC_WORDS) return false;
  px_desc d;
  std::memcpy(&d, words, sizeof(d));
  if (d.magic != PX_DESC_MAGIC || d.version != PX_DESC_VERSION) return false;
  if (d.dtype >= PX_NUM_DTYPES) return false;
  if (d.op == PX_OP_CODE) return load_code(name, d, words, nwords);
  if (d.op == PX_OP_NONE || d.op > PX_OP_MISMATCH) return false;
  if (d.op == PX_OP_TRANSFORM && d.aux >= PX_NUM_DTYPES) return false;
  if ((d.op == PX_OP_MAP || d.op == PX_OP_TRANSFORM) && d.fn >= PX_NUM_MAPOPS) return false;
  if ((d.op == PX_OP_FLAGS || d.op == PX_OP_FIND) && d.fn >= PX_NUM_PREDS) return false;
  if (d.op == PX_OP_SCATTER && d.aux > PX_SCATTER_STABLE_PARTITION) return false;
  kernels_[name] = d;
  if (debug_on()) std::fprintf(stderr, "[parallax] load %s op=%u dtype=%u\n", name.c_str(), d.op, d.dtype);
  return true;
}

// A code-object payload: header, symbol name, HSACO (descriptor.h).  The module is
// loaded once and kept for the life of the process, like the handles.
bool KernelLauncher::load_code(const std::string& name, const px_desc& d, const uint32_t* words,
                               size_t nwords) {
  if (d.aux >= PX_NUM_DTYPES || d.fn > PX_CODE_TRANSFORM) return false;
  const size_t name_len = d.fn2;
  const uint64_t code_size = d.c0;
  if (name_len == 0 || name_len > 1024 || code_size == 0) return false;
  const size_t name_words = (name_len + 3) / 4;
  const size_t code_words = static_cast<size_t>((code_size + 3) / 4);
  if (nwords < PX_DESC_WORDS + name_words + code_words) return false;
  const char* nm = reinterpret_cast<const char*>(words + PX_DESC_WORDS);
  const std::string symbol(nm, name_len);
  // copy the image to word-aligned storage of its own (the payload may be unaligned
  // for the loader's purposes only in theory; a copy also outlives the caller's words)
  std::vector<uint32_t> image(words + PX_DESC_WORDS + name_words,
                              words + PX_DESC_WORDS + name_words + code_words);
  CodeKernel k;
  if (hipModuleLoadData(&k.module, image.data()) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (hipModuleGetFunction(&k.function, k.module, symbol.c_str()) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipModuleUnload(k.module);
    return false;
  }
  code_[name] = k;
  kernels_[name] = d;
  if (debug_on())
    std::fprintf(stderr, "[parallax] load %s code object %s (%llu bytes)\n", name.c_str(),
                 symbol.c_str(), static_cast<unsigned long long>(code_size));
  return true;
}

bool KernelLauncher::run_code(const std::string& name, const void* in, void* out, uint64_t count,
                              const void* captures, size_t capture_size, bool transform) {
  auto it = code_.find(name);
  if (it == code_.end()) return false;
  if (count == 0) return true;
  void* caps = nullptr;
  if (captures && capture_size) {
    if (capture_size > dev_caps_cap_) {
      if (dev_caps_) (void)hipFree(dev_caps_);
      dev_caps_ = nullptr;
      dev_caps_cap_ = 0;
      if (hipMalloc(&dev_caps_, capture_size) != hipSuccess) {
        (void)hipGetLastError();
        dev_caps_ = nullptr;
        return false;
      }
      dev_caps_cap_ = capture_size;
    }
    if (hipMemcpyAsync(dev_caps_, captures, capture_size, hipMemcpyHostToDevice, ctx_->stream()) !=
        hipSuccess)
      return false;
    caps = dev_caps_;
  }
  // Captured pointers may reach any part of the arena: with captures, the whole used
  // arena goes to the device before the launch and comes back after it (the reference
  // synchronises the used arena around every op).
  const bool sync_arena = caps && arena_ && arena_->host_mirrored();
  if (sync_arena) arena_->flush_to_device();
  uint64_t n = count, cs = caps ? capture_size : 0;
  uint64_t hb = arena_host_base(arena_), db = arena_dev_base(arena_);
  const void* a_in = in;
  void* a_out = out;
  void* a_data = out;
  void* args_map[] = {&a_data, &n, &caps, &cs, &hb, &db};
  void* args_tr[] = {&a_in, &a_out, &n, &caps, &cs, &hb, &db};
  const uint64_t blocks = (count + 255) / 256;
  if (blocks > 0x7fffffffull) return false;
  const hipError_t e =
      hipModuleLaunchKernel(it->second.function, static_cast<unsigned>(blocks), 1, 1, 256, 1, 1, 0,
                            ctx_->stream(), transform ? args_tr : args_map, nullptr);
  if (e != hipSuccess) {
    set_last_error(PX_ERR_DEVICE, std::string("code-object launch: ") + hipGetErrorString(e));
    return false;
  }
  if (sync_arena) {
    if (hipStreamSynchronize(ctx_->stream()) != hipSuccess) return false;
    arena_->invalidate_from_device();
  }
  ++stats().transform;
  ++stats().launches;
  return true;
}

const px_desc* KernelLauncher::descriptor(const std::string& name) const {
  auto it = kernels_.find(name);
  return it == kernels_.end() ? nullptr : &it->second;
}

const px_desc* KernelLauncher::find(const std::string& name, const char* who) const {
  const px_desc* d = descriptor(name);
  if (!d) report(who, ("kernel not found: " + name).c_str());
  return d;
}

bool KernelLauncher::resolve(void* p, size_t bytes, bool read, Operand& op, int slot) {
  op = Operand{};
  op.user = p;
  op.bytes = bytes;
  if (!ctx_ || !ctx_->ok()) return false;
  if (bytes == 0) {
    op.kind = Operand::DEVICE;
    op.dev = p;
    return true;
  }
  if (!p) return false;
  if (arena_ && arena_->contains(p)) {
    op.kind = Operand::ARENA;
    op.dev = arena_->to_device(p);
    if (arena_->host_mirrored() && !arena_->contains(static_cast<char*>(p) + bytes - 1)) return false;
    return !read || arena_->flush_range(p, bytes);
  }
  if (is_device_pointer(p)) {
    op.kind = Operand::DEVICE;
    op.dev = p;
    return true;
  }
  // pageable host memory: a device staging buffer kept between calls, filled through
  // the pinned-chunk pipeline on the launcher's stream (the kernel is queued after it)
  op.kind = Operand::STAGED;
  op.dev = stager_->device_buffer(slot, bytes);
  if (!op.dev) {
    set_last_error(PX_ERR_NOMEM, "staging buffer for a host operand");
    return false;
  }
  return !read || stager_->h2d(op.dev, p, bytes, ctx_->stream());
}

bool KernelLauncher::finish(Operand& op, bool write) {
  if (wait_stream(ctx_->stream()) != hipSuccess) return false;
  if (!write || op.bytes == 0) return true;
  if (op.kind == Operand::ARENA) return arena_->invalidate_range(op.user, op.bytes);
  if (op.kind == Operand::STAGED) return stager_->d2h(op.user, op.dev, op.bytes, ctx_->stream());
  return true;
}

// A result word (or staged operand) to the caller's pointer once the launcher's stream has
// drained.  To host memory hipMemcpy returns when the copy is done; to device memory it
// may return before the copy runs, and a copy on the null stream is not ordered before
// the next call's kernels (which reuse the same device word), so that copy goes on the
// launcher's stream and is waited for.
bool KernelLauncher::copy_out(void* dst, const void* dev_src, size_t bytes) {
  if (!is_device_pointer(dst)) return hipMemcpy(dst, dev_src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  if (hipMemcpyAsync(dst, dev_src, bytes, hipMemcpyDeviceToDevice, ctx_->stream()) != hipSuccess) return false;
  return wait_stream(ctx_->stream()) == hipSuccess;
}

void KernelLauncher::release(Operand& op) {
  op = Operand{};  // staging buffers stay with the launcher for the next call
}

void KernelLauncher::sync() {
  if (ctx_) (void)wait_stream(ctx_->stream());
}

void KernelLauncher::set_profiling(bool on) {
  prof_.clear();
  prof_on_ = on;
  if (on && !ev0_) {
    (void)hipEventCreate(&ev0_);
    (void)hipEventCreate(&ev1_);
  }
}

bool KernelLauncher::profile(const std::string& key, double* total_ms, uint64_t* calls) const {
  auto it = prof_.find(key);
  if (it == prof_.end()) return false;
  if (total_ms) *total_ms = it->second.first;
  if (calls) *calls = it->second.second;
  return true;
}

void KernelLauncher::prof_begin() {
  spans_.clear();
  ev_used_ = 0;
  if (prof_on_ && ev0_) (void)hipEventRecord(ev0_, ctx_->stream());
}

// Records an event before (phase 0) and after (phase 1) one kernel launch of the call.
// Back-to-back launches share one event (the end of one is the start of the next), so
// the stream carries one marker per launch.
void KernelLauncher::prof_hook(void* self, const char* kernel, int phase) {
  auto* L = static_cast<KernelLauncher*>(self);
  if (phase == 0 && !L->spans_.empty() && L->spans_.back().b) {
    L->spans_.push_back({kernel, L->spans_.back().b, nullptr});
    return;
  }
  if (L->ev_used_ == L->ev_pool_.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    L->ev_pool_.push_back(e);
  }
  hipEvent_t e = L->ev_pool_[L->ev_used_++];
  (void)hipEventRecord(e, L->ctx_->stream());
  if (phase == 0)
    L->spans_.push_back({kernel, e, nullptr});
  else if (!L->spans_.empty())
    L->spans_.back().b = e;
}

void KernelLauncher::prof_end(const char* op, uint32_t dtype) {
  if (!prof_on_ || !ev1_) return;
  static const char* names[] = {"i32", "u32", "i64", "u64", "f32", "f64"};
  (void)hipEventRecord(ev1_, ctx_->stream());
  if (hipEventSynchronize(ev1_) != hipSuccess) return;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev0_, ev1_) != hipSuccess) return;
  const std::string dt = dtype < 6 ? names[dtype] : "?";
  auto& e = prof_[std::string(op) + ":" + dt];
  e.first += ms;
  e.second += 1;
  for (const KernelSpan& k : spans_) {
    float kms = 0.f;
    if (!k.b || hipEventElapsedTime(&kms, k.a, k.b) != hipSuccess) continue;
    auto& ke = prof_[std::string(k.kernel) + ":" + dt];
    ke.first += kms;
    ke.second += 1;
  }
  spans_.clear();
  ev_used_ = 0;
}


bool KernelLauncher::launch(const std::string& kernel_name, void* buffer, size_t count,
                            float multiplier, size_t elem_size) {
  // vector_multiply: data[i] *= multiplier, the multiplier as the capture
  const px_desc* d = find(kernel_name, "launch");
  if (!d) return false;
  if (d->dtype != PX_F32) return report("launch", "the multiplier form needs float elements"), false;
  px_desc m = *d;
  m.op = PX_OP_MAP;
  m.fn = PX_MAP_MUL;
  const std::string tmp = kernel_name + "#mul";
  kernels_[tmp] = m;
  const bool ok = launch_with_captures(tmp, buffer, count, &multiplier, sizeof(float), elem_size);
  kernels_.erase(tmp);
  return ok;
}

bool KernelLauncher::launch(const std::string& kernel_name, void* buffer, size_t count,
                            size_t elem_size) {
  return launch_with_captures(kernel_name, buffer, count, nullptr, 0, elem_size);
}

bool KernelLauncher::launch_with_captures(const std::string& kernel_name, void* buffer,
                                          size_t count, void* captures, size_t capture_size,
                                          size_t elem_size) {
  const px_desc* d = find(kernel_name, "launch");
  if (!d) return false;
  if (d->op == PX_OP_CODE) {
    if (d->fn != PX_CODE_MAP) return report("launch", "code object is not an in-place map"), false;
    if (!size_ok(d->dtype, elem_size)) return report("launch", "elem_size does not match the descriptor"), false;
    Operand op;
    bool ok = resolve(buffer, count * elem_size, true, op);
    if (ok) ok = run_code(kernel_name, op.dev, op.dev, count, captures, capture_size, false);
    ok = finish(op, ok) && ok;
    release(op);
    if (!ok) report("launch", last_error_text());
    return ok;
  }
  if (d->op != PX_OP_MAP) return report("launch", "descriptor is not an in-place map"), false;
  if (!size_ok(d->dtype, elem_size))
    return report("launch", "elem_size does not match the descriptor"), false;
  px_desc m = *d;
  m.aux = m.dtype;
  const size_t bytes = count * elem_size;
  // a gathered table captured by pointer lives in the arena mirror: migrate it
  if (m.fn == PX_MAP_GATHER && captures && capture_size >= 16 && arena_) {
    uint64_t tp = 0, sz = 0;
    std::memcpy(&tp, captures, 8);
    std::memcpy(&sz, static_cast<char*>(captures) + 8, 8);
    void* t = reinterpret_cast<void*>(tp);
    if (arena_->contains(t)) arena_->flush_range(t, sz * elem_size);
  }
  Operand op;
  bool ok = resolve(buffer, bytes, true, op);
  if (ok)
    ok = run_transform(*ctx_, m, op.dev, op.dev, count, captures, capture_size,
                       arena_host_base(arena_), arena_dev_base(arena_), arena_bytes(arena_)) == PX_OK;
  ok = finish(op, ok) && ok;
  release(op);
  if (!ok) report("launch", last_error_text());
  return ok;
}

bool KernelLauncher::launch_transform(const std::string& kernel_name, void* in_buffer,
                                      void* out_buffer, size_t count, size_t elem_size,
                                      size_t out_elem_size, void* captures, size_t capture_size) {
  const px_desc* d = find(kernel_name, "launch_transform");
  if (!d) return false;
  if (d->op == PX_OP_CODE) {
    if (d->fn != PX_CODE_TRANSFORM)
      return report("launch_transform", "code object is not a transform"), false;
    if (out_elem_size == 0) out_elem_size = elem_size;
    if (!size_ok(d->dtype, elem_size) || !size_ok(d->aux, out_elem_size))
      return report("launch_transform", "element sizes do not match the descriptor"), false;
    Operand in, out;
    bool ok = resolve(in_buffer, count * elem_size, true, in) &&
              resolve(out_buffer, count * out_elem_size, false, out, 1);
    if (ok) ok = run_code(kernel_name, in.dev, out.dev, count, captures, capture_size, true);
    ok = finish(out, ok) && ok;
    release(in);
    release(out);
    if (!ok) report("launch_transform", last_error_text());
    return ok;
  }
  px_desc m = *d;
  if (m.op == PX_OP_MAP) m.aux = m.dtype;
  if (m.op != PX_OP_TRANSFORM && m.op != PX_OP_MAP)
    return report("launch_transform", "descriptor is not a transform"), false;
  if (out_elem_size == 0) out_elem_size = elem_size;
  if (!size_ok(m.dtype, elem_size) || !size_ok(m.aux, out_elem_size))
    return report("launch_transform", "element sizes do not match the descriptor"), false;
  if (m.fn == PX_MAP_GATHER && captures && capture_size >= 16 && arena_) {
    uint64_t tp = 0, sz = 0;
    std::memcpy(&tp, captures, 8);
    std::memcpy(&sz, static_cast<char*>(captures) + 8, 8);
    void* t = reinterpret_cast<void*>(tp);
    if (arena_->contains(t)) arena_->flush_range(t, sz * out_elem_size);
  }
  Operand in, out;
  bool ok = resolve(in_buffer, count * elem_size, true, in) &&
            resolve(out_buffer, count * out_elem_size, false, out, 1);
  if (ok)
    ok = run_transform(*ctx_, m, in.dev, out.dev, count, captures, capture_size,
                       arena_host_base(arena_), arena_dev_base(arena_), arena_bytes(arena_)) == PX_OK;
  ok = finish(out, ok) && ok;
  release(in);
  release(out);
  if (!ok) report("launch_transform", last_error_text());
  return ok;
}

bool KernelLauncher::launch_reduce(const std::string& kernel_name, void* data, size_t count,
                                   size_t elem_size, void* out_result) {
  const px_desc* d = find(kernel_name, "launch_reduce");
  if (!d) return false;
  if (d->op != PX_OP_REDUCE) return report("launch_reduce", "descriptor is not a reduce"), false;
  if (!size_ok(d->dtype, elem_size) || !out_result)
    return report("launch_reduce", "bad element size or result pointer"), false;
  Operand op;
  bool ok = resolve(data, count * elem_size, true, op);
  if (ok) {
    prof_begin();
    ok = run_reduce(*ctx_, d->dtype, op.dev, count, dev_value_) == PX_OK;
    prof_end("reduce", d->dtype);
  }
  ok = finish(op, false) && ok;
  release(op);
  if (ok) ok = copy_out(out_result, dev_value_, elem_size);
  if (!ok) report("launch_reduce", last_error_text());
  return ok;
}

bool KernelLauncher::launch_transform_reduce(const std::string& map_kernel,
                                             const std::string& reduce_kernel, void* in,
                                             size_t count, size_t in_elem_size,
                                             size_t out_elem_size, void* out_result) {
  const char* who = "launch_transform_reduce";
  const px_desc* m = find(map_kernel, who);
  const px_desc* r = find(reduce_kernel, who);
  if (!m || !r) return false;
  if ((m->op != PX_OP_TRANSFORM && m->op != PX_OP_MAP) || r->op != PX_OP_REDUCE)
    return report(who, "need a transform and a reduce descriptor"), false;
  px_desc md = *m;
  if (md.op == PX_OP_MAP) md.aux = md.dtype;  // an in-place map keeps its type
  if (r->dtype != md.aux) return report(who, "reduce type differs from the map's output type"), false;
  if (!size_ok(md.dtype, in_elem_size) || !size_ok(md.aux, out_elem_size) || !out_result)
    return report(who, "bad element size or result pointer"), false;
  Operand op;
  bool ok = resolve(in, count * in_elem_size, true, op);
  if (ok) {
    prof_begin();
    if (pxk::map_reduce_supported(md)) {
      ok = run_map_reduce(*ctx_, md, op.dev, count, dev_value_) == PX_OK;
    } else {
      // two passes on the device through a temporary array of the output type
      void* tmp = nullptr;
      ok = hipMalloc(&tmp, std::max<size_t>(count * out_elem_size, 16)) == hipSuccess;
      if (ok) {
        ok = run_transform(*ctx_, md, op.dev, tmp, count, nullptr, 0, arena_host_base(arena_),
                           arena_dev_base(arena_), arena_bytes(arena_)) == PX_OK &&
             run_reduce(*ctx_, md.aux, tmp, count, dev_value_) == PX_OK;
        (void)hipStreamSynchronize(ctx_->stream());
        (void)hipFree(tmp);
      } else {
        (void)hipGetLastError();
        set_last_error(PX_ERR_NOMEM, "transform_reduce temporary");
      }
    }
    prof_end("transform_reduce", md.dtype);
  }
  ok = finish(op, false) && ok;
  release(op);
  if (ok) ok = copy_out(out_result, dev_value_, out_elem_size);
  if (!ok) report(who, last_error_text());
  return ok;
}

bool KernelLauncher::launch_scan(const std::string& scan_kernel, const std::string& add_kernel,
                                 void* data, size_t count, size_t elem_size) {
  const px_desc* s = find(scan_kernel, "launch_scan");
  const px_desc* a = find(add_kernel, "launch_scan");
  if (!s || !a) return false;
  if (s->op != PX_OP_SCAN) return report("launch_scan", "descriptor is not a scan"), false;
  if (!size_ok(s->dtype, elem_size))
    return report("launch_scan", "elem_size does not match the descriptor"), false;
  if (count <= 1) return true;  // the inclusive scan of <= 1 element is the identity
  Operand op;
  bool ok = resolve(data, count * elem_size, true, op);
  if (ok) {
    prof_begin();
    ok = run_scan(*ctx_, s->dtype, op.dev, op.dev, count, false, nullptr) == PX_OK;
    prof_end("scan", s->dtype);
  }
  ok = finish(op, ok) && ok;
  release(op);
  if (!ok) report("launch_scan", last_error_text());
  return ok;
}

bool KernelLauncher::launch_exclusive_scan(const std::string& scan_kernel,
                                           const std::string& add_kernel,
                                           const std::string& shift_kernel, void* input,
                                           void* output, size_t count, size_t elem_size,
                                           const void* init) {
  const px_desc* s = find(scan_kernel, "launch_exclusive_scan");
  const px_desc* a = find(add_kernel, "launch_exclusive_scan");
  const px_desc* h = find(shift_kernel, "launch_exclusive_scan");
  if (!s || !a || !h) return false;
  if (s->op != PX_OP_SCAN) return report("launch_exclusive_scan", "descriptor is not a scan"), false;
  if (!size_ok(s->dtype, elem_size))
    return report("launch_exclusive_scan", "elem_size does not match the descriptor"), false;
  if (count == 0) return true;
  Operand in, out;
  bool ok = resolve(input, count * elem_size, true, in) &&
            resolve(output, count * elem_size, false, out, 1);
  if (ok) {
    prof_begin();
    ok = run_scan(*ctx_, s->dtype, in.dev, out.dev, count, true, init) == PX_OK;
    prof_end("exclusive_scan", s->dtype);
  }
  ok = finish(out, ok) && ok;
  release(in);
  release(out);
  if (!ok) report("launch_exclusive_scan", last_error_text());
  return ok;
}

bool KernelLauncher::launch_sort(const std::string& kernel_name, void* data, size_t count,
                                 size_t elem_size) {
  const px_desc* d = find(kernel_name, "launch_sort");
  if (!d) return false;
  if (d->op != PX_OP_SORT) return report("launch_sort", "descriptor is not a sort"), false;
  if (!size_ok(d->dtype, elem_size))
    return report("launch_sort", "elem_size does not match the descriptor"), false;
  if (count <= 1) return true;
  Operand op;
  bool ok = resolve(data, count * elem_size, true, op);
  if (ok) {
    prof_begin();
    pxk::LaunchHook hook;
    if (prof_on_) hook = pxk::LaunchHook{&KernelLauncher::prof_hook, this};
    ok = run_sort(*ctx_, d->dtype, op.dev, count, hook) == PX_OK;
    prof_end("sort", d->dtype);
  }
  ok = finish(op, ok) && ok;
  release(op);
  if (!ok) report("launch_sort", last_error_text());
  return ok;
}

bool KernelLauncher::launch_compact(const std::string& flags_kernel, const std::string& scan_kernel,
                                    const std::string& add_kernel,
                                    const std::string& scatter_kernel, void* input, void* output,
                                    size_t count, size_t elem_size, bool elem_is_float,
                                    size_t* out_kept) {
  (void)elem_is_float;  // counts are integers here (the reference scanned float flags)
  if (out_kept) *out_kept = 0;
  const px_desc* f = find(flags_kernel, "launch_compact");
  const px_desc* s = find(scan_kernel, "launch_compact");
  const px_desc* a = find(add_kernel, "launch_compact");
  const px_desc* x = find(scatter_kernel, "launch_compact");
  if (!f || !s || !a || !x) return false;
  if (f->op != PX_OP_FLAGS || x->op != PX_OP_SCATTER)
    return report("launch_compact", "need a flags and a scatter descriptor"), false;
  if (!size_ok(f->dtype, elem_size))
    return report("launch_compact", "elem_size does not match the descriptor"), false;
  if (count == 0) return true;
  const bool partition = x->aux == PX_SCATTER_PARTITION || x->aux == PX_SCATTER_STABLE_PARTITION;
  if (partition && f->fn == PX_PRED_UNIQUE)
    return report("launch_compact", "partition by adjacent-unique flags is undefined"), false;
  // input == output compacts (or partitions) in place: one operand, staged once
  const bool in_place = input == output;
  Operand in, out;
  bool ok = resolve(input, count * elem_size, true, in) &&
            (in_place || resolve(output, count * elem_size, false, out, 1));
  Operand& dst = in_place ? in : out;
  if (ok) {
    prof_begin();
    ok = (partition ? run_partition(*ctx_, *f, in.dev, dst.dev, count, dev_word_,
                                    x->aux == PX_SCATTER_STABLE_PARTITION)
                    : run_compact(*ctx_, *f, in.dev, dst.dev, count, dev_word_)) == PX_OK;
    prof_end(partition ? "partition" : "copy_if", f->dtype);
  }
  uint64_t kept = 0;
  if (ok) ok = hipMemcpyAsync(&kept, dev_word_, 8, hipMemcpyDeviceToHost, ctx_->stream()) == hipSuccess;
  if (ok) ok = hipStreamSynchronize(ctx_->stream()) == hipSuccess;
  if (ok) {
    // only the written part of the output migrates back (the rest keeps host contents)
    dst.bytes = (partition ? count : kept) * elem_size;
    ok = finish(dst, true);
  }
  release(in);
  if (!in_place) release(out);
  if (!ok) {
    report("launch_compact", last_error_text());
    return false;
  }
  if (out_kept) *out_kept = static_cast<size_t>(kept);
  return true;
}

size_t KernelLauncher::launch_argminmax(const std::string& kernel_name, void* data, size_t count,
                                        size_t elem_size, bool is_float, bool want_max,
                                        bool want_last) {
  (void)is_float;
  const px_desc* d = find(kernel_name, "launch_argminmax");
  if (!d || count == 0 || !size_ok(d->dtype, elem_size)) return count;
  Operand op;
  bool ok = resolve(data, count * elem_size, true, op);
  uint64_t idx = count;
  if (ok) ok = run_argminmax(*ctx_, d->dtype, op.dev, count, want_max, want_last, dev_word_) == PX_OK;
  if (ok) ok = copy_out(&idx, dev_word_, 8);
  finish(op, false);
  release(op);
  return ok ? static_cast<size_t>(idx) : count;
}

size_t KernelLauncher::launch_find(const std::string& kernel_name, void* data, size_t count,
                                   size_t elem_size, bool negate, const void* value) {
  const px_desc* d = find(kernel_name, "launch_find");
  if (!d || count == 0 || !size_ok(d->dtype, elem_size)) return count;
  px_desc p = *d;
  if (value) {
    p.fn = PX_PRED_EQ;
    p.c0 = 0;
    std::memcpy(&p.c0, value, elem_size);
  }
  if (negate) p.flags ^= PX_FLAG_NEGATE;
  Operand op;
  bool ok = resolve(data, count * elem_size, true, op);
  uint64_t idx = count;
  if (ok) ok = run_find(*ctx_, p, op.dev, count, dev_word_) == PX_OK;
  if (ok) ok = copy_out(&idx, dev_word_, 8);
  finish(op, false);
  release(op);
  return ok ? static_cast<size_t>(idx) : count;
}

size_t KernelLauncher::launch_mismatch(const std::string& kernel_name, void* a, void* b,
                                       size_t count, size_t elem_size) {
  const px_desc* d = find(kernel_name, "launch_mismatch");
  if (!d || count == 0 || !size_ok(d->dtype, elem_size)) return count;
  Operand oa, ob;
  bool ok = resolve(a, count * elem_size, true, oa) && resolve(b, count * elem_size, true, ob, 1);
  uint64_t idx = count;
  if (ok) ok = run_mismatch(*ctx_, d->dtype, oa.dev, ob.dev, count, dev_word_) == PX_OK;
  if (ok) ok = copy_out(&idx, dev_word_, 8);
  finish(oa, false);
  release(oa);
  release(ob);
  return ok ? static_cast<size_t>(idx) : count;
}

bool KernelLauncher::launch_relocate_probe(void* arena_ptr, uint32_t count) {
  if (!arena_ || !arena_->host_mirrored() || !arena_->contains(arena_ptr)) return false;
  const size_t bytes = static_cast<size_t>(count) * sizeof(float);
  if (!arena_->flush_range(arena_ptr, bytes)) return false;
  // the kernel receives the HOST pointer and relocates it itself
  if (pxk::relocate_probe(reinterpret_cast<uint64_t>(arena_ptr),
                          reinterpret_cast<uint64_t>(arena_->host_base()), arena_->device_address(),
                          count, ctx_->stream()) != hipSuccess)
    return false;
  return arena_->invalidate_range(arena_ptr, bytes);
}

}  // namespace parallax
