// user_kernels.hip -- functors compiled to a gfx950 code object, as a compiler would emit
// them for funnel lambdas that no built-in map describes.  They follow the PX_OP_CODE
// kernel signatures of parallax/descriptor.h and are loaded through that payload by
// tests/test_gpu_code_object.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

// in place: x = a * x * x + b with captures {float a, float b}
extern "C" __global__ void user_poly_f32(void* data, uint64_t count, const void* caps,
                                         uint64_t cap_size, uint64_t, uint64_t) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count) return;
  float a = 1.0f, b = 0.0f;
  if (cap_size >= 8) {
    a = static_cast<const float*>(caps)[0];
    b = static_cast<const float*>(caps)[1];
  }
  float* x = static_cast<float*>(data);
  const float v = x[i];
  x[i] = __fmaf_rn(a * v, v, b);
}

// in -> out: out = sqrt((double)in) + c with captures {double c}
extern "C" __global__ void user_sqrt_u32_f64(const void* in, void* out, uint64_t count,
                                             const void* caps, uint64_t cap_size, uint64_t,
                                             uint64_t) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count) return;
  const double c = cap_size >= 8 ? static_cast<const double*>(caps)[0] : 0.0;
  static_cast<double*>(out)[i] = __builtin_sqrt(static_cast<double>(static_cast<const uint32_t*>(in)[i])) + c;
}

// in -> out: out[i] = table[in[i] mod size] through a captured arena pointer
// {const int32_t* table (host address), uint64_t size}, relocated on the device.
extern "C" __global__ void user_lookup_i32(const void* in, void* out, uint64_t count,
                                           const void* caps, uint64_t cap_size,
                                           uint64_t host_base, uint64_t dev_base) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= count || cap_size < 16) return;
  uint64_t p = static_cast<const uint64_t*>(caps)[0];
  const uint64_t size = static_cast<const uint64_t*>(caps)[1];
  if (host_base != 0 && p >= host_base) p = dev_base + (p - host_base);
  const int32_t* table = reinterpret_cast<const int32_t*>(p);
  const uint32_t k = static_cast<const uint32_t*>(in)[i];
  static_cast<int32_t*>(out)[i] = size ? table[k % size] : 0;
}
