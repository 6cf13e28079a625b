// pbf_check.hip -- device self-check of the library's own divide and square
// root (div3_shared, sqrt_cr in pbf_device.h) against the plain operations.
//
// No counterpart in the reference.  The neighbour passes use these two
// functions in place of the compiler's correctly rounded divide and square
// root; they are built to give the same bits for every input.  This kernel
// compares them on pseudo-random operands of three kinds: raw 32-bit
// patterns (every exponent, zeros, denormals, infinities, NaNs); operands
// shaped like the passes' own (divisors r + FLT_EPSILON with r < 0.13,
// numerators s d with |s| <= 1.5e5 and |d| <= 0.13, squared distances below
// 0.0169, with signed zero numerators among them); and log-uniform
// magnitudes over 2^-48 .. 2^48 (both sides of the divide's range test).
// The instances with the range tests discharged on the host
// (div3_shared_static, sqrt_corrected) are compared too, on the operands
// that meet their preconditions.
#include "pbf_device.h"

namespace pbf {

// splitmix64 (G. Steele, D. Lea, C. Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014): a bijective 64-bit mixer
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float unit(uint64_t h) { return (float)(h >> 40) * 0x1p-24f; }  // [0, 1)

__device__ __forceinline__ bool same_bits(float x, float y) { return __float_as_uint(x) == __float_as_uint(y); }
__device__ __forceinline__ bool same3(f3 a, f3 b) { return same_bits(a.x, b.x) && same_bits(a.y, b.y) && same_bits(a.z, b.z); }
// what the host guarantees before it selects div3_shared_static
__device__ __forceinline__ bool static_div_ok(f3 a, float b) {
  const unsigned ok = (unsigned)(b >= 0x1p-40f) & (unsigned)(b <= 0x1p40f) & (unsigned)(fabsf(a.x) <= 0x1p40f) &
                      (unsigned)(fabsf(a.y) <= 0x1p40f) & (unsigned)(fabsf(a.z) <= 0x1p40f);
  return ok != 0u;
}

__global__ void k_check_arithmetic(uint64_t samples, uint64_t seed, unsigned long long* __restrict__ out) {
  unsigned long long bad_div = 0, bad_sqrt = 0, fast_div = 0, fast_sqrt = 0, bad_shaped = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t base = mix64(seed);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < samples; i += stride) {
    const uint64_t h0 = mix64(base + 4 * i), h1 = mix64(base + 4 * i + 1);
    const uint64_t h2 = mix64(base + 4 * i + 2), h3 = mix64(base + 4 * i + 3);
    const uint32_t kind = (uint32_t)(i % 3);
    f3 a;
    float b, x;
    if (kind == 2) {  // log-uniform magnitudes, random signs of the numerators
      auto lu = [](uint64_t h) { return ldexpf(1.0f + unit(h), (int)((h >> 8) % 97) - 48); };
      a = {(h0 & 1 ? -1.0f : 1.0f) * lu(h0 >> 1), (h1 & 1 ? -1.0f : 1.0f) * lu(h1 >> 1),
           (h2 & 1 ? -1.0f : 1.0f) * lu(h2 >> 1)};
      b = lu(h3);
      x = lu(h3 >> 3);
    } else if (kind == 1) {  // raw bit patterns
      a = {__uint_as_float((uint32_t)h0), __uint_as_float((uint32_t)(h0 >> 32)), __uint_as_float((uint32_t)h1)};
      b = __uint_as_float((uint32_t)(h1 >> 32));
      x = __uint_as_float((uint32_t)h2);
    } else {  // shaped like the neighbour passes' operands
      const float r = 0.13f * unit(h0);
      const float s = -1.5e5f * unit(h1) * unit(h1 << 24);
      const float dx = 0.26f * unit(h2) - 0.13f, dy = 0.26f * unit(h2 << 24) - 0.13f;
      const float dz = (h3 & 7) == 0 ? 0.0f : 0.26f * unit(h3) - 0.13f;
      a = {s * dx, s * dy, s * dz};
      if ((h3 & 0x70) == 0x10) a.x = (h3 & 0x80) ? -0.0f : 0.0f;  // signed zero numerators
      b = r + 1.1920929e-7f;
      x = (dx * dx + dy * dy) + dz * dz;
    }
    const f3 q = div3_shared(a, b);
    const f3 ref = div3(a, b);
    fast_div += div3_fast(a, b) ? 1 : 0;
    const bool bad = !same3(q, ref) || (static_div_ok(a, b) && !same3(div3_shared_static(a, b), ref));
    bad_div += bad ? 1 : 0;
    bad_shaped += (bad && kind == 0) ? 1 : 0;
    fast_sqrt += sqrt_fast(x) ? 1 : 0;
    bad_sqrt += (same_bits(sqrt_cr(x), sqrtf(x)) && (!sqrt_fast(x) || same_bits(sqrt_corrected(x), sqrtf(x)))) ? 0 : 1;
  }
  // one update per wavefront
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    bad_div += __shfl_xor(bad_div, d, 64);
    bad_sqrt += __shfl_xor(bad_sqrt, d, 64);
    fast_div += __shfl_xor(fast_div, d, 64);
    fast_sqrt += __shfl_xor(fast_sqrt, d, 64);
    bad_shaped += __shfl_xor(bad_shaped, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out[0], bad_div);
    atomicAdd(&out[1], bad_sqrt);
    atomicAdd(&out[2], fast_div);
    atomicAdd(&out[3], fast_sqrt);
    atomicAdd(&out[4], bad_shaped);
  }
}

}  // namespace pbf

using namespace pbf;

extern "C" int pbf_check_arithmetic(uint64_t samples, uint64_t seed, uint64_t* out) {
  if (!out) return set_error(PBF_ERR_ARG, "null argument");
  unsigned long long* d = nullptr;
  int rc = check_hip(hipMalloc((void**)&d, 5 * sizeof(unsigned long long)), "hipMalloc");
  if (rc) return rc;
  rc = check_hip(hipMemset(d, 0, 5 * sizeof(unsigned long long)), "hipMemset");
  if (!rc && samples) {
    hipLaunchKernelGGL(k_check_arithmetic, dim3(2048), dim3(256), 0, nullptr, samples, seed, d);
    rc = check_hip(hipGetLastError(), "check_arithmetic");
  }
  if (!rc) rc = check_hip(hipMemcpy(out, d, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost), "hipMemcpy");
  hipFree(d);
  return rc;
}
