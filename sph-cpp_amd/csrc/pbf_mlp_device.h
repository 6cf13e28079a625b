// pbf_mlp_device.h -- the learned warm start's network on MFMA, shared by
// the training kernels (pbf_mlp.hip) and the warm-started density pass
// (pbf_fused.hip), which evaluates it in registers.
//
// One 16x16x16 MFMA over a wavefront (lane l, c = l % 16, g = l / 16):
//   A (M x K): lane l holds A[c][4g .. 4g+3]
//   B (K x N): lane l holds B[4g .. 4g+3][c]
//   D (M x N): lane l holds D[4g .. 4g+3][c]
// Activations are kept transposed ([units x particles], 16 particles per
// tile), so the D tiles of one layer are the B operands of the next one.
//
// Parameters (fp32, PBF_WARM_PARAMS = 4801): W1 [64][8], b1 [64], W2 [64][64]
// (row = output unit), b2 [64], w3 [64], b3.
#pragma once

#include "pbf_device.h"

namespace pbf {

constexpr int kHid = PBF_WARM_HIDDEN;
constexpr int kFeat = PBF_WARM_FEATURES;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kHid * kFeat;
constexpr int OFF_W2 = OFF_B1 + kHid;
constexpr int OFF_B2 = OFF_W2 + kHid * kHid;
constexpr int OFF_W3 = OFF_B2 + kHid;
constexpr int OFF_B3 = OFF_W3 + kHid;
constexpr int kNPar = OFF_B3 + 1;
static_assert(kNPar == PBF_WARM_PARAMS, "parameter layout");

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even
__device__ __forceinline__ short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// sum over the 16 lanes that share g
__device__ __forceinline__ float red16(float v) {
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Device copies of one set of weights.
struct MlpW {
  const short* w1;   // bf16 [64][16]: W1 with features 8..15 zero
  const short* w2;   // bf16 [64][64]: W2
  const short* w2t;  // bf16 [64][64]: W2 transposed (training only)
  const float* p;    // fp32 parameters
};

__device__ __forceinline__ void load_forward_operands(const MlpW& W, int c, int g, v4s (&a1)[4],
                                                      v4s (&a2)[4][4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) a1[t] = *(const v4s*)(W.w1 + (16 * t + c) * 16 + 4 * g);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int k = 0; k < 4; ++k) a2[t][k] = *(const v4s*)(W.w2 + (16 * t + c) * kHid + 16 * k + 4 * g);
}

// Forward pass of 16 particles (column c of the tiles).  xb: B operand of
// the features (rows 8..15 zero).  Returns the output of particle c in every lane.
__device__ __forceinline__ float mlp_forward(const MlpW& W, const v4s (&a1)[4], const v4s (&a2)[4][4], v4s xb,
                                             int g, v4f (&z1)[4], v4f (&h1)[4], v4f (&z2)[4], v4f (&h2)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const v4f acc = mfma16(a1[t], xb, v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float z = acc[j] + W.p[OFF_B1 + 16 * t + 4 * g + j];
      z1[t][j] = z;
      h1[t][j] = silu(z);
    }
  }
  v4f acc2[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc2[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const v4s hb = {f2bf(h1[k][0]), f2bf(h1[k][1]), f2bf(h1[k][2]), f2bf(h1[k][3])};
#pragma unroll
    for (int t = 0; t < 4; ++t) acc2[t] = mfma16(a2[t][k], hb, acc2[t]);
  }
  float part = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * t + 4 * g + j;
      const float z = acc2[t][j] + W.p[OFF_B2 + r];
      z2[t][j] = z;
      h2[t][j] = silu(z);
      part += W.p[OFF_W3 + r] * h2[t][j];
    }
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return part + W.p[OFF_B3];
}

}  // namespace pbf
