// pbf_mlp.hip -- learned warm start for the iterate-to-tolerance solve.
//
// No counterpart in the reference: its pressure solve starts every step's
// Jacobi iterations from lambda = 0 (Simulation/Source/particles.h:318-358).
// Here an 8-64-64-1 network with SiLU activations maps per-particle features
// of the first density pass to a correction of the previous solve's
// multiplier sums: the warm multiplier is Lambda_prev + S * net(features)
// (pbf_solve, PBF_WARM_LEARNED).
// It is evaluated and trained on the GPU: BF16 operands on MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation), FP32 master weights, Adam.
//
// The network's layout and forward pass are in pbf_mlp_device.h.  Weight
// gradients contract over particles and go through LDS.
#include <algorithm>
#include <cstring>
#include <random>
#include <vector>

#include "pbf_mlp_device.h"

namespace pbf {

constexpr int kMaxTrainWaves = 1024;

__device__ __forceinline__ v4s feature_operand(const float* __restrict__ f, bool valid, int g) {
  v4s xb = {0, 0, 0, 0};
  if (valid && g < 2) {
    const float4 t = *(const float4*)(f + 4 * g);
    xb = v4s{f2bf(t.x), f2bf(t.y), f2bf(t.z), f2bf(t.w)};
  }
  return xb;
}

// out[p] = base[p] + scale * net(feat[p]) (base may be null: 0); one
// wavefront per block, 16 particles per step.
__global__ void __launch_bounds__(64) k_mlp_forward(uint32_t n, const float* __restrict__ feat, MlpW W,
                                                    float scale, const float* __restrict__ base,
                                                    float* __restrict__ out) {
  const int l = threadIdx.x, c = l & 15, g = l >> 4;
  v4s a1[4], a2[4][4];
  load_forward_operands(W, c, g, a1, a2);
  const uint32_t groups = (n + 15) / 16;
  for (uint32_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
    const uint32_t p = gi * 16 + c;
    const bool valid = p < n;
    const v4s xb = feature_operand(feat + (uint64_t)(valid ? p : 0) * kFeat, valid, g);
    v4f z1[4], h1[4], z2[4], h2[4];
    const float y = mlp_forward(W, a1, a2, xb, g, z1, h1, z2, h2);
    if (valid && g == 0) out[p] = base ? base[p] + scale * y : scale * y;
  }
}

// Gradient of the mean squared error over `count` samples idx[0..count) for
// one batch.  Block b (one wavefront) handles sample groups b, b + grid, ...
// and writes its partial gradient to part[b][kNPar] and its partial sum of
// squared errors to lpart[b].
__global__ void __launch_bounds__(64) k_mlp_train(uint32_t count, const uint32_t* __restrict__ idx,
                                                  const float* __restrict__ feat, const float* __restrict__ tgt,
                                                  MlpW W, float inv_batch, float* __restrict__ part,
                                                  float* __restrict__ lpart) {
  __shared__ __attribute__((aligned(16))) short sA[kHid * 16];  // [unit][particle]: dZ2, then dZ1
  __shared__ __attribute__((aligned(16))) short sH[kHid * 16];  // [unit][particle]: H1
  __shared__ __attribute__((aligned(16))) short sX[16 * 16];    // [particle][feature]
  const int l = threadIdx.x, c = l & 15, g = l >> 4;
  v4s a1[4], a2[4][4], a2t[4][4];
  load_forward_operands(W, c, g, a1, a2);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int k = 0; k < 4; ++k) a2t[t][k] = *(const v4s*)(W.w2t + (16 * t + c) * kHid + 16 * k + 4 * g);
  v4f gW2[4][4], gW1[4];
  float db1[4][4], db2[4][4], dw3[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    gW1[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gW2[t][k] = v4f{0.f, 0.f, 0.f, 0.f};
      db1[t][k] = db2[t][k] = dw3[t][k] = 0.f;
    }
  }
  float db3 = 0.f, loss = 0.f;
  const uint32_t groups = (count + 15) / 16;
  for (uint32_t gi = blockIdx.x; gi < groups; gi += gridDim.x) {
    const uint32_t s = gi * 16 + c;
    const bool valid = s < count;
    const uint32_t sample = valid ? idx[s] : 0u;
    const v4s xb = feature_operand(feat + (uint64_t)sample * kFeat, valid, g);
#pragma unroll
    for (int j = 0; j < 4; ++j) sX[c * 16 + 4 * g + j] = xb[j];
    v4f z1[4], h1[4], z2[4], h2[4];
    const float out = mlp_forward(W, a1, a2, xb, g, z1, h1, z2, h2);
    const float diff = valid ? out - tgt[sample] : 0.f;
    const float dout = 2.0f * diff * inv_batch;
    if (g == 0) {
      loss += diff * diff;
      db3 += dout;
    }
    v4f dz2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * t + 4 * g + j;
        dw3[t][j] += dout * h2[t][j];
        const float d = dout * W.p[OFF_W3 + r] * dsilu(z2[t][j]);
        dz2[t][j] = d;
        db2[t][j] += d;
        sA[r * 16 + c] = f2bf(d);
        sH[r * 16 + c] = f2bf(h1[t][j]);
      }
    __syncthreads();
    // dW2 += dZ2 . H1^T (contraction over the 16 particles)
#pragma unroll
    for (int to = 0; to < 4; ++to) {
      const v4s A = *(const v4s*)(sA + (16 * to + c) * 16 + 4 * g);
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
        const v4s B = *(const v4s*)(sH + (16 * ti + c) * 16 + 4 * g);
        gW2[to][ti] = mfma16(A, B, gW2[to][ti]);
      }
    }
    // dH1 = W2^T . dZ2, dZ1 = dH1 * silu'(Z1)
    v4f dz1[4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const v4s bz = {f2bf(dz2[k][0]), f2bf(dz2[k][1]), f2bf(dz2[k][2]), f2bf(dz2[k][3])};
        acc = mfma16(a2t[ti][k], bz, acc);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        dz1[ti][j] = acc[j] * dsilu(z1[ti][j]);
        db1[ti][j] += dz1[ti][j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) sA[(16 * t + 4 * g + j) * 16 + c] = f2bf(dz1[t][j]);
    __syncthreads();
    // dW1 += dZ1 . X^T
    const v4s B1 = {sX[(4 * g + 0) * 16 + c], sX[(4 * g + 1) * 16 + c], sX[(4 * g + 2) * 16 + c],
                    sX[(4 * g + 3) * 16 + c]};
#pragma unroll
    for (int to = 0; to < 4; ++to) {
      const v4s A = *(const v4s*)(sA + (16 * to + c) * 16 + 4 * g);
      gW1[to] = mfma16(A, B1, gW1[to]);
    }
    __syncthreads();
  }
  float* P = part + (uint64_t)blockIdx.x * kNPar;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 16 * t + 4 * g + j;
      if (c < kFeat) P[OFF_W1 + r * kFeat + c] = gW1[t][j];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) P[OFF_W2 + r * kHid + 16 * ti + c] = gW2[t][ti][j];
      const float b1s = red16(db1[t][j]), b2s = red16(db2[t][j]), w3s = red16(dw3[t][j]);
      if (c == 0) {
        P[OFF_B1 + r] = b1s;
        P[OFF_B2 + r] = b2s;
        P[OFF_W3 + r] = w3s;
      }
    }
  db3 = red16(db3);
  loss = red16(loss);
  if (l == 0) {
    P[OFF_B3] = db3;
    lpart[blockIdx.x] = loss;
  }
}

// grad[i] = sum over blocks of part[b][i], in block order (deterministic);
// loss_out[0] = inv_batch * sum of lpart, and loss_acc[0] += the same.
__global__ void k_mlp_reduce(uint32_t nblk, const float* __restrict__ part, const float* __restrict__ lpart,
                             float inv_batch, float* __restrict__ grad, float* __restrict__ loss_out,
                             float* __restrict__ loss_acc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint32_t)kNPar) {
    float s = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) s += part[(uint64_t)b * kNPar + i];
    grad[i] = s;
  }
  if (i == 0) {
    float s = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) s += lpart[b];
    loss_out[0] = s * inv_batch;
    if (loss_acc) loss_acc[0] += s * inv_batch;
  }
}

// Adam (Kingma & Ba 2015, algorithm 1) on the fp32 master weights.
__global__ void k_mlp_adam(float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint32_t)kNPar) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gr;
  const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// bf16 operand copies of the weights
__global__ void k_mlp_refresh(const float* __restrict__ p, short* __restrict__ w1, short* __restrict__ w2,
                              short* __restrict__ w2t) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint32_t)(kHid * 16)) {
    const int r = i / 16, f = i % 16;
    w1[i] = f < kFeat ? f2bf(p[OFF_W1 + r * kFeat + f]) : (short)0;
  }
  if (i < (uint32_t)(kHid * kHid)) {
    const int o = i / kHid, in = i % kHid;
    const short b = f2bf(p[OFF_W2 + i]);
    w2[i] = b;
    w2t[in * kHid + o] = b;
  }
}

// samples i = 0..n-1 from particle slot i * stride: the features and the
// target (lsum - lprev) / S limited to [-kTargetLimit, kTargetLimit]
constexpr float kTargetLimit = 64.0f;
__global__ void k_take_samples(uint64_t n, uint32_t stride, const float* __restrict__ feat,
                               const float* __restrict__ lsum, const float* __restrict__ lprev, float inv_scale,
                               float* __restrict__ fo, float* __restrict__ to) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = i * stride;
  const float4* f = (const float4*)(feat + p * kFeat);
  float4* o = (float4*)(fo + i * kFeat);
  o[0] = f[0];
  o[1] = f[1];
  const float t = (lsum[p] - lprev[p]) * inv_scale;
  to[i] = t > -kTargetLimit ? (t < kTargetLimit ? t : kTargetLimit) : -kTargetLimit;
}

}  // namespace pbf

using namespace pbf;

struct pbf_warm_model {
  hipStream_t stream = nullptr;
  float* p = nullptr;  // fp32 master weights
  short *w1 = nullptr, *w2 = nullptr, *w2t = nullptr;
  float *m = nullptr, *v = nullptr, *grad = nullptr;
  float *part = nullptr, *lpart = nullptr, *loss = nullptr, *loss_acc = nullptr;
  float *feat = nullptr, *tgt = nullptr;
  uint32_t *perm = nullptr, *iota = nullptr;
  uint64_t count = 0, max_samples = 0;
  uint64_t adam_t = 0;
  bool has_params = false;
};

namespace pbf {

static MlpW model_weights(const pbf_warm_model* M) { return MlpW{M->w1, M->w2, M->w2t, M->p}; }

static void refresh(pbf_warm_model* M) {
  hipLaunchKernelGGL(k_mlp_refresh, dim3((kHid * kHid + 255) / 256), dim3(256), 0, M->stream, (const float*)M->p,
                     M->w1, M->w2, M->w2t);
}

// one batch: gradient (and loss) of samples idx[0..count)
static void batch_gradient(pbf_warm_model* M, const uint32_t* idx, uint32_t count, float* loss_acc) {
  const uint32_t groups = (count + 15) / 16;
  const uint32_t blocks = std::min<uint32_t>(groups, (uint32_t)kMaxTrainWaves);
  const float inv = 1.0f / (float)count;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(64), 0, M->stream, count, idx, (const float*)M->feat,
                     (const float*)M->tgt, model_weights(M), inv, M->part, M->lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNPar + 255) / 256), dim3(256), 0, M->stream, blocks,
                     (const float*)M->part, (const float*)M->lpart, inv, M->grad, M->loss, loss_acc);
}

// Inference into a context's warm multipliers (pbf_solve, PBF_WARM_LEARNED).

}  // namespace pbf

extern "C" {

int pbf_warm_model_create(uint64_t max_samples, pbf_warm_model** out) {
  PBF_TRACE("pbf_warm_model_create");
  if (!out) return set_error(PBF_ERR_ARG, "null argument");
  *out = nullptr;
  if (max_samples == 0 || max_samples >= (1ull << 32)) return set_error(PBF_ERR_ARG, "max_samples must be in [1, 2^32)");
  pbf_warm_model* M = new pbf_warm_model();
  M->max_samples = max_samples;
  auto A = [](void** ptr, uint64_t bytes) { return check_hip(hipMalloc(ptr, bytes), "hipMalloc"); };
  int rc;
  if ((rc = check_hip(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking), "hipStreamCreate")) ||
      (rc = A((void**)&M->p, kNPar * 4)) || (rc = A((void**)&M->w1, kHid * 16 * 2)) ||
      (rc = A((void**)&M->w2, kHid * kHid * 2)) || (rc = A((void**)&M->w2t, kHid * kHid * 2)) ||
      (rc = A((void**)&M->m, kNPar * 4)) || (rc = A((void**)&M->v, kNPar * 4)) ||
      (rc = A((void**)&M->grad, kNPar * 4)) || (rc = A((void**)&M->part, (uint64_t)kMaxTrainWaves * kNPar * 4)) ||
      (rc = A((void**)&M->lpart, kMaxTrainWaves * 4)) || (rc = A((void**)&M->loss, 4)) ||
      (rc = A((void**)&M->loss_acc, 4)) || (rc = A((void**)&M->feat, max_samples * kFeat * 4)) ||
      (rc = A((void**)&M->tgt, max_samples * 4)) || (rc = A((void**)&M->perm, max_samples * 4)) ||
      (rc = A((void**)&M->iota, max_samples * 4))) {
    pbf_warm_model_destroy(M);
    return rc;
  }
  std::vector<uint32_t> io(max_samples);
  for (uint64_t i = 0; i < max_samples; ++i) io[i] = (uint32_t)i;
  if ((rc = check_hip(hipMemcpy(M->iota, io.data(), max_samples * 4, hipMemcpyHostToDevice), "upload"))) {
    pbf_warm_model_destroy(M);
    return rc;
  }
  *out = M;
  return PBF_OK;
}

int pbf_warm_model_destroy(pbf_warm_model* M) {
  if (!M) return PBF_OK;
  if (M->stream) hipStreamSynchronize(M->stream);
  void* ptrs[] = {M->p, M->w1, M->w2, M->w2t, M->m, M->v, M->grad, M->part, M->lpart, M->loss, M->loss_acc,
                  M->feat, M->tgt, M->perm, M->iota};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (M->stream) hipStreamDestroy(M->stream);
  delete M;
  return PBF_OK;
}

int pbf_warm_model_set_params(pbf_warm_model* M, const float* params) {
  if (!M || !params) return set_error(PBF_ERR_ARG, "null argument");
  int rc = check_hip(hipMemcpyAsync(M->p, params, kNPar * 4, hipMemcpyHostToDevice, M->stream), "upload");
  if (rc) return rc;
  refresh(M);
  launch_fill_u32((uint32_t*)M->m, 0u, kNPar, M->stream);
  launch_fill_u32((uint32_t*)M->v, 0u, kNPar, M->stream);
  M->adam_t = 0;
  M->has_params = true;
  return check_hip(hipStreamSynchronize(M->stream), "set_params");
}

int pbf_warm_model_get_params(pbf_warm_model* M, float* params) {
  if (!M || !params) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  int rc = check_hip(hipMemcpyAsync(params, M->p, kNPar * 4, hipMemcpyDeviceToHost, M->stream), "download");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(M->stream), "download");
}

// Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases.
int pbf_warm_model_init(pbf_warm_model* M, uint64_t seed) {
  if (!M) return set_error(PBF_ERR_ARG, "null argument");
  std::mt19937_64 rng(seed);
  std::vector<float> p(kNPar, 0.0f);
  auto fill = [&](int off, int count, float bound) {
    std::uniform_real_distribution<float> u(-bound, bound);
    for (int i = 0; i < count; ++i) p[off + i] = u(rng);
  };
  fill(OFF_W1, kHid * kFeat, 1.0f / std::sqrt((float)kFeat));
  fill(OFF_W2, kHid * kHid, 1.0f / std::sqrt((float)kHid));
  fill(OFF_W3, kHid, 1.0f / std::sqrt((float)kHid));
  return pbf_warm_model_set_params(M, p.data());
}

int pbf_warm_model_clear_samples(pbf_warm_model* M) {
  if (!M) return set_error(PBF_ERR_ARG, "null argument");
  M->count = 0;
  return PBF_OK;
}

int pbf_warm_model_add_samples(pbf_warm_model* M, const float* features, const float* targets, uint64_t count) {
  if (!M || (count && (!features || !targets))) return set_error(PBF_ERR_ARG, "null argument");
  if (M->count + count > M->max_samples) return set_error(PBF_ERR_CAPACITY, "sample store is full");
  int rc = check_hip(hipMemcpyAsync(M->feat + M->count * kFeat, features, count * kFeat * 4, hipMemcpyHostToDevice,
                                    M->stream), "upload");
  if (rc) return rc;
  if ((rc = check_hip(hipMemcpyAsync(M->tgt + M->count, targets, count * 4, hipMemcpyHostToDevice, M->stream),
                      "upload")))
    return rc;
  if ((rc = check_hip(hipStreamSynchronize(M->stream), "upload"))) return rc;
  M->count += count;
  return PBF_OK;
}

// The features of the context's last warm-started solve and the multiplier
// sums that solve applied (scaled by 1/kWarmLambdaScale) of every stride-th
// particle, device to device.
int pbf_warm_model_add_from_context(pbf_warm_model* M, pbf_context* c, uint32_t stride) {
  PBF_TRACE("pbf_warm_model_add_from_context");
  if (!M || !c) return set_error(PBF_ERR_ARG, "null argument");
  if (stride == 0) return set_error(PBF_ERR_ARG, "stride must be >= 1");
  if (!c->feat_valid)
    return set_error(PBF_ERR_STATE,
                     "the context has no features (run a warm-started pbf_solve with PBF_OPT_KEEP_WARM_FEATURES)");
  const uint64_t n = (c->n + stride - 1) / stride;
  if (M->count + n > M->max_samples) return set_error(PBF_ERR_CAPACITY, "sample store is full");
  int rc = check_hip(hipStreamSynchronize(M->stream), "sync");
  if (rc) return rc;
  if (n) {
    const uint32_t blocks = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_take_samples, dim3(blocks), dim3(256), 0, c->stream, n, stride, (const float*)c->feat,
                       (const float*)c->lsum, (const float*)c->lsuma, 1.0f / kWarmLambdaScale,
                       M->feat + M->count * kFeat,
                       M->tgt + M->count);
  }
  if ((rc = check_hip(hipStreamSynchronize(c->stream), "add_from_context"))) return rc;
  M->count += n;
  return PBF_OK;
}

uint64_t pbf_warm_model_sample_count(const pbf_warm_model* M) { return M ? M->count : 0; }

int pbf_warm_model_gradient(pbf_warm_model* M, uint64_t first, uint32_t count, float* grad, float* loss) {
  PBF_TRACE("pbf_warm_model_gradient");
  if (!M || !grad) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  if (count == 0 || first + count > M->count) return set_error(PBF_ERR_ARG, "sample range out of bounds");
  batch_gradient(M, M->iota + first, count, nullptr);
  int rc = check_hip(hipMemcpyAsync(grad, M->grad, kNPar * 4, hipMemcpyDeviceToHost, M->stream), "download");
  if (rc) return rc;
  if (loss && (rc = check_hip(hipMemcpyAsync(loss, M->loss, 4, hipMemcpyDeviceToHost, M->stream), "download")))
    return rc;
  return check_hip(hipStreamSynchronize(M->stream), "gradient");
}

int pbf_warm_model_adam_step(pbf_warm_model* M, const float* grad, float learning_rate) {
  PBF_TRACE("pbf_warm_model_adam_step");
  if (!M || !grad) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  int rc = check_hip(hipMemcpyAsync(M->grad, grad, kNPar * 4, hipMemcpyHostToDevice, M->stream), "upload");
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f;
  ++M->adam_t;
  const float c1 = 1.0f - std::pow(b1, (float)M->adam_t), c2 = 1.0f - std::pow(b2, (float)M->adam_t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNPar + 255) / 256), dim3(256), 0, M->stream, M->p, (const float*)M->grad,
                     M->m, M->v, learning_rate, b1, b2, 1e-8f, c1, c2);
  refresh(M);
  return check_hip(hipStreamSynchronize(M->stream), "adam");
}

int pbf_warm_model_train(pbf_warm_model* M, uint32_t epochs, uint32_t batch, float learning_rate, uint64_t seed,
                         float* mean_loss) {
  PBF_TRACE("pbf_warm_model_train");
  if (!M) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  if (M->count == 0) return set_error(PBF_ERR_STATE, "no samples");
  if (batch == 0) return set_error(PBF_ERR_ARG, "batch must be positive");
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> perm(M->count);
  const float b1 = 0.9f, b2 = 0.999f;
  uint64_t batches = 0;
  launch_fill_u32((uint32_t*)M->loss_acc, 0u, 1, M->stream);
  for (uint32_t e = 0; e < epochs; ++e) {
    for (uint64_t i = 0; i < M->count; ++i) perm[i] = (uint32_t)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    // the permutation buffer is rewritten only after the previous epoch's batches ran
    int rc = check_hip(hipStreamSynchronize(M->stream), "train");
    if (rc) return rc;
    if ((rc = check_hip(hipMemcpyAsync(M->perm, perm.data(), M->count * 4, hipMemcpyHostToDevice, M->stream),
                        "upload")))
      return rc;
    for (uint64_t s = 0; s < M->count; s += batch) {
      const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, M->count - s);
      batch_gradient(M, M->perm + s, cnt, M->loss_acc);
      ++M->adam_t;
      const float c1 = 1.0f - std::pow(b1, (float)M->adam_t), c2 = 1.0f - std::pow(b2, (float)M->adam_t);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kNPar + 255) / 256), dim3(256), 0, M->stream, M->p,
                         (const float*)M->grad, M->m, M->v, learning_rate, b1, b2, 1e-8f, c1, c2);
      refresh(M);
      ++batches;
    }
  }
  float acc = 0.f;
  int rc = check_hip(hipMemcpyAsync(&acc, M->loss_acc, 4, hipMemcpyDeviceToHost, M->stream), "download");
  if (rc) return rc;
  if ((rc = check_hip(hipStreamSynchronize(M->stream), "train"))) return rc;
  if (mean_loss) *mean_loss = batches ? acc / (float)batches : 0.f;
  return PBF_OK;
}

int pbf_warm_model_predict(pbf_warm_model* M, const float* features, uint64_t count, float* out) {
  PBF_TRACE("pbf_warm_model_predict");
  if (!M || (count && (!features || !out))) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  if (!count) return PBF_OK;
  if (count >= (1ull << 32)) return set_error(PBF_ERR_ARG, "count out of range");
  float *f = nullptr, *o = nullptr;
  int rc = check_hip(hipMalloc((void**)&f, count * kFeat * 4), "hipMalloc");
  if (rc) return rc;
  if ((rc = check_hip(hipMalloc((void**)&o, count * 4), "hipMalloc"))) {
    hipFree(f);
    return rc;
  }
  rc = check_hip(hipMemcpyAsync(f, features, count * kFeat * 4, hipMemcpyHostToDevice, M->stream), "upload");
  if (!rc) {
    const uint32_t groups = (uint32_t)((count + 15) / 16);
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::min<uint32_t>(groups, 4096u)), dim3(64), 0, M->stream,
                       (uint32_t)count, (const float*)f, model_weights(M), 1.0f, nullptr, o);
    rc = check_hip(hipMemcpyAsync(out, o, count * 4, hipMemcpyDeviceToHost, M->stream), "download");
  }
  if (!rc) rc = check_hip(hipStreamSynchronize(M->stream), "predict");
  hipFree(f);
  hipFree(o);
  return rc;
}

int pbf_set_warm_gain(pbf_context* c, float gain) {
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (!(gain >= 0.0f && gain <= 4.0f)) return set_error(PBF_ERR_ARG, "gain must be in [0, 4]");
  c->warm_gain = gain;
  return PBF_OK;
}

int pbf_set_warm_model(pbf_context* c, pbf_warm_model* M) {
  PBF_TRACE("pbf_set_warm_model");
  if (!c || !M) return set_error(PBF_ERR_ARG, "null argument");
  if (!M->has_params) return set_error(PBF_ERR_STATE, "model has no parameters");
  int rc = check_hip(hipStreamSynchronize(M->stream), "sync");
  if (rc) return rc;
  if ((rc = check_hip(hipMemcpyAsync(c->mp, M->p, kNPar * 4, hipMemcpyDeviceToDevice, c->stream), "copy")) ||
      (rc = check_hip(hipMemcpyAsync(c->mw1, M->w1, kHid * 16 * 2, hipMemcpyDeviceToDevice, c->stream), "copy")) ||
      (rc = check_hip(hipMemcpyAsync(c->mw2, M->w2, kHid * kHid * 2, hipMemcpyDeviceToDevice, c->stream), "copy")))
    return rc;
  c->has_model = true;
  return check_hip(hipStreamSynchronize(c->stream), "set_warm_model");
}

}  // extern "C"
