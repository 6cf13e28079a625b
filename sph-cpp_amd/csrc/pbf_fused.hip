// pbf_fused.hip -- neighbour passes of the fused step (pbf_step).
//
// Same arithmetic, order and results as the per-pass kernels in
// pbf_passes.hip (particles.h:292-487), organised for a 64-wide wavefront:
//
//  * Latency: a neighbour walk is a chain of dependent loads (list entry, then
//    the neighbour's data).  Neighbours are taken in batches of kBatch: the
//    batch's list entries load first (entry k of a wavefront's 64 particles
//    is one 256-byte row), then all gathers are issued, then the arithmetic
//    runs in list order.  The density pass loads the next batch's entries
//    while it works on the current one.
//  * Only about half of a particle's listed neighbours lie within h (the list
//    radius is the 1.2 h bin spacing, neighbors.h:198); the others contribute
//    exact zeros.  The density pass walks the whole list once, summing the
//    entries within h as it goes, and records them in a bit mask per particle
//    (one bit per list entry); the Jacobi iteration's dp pass, which sees the
//    same positions, walks the list again but gathers and sums only the
//    masked entries.  The first XSPH pass appends the entries within h of the
//    final positions to a compact row (cnb), staged in LDS; vorticity,
//    vorticity-apply and the second XSPH pass walk only that row, in the same
//    order.
//  * The density pass writes lambda into the w word of its x* output, so the
//    dp pass gathers a neighbour's x* and lambda in one 16-byte load.
//  * The near-coincidence fix of compute_densities (particles.h:303-305) moves
//    x*_p in the middle of the walk; when it fires, the density pass redoes
//    that particle with a plain walk from its first entry and raises the
//    iteration's flag, and the lambda and dp passes of that iteration fall
//    back to plain walks for the particles next to a moved one.
#include <hip/hip_cooperative_groups.h>

#include "pbf_mlp_device.h"

namespace pbf {

constexpr int kBlock = 256;
constexpr int kBatch = 8;
static inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ f3 xyz(float4 t) { return {t.x, t.y, t.z}; }

// Entries k0 .. k0+7 of particle p's list (tiled, list_at): for each entry,
// the 64 lanes of a wavefront read one contiguous segment.  Entries at or
// past cnt are replaced by p, which keeps every gather in bounds.
__device__ __forceinline__ void load_batch(const int32_t* __restrict__ list, uint32_t stride, int k0, int cnt,
                                           uint32_t p, uint32_t (&q)[kBatch]) {
#pragma unroll
  for (int u = 0; u < kBatch; ++u) q[u] = (k0 + u < cnt) ? (uint32_t)list[list_at(k0 + u, p, stride)] : p;
}

// The pass that builds the compact lists (the first XSPH pass) runs one
// wavefront per block and stages each lane's entries in its own LDS column
// (entry k of the 64 lanes contiguous): its second loop reads them from LDS,
// and the global copy is written at the end with entry k of the 64 lanes in
// one contiguous store.
constexpr int kStageBlock = 64;

__device__ __forceinline__ void load_batch_lds(const int32_t* __restrict__ s_list, uint32_t lane, int k0, int cnt,
                                               uint32_t p, uint32_t (&q)[kBatch]) {
#pragma unroll
  for (int u = 0; u < kBatch; ++u) q[u] = (k0 + u < cnt) ? (uint32_t)s_list[(k0 + u) * kStageBlock + lane] : p;
}

__device__ __forceinline__ void flush_compact(const int32_t* __restrict__ s_list, uint32_t lane, int nc,
                                              int32_t* __restrict__ cnb, uint32_t stride, uint32_t p) {
  for (int k = 0; k < nc; ++k) cnb[list_at(k, p, stride)] = s_list[k * kStageBlock + lane];
}

// Skin lists (built with the lists, pbf_grid.hip): the entries within
// 1.1 h of the positions the lists were built from (x*_0), in list order,
// each slot packed with its place in the full list (slot | place << 25).  A
// pass may walk the skin list instead of the full one while no particle has
// moved 0.045 h or more from x*_0: an entry outside the skin list was at
// least 1.1 h away, so it is still farther than 1.01 h, well past h and any
// rounding of the within-h test, and adds nothing to any sum.  The entries
// within h come in the same order, so the sums are the same bit for bit.
// Every position a pressure update writes is checked against that move; a
// larger one (or a NaN position) sets a flag: C_DISP by iteration parity
// (read by the next iteration's passes) and C_DISP_ALL for the rest of the
// step.  The flags are written only when a move is too large, so a step at
// rest issues no atomics.
constexpr uint32_t kSkinSlotBits = 25;
constexpr uint32_t kSkinSlotMask = (1u << kSkinSlotBits) - 1u;

__device__ __forceinline__ bool skin_usable(const unsigned long long* __restrict__ counters, int flag) {
  return counters[C_SKIN] != 0ull && counters[flag] == 0ull;
}

// the squared move of a new position star from x*_0
__device__ __forceinline__ float move2(f3 nw, float4 x0) { return magsq3(sub3(nw, xyz(x0))); }

// Sets counters[a] and counters[C_DISP_ALL] when a lane's squared move m is
// not below the limit (NaN included); every lane of the wavefront calls it.
__device__ __forceinline__ void record_move(float m, unsigned long long* __restrict__ counters, int a,
                                            const KParams& K) {
  const uint64_t over = __ballot(!(m < K.disp_lim2));
  if (over != 0ull && (threadIdx.x & 63) == 0) {
    if (counters[a] == 0ull) atomicExch(&counters[a], 1ull);
    if (counters[C_DISP_ALL] == 0ull) atomicExch(&counters[C_DISP_ALL], 1ull);
  }
}

// x limited to [lo, hi] (NaN gives lo)
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return x > lo ? (x < hi ? x : hi) : lo; }

// ---------------------------------------------------------------------------
// compute_densities (particles.h:292-313) + compute_pressure_lambdas (:318-342)
// xs -> xt (x* with lambda in w), rho, lam, and the particle's within-h mask
// (bit k of wmask[p], of wmask[n + p] for k >= 64: list entry k lies within h
// and is in the sums; not read for a particle the density fix moved).  With FEAT, also the learned warm start's features of
// the particle (FeatCtl).
// Returns the particle's density error max(rho/rho0 - 1, 0).
template <bool FEAT, bool FAST, bool SR>
__device__ __forceinline__ float density_lambda_one(uint32_t p, uint32_t n, int it, const KParams& K,
                                                    const float4* __restrict__ xs, const float4* __restrict__ v,
                                                    float4* __restrict__ xt, float* __restrict__ rho,
                                                    float* __restrict__ lam, const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ nbr_count,
                                                    uint64_t* __restrict__ wmask, const FeatCtl& F,
                                                    uint32_t& hacks_out, float (&fo)[8],
                                                    const int32_t* __restrict__ snb,
                                                    const int32_t* __restrict__ snb_count, bool skin) {
  f3 xp = ld3(xs, p);
  // the walk goes over the skin list when it may (skin lists, above); the
  // plain walk after a density fix always takes the full list
  const int32_t* __restrict__ walk = skin ? snb : nbr;
  const int cnt = skin ? snb_count[p] : nbr_count[p];
  float density = K.mass_W0;
  float sum_C = 0.0f;
  f3 sg = {0.f, 0.f, 0.f};
  uint32_t hacks = 0;
  float dy_sum = 0.0f, div = 0.0f;
  const f3 vp = FEAT ? ld3(v, p) : f3{0.f, 0.f, 0.f};
  // one walk over the list: the entries within h go into the sums, in list
  // order, and into the mask.  A near-coincident entry ends the walk after
  // its batch, and the plain walk below redoes the particle; the mask of such
  // a particle is not read (the pressure update walks a moved particle's
  // full list with the test).
  uint64_t m0 = 0, m1 = 0;
  bool hack = false;
  const uint32_t qmask = skin ? kSkinSlotMask : ~0u;  // the slot of a walked entry
  uint32_t qn[kBatch];  // skin entries keep their place in the full list in the top bits
  load_batch(walk, K.stride, 0, cnt, p, qn);
  for (int k0 = 0; k0 < cnt && !hack; k0 += kBatch) {
    uint32_t q[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) q[u] = qn[u];
    float4 t[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) t[u] = gather4(xs, q[u] & qmask);
    load_batch(walk, K.stride, k0 + kBatch, cnt, p, qn);  // the next batch's entries load under this one's work
    uint32_t bits = 0;
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const f3 d = sub3(xp, xyz(t[u]));
      const float d2 = magsq3(d);
      const bool w = (k0 + u < cnt) & within_h(d2, K);
      const bool hk = w & (d2 < K.T_hack);  // particles.h:304: handled by the plain walk below
      const bool acc = w & !hk;
      bits |= (uint32_t)acc << u;
      hack |= hk;
      if (acc) {
        const float r = nb_sqrt<FAST, SR>(d2);
        density += K.mass * poly6_in(r, K);
        const f3 grad = smul3(K.neg_inv_rho0, nb_spiky<FAST, SR>(d, r, K));
        dec3(sg, grad);
        sum_C += magsq3(grad);
        if (FEAT) {  // (feature passes walk the full lists)
          dy_sum -= d.y;
          const f3 dv = sub3(vp, ld3(v, q[u]));
          div -= (dv.x * grad.x + dv.y * grad.y) + dv.z * grad.z;
        }
      }
    }
    // k0 is a multiple of kBatch = 8: a batch never straddles the two words
    if (k0 < 64)
      m0 |= (uint64_t)bits << k0;
    else
      m1 |= (uint64_t)bits << (k0 - 64);
  }
  wmask[p] = m0;
  if (K.cap > 64) wmask[n + p] = m1;
  int within = __popcll(m0) + __popcll(m1);
  if (hack) {
    // particles.h:304-305: a plain walk over the full list from its first
    // entry, in list order, moves x*_p at each near-coincident entry
    density = K.mass_W0;
    sum_C = 0.0f;
    sg = {0.f, 0.f, 0.f};
    dy_sum = 0.0f;
    div = 0.0f;
    within = 0;
    const int full = nbr_count[p];
    for (int k0 = 0; k0 < full; k0 += kBatch) {
      // a batch's positions load together; the entries are taken in order
      uint32_t q[kBatch];
      load_batch(nbr, K.stride, k0, full, p, q);
      f3 t[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) t[u] = ld3(xs, q[u]);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        f3 d = sub3(xp, t[u]);
        float d2 = magsq3(d);
        if (k0 + u >= full || !within_h(d2, K)) continue;
        if (d2 < K.T_hack) {
          dec3(xp, div3(mul3(ld3(v, p), K.dt), 50.0f));
          ++hacks;
          d = sub3(xp, t[u]);
          d2 = magsq3(d);
          if (!within_h(d2, K)) continue;
        }
        const float r = nb_sqrt<FAST>(d2);
        density += K.mass * poly6_in(r, K);
        const f3 grad = smul3(K.neg_inv_rho0, nb_spiky<FAST>(d, r, K));
        dec3(sg, grad);
        sum_C += magsq3(grad);
        if (FEAT) {
          ++within;
          dy_sum -= d.y;
          const f3 dv = sub3(vp, ld3(v, q[u]));
          div -= (dv.x * grad.x + dv.y * grad.y) + dv.z * grad.z;
        }
      }
    }
  }
  rho[p] = density;
  sum_C += magsq3(sg);
  const float constraint = density / K.rho0 - 1.0f;
  const float C_p = ((double)constraint < 0.0) ? 0.0f : constraint;
  const float lp = -C_p / (sum_C + K.eps);
  lam[p] = lp;
  xt[p] = make_float4(xp.x, xp.y, xp.z, lp);  // the dp pass gathers x* and lambda in one load
  hacks_out = hacks;
  if (FEAT) {
    const float inv_s = 1.0f / kWarmLambdaScale;
    const float dyn = dy_sum / (float)(within > 0 ? within : 1);
    fo[0] = clampf(constraint, -1.0f, 4.0f);
    fo[1] = clampf(lp * inv_s, -64.0f, 0.0f);
    fo[2] = clampf(F.lprev[p] * inv_s, -64.0f, 0.0f);
    fo[3] = clampf(sum_C * 1e-5f, 0.0f, 8.0f);
    fo[4] = clampf((float)within / 32.0f, 0.0f, 4.0f);
    fo[5] = clampf(vp.y * K.dt / K.h, -4.0f, 4.0f);
    fo[6] = clampf(dyn / K.h, -2.0f, 2.0f);
    fo[7] = clampf(div * K.mass * K.dt, -4.0f, 4.0f);
  }
  return constraint > 0.0f ? constraint : 0.0f;
}

// Max and sum over the wavefront; every lane of the wavefront must call them.
__device__ __forceinline__ float wave_max(float e) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) e = fmaxf(e, __shfl_xor(e, d, 64));
  return e;
}
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long e) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) e += __shfl_xor(e, d, 64);
  return e;
}
// a particle's error in 2^-32 units (errors above 2^31 saturate)
__device__ __forceinline__ unsigned long long fixed_error(float e) {
  return (unsigned long long)(fminf(e, 2147483648.0f) * 4294967296.0f);
}

// One block = one wavefront.  With S.ehist, the block's largest density
// error goes into slot eslot (atomicMax on the bits of a non-negative float
// orders like the float) and its fixed-point sum into S.esum[eslot].  With
// S.tol >= 0 nothing runs once the solve is done.
// The learned warm start for one wavefront's 64 particles (every lane of the
// wavefront calls it; fo holds the lane's features, zero past n): the
// features go through LDS into the B operands of the network, which runs on
// four tiles of 16 particles, and wlam[q] = lprev[q] + scale * net(q).  The
// same operands and MFMA sequence as k_mlp_forward, so the same results.
__device__ __forceinline__ void warm_network(const FeatCtl& F, uint32_t wave_first, uint32_t n, const float (&fo)[8],
                                             float* __restrict__ s_feat) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  float4* row = reinterpret_cast<float4*>(s_feat + l * 8);
  row[0] = make_float4(fo[0], fo[1], fo[2], fo[3]);
  row[1] = make_float4(fo[4], fo[5], fo[6], fo[7]);
  __syncthreads();
  const MlpW W{F.w1, F.w2, nullptr, F.mp};
  v4s a1[4], a2[4][4];
  load_forward_operands(W, c, g, a1, a2);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    v4s xb = {0, 0, 0, 0};
    if (g < 2) {
      const float4 f4 = *reinterpret_cast<const float4*>(s_feat + (16 * t + c) * 8 + 4 * g);
      xb = v4s{f2bf(f4.x), f2bf(f4.y), f2bf(f4.z), f2bf(f4.w)};
    }
    v4f z1[4], h1[4], z2[4], h2[4];
    const float y = mlp_forward(W, a1, a2, xb, g, z1, h1, z2, h2);
    const uint32_t q = wave_first + 16u * (uint32_t)t + (uint32_t)c;
    if (g == 0 && q < n) F.wlam[q] = F.lprev[q] + F.scale * y;
  }
}

// FM = 0: density and lambda only; 1: also the warm start's features (stored
// when F.feat is set); 2: also the network on them (F.wlam).  FAST: the
// PBF_OPT_FAST_MATH arithmetic.
// The pass bodies below take their block of particles (vb) and whether the
// calling thread is the one that clears the pass's flags (first): the same
// bodies run as their own kernels (vb from xcd_block()) and inside the
// persistent Jacobi kernel (vb from the calling block's share).
template <int FM, bool FAST, bool SR>
__device__ __forceinline__ void density_pass(uint32_t vb, bool first, uint32_t n, int it, int eslot,
                                             const SolveCtl& S, const FeatCtl& F,
                                             const KParams& K, const float4* __restrict__ xs,
                                             const float4* __restrict__ v, float4* __restrict__ xt,
                                             float* __restrict__ rho, float* __restrict__ lam,
                                             const int32_t* __restrict__ nbr,
                                             const int32_t* __restrict__ nbr_count,
                                             uint64_t* __restrict__ wmask,
                                             uint32_t* __restrict__ moved, uint32_t stamp,
                                             uint32_t* __restrict__ mflag,
                                             unsigned long long* __restrict__ counters,
                                             const int32_t* __restrict__ snb,
                                             const int32_t* __restrict__ snb_count) {
  if (S.tol >= 0.0f && counters[C_DONE] != 0ull) return;
  const uint32_t p = vb * blockDim.x + threadIdx.x;
  // this iteration's pressure update records its moves in the other slot
  const bool skin = FM == 0 && skin_usable(counters, C_DISP + (it & 1));
  if (first) counters[C_DISP + ((it + 1) & 1)] = 0ull;
  float e = 0.0f;
  uint32_t hacks = 0;
  float fo[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (p < n) e = density_lambda_one<FM != 0, FAST, SR>(p, n, it, K, xs, v, xt, rho, lam, nbr, nbr_count, wmask, F, hacks, fo, snb, snb_count, skin);
  if (FM != 0 && F.feat && p < n) {
    float4* o = reinterpret_cast<float4*>(F.feat + (uint64_t)p * 8);
    o[0] = make_float4(fo[0], fo[1], fo[2], fo[3]);
    o[1] = make_float4(fo[4], fo[5], fo[6], fo[7]);
  }
  if constexpr (FM == 2) {
    __shared__ __attribute__((aligned(16))) float s_feat[kBlock * 8];
    warm_network(F, p - (threadIdx.x & 63u), n, fo, s_feat + (threadIdx.x & ~63u) * 8);
  }
  // particles moved by the density fix: one set of counter updates per
  // wavefront (every lane of the wavefront is here)
  const uint64_t hm = __ballot(hacks != 0);
  if (hm) {
    const uint32_t wl = threadIdx.x & 63;
    const unsigned long long hsum = wave_sum((unsigned long long)hacks);
    unsigned long long base = 0;
    if (wl == 0) {
      base = atomicAdd(&counters[C_MOVED_N + (it & 1)], (unsigned long long)__popcll(hm));
      atomicAdd(&counters[C_HACKS], hsum);
      atomicExch(&counters[C_HACK_FLAG + (it & 1)], 1ull);
    }
    base = __shfl(base, 0, 64);
    if (hacks) {
      moved[base + (unsigned long long)__popcll(hm & ((1ull << wl) - 1ull))] = p;
      mflag[p] = stamp;
    }
  }
  if (S.ehist) {
    const unsigned long long es = wave_sum(fixed_error(e));
    e = wave_max(e);
    if ((threadIdx.x & 63) == 0) {
      if (e > 0.0f) atomicMax(&S.ehist[eslot], __float_as_uint(e));
      if (es) atomicAdd(&S.esum[eslot], es);
    }
  }
}

// (the exact and fast instances of the plain pass fit 64 registers: 8 wavefronts per SIMD)
template <int FM, bool FAST, bool SR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu((FM == 0 && (SR || FAST)) ? 8 : 1)))
k_density_lambda_c(uint32_t n, int it, int eslot, SolveCtl S, FeatCtl F,
                   KParams K, const float4* __restrict__ xs,
                   const float4* __restrict__ v, float4* __restrict__ xt,
                   float* __restrict__ rho, float* __restrict__ lam,
                   const int32_t* __restrict__ nbr,
                   const int32_t* __restrict__ nbr_count,
                   uint64_t* __restrict__ wmask,
                   uint32_t* __restrict__ moved, uint32_t stamp,
                   uint32_t* __restrict__ mflag,
                   unsigned long long* __restrict__ counters,
                   const int32_t* __restrict__ snb,
                   const int32_t* __restrict__ snb_count) {
  density_pass<FM, FAST, SR>(xcd_block(), blockIdx.x == 0 && threadIdx.x == 0, n, it, eslot, S, F, K, xs, v, xt, rho,
                         lam, nbr, nbr_count, wmask, moved, stamp, mflag, counters, snb, snb_count);
}

// A particle is dirty when it or a particle in its list was moved by the
// density fix: its fused lambda and its within-h mask may be stale.  A list
// walk visits cells computed from floating-point coordinates (x + i*s,
// neighbors.h:177-179), which lie within two cells of the particle's own cell
// on every axis, so every particle that can list q sits within two cells of
// q's cell: all particles of those 5x5x5 cells are marked, one wavefront per
// moved particle (lanes 0..24 find the 25 rows of cells, then the wavefront
// stamps each row's slot range).  Many moved particles, or a crowded
// neighbourhood, leave the work to lambda_repair's exact scan instead.
constexpr uint32_t kMaxMovedMarked = 4096;
constexpr uint32_t kMaxMarkedPerMoved = 8192;
constexpr uint32_t kMarkBlocks = 64;

// (wave, nwaves: the calling wavefront among all wavefronts that share the pass)
__device__ __forceinline__ void mark_pass(uint32_t wave, uint32_t nwaves, bool first, int it, uint32_t stamp,
                                          const SolveCtl& S, const KParams& K, const uint32_t* __restrict__ key,
                                          const uint32_t* __restrict__ start, const uint32_t* __restrict__ moved,
                                          uint32_t* __restrict__ dirty, unsigned long long* __restrict__ counters) {
  if (S.tol >= 0.0f && counters[C_DONE] != 0ull) return;
  const uint32_t nmoved = (uint32_t)counters[C_MOVED_N + (it & 1)];
  if (nmoved > kMaxMovedMarked) {
    if (first) counters[C_MOVED_ALL + (it & 1)] = 1ull;
    return;
  }
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t i = wave; i < nmoved; i += nwaves) {
    const uint32_t q = moved[i];
    const uint32_t kq = key[q];
    if (kq >= K.ncells) {  // outside the grid: no cell to start from
      if (lane == 0) counters[C_MOVED_ALL + (it & 1)] = 1ull;
      continue;
    }
    // cell coordinates in the moved particle's own grid (its scene's)
    const GridView G = grid_of(K, q);
    const uint32_t lq = kq - G.off;
    const int cz = (int)(lq / G.dxy), cy = (int)((lq % G.dxy) / G.dimx), cx = (int)(lq % G.dimx);
    uint32_t rb = 0, re = 0;
    if (lane < 25) {
      const int z = cz - 2 + (int)lane / 5, y = cy - 2 + (int)lane % 5;
      if (z >= 0 && z < (int)G.dimz && y >= 0 && y < (int)G.dimy) {
        const uint64_t row = G.off + (uint64_t)z * G.dxy + (uint64_t)y * G.dimx;
        rb = start[row + max(cx - 2, 0)];
        re = start[row + min(cx + 2, (int)G.dimx - 1) + 1];
      }
    }
    uint32_t total = re - rb;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) total += __shfl_xor(total, d, 64);
    if (total > kMaxMarkedPerMoved) {  // crowded neighbourhood: leave it to the exact scan
      if (lane == 0) counters[C_MOVED_ALL + (it & 1)] = 1ull;
      continue;
    }
    for (int r = 0; r < 25; ++r) {
      const uint32_t b = __shfl(rb, r, 64), e = __shfl(re, r, 64);
      for (uint32_t t = b + lane; t < e; t += 64) dirty[t] = stamp;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_mark_dirty(int it, uint32_t stamp, SolveCtl S, KParams K,
                                                       const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ moved,
                                                       uint32_t* __restrict__ dirty,
                                                       unsigned long long* __restrict__ counters) {
  mark_pass((blockIdx.x * blockDim.x + threadIdx.x) >> 6, (gridDim.x * blockDim.x) >> 6,
            blockIdx.x == 0 && threadIdx.x == 0, it, stamp, S, K, key, start, moved, dirty, counters);
}

// compute_pressure_lambdas (particles.h:318-342) from xt for the dirty
// particles of this iteration (plain walks).  The new lambda goes to lam and
// to the particle's xt.w; positions are read as three words, so no lane reads
// the w word that another lane writes.  When k_mark_dirty left the marking
// (many moved particles or a crowded neighbourhood), a particle is dirty
// exactly when it or an entry of its list carries this iteration's moved
// stamp; this pass finds those and stamps them in dirty for the dp pass.
__device__ __forceinline__ f3 ld3_words(const float4* a, uint32_t i) {
  const float* f = reinterpret_cast<const float*>(a + i);
  return {f[0], f[1], f[2]};
}

template <bool FAST>
__device__ __forceinline__ void repair_pass(uint32_t vb, uint32_t n, int it, uint32_t stamp, const SolveCtl& S,
                                            const KParams& K, float4* __restrict__ xt,
                                            const float* __restrict__ rho, float* __restrict__ lam,
                                            const int32_t* __restrict__ nbr,
                                            const int32_t* __restrict__ nbr_count,
                                            const uint32_t* __restrict__ mflag,
                                            uint32_t* __restrict__ dirty,
                                            const unsigned long long* __restrict__ counters) {
  if (S.tol >= 0.0f && counters[C_DONE] != 0ull) return;
  if (counters[C_HACK_FLAG + (it & 1)] == 0ull) return;
  const bool scan = counters[C_MOVED_ALL + (it & 1)] != 0ull;
  const uint32_t p = vb * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int cnt = nbr_count[p];
  if (scan) {
    bool hit = mflag[p] == stamp;
    for (int k0 = 0; k0 < cnt && !hit; k0 += kBatch) {
      uint32_t q[kBatch];
      load_batch(nbr, K.stride, k0, cnt, p, q);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) hit |= (k0 + u < cnt) & (mflag[q[u]] == stamp);
    }
    if (!hit) return;
    dirty[p] = stamp;
  } else if (dirty[p] != stamp) {
    return;
  }
  const f3 xp = ld3_words(xt, p);
  float sum_C = 0.0f;
  f3 sg = {0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    uint32_t q[kBatch];
    load_batch(nbr, K.stride, k0, cnt, p, q);
    f3 t[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) t[u] = ld3_words(xt, q[u]);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const f3 d = sub3(xp, t[u]);
      const float d2 = magsq3(d);
      if (k0 + u >= cnt || !within_h(d2, K)) continue;
      const float r = nb_sqrt<FAST>(d2);
      const f3 grad = smul3(K.neg_inv_rho0, nb_spiky<FAST>(d, r, K));
      dec3(sg, grad);
      sum_C += magsq3(grad);
    }
  }
  sum_C += magsq3(sg);
  const float constraint = rho[p] / K.rho0 - 1.0f;
  const float C_p = ((double)constraint < 0.0) ? 0.0f : constraint;
  const float lp = -C_p / (sum_C + K.eps);
  lam[p] = lp;
  reinterpret_cast<float*>(xt + p)[3] = lp;
}

template <bool FAST>
__global__ void __launch_bounds__(kBlock) k_lambda_repair(uint32_t n, int it, uint32_t stamp, SolveCtl S,
                                                          KParams K, float4* __restrict__ xt,
                                                          const float* __restrict__ rho, float* __restrict__ lam,
                                                          const int32_t* __restrict__ nbr,
                                                          const int32_t* __restrict__ nbr_count,
                                                          const uint32_t* __restrict__ mflag,
                                                          uint32_t* __restrict__ dirty,
                                                          const unsigned long long* __restrict__ counters) {
  repair_pass<FAST>(xcd_block(), n, it, stamp, S, K, xt, rho, lam, nbr, nbr_count, mflag, dirty, counters);
}

__device__ __forceinline__ float mean_error(unsigned long long sum, uint32_t n) {
  return (float)((double)sum / 4294967296.0 / (double)n);
}

// The pressure update's sum over one particle's list (particles.h:355-366):
// the entries under the within-h mask, or with plain every entry within h.
template <bool FAST, bool SR>
__device__ __forceinline__ f3 dp_walk(const int32_t* __restrict__ walk, int cnt, bool skin, bool plain, uint64_t m0,
                                      uint64_t m1, uint32_t p, f3 xp, float lp, const KParams& K,
                                      const float4* __restrict__ xt) {
  f3 dp = {0.f, 0.f, 0.f};
  const uint32_t qmask = skin ? kSkinSlotMask : ~0u;  // the slot of a walked entry
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    uint32_t q[kBatch];
    load_batch(walk, K.stride, k0, cnt, p, q);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) q[u] &= qmask;
    const uint32_t bits = (uint32_t)((k0 < 64 ? m0 >> k0 : m1 >> (k0 - 64)) & 0xFFu);
    float4 t[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) t[u] = gather4(xt, (bits >> u) & 1u ? q[u] : p);  // masked entries read p itself
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const f3 d = sub3(xp, xyz(t[u]));
      const float d2 = magsq3(d);
      if (k0 + u >= cnt || !((bits >> u) & 1u) || (plain && !within_h(d2, K))) continue;
      const float r = nb_sqrt<FAST, SR>(d2);
      acc3(dp, smul3(lp + t[u].w, nb_spiky<FAST, SR>(d, r, K)));
    }
  }
  return dp;
}

// compute_pressure_dps (particles.h:348-358) + update_position_stars
// (:363-370); on the last iteration also update_velocities (:375-385).
// In an iterate-to-tolerance solve, slot `it` first checks the error its
// density pass measured: a rejected warm start (slot 0 error above the error
// before the warm start; with the retry, also slot 1's after the damped
// start) restores xsave; an error <= tol ends the solve with x* = xt.
// Neither applies an update.
template <bool FAST, bool SR>
__device__ __forceinline__ void dp_pass(uint32_t vb, bool first, uint32_t n, int it, int last, uint32_t stamp,
                                        const SolveCtl& S,
                                        const KParams& K, const float4* __restrict__ xt,
                                        const float4* __restrict__ x, float4* __restrict__ xs,
                                        float4* __restrict__ v, const int32_t* __restrict__ nbr,
                                        const int32_t* __restrict__ nbr_count,
                                        const uint64_t* __restrict__ wmask,
                                        const uint32_t* __restrict__ dirty,
                                        unsigned long long* __restrict__ counters,
                                        const int32_t* __restrict__ snb,
                                        const int32_t* __restrict__ snb_count,
                                        const float4* __restrict__ xs0,
                                        const float* __restrict__ rho) {
  // C_DONE holds the slot that ended the solve, plus one.  Only a solve that
  // ended at an earlier slot skips this launch: blocks of the ending slot that
  // start after block 0 has written the flag must still copy xt to x*.
  if (S.tol >= 0.0f) {
    const unsigned long long done = counters[C_DONE];
    if (done != 0ull && done <= (unsigned long long)it) return;
  }
  // clear the next iteration's flags (nobody reads them until the next density pass)
  if (first) {
    counters[C_HACK_FLAG + ((it + 1) & 1)] = 0ull;
    counters[C_MOVED_N + ((it + 1) & 1)] = 0ull;
    counters[C_MOVED_ALL + ((it + 1) & 1)] = 0ull;
  }
  const uint32_t p = vb * blockDim.x + threadIdx.x;
  // what this slot does, the same for every thread: 0 the update, 1 the
  // rejected warm start's restore, 2 the end of the solve (x* = xt)
  int mode = 0;
  if (S.ehist) {
    const float e = S.mean ? mean_error(S.esum[it + 1], n) : __uint_as_float(S.ehist[it + 1]);
    const float e_pre = S.mean ? mean_error(S.esum[0], n) : __uint_as_float(S.ehist[0]);
    bool grew = false;
    // slot 0 checks the warm start; with the retry, slot 1 checks the damped one
    if (S.warm && (it == 0 || (it == 1 && S.retry && counters[C_RETRY] != 0ull))) {
      grew = e > e_pre;
      if (S.strict) {  // the other error measure must not grow either
        const float o = S.mean ? __uint_as_float(S.ehist[1]) : mean_error(S.esum[1], n);
        const float o_pre = S.mean ? __uint_as_float(S.ehist[0]) : mean_error(S.esum[0], n);
        grew = grew || o > o_pre;
      }
    }
    if (grew) {
      mode = 1;
      if (first) {
        if (S.retry && it == 0) {  // x* goes back, and k_warm_retry applies the damped start
          counters[C_RETRY] = 1ull;
          counters[C_WARM_KEPT] = 2ull;
        } else {
          counters[C_WARM_KEPT] = 0ull;
        }
      }
    } else if (S.tol >= 0.0f && e <= S.tol) {
      mode = 2;
      if (first) {
        counters[C_DONE] = (unsigned long long)(it + 1);
        counters[C_ITERS] = (unsigned long long)(it + 1);
      }
    }
  }
  const bool track = K.skin_mode && n < (1u << kSkinSlotBits);  // x*_0 is kept: check the moves
  float mv = 0.0f;
  if (p < n && mode == 1) {
    const float4 sv = S.xsave[p];
    xs[p] = sv;
    if (track) mv = move2(xyz(sv), xs0[p]);
  } else if (p < n && mode == 2) {
    const float4 t = xt[p];
    xs[p] = t;
    if (track) mv = move2(xyz(t), xs0[p]);
  } else if (p < n) {
    // a dirty particle (moved by the density fix, or next to one) walks its
    // whole list with the test; the others walk the list the density pass
    // walked (the skin list when it could) under its within-h mask
    const bool plain = counters[C_HACK_FLAG + (it & 1)] != 0ull && dirty[p] == stamp;
    const bool skin = !plain && skin_usable(counters, C_DISP + (it & 1));
    const int32_t* __restrict__ walk = skin ? snb : nbr;
    // x* and lambda of a particle come in one load (xt.w holds lambda)
    const float4 xp4 = xt[p];
    const f3 xp = xyz(xp4);
    const float lp = xp4.w;
    const int cnt = skin ? snb_count[p] : nbr_count[p];
    const uint64_t m0 = plain ? ~0ull : wmask[p];
    const uint64_t m1 = (plain || K.cap <= 64) ? ~0ull : wmask[n + p];
    // with no dirty lane in the wavefront every pair the walk takes is one
    // the density pass accepted at the same positions (static ranges)
    const f3 dp = (SR && __ballot(plain) == 0ull)
                      ? dp_walk<FAST, SR>(walk, cnt, skin, false, m0, m1, p, xp, lp, K, xt)
                      : dp_walk<FAST, false>(walk, cnt, skin, plain, m0, m1, p, xp, lp, K, xt);
    f3 nw = add3(xp, smul3(K.inv_rho0, dp));
    boundary(nw, K, p);
    st3(xs, p, nw);
    if (track) mv = move2(nw, xs0[p]);
    if (S.lsum) S.lsum[p] += lp;
    if (last) {  // the density rides in v.w for the XSPH passes' gathers
      f3 vel = div3(sub3(nw, ld3(x, p)), K.dt);
      if ((double)magsq3(vel) < K.vthr) vel = {0.f, 0.f, 0.f};
      v[p] = make_float4(vel.x, vel.y, vel.z, rho[p]);
    }
  }
  if (track) record_move(mv, counters, C_DISP + ((it + 1) & 1), K);
}

template <bool FAST, bool SR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) k_dp_update_c(uint32_t n, int it, int last, uint32_t stamp, SolveCtl S,
                                                        KParams K, const float4* __restrict__ xt,
                                                        const float4* __restrict__ x, float4* __restrict__ xs,
                                                        float4* __restrict__ v, const int32_t* __restrict__ nbr,
                                                        const int32_t* __restrict__ nbr_count,
                                                        const uint64_t* __restrict__ wmask,
                                                        const uint32_t* __restrict__ dirty,
                                                        unsigned long long* __restrict__ counters,
                                                        const int32_t* __restrict__ snb,
                                                        const int32_t* __restrict__ snb_count,
                                                        const float4* __restrict__ xs0,
                                                        const float* __restrict__ rho) {
  dp_pass<FAST, SR>(xcd_block(), blockIdx.x == 0 && threadIdx.x == 0, n, it, last, stamp, S, K, xt, x, xs, v, nbr,
                nbr_count, wmask, dirty, counters, snb, snb_count, xs0, rho);
}

// Warm start of an iterate-to-tolerance solve: one compute_pressure_dps +
// update_position_stars pass (particles.h:348-370) with the given
// multipliers in place of lambda, over the full lists with the within-h test.
// xt holds the position stars after the first density pass; they are kept in
// xsave for the fallback.
template <bool FAST>
__global__ void __launch_bounds__(kBlock) k_warm_dp(uint32_t n, KParams K, const float4* __restrict__ xt,
                                                    const float* __restrict__ wlam,
                                                    const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ nbr_count, float4* __restrict__ xs,
                                                    float4* __restrict__ xsave,
                                                    unsigned long long* __restrict__ counters,
                                                    const float4* __restrict__ xs0) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  const bool track = K.skin_mode && n < (1u << kSkinSlotBits);
  float mv = 0.0f;
  if (p < n) {
    const float4 xp4 = xt[p];
    const f3 xp = xyz(xp4);
    const float lp = wlam[p];
    f3 dp = {0.f, 0.f, 0.f};
    const int cnt = nbr_count[p];
    for (int k0 = 0; k0 < cnt; k0 += kBatch) {
      uint32_t q[kBatch];
      load_batch(nbr, K.stride, k0, cnt, p, q);
      float4 t[kBatch];
      float lq[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        t[u] = xt[q[u]];
        lq[u] = wlam[q[u]];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const f3 d = sub3(xp, xyz(t[u]));
        const float d2 = magsq3(d);
        if (k0 + u >= cnt || !within_h(d2, K)) continue;
        const float r = nb_sqrt<FAST>(d2);
        acc3(dp, smul3(lp + lq[u], nb_spiky<FAST>(d, r, K)));
      }
    }
    f3 nw = add3(xp, smul3(K.inv_rho0, dp));
    boundary(nw, K, p);
    st3(xs, p, nw);
    xsave[p] = xp4;
    // both x* (kept) and xt (restored on a rejection) count as moves
    if (track) mv = fmaxf(move2(nw, xs0[p]), move2(xp, xs0[p]));
  }
  // slot 0 reads parity 0
  if (track) record_move(mv, counters, C_DISP + 0, K);
}

// After the Jacobi loop: the kept warm-start multipliers join the sum.
__global__ void k_add_warm(uint32_t n, const float* __restrict__ wlam, float* __restrict__ lsum,
                           const unsigned long long* __restrict__ counters) {
  const unsigned long long kept = counters[C_WARM_KEPT];
  if (kept == 0ull) return;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) lsum[p] += kept == 2ull ? kWarmRetryScale * wlam[p] : wlam[p];
}

// PBF_OPT_WARM_RETRY, after slot 0's pressure update: when that update put x*
// back because the warm start grew the error (C_RETRY), x* becomes the
// position stars before the warm start plus the update with the warm
// multipliers times kWarmRetryScale (k_warm_dp with scaled multipliers, from
// the saved positions); slot 1 then checks it against the same error.
template <bool FAST>
__global__ void __launch_bounds__(kBlock) k_warm_retry(uint32_t n, KParams K, const float4* __restrict__ xsave,
                                                       const float* __restrict__ wlam,
                                                       const int32_t* __restrict__ nbr,
                                                       const int32_t* __restrict__ nbr_count,
                                                       float4* __restrict__ xs,
                                                       unsigned long long* __restrict__ counters,
                                                       const float4* __restrict__ xs0) {
  if (counters[C_RETRY] == 0ull) return;
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  const bool track = K.skin_mode && n < (1u << kSkinSlotBits);
  float mv = 0.0f;
  if (p < n) {
    const f3 xp = ld3(xsave, p);
    const float lp = kWarmRetryScale * wlam[p];
    f3 dp = {0.f, 0.f, 0.f};
    const int cnt = nbr_count[p];
    for (int k0 = 0; k0 < cnt; k0 += kBatch) {
      uint32_t q[kBatch];
      load_batch(nbr, K.stride, k0, cnt, p, q);
      float4 t[kBatch];
      float lq[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        t[u] = gather4(xsave, q[u]);
        lq[u] = kWarmRetryScale * wlam[q[u]];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const f3 d = sub3(xp, xyz(t[u]));
        const float d2 = magsq3(d);
        if (k0 + u >= cnt || !within_h(d2, K)) continue;
        const float r = nb_sqrt<FAST>(d2);
        acc3(dp, smul3(lp + lq[u], nb_spiky<FAST>(d, r, K)));
      }
    }
    f3 nw = add3(xp, smul3(K.inv_rho0, dp));
    boundary(nw, K, p);
    st3(xs, p, nw);
    if (track) mv = move2(nw, xs0[p]);
  }
  // slot 1 reads parity 1
  if (track) record_move(mv, counters, C_DISP + 1, K);
}

// XSPH sum (particles.h:441-444) over a compact row, from global memory
// (FROM_LDS = false: cnb, tiled like the lists) or from the lane's LDS column
template <bool FROM_LDS>
__device__ __forceinline__ f3 xsph_sum(const int32_t* __restrict__ list, uint32_t n, uint32_t lane, int nc,
                                       uint32_t p, f3 xp, f3 vp, const KParams& K, const float4* __restrict__ xs,
                                       const float4* __restrict__ vin) {
  f3 dv = {0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < nc; k0 += kBatch) {
    uint32_t q[kBatch];
    if (FROM_LDS)
      load_batch_lds(list, lane, k0, nc, p, q);
    else
      load_batch(list, K.stride, k0, nc, p, q);
    float4 t[kBatch], vq[kBatch];  // vq.w: the neighbour's density
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      t[u] = gather4(xs, q[u]);
      vq[u] = gather4(vin, q[u]);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (k0 + u >= nc) continue;
      const float r = sqrt_cr(magsq3(sub3(xp, xyz(t[u]))));
      acc3(dv, div3(mul3(sub3(xyz(vq[u]), vp), poly6_in(r, K)), vq[u].w));
    }
  }
  return dv;
}

// ---------------------------------------------------------------------------
// apply_viscosity (particles.h:436-447), first call: walks the whole list and
// leaves the compact rows of the final positions for the passes that follow.
__global__ void __launch_bounds__(kStageBlock) k_viscosity_first(uint32_t n, KParams K, const float4* __restrict__ xs,
                                                            const float4* __restrict__ vin,
                                                            float4* __restrict__ vout,
                                                            const int32_t* __restrict__ nbr,
                                                            const int32_t* __restrict__ nbr_count,
                                                            int32_t* __restrict__ cnb,
                                                            int32_t* __restrict__ cnb_count,
                                                            const unsigned long long* __restrict__ counters,
                                                            const int32_t* __restrict__ snb,
                                                            const int32_t* __restrict__ snb_count,
                                                            unsigned long long* __restrict__ skin_hint) {
  extern __shared__ int32_t s_list[];  // [cap + 1][kStageBlock]
  const uint32_t lane = threadIdx.x;
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  // the next step builds skin lists only if this one's moves kept them usable to the end
  if (K.skin_mode && blockIdx.x == 0 && lane == 0) *skin_hint = counters[C_DISP_ALL] == 0ull ? 1ull : 0ull;
  if (p >= n) return;
  const f3 xp = ld3(xs, p);
  const f3 vp = ld3(vin, p);
  // the skin list holds every entry within h while no particle has moved
  // 0.045 h over the step (skin lists, above)
  const bool skin = skin_usable(counters, C_DISP_ALL);
  const int32_t* __restrict__ walk = skin ? snb : nbr;
  const int cnt = skin ? snb_count[p] : nbr_count[p];
  const uint32_t qmask = skin ? kSkinSlotMask : ~0u;  // the slot of a walked entry
  int nc = 0;
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    uint32_t q[kBatch];
    load_batch(walk, K.stride, k0, cnt, p, q);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) q[u] &= qmask;
    float4 t[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) t[u] = gather4(xs, q[u]);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const float d2 = magsq3(sub3(xp, xyz(t[u])));
      const bool acc = (k0 + u < cnt) & within_h(d2, K);
      s_list[nc * kStageBlock + lane] = (int32_t)q[u];  // a rejected entry is overwritten by the next accepted one
      nc += acc ? 1 : 0;
    }
  }
  cnb_count[p] = nc;
  flush_compact(s_list, lane, nc, cnb, K.stride, p);
  const f3 dv = xsph_sum<true>(s_list, n, lane, nc, p, xp, vp, K, xs, vin);
  f3 nv = vp;
  acc3(nv, smul3(K.visc, dv));
  vout[p] = make_float4(nv.x, nv.y, nv.z, vin[p].w);  // the density stays in w
}

// apply_viscosity, second call, fused with update_positions (particles.h:390-394)
__global__ void __launch_bounds__(kBlock) k_viscosity_last(uint32_t n, KParams K, const float4* __restrict__ xs,
                                                           const float4* __restrict__ vin, float4* __restrict__ vout,
                                                           float4* __restrict__ x, const int32_t* __restrict__ cnb,
                                                           const int32_t* __restrict__ cnb_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 xp = ld3(xs, p);
  const f3 vp = ld3(vin, p);
  const f3 dv = xsph_sum<false>(cnb, n, 0, cnb_count[p], p, xp, vp, K, xs, vin);
  f3 nv = vp;
  acc3(nv, smul3(K.visc, dv));
  st3(vout, p, nv);
  x[p] = make_float4(xp.x, xp.y, xp.z, 0.f);
}

// compute_vorticity (particles.h:452-466) -> scr (omega, and |omega| in w)
template <bool SR>
__global__ void __launch_bounds__(kBlock) k_vorticity_c(uint32_t n, KParams K, const float4* __restrict__ xs,
                                                        const float4* __restrict__ vin, float4* __restrict__ scr,
                                                        const int32_t* __restrict__ cnb,
                                                        const int32_t* __restrict__ cnb_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 xp = ld3(xs, p);
  const f3 vp = ld3(vin, p);
  const int cnt = cnb_count[p];
  f3 w = {0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    uint32_t q[kBatch];
    load_batch(cnb, K.stride, k0, cnt, p, q);
    float4 t[kBatch], vq[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      t[u] = gather4(xs, q[u]);
      vq[u] = gather4(vin, q[u]);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (k0 + u >= cnt) continue;
      const f3 d = sub3(xp, xyz(t[u]));
      const float r = sqrtf(magsq3(d));
      acc3(w, cross3(sub3(xyz(vq[u]), vp), nb_spiky<false, SR>(d, r, K)));
    }
  }
  scr[p] = make_float4(w.x, w.y, w.z, sqrtf(magsq3(w)));  // |omega| for apply_vorticity's gathers
}

// apply_vorticity (particles.h:471-487), v in place (own particle only)
template <bool SR>
__global__ void __launch_bounds__(kBlock) k_apply_vorticity_c(uint32_t n, KParams K, const float4* __restrict__ xs,
                                                              const float4* __restrict__ scr,
                                                              float4* __restrict__ v,
                                                              const int32_t* __restrict__ cnb,
                                                              const int32_t* __restrict__ cnb_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 xp = ld3(xs, p);
  const int cnt = cnb_count[p];
  f3 eta = {0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < cnt; k0 += kBatch) {
    uint32_t q[kBatch];
    load_batch(cnb, K.stride, k0, cnt, p, q);
    float4 t[kBatch], wq[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      t[u] = gather4(xs, q[u]);
      wq[u] = gather4(scr, q[u]);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (k0 + u >= cnt) continue;
      const f3 d = sub3(xp, xyz(t[u]));
      const float r = sqrtf(magsq3(d));
      const float wm = wq[u].w;  // |omega_q|, computed by the vorticity pass
      acc3(eta, smul3(wm, nb_spiky<false, SR>(d, r, K)));
    }
  }
  const f3 N = div3(eta, sqrtf(magsq3(eta)) + K.r_eps);
  const float4 v4 = v[p];
  f3 vp = xyz(v4);
  acc3(vp, mul3(mul3(cross3(N, ld3(scr, p)), K.vort), K.dt));
  v[p] = make_float4(vp.x, vp.y, vp.z, v4.w);  // the density stays in w for the second XSPH pass
}

// ---- launchers ----
#define PBF_LAUNCH(c, name, kern, ...)                                                                    \
  do {                                                                                                    \
    const uint32_t n_ = (uint32_t)(c)->n;                                                                 \
    if (!n_) return;                                                                                      \
    KernelTimer* t_ = timer_begin((c), name);                                                             \
    const uint32_t b_ = (c)->tune.block_size;                                                             \
    hipLaunchKernelGGL(kern, dim3((n_ + b_ - 1) / b_), dim3(b_), 0, (c)->stream, n_, __VA_ARGS__);         \
    timer_end((c), t_);                                                                                   \
  } while (0)

#define PBF_LAUNCH_STAGED(c, name, kern, ...)                                                              \
  do {                                                                                                      \
    const uint32_t n_ = (uint32_t)(c)->n;                                                                   \
    if (!n_) return;                                                                                        \
    KernelTimer* t_ = timer_begin((c), name);                                                               \
    hipLaunchKernelGGL(kern, dim3((n_ + kStageBlock - 1) / kStageBlock), dim3(kStageBlock),                 \
                       (uint32_t)(kStageBlock * ((c)->K.cap + 1) * sizeof(int32_t)), (c)->stream, n_, __VA_ARGS__); \
    timer_end((c), t_);                                                                                     \
  } while (0)

// the kernel template KT instantiated for the context's arithmetic (PBF_OPT_FAST_MATH)
#define PBF_LAUNCH_ARITH(launch, c, name, KT, ...)        \
  do {                                                    \
    if ((c)->K.fast)                                      \
      launch(c, name, (KT<true>), __VA_ARGS__);           \
    else                                                  \
      launch(c, name, (KT<false>), __VA_ARGS__);          \
  } while (0)
// the same with the static-range instance of the exact arithmetic (KParams::static_ranges)
#define PBF_LAUNCH_ARITH_SR(launch, c, name, KT, ...)     \
  do {                                                    \
    if ((c)->K.fast)                                      \
      launch(c, name, (KT<true, false>), __VA_ARGS__);    \
    else if ((c)->K.static_ranges)                        \
      launch(c, name, (KT<false, true>), __VA_ARGS__);    \
    else                                                  \
      launch(c, name, (KT<false, false>), __VA_ARGS__);   \
  } while (0)

template <int FM>
static void launch_density_fm(pbf_context* c, const char* name, int it, int eslot, const SolveCtl& S,
                              const FeatCtl& F) {
  if (c->K.fast)
    PBF_LAUNCH(c, name, (k_density_lambda_c<FM, true, false>), it, eslot, S, F, c->K, (const float4*)c->xs,
               (const float4*)c->v, c->xt, c->rho, c->lam, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count,
               c->wmask, c->moved, c->dirty_stamp, c->mflag, c->counters, (const int32_t*)c->snb,
               (const int32_t*)c->snb_count);
  else if (c->K.static_ranges)
    PBF_LAUNCH(c, name, (k_density_lambda_c<FM, false, true>), it, eslot, S, F, c->K, (const float4*)c->xs,
               (const float4*)c->v, c->xt, c->rho, c->lam, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count,
               c->wmask, c->moved, c->dirty_stamp, c->mflag, c->counters, (const int32_t*)c->snb,
               (const int32_t*)c->snb_count);
  else
    PBF_LAUNCH(c, name, (k_density_lambda_c<FM, false, false>), it, eslot, S, F, c->K, (const float4*)c->xs,
               (const float4*)c->v, c->xt, c->rho, c->lam, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count,
               c->wmask, c->moved, c->dirty_stamp, c->mflag, c->counters, (const int32_t*)c->snb,
               (const int32_t*)c->snb_count);
}

void launch_density_lambda_masked(pbf_context* c, int it, int eslot, const SolveCtl& S, const FeatCtl& F) {
  if (F.wlam)
    launch_density_fm<2>(c, "density_warm_network", it, eslot, S, F);
  else if (F.feat)
    launch_density_fm<1>(c, "density_features", it, eslot, S, F);
  else
    launch_density_fm<0>(c, "density_lambda", it, eslot, S, F);
}

void launch_mark_dirty(pbf_context* c, int it, const SolveCtl& S) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "mark_dirty");
  hipLaunchKernelGGL(k_mark_dirty, dim3(kMarkBlocks), dim3(kBlock), 0, c->stream, it, c->dirty_stamp, S, c->K,
                     (const uint32_t*)c->key, (const uint32_t*)c->cell_start, (const uint32_t*)c->moved, c->dirty,
                     c->counters);
  timer_end(c, t);
}

void launch_lambda_repair(pbf_context* c, int it, const SolveCtl& S) {
  PBF_LAUNCH_ARITH(PBF_LAUNCH, c, "lambda_repair", k_lambda_repair, it, c->dirty_stamp, S, c->K, c->xt,
             (const float*)c->rho, c->lam, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count,
             (const uint32_t*)c->mflag, c->dirty, (const unsigned long long*)c->counters);
}

void launch_dp_update_masked(pbf_context* c, int it, int last, const SolveCtl& S) {
  PBF_LAUNCH_ARITH_SR(PBF_LAUNCH, c, "dp_update", k_dp_update_c, it, last, c->dirty_stamp, S, c->K, (const float4*)c->xt,
             (const float4*)c->x, c->xs, c->v, (const int32_t*)c->nbr,
             (const int32_t*)c->nbr_count, (const uint64_t*)c->wmask, (const uint32_t*)c->dirty, c->counters,
             (const int32_t*)c->snb, (const int32_t*)c->snb_count, (const float4*)c->xs0, (const float*)c->rho);
}

void launch_warm_retry(pbf_context* c) {
  PBF_LAUNCH_ARITH(PBF_LAUNCH, c, "warm_retry", k_warm_retry, c->K, (const float4*)c->xsave,
                   (const float*)c->wlam, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count, c->xs, c->counters,
                   (const float4*)c->xs0);
}

void launch_warm_dp(pbf_context* c) {
  PBF_LAUNCH_ARITH(PBF_LAUNCH, c, "warm_dp", k_warm_dp, c->K, (const float4*)c->xt, (const float*)c->wlam, (const int32_t*)c->nbr,
             (const int32_t*)c->nbr_count, c->xs, c->xsave, c->counters, (const float4*)c->xs0);
}

// ---------------------------------------------------------------------------
// The Jacobi iterations of a step or the slots of a solve in one cooperative
// launch (PBF_OPT_PERSISTENT): each iteration runs the density, dirty-marking,
// lambda-repair and pressure-update bodies above over the whole particle
// range, separated by grid-wide barriers instead of kernel boundaries, and a
// solve leaves the loop on the device once its slot met the tolerance (every
// block reads the done flag after the same barrier).  Each block takes a
// contiguous share of the 256-particle blocks, its share placed on its XCD
// (xcd_block()).  The host launches it with one 256-particle block per block
// and at most one block per CU (persistent_fits), far inside what the chip
// holds at once, so every block is resident whatever else the device runs.
struct JacobiArgs {
  uint32_t n, iters;
  int solve, last_vel;  // solve: slot it's errors go to slot it + 1; last_vel: velocities in the last iteration
  uint32_t stamp0;      // stamp of the first iteration (then one more per iteration)
  SolveCtl S;
  KParams K;
  float4 *xs, *v, *xt;
  const float4 *x, *xs0;
  float *rho, *lam;
  const int32_t *nbr, *nbr_count, *snb, *snb_count;
  uint64_t* wmask;
  uint32_t *moved, *mflag, *dirty;
  const uint32_t *key, *start;
  unsigned long long* counters;
};

template <bool FAST, bool SR>
__global__ void __launch_bounds__(kBlock) k_jacobi_persistent(JacobiArgs A) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const uint32_t nvb = (A.n + blockDim.x - 1) / blockDim.x;
  const uint32_t per = (nvb + gridDim.x - 1) / gridDim.x;
  const uint32_t vb0 = xcd_block() * per, vb1 = min(nvb, vb0 + per);
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
  const FeatCtl F0{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0f, nullptr};
  for (uint32_t it = 0; it < A.iters; ++it) {
    if (A.S.tol >= 0.0f && __atomic_load_n(&A.counters[C_DONE], __ATOMIC_RELAXED) != 0ull) break;
    const uint32_t stamp = A.stamp0 + it;
    const int eslot = A.solve ? (int)it + 1 : 0;
    for (uint32_t vb = vb0; vb < vb1; ++vb)
      density_pass<0, FAST, SR>(vb, vb == 0 && threadIdx.x == 0, A.n, (int)it, eslot, A.S, F0, A.K, A.xs, A.v, A.xt,
                            A.rho, A.lam, A.nbr, A.nbr_count, A.wmask, A.moved, stamp, A.mflag, A.counters, A.snb,
                            A.snb_count);
    grid.sync();
    mark_pass(wave, nwaves, wave == 0 && threadIdx.x == 0, (int)it, stamp, A.S, A.K, A.key, A.start, A.moved,
              A.dirty, A.counters);
    grid.sync();
    for (uint32_t vb = vb0; vb < vb1; ++vb)
      repair_pass<FAST>(vb, A.n, (int)it, stamp, A.S, A.K, A.xt, A.rho, A.lam, A.nbr, A.nbr_count, A.mflag,
                        A.dirty, A.counters);
    grid.sync();
    const int last = A.last_vel && it + 1 == A.iters;
    for (uint32_t vb = vb0; vb < vb1; ++vb)
      dp_pass<FAST, SR>(vb, vb == 0 && threadIdx.x == 0, A.n, (int)it, last, stamp, A.S, A.K, A.xt, A.x, A.xs, A.v,
                    A.nbr, A.nbr_count, A.wmask, A.dirty, A.counters, A.snb, A.snb_count, A.xs0, A.rho);
    grid.sync();
  }
}

// Blocks of one cooperative launch of the persistent kernel on the context's
// device (0: cooperative launches unavailable), found once per context.
// the persistent kernel's instance for the context's arithmetic
static const void* persistent_kernel(const pbf_context* c) {
  if (c->K.fast) return (const void*)k_jacobi_persistent<true, false>;
  if (c->K.static_ranges) return (const void*)k_jacobi_persistent<false, true>;
  return (const void*)k_jacobi_persistent<false, false>;
}

static uint32_t persistent_blocks(pbf_context* c) {
  int& blocks = c->coop_blocks[c->K.fast ? 1 : (c->K.static_ranges ? 2 : 0)];
  if (blocks >= 0) return (uint32_t)blocks;
  blocks = 0;
  int coop = 0, cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) != hipSuccess || !coop)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) return 0;
  c->coop_cus = cus;
  const void* kern = persistent_kernel(c);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess) return 0;
  blocks = std::max(0, per_cu * cus);
  return (uint32_t)blocks;
}

uint32_t persistent_capacity(pbf_context* c) { return persistent_blocks(c); }

// The cooperative launch is used for at most one block per CU (particles up
// to 256 x CUs): a launch of the full co-resident capacity (5 blocks per CU at
// this kernel's registers) did not finish on the 1M-particle dam break.
bool persistent_fits(pbf_context* c) {
  const uint32_t cap = persistent_blocks(c);
  const uint64_t nvb = (c->n + kBlock - 1) / kBlock;
  return cap > 0 && nvb <= (uint64_t)std::min<uint32_t>((uint32_t)c->coop_cus, cap);
}

bool persistent_available(pbf_context* c) { return persistent_blocks(c) > 0; }

int launch_jacobi_persistent(pbf_context* c, uint32_t iters, bool solve, bool last_vel, const SolveCtl& S) {
  const uint32_t n = (uint32_t)c->n;
  if (!n || !iters) return PBF_OK;
  const uint32_t nvb = (n + kBlock - 1) / kBlock;
  const uint32_t blocks = std::min<uint32_t>(persistent_blocks(c), nvb);
  if (!blocks) return set_error(PBF_ERR_UNSUPPORTED, "cooperative launches are not available on this device");
  JacobiArgs A{n, iters, solve ? 1 : 0, last_vel ? 1 : 0, c->dirty_stamp + 1, S, c->K, c->xs, c->v, c->xt,
               (const float4*)c->x, (const float4*)c->xs0, c->rho, c->lam, (const int32_t*)c->nbr,
               (const int32_t*)c->nbr_count, (const int32_t*)c->snb, (const int32_t*)c->snb_count, c->wmask,
               c->moved, c->mflag, c->dirty, (const uint32_t*)c->key, (const uint32_t*)c->cell_start,
               c->counters};
  void* args[] = {&A};
  KernelTimer* t = timer_begin(c, "jacobi_persistent");
  const void* kern = persistent_kernel(c);
  c->coop_grid = blocks;
  const hipError_t e = hipLaunchCooperativeKernel(kern, dim3(blocks), dim3(kBlock), args, 0, c->stream);
  timer_end(c, t);
  c->dirty_stamp += iters;
  return check_hip(e, "jacobi_persistent");
}

// The record of the solve that just ran (pbf_get_solve_log): slots used,
// whether the warm start was kept, the last used slot's errors and the errors
// before the warm start.  One thread: the ring index advances on the device,
// so graph replays record every solve.
__global__ void k_log_solve(uint32_t n, const unsigned long long* __restrict__ counters,
                            const uint32_t* __restrict__ ehist, const unsigned long long* __restrict__ esum,
                            pbf_solve_record* __restrict__ log, unsigned long long* __restrict__ count) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const unsigned long long idx = *count;
  const uint32_t k = (uint32_t)counters[C_ITERS];  // 1..kMaxSolveSlots
  const uint32_t nn = n ? n : 1u;
  pbf_solve_record r;
  r.index = idx;
  r.iterations = k;
  r.warm_start_kept = (int32_t)counters[C_WARM_KEPT];
  r.max_error = __uint_as_float(ehist[k]);
  r.mean_error = mean_error(esum[k], nn);
  r.max_error_before_warm_start = __uint_as_float(ehist[0]);
  r.mean_error_before_warm_start = mean_error(esum[0], nn);
  log[idx % PBF_SOLVE_LOG_SIZE] = r;
  *count = idx + 1;
}

void launch_log_solve(pbf_context* c) {
  hipLaunchKernelGGL(k_log_solve, dim3(1), dim3(64), 0, c->stream, (uint32_t)c->n,
                     (const unsigned long long*)c->counters, (const uint32_t*)c->ehist,
                     (const unsigned long long*)c->esum, c->slog, c->slog_count);
}

void launch_add_warm(pbf_context* c) {
  PBF_LAUNCH(c, "add_warm", k_add_warm, (const float*)c->wlam, c->lsum, (const unsigned long long*)c->counters);
}

void launch_viscosity_first(pbf_context* c) {
  PBF_LAUNCH_STAGED(c, "viscosity", k_viscosity_first, c->K, (const float4*)c->xs,
             (const float4*)c->v, c->va, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count, c->cnb,
             c->cnb_count, (const unsigned long long*)c->counters, (const int32_t*)c->snb,
             (const int32_t*)c->snb_count, c->skin_hint);
}

// the vorticity passes stay exact in both arithmetic modes; the static-range
// instance of the divide where KParams::static_ranges holds
#define PBF_LAUNCH_SR(c, name, KT, ...)                                     \
  do {                                                                      \
    if ((c)->K.static_ranges)                                               \
      PBF_LAUNCH(c, name, (KT<true>), __VA_ARGS__);                         \
    else                                                                    \
      PBF_LAUNCH(c, name, (KT<false>), __VA_ARGS__);                        \
  } while (0)

void launch_vorticity_masked(pbf_context* c) {
  PBF_LAUNCH_SR(c, "vorticity", k_vorticity_c, c->K, (const float4*)c->xs, (const float4*)c->va, c->scr,
                (const int32_t*)c->cnb, (const int32_t*)c->cnb_count);
}

void launch_apply_vorticity_masked(pbf_context* c) {
  PBF_LAUNCH_SR(c, "apply_vorticity", k_apply_vorticity_c, c->K, (const float4*)c->xs, (const float4*)c->scr,
                c->va, (const int32_t*)c->cnb, (const int32_t*)c->cnb_count);
}

void launch_viscosity_last(pbf_context* c) {
  PBF_LAUNCH(c, "viscosity_positions", k_viscosity_last, c->K, (const float4*)c->xs,
             (const float4*)c->va, c->v, c->x, (const int32_t*)c->cnb, (const int32_t*)c->cnb_count);
}

}  // namespace pbf
