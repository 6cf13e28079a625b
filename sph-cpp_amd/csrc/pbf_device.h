// pbf_device.h -- device arithmetic of the PBF step.
//
// Every function keeps the operation order and rounding of the reference's
// expressions (Common/vec.h, Simulation/Source/kernels.h, particles.h) so
// that, compiled with -ffp-contract=off and HIP's default correctly rounded
// single-precision divide and square root, results are bit-identical to the
// same expressions on the CPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_internal.h"

namespace pbf {

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 ld3(const float4* a, uint32_t i) {
  const float4 t = a[i];
  return {t.x, t.y, t.z};
}
__device__ __forceinline__ void st3(float4* a, uint32_t i, f3 v) { a[i] = make_float4(v.x, v.y, v.z, 0.f); }
// Element i of a particle array, addressed as the array plus a 32-bit byte
// offset (slots stay below 2^28, pbf_create): a gather then takes the base
// from scalar registers and one vector register for its offset, instead of
// a 64-bit vector address.
__device__ __forceinline__ float4 gather4(const float4* __restrict__ a, uint32_t i) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a) + (i << 4));
}

// vec.h operator-, operator+, vec*scalar, scalar*vec, vec/scalar, +=, -=
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 mul3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 smul3(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 div3(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ void acc3(f3& a, f3 b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
}
__device__ __forceinline__ void dec3(f3& a, f3 b) {
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
}
// vec.h:500-508 dot from 0: for a square the leading 0 + x*x is exact, so
// magnitude_squared is (x*x + y*y) + z*z.
__device__ __forceinline__ float magsq3(f3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
// vec.h:620-630
__device__ __forceinline__ f3 cross3(f3 l, f3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}

// "r is within h", i.e. !(sqrtf(d2) > h), decided on d2 alone: sqrt with
// round-to-nearest is monotone, so sqrtf(d2) > h exactly when d2 >= T_h.
// NaN compares false and is kept, like the reference's `r_mag > h` test.
__device__ __forceinline__ bool within_h(float d2, const KParams& K) { return !(d2 >= K.T_h); }

// Poly6<float,3>::operator() for r <= h (kernels.h:69)
__device__ __forceinline__ float poly6_in(float r, const KParams& K) {
  const float a = K.hh - r * r;
  return K.poly6_norm * a * a * a;
}
// Three correctly rounded quotients a/b with one divisor.  A correctly
// rounded divide on gfx950 is a Markstein division (P. Markstein, IBM J. Res.
// Dev. 34(1), 1990): reciprocal estimate, one Newton step
// r1 = r + r (1 - b r), then q = a r1 corrected twice with the exact residual
// a - b q (FMA); range scaling and a fix-up step only change anything for
// zero, infinite or NaN operands, denormal operands or quotients, and
// operands 96 or more binades apart.  For b and |a| in [2^-40, 2^40] (or
// a == 0) none of that happens and this computes exactly those steps,
// sharing the divisor's reciprocal and its Newton step between the three
// numerators; anything else takes the plain divide, so every result equals
// a / b (pbf_check_arithmetic compares the two on the device).
__device__ __forceinline__ bool div_in_range(float a) {
  const float m = fabsf(a);
  return (a == 0.0f) | ((m >= 0x1p-40f) & (m <= 0x1p40f));
}
// Each correction adds (b q - a)(-r1), i.e. the residual (a - b q) r1 (round
// to nearest is symmetric, so b q - a rounds to minus a - b q); written this
// way a zero numerator keeps its sign through both steps (b > 0 here: the
// quotient of +-0 is +-0), where a - b q would turn -0 into +0.
__device__ __forceinline__ float div_markstein(float a, float b, float r1) {
  const float q = a * r1;
  const float q1 = __builtin_fmaf(__builtin_fmaf(b, q, -a), -r1, q);
  return __builtin_fmaf(__builtin_fmaf(b, q1, -a), -r1, q1);
}
__device__ __forceinline__ bool div3_fast(f3 a, float b) {  // one condition, evaluated without branches
  const unsigned ok = (unsigned)(b >= 0x1p-40f) & (unsigned)(b <= 0x1p40f) & (unsigned)div_in_range(a.x) &
                      (unsigned)div_in_range(a.y) & (unsigned)div_in_range(a.z);
  return ok != 0u;
}
__device__ __forceinline__ f3 div3_markstein(f3 a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float r1 = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
  return {div_markstein(a.x, b, r1), div_markstein(a.y, b, r1), div_markstein(a.z, b, r1)};
}
__device__ __forceinline__ f3 div3_shared(f3 a, float b) {
  if (__builtin_expect(div3_fast(a, b), 1)) return div3_markstein(a, b);
  return div3(a, b);
}

// Correctly rounded square root.  On gfx950 it is the hardware estimate s
// corrected to its neighbour s -+ 1 ulp by the sign of the exact residual
// x - s' s (FMA), wrapped in a rescaling for x < 2^-96 and a pass-through of
// +-0 and +inf.  For 2^-96 <= x <= 2^96 this computes exactly those steps
// without the wrapper; other x take the plain sqrtf, so every result equals
// sqrtf(x).
__device__ __forceinline__ bool sqrt_fast(float x) { return (x >= 0x1p-96f) & (x <= 0x1p96f); }
__device__ __forceinline__ float sqrt_corrected(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  const float r = __builtin_fmaf(-sm, s, x) <= 0.0f ? sm : s;
  return __builtin_fmaf(-sp, s, x) > 0.0f ? sp : r;
}
__device__ __forceinline__ float sqrt_cr(float x) {
  if (__builtin_expect(sqrt_fast(x), 1)) return sqrt_corrected(x);
  return sqrtf(x);
}

// The same two operations inside the pressure solve's within-h sums when the
// host has shown (KParams::static_ranges) that every pair those sums accept,
// T_hack <= d2 < T_h, keeps the square root's operand in [2^-96, 2^96], the
// divisor r + r_eps in [2^-40, 2^40] and each numerator |s d_i| at or below
// 2^40.  Of the range tests only "|a_i| >= 2^-40 or a_i == 0" is then left,
// one unsigned minimum over (bits << 1) - 1 of the three numerators (zero
// wraps to the top).  The same branches are taken as with the full tests, so
// the results are the same bit for bit.
__device__ __forceinline__ bool div3_fast_static(f3 a) {
  constexpr uint32_t lim = (0x2B800000u << 1) - 1u;  // the bits of 2^-40, shifted likewise
  const uint32_t tx = (__float_as_uint(a.x) << 1) - 1u;
  const uint32_t ty = (__float_as_uint(a.y) << 1) - 1u;
  const uint32_t tz = (__float_as_uint(a.z) << 1) - 1u;
  return min(min(tx, ty), tz) >= lim;
}
__device__ __forceinline__ f3 div3_shared_static(f3 a, float b) {
  if (__builtin_expect(div3_fast_static(a), 1)) return div3_markstein(a, b);
  return div3(a, b);
}

// Del_Spikey<float,3>::operator() for r <= h, with r = |d| (kernels.h:171)
__device__ __forceinline__ f3 spiky_in(f3 d, float r, const KParams& K) {
  const float s = K.spiky_norm * (K.h - r) * (K.h - r);
  return div3_shared(smul3(s, d), r + K.r_eps);
}
// The pressure-solve sums of the fused step in either arithmetic: FAST (the
// PBF_OPT_FAST_MATH mode) takes the hardware square root and reciprocal
// (v_sqrt_f32, v_rcp_f32, within about 1 ulp) in place of the correctly
// rounded square root and divide; otherwise the operations are exactly the
// reference's.  (The XSPH and vorticity passes stay exact in both modes: with
// the hardware operations they measured slower, 110 -> 132 us for the
// vorticity pass at 1M, from a different register allocation.)
// SR: KParams::static_ranges holds and, for nb_sqrt, the pair is one the
// density sums accept (d2 >= T_hack); nb_spiky's divide needs only r <= h.
template <bool FAST, bool SR = false>
__device__ __forceinline__ float nb_sqrt(float d2) {
  if constexpr (FAST) return __builtin_amdgcn_sqrtf(d2);
  else if constexpr (SR) return sqrt_corrected(d2);
  else return sqrt_cr(d2);
}
template <bool FAST, bool SR = false>
__device__ __forceinline__ f3 nb_spiky(f3 d, float r, const KParams& K) {
  if constexpr (FAST) {
    const float s = K.spiky_norm * (K.h - r) * (K.h - r) * __builtin_amdgcn_rcpf(r + K.r_eps);
    return smul3(s, d);
  } else if constexpr (SR) {
    const float s = K.spiky_norm * (K.h - r) * (K.h - r);  // as spiky_in
    return div3_shared_static(smul3(s, d), r + K.r_eps);
  } else {
    return spiky_in(d, r, K);
  }
}

// C_Spline<float,3>::operator() (kernels.h:220-228); the first branch is in double
__device__ __forceinline__ float cspline(float r, const KParams& K) {
  const float h = K.h;
  if (r > h) return 0.0f;
  if (r <= h * 0.5f) {
    const double t = 2.0 * (double)(h - r) * (double)(h - r) * (double)(h - r) * (double)r * (double)r * (double)r;
    const float h6 = h * h * h * h * h * h / 64.0f;
    return (float)((double)K.cspline_norm * (t - (double)h6));
  }
  return K.cspline_norm * (h - r) * (h - r) * (h - r) * r * r * r;
}

// The grid and box a slot's particle lives in: the context's own, or its
// scene's in a batched context.
struct GridView {
  uint32_t dimx, dimy, dimz, off, ncells;
  uint64_t dxy;
  float bmax[3];
};
__device__ __forceinline__ GridView grid_of(const KParams& K, uint32_t p) {
  if (!K.batched) return GridView{K.dimx, K.dimy, K.dimz, 0u, K.ncells, K.dxy, {K.bmax[0], K.bmax[1], K.bmax[2]}};
  const SceneGeo g = K.scenes[K.sid[p]];
  return GridView{g.dimx, g.dimy, g.dimz, g.off, g.ncells, g.dxy, {g.bmax[0], g.bmax[1], g.bmax[2]}};
}

// apply_boundary_conditions (particles.h:510-526)
__device__ __forceinline__ void boundary(f3& p, const KParams& K, const GridView& G) {
  const float mover_radius = 0.2f;
  const f3 mc = {K.mover[0], K.mover[1], K.mover[2]};
  const float dr_squared = magsq3(sub3(p, mc));
  if (dr_squared < mover_radius * mover_radius) {
    const float dr = sqrtf(dr_squared);
    acc3(p, div3(smul3(mover_radius - dr, sub3(p, mc)), dr));
  }
  if (p.x < K.bmin[0]) p.x = K.bmin[0]; else if (p.x > G.bmax[0]) p.x = G.bmax[0];
  if (p.y < K.bmin[1]) p.y = K.bmin[1]; else if (p.y > G.bmax[1]) p.y = G.bmax[1];
  if (p.z < K.bmin[2]) p.z = K.bmin[2]; else if (p.z > G.bmax[2]) p.z = G.bmax[2];
}
__device__ __forceinline__ void boundary(f3& p, const KParams& K, uint32_t slot) { boundary(p, K, grid_of(K, slot)); }

// calculate_bin_id one axis (neighbors.h:110-111): floor((p + s) / s)
__device__ __forceinline__ float bin_coord(float p, float s) { return floorf((p + s) / s); }

// Global key of a cell given per-axis coordinates in grid G; anything that
// is not a cell of G (negative, NaN, past the last cell) maps to K.ncells
// ("outside"), so a walk never leaves its scene.
__device__ __forceinline__ uint32_t cell_key(float fx, float fy, float fz, const GridView& G, const KParams& K) {
  if (!(fx >= 0.f) || !(fy >= 0.f) || !(fz >= 0.f)) return K.ncells;
  if (fx > 4.0e9f || fy > 4.0e9f || fz > 4.0e9f) return K.ncells;
  const uint64_t bx = (uint64_t)fx, by = (uint64_t)fy, bz = (uint64_t)fz;
  const uint64_t k = G.dxy * bz + (by * (uint64_t)G.dimx + bx);
  return k < (uint64_t)G.ncells ? G.off + (uint32_t)k : K.ncells;
}

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD works on
// one contiguous range of particles (neighbouring cells share its L2).
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = nb / 8u, rem = nb % 8u;
  const uint32_t xcd = b % 8u, idx = b / 8u;
  // XCDs 0..rem-1 get per+1 blocks, the rest get per
  const uint32_t base = xcd < rem ? xcd * (per + 1u) : rem * (per + 1u) + (xcd - rem) * per;
  return base + idx;
}

}  // namespace pbf
