// pbf_passes.hip -- per-particle and neighbour-walk passes of the PBF step.
//
// Reference: Simulation/Source/particles.h:263-487 (one for_each_index lambda
// per pass) and Simulation/Source/main.cpp:71-124 (their order).
//
// Semantics on the device (the oracle's ORC_JACOBI mode): a pass reads the
// state as it was when the pass began and writes separate buffers, so the
// result does not depend on thread scheduling.  Where the reference writes
// in place while other particles read (the density fix at particles.h:303-305,
// XSPH at :439-446), this is the one place the device differs from the
// reference's serial CPU backend.
//
// Neighbour walks keep the list order and skip neighbours with r > h: every
// kernel term is exactly zero there (kernels.h:66,168) and adding a zero to a
// sum that started at +0 leaves it unchanged, so skipping is bit-exact for
// finite states.
#include "pbf_device.h"

namespace pbf {

constexpr int kBlock = 256;
static inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ bool in_span(uint32_t o, uint32_t b, uint32_t e) { return o >= b && o < e; }

// apply_external_forces (particles.h:263-270) and predict_positions (:275-287)
__global__ void __launch_bounds__(kBlock) k_forces_predict(uint32_t n, uint32_t b, uint32_t e, int flags, KParams K,
                                                           const uint32_t* __restrict__ oid,
                                                           const float4* __restrict__ x, float4* __restrict__ v,
                                                           float4* __restrict__ xs,
                                                           unsigned long long* __restrict__ counters) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if ((flags & F_CLEAR) && p < (uint32_t)C_NUM) counters[p] = 0ull;  // a fused step's counters start at zero
  if (p >= n || !in_span(oid[p], b, e)) return;
  f3 vp = ld3(v, p);
  if (flags & F_FORCES) {
    vp.y += K.g * K.dt;
    st3(v, p, vp);
  }
  if (flags & F_PREDICT) {
    f3 ps = add3(ld3(x, p), mul3(vp, K.dt));
    boundary(ps, K, p);
    st3(xs, p, ps);
  }
}

void launch_forces_predict(pbf_context* c, int flags, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "forces_predict");
  hipLaunchKernelGGL(k_forces_predict, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e,
                     flags, c->K, (const uint32_t*)c->oid, (const float4*)c->x, c->v, c->xs, c->counters);
  timer_end(c, t);
}

// compute_densities (particles.h:292-313).  Reads xs, writes xt (the position
// star with the near-coincidence fix applied, particles.h:303-305) and rho;
// the API swaps xs and xt afterwards.
__global__ void __launch_bounds__(kBlock) k_density(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                    const uint32_t* __restrict__ oid, const float4* __restrict__ xs,
                                                    const float4* __restrict__ v, float4* __restrict__ xt,
                                                    float* __restrict__ rho, const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ nbr_count,
                                                    unsigned long long* __restrict__ counters) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  f3 xp = ld3(xs, p);
  if (!in_span(oid[p], b, e)) {
    st3(xt, p, xp);
    return;
  }
  float density = K.mass_W0;
  uint32_t hacks = 0;
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 xq = ld3(xs, q);
    f3 d = sub3(xp, xq);
    float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    if (d2 < K.T_hack) {  // magnitude(x*_p - x*_q) < 1e-8 (particles.h:304)
      const f3 vp = ld3(v, p);
      dec3(xp, div3(mul3(vp, K.dt), 50.0f));
      ++hacks;
      d = sub3(xp, xq);
      d2 = magsq3(d);
      if (!within_h(d2, K)) continue;
    }
    density += K.mass * poly6_in(sqrtf(d2), K);
  }
  rho[p] = density;
  st3(xt, p, xp);
  if (hacks) atomicAdd(&counters[C_HACKS], (unsigned long long)hacks);
}

void launch_density(pbf_context* c, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "density");
  hipLaunchKernelGGL(k_density, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     (const uint32_t*)c->oid, (const float4*)c->xs, (const float4*)c->v, c->xt, c->rho,
                     (const int32_t*)c->nbr, (const int32_t*)c->nbr_count, c->counters);
  timer_end(c, t);
}

// compute_pressure_lambdas (particles.h:318-342) from positions `pos`
__global__ void __launch_bounds__(kBlock) k_lambda(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                   const uint32_t* __restrict__ oid, const float4* __restrict__ pos,
                                                   const float* __restrict__ rho, float* __restrict__ lam,
                                                   const int32_t* __restrict__ nbr,
                                                   const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  const f3 xp = ld3(pos, p);
  float sum_C = 0.0f;
  f3 sg = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(pos, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    const f3 grad = smul3(K.neg_inv_rho0, spiky_in(d, r, K));
    dec3(sg, grad);
    sum_C += magsq3(grad);
  }
  sum_C += magsq3(sg);
  const float constraint = rho[p] / K.rho0 - 1.0f;
  const float C_p = ((double)constraint < 0.0) ? 0.0f : constraint;
  lam[p] = -C_p / (sum_C + K.eps);
}

void launch_lambda(pbf_context* c, const float4* pos, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "lambda");
  hipLaunchKernelGGL(k_lambda, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     (const uint32_t*)c->oid, pos, (const float*)c->rho, c->lam, (const int32_t*)c->nbr,
                     (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

// compute_pressure_dps (particles.h:348-358): dp sum in list order
__device__ __forceinline__ f3 pressure_dp(uint32_t p, uint32_t n, const f3 xp, const float lp, const KParams& K,
                                          const float4* __restrict__ pos, const float* __restrict__ lam,
                                          const int32_t* __restrict__ nbr, int cnt) {
  f3 dp = {0.f, 0.f, 0.f};
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(pos, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    acc3(dp, smul3(lp + lam[q], spiky_in(d, r, K)));
  }
  return smul3(K.inv_rho0, dp);
}

__global__ void __launch_bounds__(kBlock) k_dp(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                               const uint32_t* __restrict__ oid, const float4* __restrict__ pos,
                                               const float* __restrict__ lam, float4* __restrict__ scr,
                                               const int32_t* __restrict__ nbr,
                                               const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  st3(scr, p, pressure_dp(p, n, ld3(pos, p), lam[p], K, pos, lam, nbr, nbr_count[p]));
}

void launch_dp(pbf_context* c, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "dp");
  hipLaunchKernelGGL(k_dp, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     (const uint32_t*)c->oid, (const float4*)c->xs, (const float*)c->lam, c->scr,
                     (const int32_t*)c->nbr, (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

// update_position_stars (particles.h:363-370): x* += scratch, boundary
__global__ void __launch_bounds__(kBlock) k_update_position_stars(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                                  const uint32_t* __restrict__ oid,
                                                                  const float4* __restrict__ scr,
                                                                  float4* __restrict__ xs) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  f3 nw = add3(ld3(xs, p), ld3(scr, p));
  boundary(nw, K, p);
  st3(xs, p, nw);
}

void launch_update_position_stars(pbf_context* c, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "update_position_stars");
  hipLaunchKernelGGL(k_update_position_stars, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b,
                     (uint32_t)e, c->K, (const uint32_t*)c->oid, (const float4*)c->scr, c->xs);
  timer_end(c, t);
}

// update_velocities (particles.h:375-385)
// (the density rides in v.w, where the fused step's XSPH passes gather it)
__global__ void __launch_bounds__(kBlock) k_update_velocities(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                              const uint32_t* __restrict__ oid,
                                                              const float4* __restrict__ x,
                                                              const float4* __restrict__ xs,
                                                              const float* __restrict__ rho, float4* __restrict__ v) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  f3 vel = div3(sub3(ld3(xs, p), ld3(x, p)), K.dt);
  if ((double)magsq3(vel) < K.vthr) vel = {0.f, 0.f, 0.f};
  v[p] = make_float4(vel.x, vel.y, vel.z, rho[p]);
}

void launch_update_velocities(pbf_context* c, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "update_velocities");
  hipLaunchKernelGGL(k_update_velocities, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b,
                     (uint32_t)e, c->K, (const uint32_t*)c->oid, (const float4*)c->x, (const float4*)c->xs,
                     (const float*)c->rho, c->v);
  timer_end(c, t);
}

// update_positions (particles.h:390-394)
__global__ void __launch_bounds__(kBlock) k_update_positions(uint32_t n, uint32_t b, uint32_t e,
                                                             const uint32_t* __restrict__ oid,
                                                             const float4* __restrict__ xs, float4* __restrict__ x) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  x[p] = xs[p];
}

void launch_update_positions(pbf_context* c, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "update_positions");
  hipLaunchKernelGGL(k_update_positions, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e,
                     (const uint32_t*)c->oid, (const float4*)c->xs, c->x);
  timer_end(c, t);
}

// apply_viscosity (particles.h:436-447), XSPH from vin into vout; with
// write_positions also update_positions (:390-394) for the same particle.
__global__ void __launch_bounds__(kBlock) k_viscosity(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                      int write_positions, const uint32_t* __restrict__ oid,
                                                      const float4* __restrict__ xs, const float* __restrict__ rho,
                                                      const float4* __restrict__ vin, float4* __restrict__ vout,
                                                      float4* __restrict__ x, const int32_t* __restrict__ nbr,
                                                      const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 vp = ld3(vin, p);
  if (!in_span(oid[p], b, e)) {
    st3(vout, p, vp);
    return;
  }
  const f3 xp = ld3(xs, p);
  f3 dv = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(xs, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    acc3(dv, div3(mul3(sub3(ld3(vin, q), vp), poly6_in(r, K)), rho[q]));
  }
  f3 nv = vp;
  acc3(nv, smul3(K.visc, dv));
  st3(vout, p, nv);
  if (write_positions) x[p] = make_float4(xp.x, xp.y, xp.z, 0.f);
}

void launch_viscosity(pbf_context* c, const float4* vin, float4* vout, int write_positions, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, write_positions ? "viscosity_positions" : "viscosity");
  hipLaunchKernelGGL(k_viscosity, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     write_positions, (const uint32_t*)c->oid, (const float4*)c->xs, (const float*)c->rho, vin, vout,
                     c->x, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

// compute_vorticity (particles.h:452-466): omega -> scr
__global__ void __launch_bounds__(kBlock) k_vorticity(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                      const uint32_t* __restrict__ oid,
                                                      const float4* __restrict__ xs, const float4* __restrict__ vin,
                                                      float4* __restrict__ scr, const int32_t* __restrict__ nbr,
                                                      const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  const f3 xp = ld3(xs, p);
  const f3 vp = ld3(vin, p);
  f3 w = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(xs, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    const f3 del = spiky_in(d, r, K);
    acc3(w, cross3(sub3(ld3(vin, q), vp), del));
  }
  st3(scr, p, w);
}

void launch_vorticity(pbf_context* c, const float4* vin, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "vorticity");
  hipLaunchKernelGGL(k_vorticity, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     (const uint32_t*)c->oid, (const float4*)c->xs, vin, c->scr, (const int32_t*)c->nbr,
                     (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

// apply_vorticity (particles.h:471-487): reads |omega_q| of neighbours, own v
__global__ void __launch_bounds__(kBlock) k_apply_vorticity(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                            const uint32_t* __restrict__ oid,
                                                            const float4* __restrict__ xs,
                                                            const float4* __restrict__ scr, float4* __restrict__ v,
                                                            const int32_t* __restrict__ nbr,
                                                            const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  const f3 xp = ld3(xs, p);
  f3 eta = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(xs, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    const f3 del = spiky_in(d, r, K);
    const float wm = sqrtf(magsq3(ld3(scr, q)));
    acc3(eta, smul3(wm, del));
  }
  const f3 N = div3(eta, sqrtf(magsq3(eta)) + K.r_eps);
  f3 vp = ld3(v, p);
  acc3(vp, mul3(mul3(cross3(N, ld3(scr, p)), K.vort), K.dt));
  st3(v, p, vp);
}

void launch_apply_vorticity(pbf_context* c, float4* vio, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "apply_vorticity");
  hipLaunchKernelGGL(k_apply_vorticity, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e,
                     c->K, (const uint32_t*)c->oid, (const float4*)c->xs, (const float4*)c->scr, vio,
                     (const int32_t*)c->nbr, (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

// apply_surface_tension (particles.h:400-431).  Pass 1: colour field -> scr.
__global__ void __launch_bounds__(kBlock) k_color(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                  const uint32_t* __restrict__ oid, const float4* __restrict__ xs,
                                                  const float* __restrict__ rho, float4* __restrict__ scr,
                                                  const int32_t* __restrict__ nbr,
                                                  const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  const f3 xp = ld3(xs, p);
  f3 col = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 d = sub3(xp, ld3(xs, q));
    const float d2 = magsq3(d);
    if (!within_h(d2, K)) continue;
    const float r = sqrtf(d2);
    acc3(col, div3(spiky_in(d, r, K), rho[q]));
  }
  st3(scr, p, smul3(K.h, col));
}

// Pass 2: cohesion + curvature force into own v.  The r > h shortcut does
// not apply here: curvature (scratch_p - scratch_q) has no kernel factor.
__global__ void __launch_bounds__(kBlock) k_tension(uint32_t n, uint32_t b, uint32_t e, KParams K,
                                                    const uint32_t* __restrict__ oid, const float4* __restrict__ xs,
                                                    const float* __restrict__ rho, const float4* __restrict__ scr,
                                                    float4* __restrict__ v, const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ nbr_count) {
  const uint32_t p = xcd_block() * blockDim.x + threadIdx.x;
  if (p >= n || !in_span(oid[p], b, e)) return;
  const f3 xp = ld3(xs, p);
  const f3 sp = ld3(scr, p);
  const float rp = rho[p];
  f3 f = {0.f, 0.f, 0.f};
  const int cnt = nbr_count[p];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t q = (uint32_t)nbr[list_at(k, p, K.stride)];
    const f3 r = sub3(xp, ld3(xs, q));
    float r_mag = sqrtf(magsq3(r));
    if ((double)r_mag < K.st_rmin) r_mag = (float)K.st_rmin;
    const f3 cohesion = div3(smul3(-K.gamma * cspline(r_mag, K), r), r_mag);
    const f3 curvature = smul3(-K.gamma, sub3(sp, ld3(scr, q)));
    const float Kq = (float)(2.0 * (double)K.rho0 / (double)(rp + rho[q]));
    acc3(f, smul3(Kq, add3(cohesion, curvature)));
  }
  f3 vp = ld3(v, p);
  acc3(vp, mul3(div3(f, rp), K.dt));
  v[p] = make_float4(vp.x, vp.y, vp.z, rp);  // the density stays in v.w for the XSPH passes
}

void launch_surface_tension(pbf_context* c, uint64_t cb, uint64_t ce, uint64_t b, uint64_t e) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "surface_tension");
  hipLaunchKernelGGL(k_color, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)cb, (uint32_t)ce, c->K,
                     (const uint32_t*)c->oid, (const float4*)c->xs, (const float*)c->rho, c->scr,
                     (const int32_t*)c->nbr, (const int32_t*)c->nbr_count);
  hipLaunchKernelGGL(k_tension, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (uint32_t)b, (uint32_t)e, c->K,
                     (const uint32_t*)c->oid, (const float4*)c->xs, (const float*)c->rho, (const float4*)c->scr,
                     c->v, (const int32_t*)c->nbr, (const int32_t*)c->nbr_count);
  timer_end(c, t);
}

__global__ void k_copy_f4(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
void launch_copy_f4(const float4* src, float4* dst, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_f4, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, n);
}

__global__ void k_copy_f1(const float* __restrict__ src, float* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
void launch_copy_f1(const float* src, float* dst, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_f1, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, n);
}
__global__ void k_scale_f1(const float* __restrict__ src, float* __restrict__ dst, float scale, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] * scale;
}
void launch_scale_f1(const float* src, float* dst, float scale, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_scale_f1, dim3(nblocks(n)), dim3(kBlock), 0, s, src, dst, scale, n);
}

// ---- original-order transfers through the staging buffer ----
__global__ void k_gather_out_f4(uint32_t n, const uint32_t* __restrict__ oid, const float4* __restrict__ src,
                                float* __restrict__ dst, int comps) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t o = oid[s];
  const float4 t = src[s];
  dst[(uint64_t)o * comps + 0] = t.x;
  if (comps > 1) dst[(uint64_t)o * comps + 1] = t.y;
  if (comps > 2) dst[(uint64_t)o * comps + 2] = t.z;
}
__global__ void k_gather_out_f1(uint32_t n, const uint32_t* __restrict__ oid, const float* __restrict__ src,
                                float* __restrict__ dst) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[oid[s]] = src[s];
}
__global__ void k_scatter_in_f4(uint32_t n, const uint32_t* __restrict__ oid, const float* __restrict__ src3,
                                float4* __restrict__ dst) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t o = oid[s];
  dst[s] = make_float4(src3[o * 3], src3[o * 3 + 1], src3[o * 3 + 2], 0.f);
}
__global__ void k_scatter_in_f1(uint32_t n, const uint32_t* __restrict__ oid, const float* __restrict__ src,
                                float* __restrict__ dst) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[s] = src[oid[s]];
}

void launch_gather_out_f4(pbf_context* c, const float4* src, float* dst3, int comps) {
  const uint32_t n = (uint32_t)c->n;
  if (n) hipLaunchKernelGGL(k_gather_out_f4, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (const uint32_t*)c->oid,
                            src, dst3, comps);
}
__global__ void k_gather_out_rows(uint32_t n, const uint32_t* __restrict__ oid, const float* __restrict__ src,
                                  float* __restrict__ dst, int width) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t o = oid[s];
  for (int k = 0; k < width; ++k) dst[o * width + k] = src[(uint64_t)s * width + k];
}
void launch_gather_out_rows(pbf_context* c, const float* src, float* dst, int width) {
  const uint32_t n = (uint32_t)c->n;
  if (n) hipLaunchKernelGGL(k_gather_out_rows, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n,
                            (const uint32_t*)c->oid, src, dst, width);
}
void launch_gather_out_f1(pbf_context* c, const float* src, float* dst) {
  const uint32_t n = (uint32_t)c->n;
  if (n) hipLaunchKernelGGL(k_gather_out_f1, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (const uint32_t*)c->oid,
                            src, dst);
}
void launch_scatter_in_f4(pbf_context* c, const float* src3, float4* dst) {
  const uint32_t n = (uint32_t)c->n;
  if (n) hipLaunchKernelGGL(k_scatter_in_f4, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (const uint32_t*)c->oid,
                            src3, dst);
}
void launch_scatter_in_f1(pbf_context* c, const float* src, float* dst) {
  const uint32_t n = (uint32_t)c->n;
  if (n) hipLaunchKernelGGL(k_scatter_in_f1, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (const uint32_t*)c->oid,
                            src, dst);
}

}  // namespace pbf
