// pbf_api.hip -- C ABI of the PBF library: contexts, buffers, passes, step.
//
// Reference call surface: Common/sim_algorithms.h (primitives),
// Simulation/Source/particles.h and neighbors.h (passes), Common/array.h and
// Common/managed_allocation.h (storage, replaced by device buffers),
// Common/parameters.h:123-148 (derived constants), Simulation/Source/main.cpp:71-124
// (step order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "pbf_internal.h"

namespace pbf {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return PBF_OK;
  return set_error(PBF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Events of finished launches are folded into the totals once a timer holds
// kMaxPendingEvents pairs, so a long run keeps a bounded number of events
// (the oldest launches have long finished: waiting on them costs nothing);
// their events are kept for later launches.
constexpr size_t kMaxPendingEvents = 64;

static void fold_oldest(KernelTimer& t, size_t count) {
  for (size_t i = 0; i < count && i < t.pending.size(); ++i) {
    auto& pr = t.pending[i];
    float ms = 0.f;
    // a pair whose time cannot be read is left out of the totals (the launch
    // count then shows the gap)
    if (hipEventSynchronize(pr.second) == hipSuccess && hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      t.total_ms += ms;
      t.launches++;
    }
    t.spare.push_back(pr.first);
    t.spare.push_back(pr.second);
  }
  t.pending.erase(t.pending.begin(), t.pending.begin() + std::min(count, t.pending.size()));
}

KernelTimer* timer_begin(pbf_context* c, const char* name) {
  if (c->capturing) return nullptr;
  for (auto& t : c->timers) {
    if (t.name == name) {
      if (t.pending.size() >= kMaxPendingEvents) fold_oldest(t, kMaxPendingEvents / 2);
      hipEvent_t ev[2] = {nullptr, nullptr};
      for (hipEvent_t& e : ev) {
        if (t.spare.empty()) {
          if (hipEventCreate(&e) != hipSuccess) e = nullptr;
        } else {
          e = t.spare.back();
          t.spare.pop_back();
        }
      }
      const hipEvent_t a = ev[0], b = ev[1];
      if (!a || !b || hipEventRecord(a, c->stream) != hipSuccess) {
        // no events for this launch: it runs untimed
        for (hipEvent_t e : ev)
          if (e) t.spare.push_back(e);
        return nullptr;
      }
      t.pending.push_back({a, b});
      return &t;
    }
  }
  return nullptr;
}

void timer_end(pbf_context* c, KernelTimer* t) {
  if (t) hipEventRecord(t->pending.back().second, c->stream);
}

static void timers_resolve(pbf_context* c) {
  for (auto& t : c->timers) fold_oldest(t, t.pending.size());
}

// resolve, then release every event the timers hold
static void timers_release(pbf_context* c) {
  timers_resolve(c);
  for (auto& t : c->timers) {
    for (hipEvent_t e : t.spare) hipEventDestroy(e);
    t.spare.clear();
  }
}

// smallest float t with sqrtf(t) > h (sqrt with round-to-nearest is monotone)
static float threshold_sqrt_gt(float h) {
  float t = h * h;
  while (t > 0.f && std::sqrt(std::nextafter(t, 0.f)) > h) t = std::nextafter(t, 0.f);
  while (!(std::sqrt(t) > h)) t = std::nextafter(t, std::numeric_limits<float>::infinity());
  while (t > 0.f && std::sqrt(std::nextafter(t, 0.f)) > h) t = std::nextafter(t, 0.f);
  return t;
}

// smallest float t with (double)sqrtf(t) >= 1e-8 (particles.h:304 compares in double)
static float threshold_hack() {
  float t = (float)(1e-8 * 1e-8);
  while (t > 0.f && (double)std::sqrt(std::nextafter(t, 0.f)) >= 0.00000001) t = std::nextafter(t, 0.f);
  while (!((double)std::sqrt(t) >= 0.00000001)) t = std::nextafter(t, std::numeric_limits<float>::infinity());
  while (t > 0.f && (double)std::sqrt(std::nextafter(t, 0.f)) >= 0.00000001) t = std::nextafter(t, 0.f);
  return t;
}

static int compute_kparams(const pbf_params& P, KParams& K, uint64_t dims[3]) {
  const float h = P.smoothing_radius, s = P.neighbor_bin_spacing;
  if (!(h > 0.f) || !(s > 0.f) || !std::isfinite(h) || !std::isfinite(s))
    return set_error(PBF_ERR_ARG, "smoothing_radius and neighbor_bin_spacing must be positive");
  if (P.boundary_min[0] != 0.f || P.boundary_min[1] != 0.f || P.boundary_min[2] != 0.f)
    return set_error(PBF_ERR_UNSUPPORTED,
                     "boundary_min must be the origin: bin ids add the spacing instead of subtracting the minimum "
                     "(neighbors.h:110-113)");
  if (P.max_neighbors < 1 || P.max_neighbors > 128)
    return set_error(PBF_ERR_ARG, "max_neighbors must be in [1, 128]");
  if (!(P.time_step != 0.f)) return set_error(PBF_ERR_ARG, "time_step must be non-zero");
  std::memset(&K, 0, sizeof(K));
  K.h = h;
  K.hh = h * h;
  K.s = s;
  K.ss = s * s;
  K.T_h = threshold_sqrt_gt(h);
  K.T_hack = threshold_hack();
  // kernels.h:62, :160-161, :218 -- norms computed in double, then cast
  K.poly6_norm = static_cast<float>(315.0 / (64.0 * M_PI * std::pow((double)h, 9.0)));
  K.spiky_norm = static_cast<float>(-45.0 / (M_PI * std::pow((double)h, 6.0)));
  K.cspline_norm = static_cast<float>(32.0 / (M_PI * std::pow((double)h, 9.0)));
  K.r_eps = std::numeric_limits<float>::epsilon();
  {
    // every pair the within-h sums accept has T_hack <= d2 < T_h, so
    // sqrt_cr(d2) <= h, a divisor r + r_eps in [r_eps, h + r_eps] and
    // numerators |spiky_norm (h - r)^2 d_i| <= |spiky_norm| h^3 (up to
    // rounding): where those stay inside the exact operations' fast ranges
    // the kernels skip the tests (div3_fast_static)
    const double hd = h;
    const bool sq = K.T_hack >= 0x1p-96f && K.T_h <= 0x1p96f;
    const bool dv = K.r_eps >= 0x1p-40f && hd + K.r_eps <= 0x1p40;
    const bool nm = std::fabs((double)K.spiky_norm) * hd * hd * hd * 1.001 <= 0x1p40;
    K.static_ranges = (sq && dv && nm) ? 1 : 0;
  }
  K.mass = P.rest_mass;
  {
    const float r = 0.0f;  // W(0), particles.h:294
    const float a = h * h - r * r;
    K.mass_W0 = P.rest_mass * (K.poly6_norm * a * a * a);
  }
  K.rho0 = P.rest_density;
  K.neg_inv_rho0 = (float)-1.0 / P.rest_density;  // particles.h:332
  K.inv_rho0 = (float)1.0 / P.rest_density;       // particles.h:356
  K.g = P.gravity;
  K.dt = P.time_step;
  K.g_dt = P.gravity * P.time_step;
  K.eps = P.lambda_epsilon;
  K.visc = P.visc_c;
  K.vort = P.vorticity_coef;
  K.gamma = P.gamma;
  K.vthr = 0.000001 * (double)P.max_speed;         // particles.h:380
  K.st_rmin = (double)P.smoothing_radius * 0.000001;  // particles.h:420
  for (int i = 0; i < 3; ++i) {
    K.bmin[i] = P.boundary_min[i];
    K.bmax[i] = P.boundary_max[i];
    K.mover[i] = P.mover_center[i];
  }
  // neighbors.h:69-71: ceil(extent / spacing) + 2 in float, then size_t
  for (int i = 0; i < 3; ++i) {
    const float ext = P.boundary_max[i] - P.boundary_min[i];
    const float d = std::ceil(ext / s) + 2.0f;
    if (!(d >= 1.f) || !(d < 4.0e9f)) return set_error(PBF_ERR_ARG, "boundary / spacing gives no valid grid");
    dims[i] = (uint64_t)d;
  }
  const double cells = (double)dims[0] * (double)dims[1] * (double)dims[2];
  if (cells > 4.0e9) return set_error(PBF_ERR_UNSUPPORTED, "grid has more than 4e9 cells");
  K.dimx = (uint32_t)dims[0];
  K.dimy = (uint32_t)dims[1];
  K.dimz = (uint32_t)dims[2];
  K.dxy = dims[0] * dims[1];
  K.ncells = (uint32_t)(dims[0] * dims[1] * dims[2]);
  K.cap = (int)P.max_neighbors;
  K.stride = (K.cap + 7) / 8 * 8;
  K.count_overflow = (P.options & PBF_OPT_COUNT_OVERFLOW) ? 1 : 0;
  K.fast = (P.options & PBF_OPT_FAST_MATH) ? 1 : 0;
  K.skin2 = (1.1f * h) * (1.1f * h);
  K.disp_lim2 = (0.045f * h) * (0.045f * h);
  K.skin_mode = (P.options & PBF_OPT_FULL_LISTS) ? 0 : 1;
  K.tile_lists = (P.options & PBF_OPT_LANE_LISTS) ? 0 : 1;
  K.batched = 0;
  K.scenes = nullptr;
  K.sid = nullptr;
  K.heavy_budget = 768u;
  return PBF_OK;
}

// Batched contexts keep their scene grids when parameters are replaced.
static void keep_scenes(const pbf_context* c, KParams& K) {
  if (c->nscenes <= 1 && !c->K.batched) return;
  const SceneGeo& g0 = c->scene_geo[0];
  K.dimx = g0.dimx;
  K.dimy = g0.dimy;
  K.dimz = g0.dimz;
  K.dxy = g0.dxy;
  for (int i = 0; i < 3; ++i) K.bmax[i] = g0.bmax[i];
  K.ncells = c->K.ncells;
  K.batched = 1;
  K.scenes = c->scenes;
  K.sid = c->sid;
}

template <class T>
static int dalloc(T** p, uint64_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return check_hip(hipMalloc((void**)p, count * sizeof(T)), "hipMalloc");
}

static void dfree(void* p) {
  if (p) hipFree(p);
}

static int alloc_grid(pbf_context* c) {
  const uint64_t need = c->ncells + 2;
  if (need <= c->grid_alloc) return PBF_OK;
  dfree(c->cell_start);
  dfree(c->cbox);
  c->cell_start = nullptr;
  c->cbox = nullptr;
  int rc;
  if ((rc = dalloc(&c->cell_start, need)) || (rc = dalloc(&c->cbox, 6 * need))) return rc;
  c->grid_alloc = need;
  return PBF_OK;
}

static void free_all(pbf_context* c) {
  void* ptrs[] = {c->x,    c->xs,   c->v,    c->xa,   c->xsa,  c->va,   c->xt,         c->scr,
                  c->scra, c->rho,  c->lam,  c->rhoa, c->lama, c->oid,  c->oida,       c->key,
                  c->keya, c->slot, c->tmp,  c->tmpo, c->nbr,  c->nbr_count, c->counters, c->stage,
                  c->stage64, c->cell_start, c->scan_partials, c->sk[0], c->sk[1], c->sv[0], c->sv[1],
                  c->rhist, c->roffs, c->rpart, c->heavy, c->cnb, c->cnb_count, c->moved, c->dirty,
                  c->ehist, c->lsum, c->lsuma, c->wlam, c->wgiven, c->xsave, c->esum, c->feat, c->mw1,
                  c->mw2, c->mp, c->scenes, c->sid, c->sida, c->cbox, c->chbox, c->wmask, c->mflag,
                  c->slog, c->slog_count, c->snb, c->snb_count, c->xs0, c->skin_hint};
  for (void* p : ptrs) dfree(p);
}

static bool span_ok(const pbf_context* c, pbf_index_span s) { return s.begin <= s.end && s.end <= c->n; }

static int need_lists(pbf_context* c) {
  if (!c->lists_valid) return set_error(PBF_ERR_STATE, "neighbour pass before find_neighbors (or after add/remove)");
  return PBF_OK;
}

// iota of original ids for freshly added slots
__global__ void k_iota(uint32_t* p, uint64_t first, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[first + i] = (uint32_t)(first + i);
}
__global__ void k_widen3(const float* __restrict__ src, float4* __restrict__ dst, uint64_t first, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[first + i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.f);
}
__global__ void k_zero_f4(float4* p, uint64_t first, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[first + i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_zero_f1(float* p, uint64_t first, uint64_t count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[first + i] = 0.f;
}
// keep-flags for remove(): 1 when the original index survives
__global__ void k_keep(uint32_t n, const uint32_t* __restrict__ oid, uint32_t newn, uint32_t* __restrict__ flag) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) flag[s] = oid[s] < newn ? 1u : 0u;
}
__global__ void k_compact_perm(uint32_t n, const uint32_t* __restrict__ oid, uint32_t newn,
                               const uint32_t* __restrict__ pos, uint32_t* __restrict__ perm) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && oid[s] < newn) perm[pos[s]] = s;
}

// neighbour search body shared by pbf_find_neighbors and pbf_step
static void find_neighbors_impl(pbf_context* c, uint64_t bin_end, uint64_t fb, uint64_t fe, bool skin = false);

void drop_graph(pbf_context* c) {
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

static void drop_solve_graph(pbf_context* c) {
  if (c->sgraph_exec) hipGraphExecDestroy(c->sgraph_exec);
  if (c->sgraph) hipGraphDestroy(c->sgraph);
  c->sgraph_exec = nullptr;
  c->sgraph = nullptr;
}

// The buffers whose roles the host swaps (reorder, per-pass calls): a graph
// is valid only while they are where they were at capture.
static std::vector<void*> buffer_sig(const pbf_context* c) {
  return {c->x, c->xs, c->v, c->xa, c->xsa, c->va, c->xt, c->scr, c->scra, c->rho, c->lam, c->rhoa, c->lama,
          c->oid, c->oida, c->key, c->keya, c->lsum, c->lsuma, c->sid, c->sida, (void*)(uintptr_t)c->track_lsum};
}

// The grid of every scene of a batched context (boxes: 3 extents per scene)
// and the total cell count; scene s takes the keys [off_s, off_s + ncells_s).
static int build_scene_geo(const pbf_params& params, const float* boxes, uint32_t num_scenes,
                           std::vector<SceneGeo>& geo, uint64_t& total) {
  geo.assign(num_scenes, SceneGeo{});
  total = 0;
  for (uint32_t sc = 0; sc < num_scenes; ++sc) {
    pbf_params ps = params;
    for (int i = 0; i < 3; ++i) {
      if (!(boxes[3 * sc + i] > 0.0f)) return set_error(PBF_ERR_ARG, "scene boxes must be positive");
      ps.boundary_max[i] = boxes[3 * sc + i];
    }
    KParams Ks;
    uint64_t d[3];
    int rc = compute_kparams(ps, Ks, d);
    if (rc) return rc;
    SceneGeo& g = geo[sc];
    std::memset(&g, 0, sizeof(g));
    g.dimx = Ks.dimx;
    g.dimy = Ks.dimy;
    g.dimz = Ks.dimz;
    g.off = (uint32_t)total;
    g.ncells = Ks.ncells;
    g.dxy = Ks.dxy;
    for (int i = 0; i < 3; ++i) g.bmax[i] = Ks.bmax[i];
    total += Ks.ncells;
    if (total >= 4000000000ull) return set_error(PBF_ERR_UNSUPPORTED, "scenes hold more than 4e9 cells");
  }
  return PBF_OK;
}

}  // namespace pbf

using namespace pbf;

// ---------------------------------------------------------------------------
extern "C" {

const char* pbf_last_error(void) { return g_last_error.c_str(); }

int pbf_params_derive(pbf_params* p) {
  PBF_TRACE("pbf_params_derive");
  if (!p) return set_error(PBF_ERR_ARG, "null params");
  // parameters.h:123-148
  if (!(p->particle_rest_spacing > 0.f)) return set_error(PBF_ERR_ARG, "particle_rest_spacing must be positive");
  p->particle_radius = (float)(p->particle_rest_spacing / 2.0);
  if (p->smoothing_radius <= 0.0) p->smoothing_radius = (float)(1.8 * p->particle_rest_spacing);
  if (p->neighbor_bin_spacing <= 0.0) p->neighbor_bin_spacing = (float)(1.2 * p->smoothing_radius);
  p->rest_mass = 1.0f;
  const float particle_volume = std::pow(p->particle_rest_spacing, (float)3);
  p->rest_density = p->rest_mass / particle_volume;
  p->max_speed = (float)(0.5 * p->smoothing_radius * p->solve_step_count / p->time_step);
  if (p->max_neighbors == 0) p->max_neighbors = 60;  // MAX_NEIGHBORS (neighbors.h:39)
  return PBF_OK;
}

int pbf_create(const pbf_params* params, pbf_context** out) {
  PBF_TRACE("pbf_create");
  if (!params || !out) return set_error(PBF_ERR_ARG, "null argument");
  *out = nullptr;
  // slots below 2^28: the neighbour passes address particle arrays with
  // 32-bit byte offsets (gather4)
  if (params->max_particles_local == 0 || params->max_particles_local >= (1ull << 28))
    return set_error(PBF_ERR_ARG, "max_particles_local must be in [1, 2^28)");
  pbf_context* c = new pbf_context();
  c->P = *params;
  int rc = compute_kparams(c->P, c->K, c->dims);
  if (rc) {
    delete c;
    return rc;
  }
  c->ncells = c->K.ncells;
  c->cap = c->K.cap;
  c->cap_alloc = c->K.cap;
  c->capacity = params->max_particles_local;
  if ((rc = check_hip(hipGetDevice(&c->device), "hipGetDevice")) ||
      (rc = check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate"))) {
    delete c;
    return rc;
  }
  const uint64_t N = c->capacity;
  // position buffers carry kRowPad rows of padding: the neighbour walk loads
  // candidates 8 at a time past the end of a cell's range
  const uint64_t NP = N + kRowPad;
  if ((rc = dalloc(&c->x, NP)) || (rc = dalloc(&c->xs, NP)) || (rc = dalloc(&c->v, NP)) ||
      (rc = dalloc(&c->xa, NP)) || (rc = dalloc(&c->xsa, NP)) || (rc = dalloc(&c->va, NP)) ||
      (rc = dalloc(&c->xt, NP)) ||
      (rc = dalloc(&c->scr, N)) || (rc = dalloc(&c->scra, N)) || (rc = dalloc(&c->rho, N)) ||
      (rc = dalloc(&c->lam, N)) || (rc = dalloc(&c->rhoa, N)) || (rc = dalloc(&c->lama, N)) ||
      (rc = dalloc(&c->oid, N)) || (rc = dalloc(&c->oida, N)) || (rc = dalloc(&c->key, N)) ||
      (rc = dalloc(&c->keya, N)) || (rc = dalloc(&c->slot, N)) || (rc = dalloc(&c->tmp, N + 1)) ||
      (rc = dalloc(&c->tmpo, N + 1)) || (rc = dalloc(&c->nbr, (N + 63) / 64 * 64 * (uint64_t)c->K.stride)) ||
      (rc = dalloc(&c->nbr_count, N)) || (rc = dalloc(&c->counters, (uint64_t)C_NUM)) ||
      (rc = dalloc(&c->stage, 4 * N)) || (rc = dalloc(&c->stage64, N + 1)) || (rc = alloc_grid(c)) ||
      (rc = dalloc(&c->sk[0], N + 1)) || (rc = dalloc(&c->sk[1], N + 1)) || (rc = dalloc(&c->sv[0], N + 1)) ||
      (rc = dalloc(&c->sv[1], N + 1)) || (rc = dalloc(&c->rhist, radix_hist_size(N))) ||
      (rc = dalloc(&c->roffs, radix_hist_size(N) + 1)) ||
      (rc = dalloc(&c->rpart, scan_partials_size(radix_hist_size(N) + 1))) || (rc = dalloc(&c->heavy, N)) ||
      (rc = dalloc(&c->chbox, 2 * ((N + 63) / 64))) || (rc = dalloc(&c->wmask, 2 * N)) ||
      (rc = dalloc(&c->cnb, (N + 63) / 64 * 64 * (uint64_t)c->K.stride)) || (rc = dalloc(&c->cnb_count, N)) ||
      (rc = dalloc(&c->snb, (N + 63) / 64 * 64 * (uint64_t)c->K.stride)) || (rc = dalloc(&c->snb_count, N)) ||
      (rc = dalloc(&c->xs0, NP)) || (rc = dalloc(&c->skin_hint, 1)) ||
      (rc = dalloc(&c->moved, N)) || (rc = dalloc(&c->dirty, N)) || (rc = dalloc(&c->mflag, N)) ||
      (rc = dalloc(&c->ehist, (uint64_t)kMaxSolveSlots + 1)) || (rc = dalloc(&c->lsum, N)) ||
      (rc = dalloc(&c->lsuma, N)) || (rc = dalloc(&c->wlam, N)) || (rc = dalloc(&c->wgiven, N)) ||
      (rc = dalloc(&c->xsave, N)) || (rc = dalloc(&c->esum, (uint64_t)kMaxSolveSlots + 1)) ||
      (rc = dalloc(&c->feat, N * PBF_WARM_FEATURES)) || (rc = dalloc(&c->mw1, (uint64_t)PBF_WARM_HIDDEN * 16)) ||
      (rc = dalloc(&c->mw2, (uint64_t)PBF_WARM_HIDDEN * PBF_WARM_HIDDEN)) ||
      (rc = dalloc(&c->mp, (uint64_t)PBF_WARM_PARAMS)) ||
      (rc = dalloc(&c->scan_partials, scan_partials_size(N + 1))) ||
      (rc = dalloc(&c->slog, (uint64_t)PBF_SOLVE_LOG_SIZE)) || (rc = dalloc(&c->slog_count, 1)) ||
      (rc = check_hip(hipHostMalloc((void**)&c->done_host, sizeof(unsigned long long)), "hipHostMalloc")) ||
      (rc = check_hip(hipEventCreate(&c->region_a), "hipEventCreate")) ||
      (rc = check_hip(hipEventCreate(&c->region_b), "hipEventCreate"))) {
    free_all(c);
    if (c->done_host) (void)hipHostFree(c->done_host);
    if (c->region_a) (void)hipEventDestroy(c->region_a);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  launch_fill_u64(c->counters, 0ull, C_NUM, c->stream);
  launch_fill_u32((uint32_t*)c->nbr_count, 0u, N, c->stream);
  launch_fill_u32(c->dirty, 0u, N, c->stream);
  launch_fill_u32(c->mflag, 0u, N, c->stream);
  launch_fill_u32((uint32_t*)c->lsum, 0u, N, c->stream);
  launch_fill_u32(c->ehist, 0u, (uint64_t)kMaxSolveSlots + 1, c->stream);
  launch_fill_u64(c->esum, 0ull, (uint64_t)kMaxSolveSlots + 1, c->stream);
  launch_fill_u64(c->slog_count, 0ull, 1, c->stream);
  launch_fill_u64(c->skin_hint, 1ull, 1, c->stream);
  if ((rc = check_hip(hipStreamSynchronize(c->stream), "pbf_create"))) {
    pbf_destroy(c);
    return rc;
  }
  *out = c;
  return PBF_OK;
}

int pbf_destroy(pbf_context* c) {
  PBF_TRACE("pbf_destroy");
  if (!c) return PBF_OK;
  hipStreamSynchronize(c->stream);
  drop_graph(c);
  drop_solve_graph(c);
  timers_release(c);
  free_all(c);
  if (c->done_host) hipHostFree(c->done_host);
  if (c->region_a) hipEventDestroy(c->region_a);
  if (c->region_b) hipEventDestroy(c->region_b);
  hipStreamDestroy(c->stream);
  delete c;
  return PBF_OK;
}

int pbf_set_params(pbf_context* c, const pbf_params* params) {
  PBF_TRACE("pbf_set_params");
  if (!c || !params) return set_error(PBF_ERR_ARG, "null argument");
  if (params->max_particles_local != c->capacity)
    return set_error(PBF_ERR_ARG, "capacity is fixed at pbf_create");
  if ((int)params->max_neighbors > c->cap_alloc)
    return set_error(PBF_ERR_ARG, "max_neighbors may not exceed the cap the context was created with");
  KParams K;
  uint64_t dims[3];
  int rc = compute_kparams(*params, K, dims);
  if (rc) return rc;
  if (c->K.batched) {
    if (K.s != c->K.s) {
      // new bin spacing: every scene's grid is derived again from its box
      std::vector<float> boxes(3 * (size_t)c->nscenes);
      for (uint32_t sc = 0; sc < c->nscenes; ++sc)
        for (int i = 0; i < 3; ++i) boxes[3 * sc + i] = c->scene_geo[sc].bmax[i];
      std::vector<SceneGeo> geo;
      uint64_t total = 0;
      if ((rc = build_scene_geo(*params, boxes.data(), c->nscenes, geo, total))) return rc;
      if ((rc = check_hip(hipStreamSynchronize(c->stream), "sync")) ||
          (rc = check_hip(hipMemcpy(c->scenes, geo.data(), geo.size() * sizeof(SceneGeo), hipMemcpyHostToDevice),
                          "upload")))
        return rc;
      c->scene_geo = geo;
      c->K.ncells = (uint32_t)total;
      c->dims[0] = geo[0].dimx;
      c->dims[1] = geo[0].dimy;
      c->dims[2] = geo[0].dimz;
    }
    keep_scenes(c, K);
    for (int i = 0; i < 3; ++i) dims[i] = c->dims[i];
  }
  const bool grid_changed = K.ncells != c->K.ncells || K.s != c->K.s || K.dimx != c->K.dimx ||
                            K.dimy != c->K.dimy || K.cap != c->K.cap;
  K.heavy_budget = c->tune.heavy_budget;
  c->P = *params;
  c->K = K;
  c->cap = K.cap;
  ++c->version;
  for (int i = 0; i < 3; ++i) c->dims[i] = dims[i];
  c->ncells = K.ncells;
  if (grid_changed) {
    c->lists_valid = false;
    if ((rc = check_hip(hipStreamSynchronize(c->stream), "sync")) || (rc = alloc_grid(c))) return rc;
  }
  return PBF_OK;
}

int pbf_get_params(const pbf_context* c, pbf_params* out) {
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  *out = c->P;
  return PBF_OK;
}

uint64_t pbf_local_count(const pbf_context* c) { return c ? c->n : 0; }
uint64_t pbf_available(const pbf_context* c) { return c ? c->capacity - c->n : 0; }
uint64_t pbf_max_local_count(const pbf_context* c) { return c ? c->capacity : 0; }

int pbf_add(pbf_context* c, const float* pos, const float* pos_star, const float* vel, uint64_t count) {
  PBF_TRACE("pbf_add");
  if (!c || (count && (!pos || !pos_star || !vel))) return set_error(PBF_ERR_ARG, "null argument");
  if (c->n + count > c->capacity) return set_error(PBF_ERR_CAPACITY, "Not enough capacity to push_back");
  if (!count) return PBF_OK;
  const uint64_t first = c->n;
  const uint32_t nb = (uint32_t)((count + 255) / 256);
  // one contiguous upload per array into the staging buffer, then a kernel
  // widens Vec<float,3> to the float4 layout (a pitched host-to-device copy
  // of 12-byte rows is far slower)
  const float* src[3] = {pos, pos_star, vel};
  float4* dst[3] = {c->x, c->xs, c->v};
  for (int a = 0; a < 3; ++a) {
    int rc = check_hip(hipMemcpyAsync(c->stage, src[a], count * 3 * sizeof(float), hipMemcpyHostToDevice,
                                      c->stream), "add");
    if (rc) return rc;
    hipLaunchKernelGGL(k_widen3, dim3(nb), dim3(256), 0, c->stream, (const float*)c->stage, dst[a], first, count);
  }
  hipLaunchKernelGGL(k_iota, dim3(nb), dim3(256), 0, c->stream, c->oid, first, count);
  hipLaunchKernelGGL(k_zero_f4, dim3(nb), dim3(256), 0, c->stream, c->scr, first, count);
  hipLaunchKernelGGL(k_zero_f1, dim3(nb), dim3(256), 0, c->stream, c->rho, first, count);
  hipLaunchKernelGGL(k_zero_f1, dim3(nb), dim3(256), 0, c->stream, c->lam, first, count);
  hipLaunchKernelGGL(k_zero_f1, dim3(nb), dim3(256), 0, c->stream, c->lsum, first, count);
  hipLaunchKernelGGL(k_zero_f1, dim3(nb), dim3(256), 0, c->stream, (float*)c->nbr_count, first, count);
  if (c->K.batched) launch_fill_u32(c->sid + first, 0u, count, c->stream);
  c->n += count;
  c->lists_valid = false;
  return check_hip(hipStreamSynchronize(c->stream), "pbf_add");
}

int pbf_add_to_scene(pbf_context* c, uint32_t scene, const float* pos, const float* pos_star, const float* vel,
                     uint64_t count) {
  PBF_TRACE("pbf_add_to_scene");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (scene >= c->nscenes) return set_error(PBF_ERR_ARG, "scene index out of range");
  const uint64_t first = c->n;
  int rc = pbf_add(c, pos, pos_star, vel, count);
  if (rc || !count || !c->K.batched) return rc;
  launch_fill_u32(c->sid + first, scene, count, c->stream);
  return check_hip(hipStreamSynchronize(c->stream), "pbf_add_to_scene");
}

// Independent scenes in one context: each scene's box is [0, boxes[3s..3s+2]]
// and its grid is derived from it as for a single scene; all other
// parameters are shared.
int pbf_create_batched(const pbf_params* params, uint32_t num_scenes, const float* boxes, pbf_context** out) {
  PBF_TRACE("pbf_create_batched");
  if (!params || !boxes || !out) return set_error(PBF_ERR_ARG, "null argument");
  *out = nullptr;
  if (num_scenes == 0) return set_error(PBF_ERR_ARG, "num_scenes must be positive");
  std::vector<SceneGeo> geo;
  uint64_t total = 0;
  if (int rc = build_scene_geo(*params, boxes, num_scenes, geo, total)) return rc;
  pbf_params p0 = *params;
  for (int i = 0; i < 3; ++i) p0.boundary_max[i] = boxes[i];
  pbf_context* c = nullptr;
  int rc = pbf_create(&p0, &c);
  if (rc) return rc;
  const uint64_t N = c->capacity;
  c->nscenes = num_scenes;
  c->scene_geo = geo;
  if ((rc = dalloc(&c->scenes, (uint64_t)num_scenes)) || (rc = dalloc(&c->sid, N)) || (rc = dalloc(&c->sida, N)) ||
      (rc = check_hip(hipMemcpy(c->scenes, geo.data(), num_scenes * sizeof(SceneGeo), hipMemcpyHostToDevice),
                      "upload"))) {
    pbf_destroy(c);
    return rc;
  }
  c->K.ncells = (uint32_t)total;
  c->ncells = total;
  c->K.batched = 1;
  c->K.scenes = c->scenes;
  c->K.sid = c->sid;
  if ((rc = alloc_grid(c))) {
    pbf_destroy(c);
    return rc;
  }
  *out = c;
  return PBF_OK;
}

uint32_t pbf_num_scenes(const pbf_context* c) { return c ? c->nscenes : 0; }

int pbf_remove(pbf_context* c, uint64_t count) {
  PBF_TRACE("pbf_remove");
  if (!c) return set_error(PBF_ERR_ARG, "null argument");
  if (count > c->n) return set_error(PBF_ERR_ARG, "remove: count exceeds local count");
  if (!count) return PBF_OK;
  const uint32_t n = (uint32_t)c->n, newn = (uint32_t)(c->n - count);
  const uint32_t nb = (n + 255) / 256;
  // flags -> positions (exclusive scan) -> perm -> gather
  hipLaunchKernelGGL(k_keep, dim3(nb), dim3(256), 0, c->stream, n, (const uint32_t*)c->oid, newn, c->tmp);
  launch_exclusive_scan(c->tmp, c->tmpo, n, c->scan_partials, c->stream);
  hipLaunchKernelGGL(k_compact_perm, dim3(nb), dim3(256), 0, c->stream, n, (const uint32_t*)c->oid, newn,
                     (const uint32_t*)c->tmpo, c->slot);
  c->perm = c->slot;
  c->n = newn;
  launch_reorder(c);
  c->lists_valid = false;
  return check_hip(hipStreamSynchronize(c->stream), "pbf_remove");
}

// ---- getters / setters in original order ----
static int get_vec(pbf_context* c, const float4* src, float* out) {
  PBF_TRACE("pbf_get_state");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->n) return PBF_OK;
  launch_gather_out_f4(c, src, c->stage, 3);
  int rc = check_hip(hipMemcpyAsync(out, c->stage, c->n * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                     "download");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(c->stream), "download");
}
static int get_scalar(pbf_context* c, const float* src, float* out) {
  PBF_TRACE("pbf_get_state");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->n) return PBF_OK;
  launch_gather_out_f1(c, src, c->stage);
  int rc = check_hip(hipMemcpyAsync(out, c->stage, c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                     "download");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(c->stream), "download");
}
static int set_vec(pbf_context* c, float4* dst, const float* in) {
  PBF_TRACE("pbf_set_state");
  if (!c || !in) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->n) return PBF_OK;
  int rc = check_hip(hipMemcpyAsync(c->stage, in, c->n * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream),
                     "upload");
  if (rc) return rc;
  launch_scatter_in_f4(c, c->stage, dst);
  return check_hip(hipStreamSynchronize(c->stream), "upload");
}
static int set_scalar(pbf_context* c, float* dst, const float* in) {
  PBF_TRACE("pbf_set_state");
  if (!c || !in) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->n) return PBF_OK;
  int rc = check_hip(hipMemcpyAsync(c->stage, in, c->n * sizeof(float), hipMemcpyHostToDevice, c->stream), "upload");
  if (rc) return rc;
  launch_scatter_in_f1(c, c->stage, dst);
  return check_hip(hipStreamSynchronize(c->stream), "upload");
}

int pbf_get_positions(pbf_context* c, float* out) { return get_vec(c, c ? c->x : nullptr, out); }
int pbf_get_position_stars(pbf_context* c, float* out) { return get_vec(c, c ? c->xs : nullptr, out); }
int pbf_get_velocities(pbf_context* c, float* out) { return get_vec(c, c ? c->v : nullptr, out); }
int pbf_get_scratch(pbf_context* c, float* out) { return get_vec(c, c ? c->scr : nullptr, out); }
int pbf_get_densities(pbf_context* c, float* out) { return get_scalar(c, c ? c->rho : nullptr, out); }
int pbf_get_lambdas(pbf_context* c, float* out) { return get_scalar(c, c ? c->lam : nullptr, out); }
int pbf_set_positions(pbf_context* c, const float* in) { return set_vec(c, c ? c->x : nullptr, in); }
int pbf_set_position_stars(pbf_context* c, const float* in) { return set_vec(c, c ? c->xs : nullptr, in); }
int pbf_set_velocities(pbf_context* c, const float* in) { return set_vec(c, c ? c->v : nullptr, in); }
int pbf_set_scratch(pbf_context* c, const float* in) { return set_vec(c, c ? c->scr : nullptr, in); }
int pbf_set_densities(pbf_context* c, const float* in) { return set_scalar(c, c ? c->rho : nullptr, in); }
int pbf_set_lambdas(pbf_context* c, const float* in) { return set_scalar(c, c ? c->lam : nullptr, in); }

// ---- passes ----
#define PBF_CHECK_CTX_SPAN(c, s)                                            \
  do {                                                                      \
    if (!(c)) return set_error(PBF_ERR_ARG, "null context");                \
    if (!span_ok((c), (s))) return set_error(PBF_ERR_ARG, "span out of range"); \
  } while (0)

int pbf_apply_external_forces(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_apply_external_forces");
  PBF_CHECK_CTX_SPAN(c, s);
  launch_forces_predict(c, F_FORCES, s.begin, s.end);
  return check_hip(hipGetLastError(), "apply_external_forces");
}

int pbf_predict_positions(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_predict_positions");
  PBF_CHECK_CTX_SPAN(c, s);
  launch_forces_predict(c, F_PREDICT, s.begin, s.end);
  return check_hip(hipGetLastError(), "predict_positions");
}

int pbf_find_neighbors(pbf_context* c, pbf_index_span bin_span, pbf_index_span fill_span) {
  PBF_TRACE("pbf_find_neighbors");
  PBF_CHECK_CTX_SPAN(c, bin_span);
  PBF_CHECK_CTX_SPAN(c, fill_span);
  if (bin_span.begin != 0)
    return set_error(PBF_ERR_ARG, "bin span must start at 0: sort_bins sorts [0, count) (neighbors.h:130-132)");
  launch_fill_u64(c->counters, 0ull, 3, c->stream);  // overflow, outside, hacks
  find_neighbors_impl(c, bin_span.end, fill_span.begin, fill_span.end);
  return check_hip(hipGetLastError(), "find_neighbors");
}

int pbf_compute_densities(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_compute_densities");
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_density(c, s.begin, s.end);
  std::swap(c->xs, c->xt);  // the pass's x* (with the near-coincidence fix) becomes current
  return check_hip(hipGetLastError(), "compute_densities");
}

int pbf_compute_pressure_lambdas(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_compute_pressure_lambdas");
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_lambda(c, c->xs, s.begin, s.end);
  return check_hip(hipGetLastError(), "compute_pressure_lambdas");
}

int pbf_compute_pressure_dps(pbf_context* c, pbf_index_span s, int substep) {
  PBF_TRACE("pbf_compute_pressure_dps");
  (void)substep;  // unused by the reference as well (particles.h:348)
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_dp(c, s.begin, s.end);
  return check_hip(hipGetLastError(), "compute_pressure_dps");
}

int pbf_update_position_stars(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_update_position_stars");
  PBF_CHECK_CTX_SPAN(c, s);
  launch_update_position_stars(c, s.begin, s.end);
  return check_hip(hipGetLastError(), "update_position_stars");
}

int pbf_update_velocities(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_update_velocities");
  PBF_CHECK_CTX_SPAN(c, s);
  launch_update_velocities(c, s.begin, s.end);
  return check_hip(hipGetLastError(), "update_velocities");
}

int pbf_apply_surface_tension(pbf_context* c, pbf_index_span cs, pbf_index_span ts) {
  PBF_TRACE("pbf_apply_surface_tension");
  PBF_CHECK_CTX_SPAN(c, cs);
  PBF_CHECK_CTX_SPAN(c, ts);
  if (int rc = need_lists(c)) return rc;
  launch_surface_tension(c, cs.begin, cs.end, ts.begin, ts.end);
  return check_hip(hipGetLastError(), "apply_surface_tension");
}

int pbf_apply_viscosity(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_apply_viscosity");
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_viscosity(c, c->v, c->va, 0, s.begin, s.end);
  std::swap(c->v, c->va);
  return check_hip(hipGetLastError(), "apply_viscosity");
}

int pbf_compute_vorticity(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_compute_vorticity");
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_vorticity(c, c->v, s.begin, s.end);
  return check_hip(hipGetLastError(), "compute_vorticity");
}

int pbf_apply_vorticity(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_apply_vorticity");
  PBF_CHECK_CTX_SPAN(c, s);
  if (int rc = need_lists(c)) return rc;
  launch_apply_vorticity(c, c->v, s.begin, s.end);
  return check_hip(hipGetLastError(), "apply_vorticity");
}

int pbf_update_positions(pbf_context* c, pbf_index_span s) {
  PBF_TRACE("pbf_update_positions");
  PBF_CHECK_CTX_SPAN(c, s);
  launch_update_positions(c, s.begin, s.end);
  return check_hip(hipGetLastError(), "update_positions");
}

// The part of the step after the pressure solve (Simulation/Source/main.cpp:108-124).
static void step_tail(pbf_context* c) {
  const uint64_t n = c->n;
  if (c->P.gamma != 0.f) launch_surface_tension(c, 0, n, 0, n);
  launch_viscosity_first(c);         // XSPH 1: v -> va, compact lists of the final positions
  launch_vorticity_masked(c);        // omega -> scr
  launch_apply_vorticity_masked(c);  // va += ...
  launch_viscosity_last(c);          // XSPH 2: va -> v, x = x*
}

// Whether the Jacobi iterations go to one cooperative launch: asked for, not
// while a graph is being captured, and the particles fit one launch.
static bool persistent_jacobi(pbf_context* c) {
  return (c->P.options & PBF_OPT_PERSISTENT) && !c->capturing && c->n && persistent_fits(c);
}

// PBF_ERR_UNSUPPORTED when the persistent Jacobi kernel is asked for and the
// device cannot launch it.
static int check_persistent(pbf_context* c) {
  if ((c->P.options & PBF_OPT_PERSISTENT) && !persistent_available(c))
    return set_error(PBF_ERR_UNSUPPORTED, "PBF_OPT_PERSISTENT: no cooperative launches on this device");
  return PBF_OK;
}

// Simulation/Source/main.cpp:71-124 on one domain, fused kernels.
static void step_body(pbf_context* c) {
  const uint64_t n = c->n;
  {
    // the multiplier sums move with the reorder inside find_neighbors
    if (c->capturing) {  // stamps repeat on every replay
      launch_fill_u32(c->dirty, 0u, n, c->stream);
      launch_fill_u32(c->mflag, 0u, n, c->stream);
    }
    if (!n) launch_fill_u64(c->counters, 0ull, C_NUM, c->stream);
    launch_forces_predict(c, F_FORCES | F_PREDICT | F_CLEAR, 0, n);  // (counters zeroed there)
    find_neighbors_impl(c, n, 0, n, true);
    if (c->track_lsum) launch_fill_u32((uint32_t*)c->lsum, 0u, n, c->stream);
    const SolveCtl S{-1.0f, 0, 0, nullptr, nullptr, c->track_lsum ? c->lsum : nullptr, nullptr};
    const uint32_t I = c->P.solve_step_count;
    c->coop_grid = 0;
    if (I && persistent_jacobi(c)) launch_jacobi_persistent(c, I, false, true, S);
    else for (uint32_t it = 0; it < I; ++it) {
      ++c->dirty_stamp;
      launch_density_lambda_masked(c, (int)it, 0, S);       // xs -> xt, rho, lam, compact lists
      launch_mark_dirty(c, (int)it, S);                     // only if the density fix fired
      launch_lambda_repair(c, (int)it, S);
      launch_dp_update_masked(c, (int)it, it + 1 == I, S);  // xt, lam -> xs (+ v)
    }
    if (I == 0) launch_update_velocities(c, 0, n);
    c->last_iterations = I;
    c->last_was_solve = false;
    step_tail(c);
  }
}

int pbf_step(pbf_context* c, uint32_t nsteps) {
  PBF_TRACE("pbf_step");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (int rc = check_persistent(c)) return rc;
  for (uint32_t st = 0; st < nsteps; ++st) step_body(c);
  return check_hip(hipGetLastError(), "pbf_step");
}

// Two fused steps captured into one HIP graph and replayed.  Two, because a
// step ends by swapping the reordered buffers on the host: after two steps
// every buffer is back where the captured launches expect it.  An odd step
// count ends with one ordinary step.  The graph is captured again after
// anything that changes a step's launches (particle count, parameters).
int pbf_step_graph(pbf_context* c, uint32_t nsteps) {
  PBF_TRACE("pbf_step_graph");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (int rc = check_persistent(c)) return rc;  // an odd count's last step
  int rc;
  if (nsteps >= 2 && (!c->graph_exec || c->graph_n != c->n || c->graph_version != c->version ||
                      c->graph_sig != buffer_sig(c))) {
    drop_graph(c);
    c->capturing = true;
    rc = check_hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!rc) {
      step_body(c);
      step_body(c);
    }
    const int rc2 = check_hip(hipStreamEndCapture(c->stream, &c->graph), "hipStreamEndCapture");
    c->capturing = false;
    if (rc || (rc = rc2)) {
      drop_graph(c);
      return rc;
    }
    if ((rc = check_hip(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
      drop_graph(c);
      return rc;
    }
    c->graph_n = c->n;
    c->graph_version = c->version;
    c->graph_sig = buffer_sig(c);
  }
  uint32_t st = 0;
  for (; st + 2 <= nsteps; st += 2)
    if ((rc = check_hip(hipGraphLaunch(c->graph_exec, c->stream), "hipGraphLaunch"))) return rc;
  if (st < nsteps) step_body(c);
  return check_hip(hipGetLastError(), "pbf_step_graph");
}

static int solve_check_args(const pbf_context* c, float tolerance, uint32_t max_iterations, int warm_start) {
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (!(tolerance >= 0.0f)) return set_error(PBF_ERR_ARG, "tolerance must be >= 0");
  if (max_iterations < 1 || max_iterations > kMaxSolveSlots)
    return set_error(PBF_ERR_ARG, "max_iterations must be in [1, PBF_MAX_SOLVE_ITERATIONS]");
  if (warm_start < PBF_WARM_NONE || warm_start > PBF_WARM_LEARNED)
    return set_error(PBF_ERR_ARG, "unknown warm start");
  if (warm_start == PBF_WARM_GIVEN && !c->has_given)
    return set_error(PBF_ERR_STATE, "PBF_WARM_GIVEN without pbf_set_warm_lambdas");
  if (warm_start == PBF_WARM_LEARNED && !c->has_model)
    return set_error(PBF_ERR_STATE, "PBF_WARM_LEARNED without pbf_set_warm_model");
  return PBF_OK;
}

// One step with an iterate-to-tolerance pressure solve, queued on the stream
// (arguments checked by solve_check_args).  Every slot's kernels are queued;
// once a slot measures an error <= tol the later slots' kernels return at
// once.  With host_chunks the host reads the done flag after every 8 slots
// and stops queueing once the solve has ended (a finished solve's kernels
// still cost their dispatch); without it nothing waits for the device, which
// is what graph capture needs.
static int solve_body(pbf_context* c, float tolerance, uint32_t max_iterations, int warm_start, bool host_chunks) {
  const uint64_t n = c->n;
  const int warm = warm_start != PBF_WARM_NONE;
  const bool had_lsum = c->track_lsum;
  c->track_lsum = true;
  launch_fill_u64(c->counters, 0ull, C_NUM, c->stream);
  if (c->capturing) {  // stamps repeat on every replay
    launch_fill_u32(c->dirty, 0u, n, c->stream);
    launch_fill_u32(c->mflag, 0u, n, c->stream);
  }
  launch_fill_u64(c->counters + C_ITERS, (unsigned long long)max_iterations, 1, c->stream);
  launch_fill_u64(c->counters + C_WARM_KEPT, (unsigned long long)warm, 1, c->stream);
  launch_fill_u32(c->ehist, 0u, (uint64_t)max_iterations + 1, c->stream);
  launch_fill_u64(c->esum, 0ull, (uint64_t)max_iterations + 1, c->stream);
  launch_forces_predict(c, F_FORCES | F_PREDICT, 0, n);
  find_neighbors_impl(c, n, 0, n, true);  // carries the previous sums to the new order
  // lsuma: the previous sums; lsum: this solve's
  std::swap(c->lsum, c->lsuma);
  if (!had_lsum) launch_fill_u32((uint32_t*)c->lsuma, 0u, n, c->stream);
  launch_fill_u32((uint32_t*)c->lsum, 0u, n, c->stream);
  if (warm_start == PBF_WARM_PREVIOUS) launch_copy_f1(c->lsuma, c->wlam, n, c->stream);
  else if (warm_start == PBF_WARM_GIVEN) launch_scatter_in_f1(c, c->wgiven, c->wlam);
  const int mean = (c->P.options & PBF_OPT_SOLVE_MEAN_ERROR) ? 1 : 0;
  const int strict = (c->P.options & PBF_OPT_WARM_STRICT) ? 1 : 0;
  const int retry = (warm && (c->P.options & PBF_OPT_WARM_RETRY)) ? 1 : 0;
  const SolveCtl S{tolerance, warm, mean, c->ehist, c->esum, c->lsum, c->xsave, strict, retry};
  if (warm) {
    // first density pass (its error is the reference for the fallback; it
    // writes the features), then the warm update
    ++c->dirty_stamp;
    // writes the features, or with the learned warm start evaluates the
    // network on them in the same pass), then the warm update
    const SolveCtl S0{-1.0f, 0, mean, c->ehist, c->esum, nullptr, nullptr};
    const bool keep = (c->P.options & PBF_OPT_KEEP_WARM_FEATURES) != 0;
    FeatCtl F{keep ? c->feat : nullptr, c->lsuma, nullptr, nullptr, nullptr, 0.0f, nullptr};
    if (warm_start == PBF_WARM_LEARNED) {
      F.w1 = c->mw1;
      F.w2 = c->mw2;
      F.mp = c->mp;
      F.scale = c->warm_gain * kWarmLambdaScale;
      F.wlam = c->wlam;
    }
    launch_density_lambda_masked(c, 1, 0, S0, F);
    c->feat_valid = keep;
    launch_warm_dp(c);
  }
  constexpr uint32_t kChunk = 8;
  c->coop_grid = 0;
  if (persistent_jacobi(c) && !retry) launch_jacobi_persistent(c, max_iterations, true, false, S);
  else for (uint32_t it = 0; it < max_iterations; ++it) {
    ++c->dirty_stamp;
    launch_density_lambda_masked(c, (int)it, (int)it + 1, S);
    launch_mark_dirty(c, (int)it, S);
    launch_lambda_repair(c, (int)it, S);
    launch_dp_update_masked(c, (int)it, 0, S);
    if (retry && it == 0) launch_warm_retry(c);  // (does nothing unless slot 0 rejected the warm start)
    if (host_chunks && (it + 1) % kChunk == 0 && it + 1 < max_iterations) {
      int rc = check_hip(hipMemcpyAsync(c->done_host, c->counters + C_DONE, sizeof(unsigned long long),
                                        hipMemcpyDeviceToHost, c->stream), "pbf_solve");
      if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "pbf_solve");
      if (rc) return rc;
      if (*c->done_host) break;
    }
  }
  if (warm) launch_add_warm(c);
  launch_update_velocities(c, 0, n);
  launch_log_solve(c);
  c->last_iterations = max_iterations;
  c->last_was_solve = true;
  step_tail(c);
  return PBF_OK;
}

// One step with an iterate-to-tolerance pressure solve, queued slot chunk by
// slot chunk (the host stops queueing once the device reports the solve done).
int pbf_solve(pbf_context* c, float tolerance, uint32_t max_iterations, int warm_start) {
  PBF_TRACE("pbf_solve");
  int rc = solve_check_args(c, tolerance, max_iterations, warm_start);
  if (rc) return rc;
  if ((rc = check_persistent(c))) return rc;
  if ((rc = solve_body(c, tolerance, max_iterations, warm_start, true))) return rc;
  return check_hip(hipGetLastError(), "pbf_solve");
}

// Two iterate-to-tolerance steps captured into one HIP graph and replayed
// (two for the same reason as pbf_step_graph: each step swaps the reordered
// buffers and the multiplier sums on the host).  Nothing in a captured step
// depends on the host once it is queued.
int pbf_solve_graph(pbf_context* c, float tolerance, uint32_t max_iterations, int warm_start, uint32_t nsteps) {
  PBF_TRACE("pbf_solve_graph");
  int rc = solve_check_args(c, tolerance, max_iterations, warm_start);
  if (rc) return rc;
  if ((rc = check_persistent(c))) return rc;  // an odd count's last step
  if (!c->track_lsum) {
    // the captured steps read the previous sums: start them at zero here, so
    // that both captured steps see kept sums
    if (c->n) launch_fill_u32((uint32_t*)c->lsum, 0u, c->n, c->stream);
    c->track_lsum = true;
  }
  const bool stale = !c->sgraph_exec || c->sgraph_n != c->n || c->sgraph_version != c->version ||
                     c->sgraph_sig != buffer_sig(c) || c->sgraph_max_it != max_iterations ||
                     c->sgraph_warm != warm_start || std::memcmp(&c->sgraph_tol, &tolerance, 4) != 0 ||
                     std::memcmp(&c->sgraph_gain, &c->warm_gain, 4) != 0;
  if (nsteps >= 2 && stale) {
    drop_solve_graph(c);
    c->capturing = true;
    rc = check_hip(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!rc) rc = solve_body(c, tolerance, max_iterations, warm_start, false);
    if (!rc) rc = solve_body(c, tolerance, max_iterations, warm_start, false);
    const int rc2 = check_hip(hipStreamEndCapture(c->stream, &c->sgraph), "hipStreamEndCapture");
    c->capturing = false;
    if (rc || (rc = rc2)) {
      drop_solve_graph(c);
      return rc;
    }
    if ((rc = check_hip(hipGraphInstantiate(&c->sgraph_exec, c->sgraph, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
      drop_solve_graph(c);
      return rc;
    }
    c->sgraph_n = c->n;
    c->sgraph_version = c->version;
    c->sgraph_sig = buffer_sig(c);
    c->sgraph_max_it = max_iterations;
    c->sgraph_warm = warm_start;
    c->sgraph_tol = tolerance;
    c->sgraph_gain = c->warm_gain;
  }
  uint32_t st = 0;
  for (; st + 2 <= nsteps; st += 2)  // nsteps >= 2: the graph is current
    if ((rc = check_hip(hipGraphLaunch(c->sgraph_exec, c->stream), "hipGraphLaunch"))) return rc;
  if (st < nsteps && (rc = solve_body(c, tolerance, max_iterations, warm_start, false))) return rc;
  if (nsteps && warm_start != PBF_WARM_NONE) c->feat_valid = (c->P.options & PBF_OPT_KEEP_WARM_FEATURES) != 0;
  if (nsteps) {
    c->last_was_solve = true;
    c->last_iterations = max_iterations;
  }
  return check_hip(hipGetLastError(), "pbf_solve_graph");
}

int pbf_get_solve_log(pbf_context* c, pbf_solve_record* out, uint32_t count, uint32_t* returned) {
  PBF_TRACE("pbf_get_solve_log");
  if (!c || !returned || (count && !out)) return set_error(PBF_ERR_ARG, "null argument");
  *returned = 0;
  unsigned long long total = 0;
  int rc = check_hip(hipMemcpyAsync(&total, c->slog_count, sizeof(total), hipMemcpyDeviceToHost, c->stream),
                     "solve log");
  if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "solve log");
  if (rc) return rc;
  const uint64_t m = std::min<uint64_t>({(uint64_t)count, (uint64_t)total, (uint64_t)PBF_SOLVE_LOG_SIZE});
  // records total-m .. total-1 sit at (index % size): at most two runs of the ring
  uint64_t done = 0;
  while (done < m) {
    const uint64_t idx = total - m + done;
    const uint64_t at = idx % PBF_SOLVE_LOG_SIZE;
    const uint64_t run = std::min<uint64_t>(m - done, PBF_SOLVE_LOG_SIZE - at);
    if ((rc = check_hip(hipMemcpyAsync(out + done, c->slog + at, run * sizeof(pbf_solve_record),
                                       hipMemcpyDeviceToHost, c->stream), "solve log")))
      return rc;
    done += run;
  }
  if ((rc = check_hip(hipStreamSynchronize(c->stream), "solve log"))) return rc;
  *returned = (uint32_t)m;
  return PBF_OK;
}

int pbf_get_warm_features(pbf_context* c, float* out) {
  PBF_TRACE("pbf_get_warm_features");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->feat_valid)
    return set_error(PBF_ERR_STATE, "no warm-started pbf_solve with PBF_OPT_KEEP_WARM_FEATURES has run");
  if (!c->n) return PBF_OK;
  float* tmp = nullptr;
  int rc = check_hip(hipMallocAsync((void**)&tmp, c->n * PBF_WARM_FEATURES * sizeof(float), c->stream), "alloc");
  if (rc) return rc;
  launch_gather_out_rows(c, c->feat, tmp, PBF_WARM_FEATURES);
  rc = check_hip(hipMemcpyAsync(out, tmp, c->n * PBF_WARM_FEATURES * sizeof(float), hipMemcpyDeviceToHost,
                                c->stream), "download");
  hipFreeAsync(tmp, c->stream);
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(c->stream), "download");
}

int pbf_set_warm_lambdas(pbf_context* c, const float* lambdas) {
  PBF_TRACE("pbf_set_warm_lambdas");
  if (!c || !lambdas) return set_error(PBF_ERR_ARG, "null argument");
  if (c->n) {
    int rc = check_hip(hipMemcpyAsync(c->wgiven, lambdas, c->n * sizeof(float), hipMemcpyHostToDevice, c->stream),
                       "upload");
    if (rc) return rc;
    if ((rc = check_hip(hipStreamSynchronize(c->stream), "upload"))) return rc;
  }
  c->has_given = true;
  return PBF_OK;
}

int pbf_set_lambda_sums(pbf_context* c, const float* in) {
  PBF_TRACE("pbf_set_lambda_sums");
  if (!c || !in) return set_error(PBF_ERR_ARG, "null argument");
  if (c->n) {
    int rc = check_hip(hipMemcpyAsync(c->stage, in, c->n * sizeof(float), hipMemcpyHostToDevice, c->stream),
                       "upload");
    if (rc) return rc;
    launch_scatter_in_f1(c, c->stage, c->lsum);
    if ((rc = check_hip(hipStreamSynchronize(c->stream), "upload"))) return rc;
  }
  c->track_lsum = true;
  return PBF_OK;
}

int pbf_get_lambda_sums(pbf_context* c, float* out) {
  PBF_TRACE("pbf_get_lambda_sums");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->track_lsum) return set_error(PBF_ERR_STATE, "no pbf_solve has run");
  if (!c->n) return PBF_OK;
  launch_gather_out_f1(c, c->lsum, c->stage);
  int rc = check_hip(hipMemcpyAsync(out, c->stage, c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream),
                     "download");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(c->stream), "download");
}

int pbf_get_solve_info(pbf_context* c, pbf_solve_info* out) {
  PBF_TRACE("pbf_get_solve_info");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (!c->last_was_solve) return set_error(PBF_ERR_STATE, "the last step was not a pbf_solve");
  unsigned long long h[C_NUM];
  static_assert(kMaxSolveSlots + 1 <= 1024, "history size");
  uint32_t e[kMaxSolveSlots + 1];
  unsigned long long es[kMaxSolveSlots + 1];
  int rc = check_hip(hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream), "solve info");
  if (rc) return rc;
  if ((rc = check_hip(hipMemcpyAsync(e, c->ehist, sizeof(e), hipMemcpyDeviceToHost, c->stream), "solve info")) ||
      (rc = check_hip(hipMemcpyAsync(es, c->esum, sizeof(es), hipMemcpyDeviceToHost, c->stream), "solve info")))
    return rc;
  if ((rc = check_hip(hipStreamSynchronize(c->stream), "solve info"))) return rc;
  std::memset(out, 0, sizeof(*out));
  out->iterations = (uint32_t)h[C_ITERS];
  out->warm_start_kept = (int32_t)h[C_WARM_KEPT];
  const double n = (double)(c->n ? c->n : 1);
  auto mean = [&](unsigned long long s) { return (float)((double)s / 4294967296.0 / n); };
  std::memcpy(&out->max_error_before_warm_start, &e[0], 4);
  out->mean_error_before_warm_start = mean(es[0]);
  for (uint32_t k = 0; k < out->iterations && k < kMaxSolveSlots; ++k) {
    std::memcpy(&out->max_error_history[k], &e[k + 1], 4);
    out->mean_error_history[k] = mean(es[k + 1]);
  }
  return PBF_OK;
}

int pbf_set_tuning(pbf_context* c, const pbf_tuning* t) {
  PBF_TRACE("pbf_set_tuning");
  if (!c || !t) return set_error(PBF_ERR_ARG, "null argument");
  if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
    return set_error(PBF_ERR_ARG, "block_size must be 64, 128 or 256");
  if (t->heavy_budget < 64 || t->heavy_budget > 65536) return set_error(PBF_ERR_ARG, "heavy_budget out of range");
  if (t->heavy_blocks < 1 || t->heavy_blocks > 65536) return set_error(PBF_ERR_ARG, "heavy_blocks out of range");
  c->tune = *t;
  c->K.heavy_budget = t->heavy_budget;
  ++c->version;
  return PBF_OK;
}

int pbf_get_tuning(const pbf_context* c, pbf_tuning* out) {
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  *out = c->tune;
  return PBF_OK;
}

// Times every candidate from the same saved state (positions, velocities,
// original ids, and the scene ids and multiplier sums where kept), then
// restores that state: the next step sorts by (cell, original id) and
// recomputes everything else, so it is the step an untuned run would take.
// The state a step starts from (positions, velocities, original ids, and
// the scene ids and multiplier sums where kept), saved on the device so that
// candidates can each be timed from it: the next step sorts by (cell,
// original id) and recomputes everything else, so after restore() it is the
// step an untuned run would take.
struct SavedState {
  pbf_context* c = nullptr;
  uint64_t n = 0;
  float4 *x = nullptr, *v = nullptr;
  uint32_t *oid = nullptr, *sid = nullptr;
  float* lsum = nullptr;
  ~SavedState() {
    dfree(x);
    dfree(v);
    dfree(oid);
    dfree(sid);
    dfree(lsum);
  }
  int copy(void* d, const void* s, uint64_t bytes) {
    return check_hip(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream), "autotune copy");
  }
  int save(pbf_context* ctx) {
    c = ctx;
    n = c->n;
    int r;
    if ((r = dalloc(&x, n)) || (r = dalloc(&v, n)) || (r = dalloc(&oid, n)) || (r = dalloc(&sid, n)) ||
        (r = dalloc(&lsum, n)))
      return r;
    r = copy(x, c->x, n * 16);
    if (!r) r = copy(v, c->v, n * 16);
    if (!r) r = copy(oid, c->oid, n * 4);
    if (!r && c->K.batched) r = copy(sid, c->sid, n * 4);
    if (!r && c->track_lsum) r = copy(lsum, c->lsum, n * 4);
    return r;
  }
  int restore() {
    int r = copy(c->x, x, n * 16);
    if (!r) r = copy(c->v, v, n * 16);
    if (!r) r = copy(c->oid, oid, n * 4);
    if (!r && c->K.batched) r = copy(c->sid, sid, n * 4);
    if (!r && c->track_lsum) r = copy(c->lsum, lsum, n * 4);
    c->lists_valid = false;
    return r;
  }
};

// Mean ms of `steps` fused steps from the saved state (one untimed step first).
static int time_steps(pbf_context* c, SavedState& S, uint32_t steps, hipEvent_t a, hipEvent_t b, double* ms_out) {
  int rc;
  if ((rc = check_persistent(c))) return rc;
  if ((rc = S.restore())) return rc;
  step_body(c);  // first launches of this setting, untimed
  if ((rc = S.restore())) return rc;
  hipEventRecord(a, c->stream);
  for (uint32_t st = 0; st < steps; ++st) step_body(c);
  hipEventRecord(b, c->stream);
  if ((rc = check_hip(hipEventSynchronize(b), "autotune"))) return rc;
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  *ms_out = (double)ms / steps;
  return PBF_OK;
}

int pbf_autotune(pbf_context* c, uint32_t steps, pbf_tuning* best, double* best_ms) {
  PBF_TRACE("pbf_autotune");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (steps == 0) return set_error(PBF_ERR_ARG, "steps must be positive");
  if (!c->n) return set_error(PBF_ERR_STATE, "no particles");
  SavedState S;
  int rc = S.save(c);
  const pbf_tuning start = c->tune;
  pbf_tuning best_t = start;
  double best_v = 1e300;
  hipEvent_t a = nullptr, b = nullptr;
  if (!rc) rc = check_hip(hipEventCreate(&a), "hipEventCreate");
  if (!rc) rc = check_hip(hipEventCreate(&b), "hipEventCreate");
  const uint32_t blocks[] = {64u, 128u, 256u};
  const uint32_t budgets[] = {384u, 768u, 1536u};
  const uint32_t hblocks[] = {2048u, 8192u, 16384u};
  for (uint32_t bs : blocks)
    for (uint32_t hb : budgets)
      for (uint32_t hk : hblocks) {
        if (rc) break;
        const pbf_tuning t{bs, hb, hk};
        double ms = 0.0;
        if ((rc = pbf_set_tuning(c, &t)) || (rc = time_steps(c, S, steps, a, b, &ms))) break;
        if (ms < best_v) {
          best_v = ms;
          best_t = t;
        }
      }
  if (!rc) rc = S.restore();
  if (!rc) rc = pbf_set_tuning(c, &best_t);
  if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "autotune");
  if (rc) pbf_set_tuning(c, &start);
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (rc) return rc;
  if (best) *best = best_t;
  if (best_ms) *best_ms = best_v;
  return PBF_OK;
}

static int download_u32(pbf_context* c, const uint32_t* src, uint64_t count, std::vector<uint32_t>& out);

// Largest list length of the last list build (host read of the counts).
static int max_list_length(pbf_context* c, uint32_t* out) {
  std::vector<uint32_t> cnt;
  if (int rc = download_u32(c, (const uint32_t*)c->nbr_count, c->n, cnt)) return rc;
  uint32_t m = 0;
  for (uint32_t v : cnt) m = std::max(m, v);
  *out = m;
  return PBF_OK;
}

int pbf_autotune_layout(pbf_context* c, uint32_t steps, pbf_layout* best) {
  PBF_TRACE("pbf_autotune_layout");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  if (steps == 0) return set_error(PBF_ERR_ARG, "steps must be positive");
  if (!c->n) return set_error(PBF_ERR_STATE, "no particles");
  SavedState S;
  int rc = S.save(c);
  const pbf_params P0 = c->P;
  const float h = P0.smoothing_radius;
  pbf_layout best_l{P0.neighbor_bin_spacing, P0.max_neighbors, 1e300, 0u};
  bool found = false;
  hipEvent_t a = nullptr, b = nullptr;
  if (!rc) rc = check_hip(hipEventCreate(&a), "hipEventCreate");
  if (!rc) rc = check_hip(hipEventCreate(&b), "hipEventCreate");
  const float factors[] = {1.0f, 1.1f, 1.2f};
  for (float f : factors) {
    if (rc) break;
    const float spacing = h * f;
    // the longest list at this spacing with the largest cap the context holds,
    // the cap's rejections counted
    pbf_params P = P0;
    P.neighbor_bin_spacing = spacing;
    P.max_neighbors = (uint32_t)c->cap_alloc;
    P.options |= PBF_OPT_COUNT_OVERFLOW;
    if ((rc = pbf_set_params(c, &P)) || (rc = S.restore())) break;
    launch_fill_u64(c->counters, 0ull, C_NUM, c->stream);
    launch_forces_predict(c, F_FORCES | F_PREDICT, 0, c->n);
    find_neighbors_impl(c, c->n, 0, c->n);
    unsigned long long over = 0;
    uint32_t longest = 0;
    if ((rc = check_hip(hipMemcpyAsync(&over, c->counters + C_OVERFLOW, sizeof(over), hipMemcpyDeviceToHost,
                                       c->stream), "autotune")) ||
        (rc = max_list_length(c, &longest)))
      break;
    if (over != 0) continue;  // this spacing needs a longer list than the context holds
    // room for 25% longer lists than the state's longest, in rows of 8
    uint32_t cap = (uint32_t)std::ceil(1.25 * (double)longest);
    cap = std::min<uint32_t>((cap + 7u) / 8u * 8u, (uint32_t)c->cap_alloc);
    cap = std::max<uint32_t>(cap, std::max<uint32_t>(longest, 1u));
    P.max_neighbors = cap;
    P.options = P0.options;
    double ms = 0.0;
    if ((rc = pbf_set_params(c, &P)) || (rc = time_steps(c, S, steps, a, b, &ms))) break;
    if (ms < best_l.ms_per_step) {
      best_l = pbf_layout{spacing, cap, ms, longest};
      found = true;
    }
  }
  pbf_params chosen = P0;
  if (!rc && found) {
    chosen.neighbor_bin_spacing = best_l.neighbor_bin_spacing;
    chosen.max_neighbors = best_l.max_neighbors;
  }
  // whatever happened above, leave the context on a valid layout and its state
  const int rc2 = pbf_set_params(c, rc ? &P0 : &chosen);
  const int rc3 = S.restore();
  if (!rc) rc = rc2 ? rc2 : rc3;
  if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "autotune_layout");
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (rc) return rc;
  if (!found) return set_error(PBF_ERR_CAPACITY, "no bin spacing keeps every neighbour within the list cap");
  if (best) *best = best_l;
  return PBF_OK;
}

int pbf_sync(pbf_context* c) {
  PBF_TRACE("pbf_sync");
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  return check_hip(hipStreamSynchronize(c->stream), "pbf_sync");
}

int pbf_get_stats(pbf_context* c, pbf_stats* out) {
  PBF_TRACE("pbf_get_stats");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  unsigned long long h[C_NUM];
  int rc = check_hip(hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream), "stats");
  if (rc) return rc;
  std::vector<float> rho(c->n);
  if (c->n && (rc = check_hip(hipMemcpyAsync(rho.data(), c->rho, c->n * 4, hipMemcpyDeviceToHost, c->stream),
                              "stats")))
    return rc;
  if ((rc = check_hip(hipStreamSynchronize(c->stream), "stats"))) return rc;
  out->neighbor_overflow = h[C_OVERFLOW];
  out->outside_grid = h[C_OUTSIDE];
  out->density_hacks = h[C_HACKS];
  out->tile_listed = h[C_TILE];
  out->persistent_grid = c->coop_grid;
  out->persistent_capacity = (c->P.options & PBF_OPT_PERSISTENT) ? persistent_capacity(c) : 0;
  out->iterations = c->last_was_solve ? h[C_ITERS] : c->last_iterations;
  double m = 0.0;
  for (float r : rho) {
    const float e = r / c->P.rest_density - 1.0f;
    if ((double)e > m) m = (double)e;
  }
  out->max_density_error = m;
  return PBF_OK;
}

// ---- introspection ----
int pbf_grid_dims(const pbf_context* c, uint64_t dims[3]) {
  PBF_TRACE("pbf_grid_dims");
  if (!c || !dims) return set_error(PBF_ERR_ARG, "null argument");
  for (int i = 0; i < 3; ++i) dims[i] = c->dims[i];
  return PBF_OK;
}
uint64_t pbf_num_cells(const pbf_context* c) { return c ? c->ncells : 0; }

static int download_u32(pbf_context* c, const uint32_t* src, uint64_t count, std::vector<uint32_t>& out) {
  out.resize(count);
  if (!count) return PBF_OK;
  int rc = check_hip(hipMemcpyAsync(out.data(), src, count * 4, hipMemcpyDeviceToHost, c->stream), "download");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(c->stream), "download");
}

int pbf_get_bin_ids(pbf_context* c, uint64_t* out) {
  PBF_TRACE("pbf_get_bin_ids");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (int rc = need_lists(c)) return rc;
  std::vector<uint32_t> k;
  if (int rc = download_u32(c, c->key, c->n, k)) return rc;
  for (uint64_t i = 0; i < c->n; ++i) out[i] = k[i];
  return PBF_OK;
}

int pbf_get_particle_ids(pbf_context* c, uint64_t* out) {
  PBF_TRACE("pbf_get_particle_ids");
  if (!c || !out) return set_error(PBF_ERR_ARG, "null argument");
  if (int rc = need_lists(c)) return rc;
  std::vector<uint32_t> o;
  if (int rc = download_u32(c, c->oid, c->n, o)) return rc;
  for (uint64_t i = 0; i < c->n; ++i) out[i] = o[i];
  return PBF_OK;
}

int pbf_get_bin_bounds(pbf_context* c, uint64_t* begin, uint64_t* end) {
  PBF_TRACE("pbf_get_bin_bounds");
  if (!c || !begin || !end) return set_error(PBF_ERR_ARG, "null argument");
  if (int rc = need_lists(c)) return rc;
  std::vector<uint32_t> st;
  if (int rc = download_u32(c, c->cell_start, c->ncells + 1, st)) return rc;
  for (uint64_t i = 0; i < c->ncells; ++i) {
    begin[i] = st[i];
    end[i] = st[i + 1];
  }
  return PBF_OK;
}

int pbf_get_neighbor_lists(pbf_context* c, uint64_t* out, int32_t* counts) {
  PBF_TRACE("pbf_get_neighbor_lists");
  if (!c || !out || !counts) return set_error(PBF_ERR_ARG, "null argument");
  if (int rc = need_lists(c)) return rc;
  const uint64_t n = c->n, K = (uint64_t)c->cap;
  std::vector<uint32_t> o, cnt, lists;
  if (int rc = download_u32(c, c->oid, n, o)) return rc;
  if (int rc = download_u32(c, (const uint32_t*)c->nbr_count, n, cnt)) return rc;
  if (int rc = download_u32(c, (const uint32_t*)c->nbr, (n + 63) / 64 * 64 * c->K.stride, lists)) return rc;
  std::memset(out, 0, n * K * sizeof(uint64_t));
  for (uint64_t p = 0; p < n; ++p) {
    const uint64_t op = o[p];
    const int32_t cp = (int32_t)cnt[p];
    counts[op] = cp;
    for (int32_t k = 0; k < cp; ++k) out[op * K + k] = o[lists[list_at((uint32_t)k, (uint32_t)p, c->K.stride)]];
  }
  return PBF_OK;
}

// ---- primitives ----
int pbf_sort_by_key_u32(uint32_t* d_keys, uint32_t* d_values, uint64_t n, uint32_t key_bits) {
  PBF_TRACE("pbf_sort_by_key_u32");
  if (n && (!d_keys || !d_values)) return set_error(PBF_ERR_ARG, "null argument");
  int rc = radix_sort_u32(d_keys, d_values, n, key_bits, 0);
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(0), "sort_by_key");
}

int pbf_lower_bound_counting(const uint32_t* d_sorted, uint64_t n, pbf_index_span search, uint32_t* d_out) {
  PBF_TRACE("pbf_lower_bound_counting");
  if (search.end < search.begin || (search.end > search.begin && (!d_out || (n && !d_sorted))))
    return set_error(PBF_ERR_ARG, "bad argument");
  launch_bounds_counting(d_sorted, n, search.begin, search.end, d_out, 0, 0);
  return check_hip(hipStreamSynchronize(0), "lower_bound");
}

int pbf_upper_bound_counting(const uint32_t* d_sorted, uint64_t n, pbf_index_span search, uint32_t* d_out) {
  PBF_TRACE("pbf_upper_bound_counting");
  if (search.end < search.begin || (search.end > search.begin && (!d_out || (n && !d_sorted))))
    return set_error(PBF_ERR_ARG, "bad argument");
  launch_bounds_counting(d_sorted, n, search.begin, search.end, d_out, 1, 0);
  return check_hip(hipStreamSynchronize(0), "upper_bound");
}

int pbf_device_alloc(uint64_t bytes, void** out) {
  if (!out) return set_error(PBF_ERR_ARG, "null argument");
  return check_hip(hipMalloc(out, bytes ? bytes : 1), "hipMalloc");
}
int pbf_device_free(void* p) { return p ? check_hip(hipFree(p), "hipFree") : PBF_OK; }
int pbf_memcpy_to_device(void* dst, const void* src, uint64_t bytes) {
  PBF_TRACE("pbf_memcpy_to_device");
  return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
}
int pbf_memcpy_to_host(void* dst, const void* src, uint64_t bytes) {
  PBF_TRACE("pbf_memcpy_to_host");
  return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
}
int pbf_memcpy_device(void* dst, const void* src, uint64_t bytes) {
  PBF_TRACE("pbf_memcpy_device");
  return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy");
}

// ---- measurement ----
int pbf_profile_kernels(pbf_context* c, const char* names) {
  if (!c || !names) return set_error(PBF_ERR_ARG, "null argument");
  hipStreamSynchronize(c->stream);
  timers_release(c);
  c->timers.clear();
  ++c->version;
  std::string all(names);
  size_t pos = 0;
  while (pos <= all.size()) {
    size_t comma = all.find(',', pos);
    if (comma == std::string::npos) comma = all.size();
    const std::string nm = all.substr(pos, comma - pos);
    if (!nm.empty()) {
      KernelTimer t;
      t.name = nm;
      c->timers.push_back(t);
    }
    pos = comma + 1;
  }
  return PBF_OK;
}

int pbf_kernel_time(pbf_context* c, const char* name, uint64_t* launches, double* total_ms) {
  if (!c || !name || !launches || !total_ms) return set_error(PBF_ERR_ARG, "null argument");
  timers_resolve(c);
  for (auto& t : c->timers) {
    if (t.name == name) {
      *launches = t.launches;
      *total_ms = t.total_ms;
      return PBF_OK;
    }
  }
  return set_error(PBF_ERR_ARG, std::string("kernel not profiled: ") + name);
}

int pbf_region_begin(pbf_context* c) {
  if (!c) return set_error(PBF_ERR_ARG, "null context");
  return check_hip(hipEventRecord(c->region_a, c->stream), "hipEventRecord");
}

int pbf_region_end(pbf_context* c, double* ms) {
  if (!c || !ms) return set_error(PBF_ERR_ARG, "null argument");
  int rc = check_hip(hipEventRecord(c->region_b, c->stream), "hipEventRecord");
  if (rc) return rc;
  if ((rc = check_hip(hipEventSynchronize(c->region_b), "hipEventSynchronize"))) return rc;
  float f = 0.f;
  rc = check_hip(hipEventElapsedTime(&f, c->region_a, c->region_b), "hipEventElapsedTime");
  *ms = f;
  return rc;
}

}  // extern "C"

namespace pbf {

static void find_neighbors_impl(pbf_context* c, uint64_t bin_end, uint64_t fb, uint64_t fe, bool skin) {
  launch_bin(c, bin_end);  // (clears the list build's counters)
  launch_sort(c);
  // a fused step recomputes densities, multipliers and the scratch vector
  // before reading them (with at least one Jacobi iteration): they stay unsorted
  launch_reorder(c, !(skin && c->P.solve_step_count > 0));
  launch_fill_neighbors(c, fb, fe, skin);
  c->lists_valid = true;
}

}  // namespace pbf

extern "C" int pbf_append_positions_record(pbf_context* c, const char* path) {
  PBF_TRACE("pbf_append_positions_record");
  if (!c || !path) return set_error(PBF_ERR_ARG, "null argument");
  std::vector<float> pos(c->n * 3);
  if (int rc = pbf_get_positions(c, pos.data())) return rc;
  FILE* f = std::fopen(path, "ab");
  if (!f) return set_error(PBF_ERR_ARG, std::string("cannot open ") + path);
  const int64_t local_bytes = (int64_t)(c->n * 3 * sizeof(float));
  const int64_t header[3] = {local_bytes, local_bytes, 0};  // global, local, offset (one domain)
  bool ok = std::fwrite(header, sizeof(header), 1, f) == 1;
  if (ok && local_bytes) ok = std::fwrite(pos.data(), (size_t)local_bytes, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? PBF_OK : set_error(PBF_ERR_ARG, std::string("write failed: ") + path);
}
