// pbf_internal.h -- shared declarations of the HIP PBF library (not part of the ABI).
//
// Device data layout (one context, n particles, C cells, K = list cap):
//   x, xs, v         float4[n]   position, position star, velocity, stored in cell order (sorted by
//                                (cell key, original index)); from the velocity update on, v.w holds the
//                                density of the last density pass (gathered by the XSPH passes)
//   xa, xsa, va      float4[n]   reorder targets, swapped with the above
//   xt               float4[n]   position stars after the density pass (hack applied)
//   scr              float4[n]   scratch vector: dp, vorticity or colour field
//   rho, lam         float[n]
//   oid, oida        uint32[n]   original index of each sorted slot
//   key, keya        uint32[n]   cell key of each slot
//   slot             uint32[n]   inverse permutation (original index -> slot)
//   sk[2], sv[2]     uint32[n]   radix-sort ping-pong buffers (cell key, slot)
//   rhist, roffs     uint32[256 * tiles]  radix digit histograms and their scan
//   cell_start       uint32[C+2] first sorted slot of every cell (C = the outside bucket)
//   heavy            uint32[n]   particles whose 27 cells hold too many candidates
//                                for one lane (built by a whole wavefront instead)
//   nbr              int32[n/64][K][64] neighbour lists tiled by groups of 64 particles:
//                                nbr[list_at(k, p, K)] (entry k of a wavefront's 64 particles
//                                is one contiguous 256-byte read, its whole list one block)
//   nbr_count        int32[n]
//   cnb, cnb_count   int32[n/64][K][64], int32[n]  compact lists of the entries within h
//                                (same tiling; built by the first XSPH pass)
//   snb, snb_count   int32[n/64][K][64], int32[n]  skin lists: the entries within 1.1 h of the
//                                positions the lists were built from (xs0), in list order, each
//                                slot packed with its place in the full list (bits 25-31)
//   counters         u64[16]     overflow, outside, hacks, hack flags, solve state, ...
//   ehist            uint32[kMaxSolveSlots+1]  max density error per solve slot (float bits)
//   lsum, lsuma      float[n]    multipliers summed over the updates of the last solve
//                                (carried through the reorder while a warm start may use them)
//   wlam, wgiven     float[n]    warm-start multipliers (slot order / original order)
//   xsave            float4[n]   position stars before a warm start (for its fallback)
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pbf.h"

namespace pbf {

// One scene of a batched context: its box (from the origin) and its grid,
// whose cells take the global keys [off, off + ncells).
struct SceneGeo {
  uint32_t dimx, dimy, dimz, off;
  uint32_t ncells, pad;
  uint64_t dxy;
  float bmax[3], pad2;
};

// Constants of one step, computed on the host exactly as the reference
// computes them (kernels.h norms in double, parameters.h derivations).
struct KParams {
  float h, hh, s, ss;            // smoothing radius, h*h, bin spacing, s*s
  float T_h;                     // smallest float d2 with sqrtf(d2) > h
  float T_hack;                  // smallest float d2 with (double)sqrtf(d2) >= 1e-8
  float poly6_norm, spiky_norm, cspline_norm, r_eps;
  float mass, mass_W0;           // rest mass, rest_mass * W(0)
  float rho0, neg_inv_rho0, inv_rho0;
  float g, dt, g_dt;
  float eps, visc, vort, gamma;
  double vthr;                   // 0.000001 * max_speed (particles.h:380), in double
  double st_rmin;                // h * 0.000001 (particles.h:420), in double
  float bmin[3], bmax[3], mover[3];
  uint32_t dimx, dimy, dimz;
  uint32_t ncells;
  uint64_t dxy;
  int cap;
  int stride;                    // list row length: cap rounded up to a multiple of 8
  int count_overflow;            // PBF_OPT_COUNT_OVERFLOW
  float skin2;                   // (1.1 h)^2: the radius of the skin lists
  float disp_lim2;               // (0.045 h)^2: squared moves below it keep the skin lists usable
  int skin_mode;                 // skin lists allowed (no PBF_OPT_FULL_LISTS)
  int tile_lists;                // the cell-tile list walk (no PBF_OPT_LANE_LISTS)
  int fast;                      // PBF_OPT_FAST_MATH: hardware sqrt and reciprocal in the fused step
  int static_ranges;             // the exact sums' range tests hold for every accepted pair (nb_sqrt, nb_spiky)
  uint32_t heavy_budget;         // candidates one lane scans before a wavefront takes over (see walk_list)
  // batched scenes: ncells is the total over the scenes, sid the scene of each slot
  int batched;
  const SceneGeo* scenes;
  const uint32_t* sid;
};

enum Counter {
  C_OVERFLOW = 0,
  C_OUTSIDE = 1,
  C_HACKS = 2,
  C_HACK_FLAG = 3,  // two flags (3, 4) used on alternate Jacobi iterations
  C_HEAVY = 5,      // number of entries in the heavy list
  C_EMAX = 6,       // float bits of the max density error (atomicMax on non-negative floats)
  C_MOVED_N = 7,    // two counts (7, 8): particles moved by the density fix, per iteration parity
  C_MOVED_ALL = 9,  // two flags (9, 10): too many moved particles or marked cells -> treat all as dirty
  C_DONE = 11,      // iterate-to-tolerance: 0, or the slot whose error fell below the tolerance plus one
  C_ITERS = 12,     // iterate-to-tolerance: slots used
  C_WARM_KEPT = 13, // the warm-start positions were kept (2: the damped retry's, PBF_OPT_WARM_RETRY)
  C_DISP = 14,      // two flags (14, 15): some particle has moved too far for the skin lists by the
                    // end of the previous Jacobi iteration (parity of the iteration that reads it)
  C_SKIN = 16,      // 1: this step's skin lists are built
  C_DISP_ALL = 17,  // the same over the whole step so far
  C_TILE = 18,      // particles listed by the cell-tile walk (counted with PBF_OPT_COUNT_OVERFLOW)
  C_RETRY = 19,     // the warm start grew the error at slot 0 and its damped retry was applied
  C_NUM = 24
};

constexpr uint32_t kMaxSolveSlots = PBF_MAX_SOLVE_ITERATIONS;
constexpr uint64_t kRowPad = 8;  // padding rows of the float4 particle buffers
constexpr float kWarmLambdaScale = PBF_WARM_LAMBDA_SCALE;

// Per-launch control of the Jacobi loop kernels.  A plain fixed-iteration
// step passes tol < 0 and null pointers.
struct SolveCtl {
  float tol;                   // >= 0: stop once the slot's density error is <= tol
  int warm;                    // a warm-start update was applied before slot 0
  int mean;                    // compare the mean error (else the maximum)
  uint32_t* ehist;             // max error as float bits: [0] before the warm start, [k+1] slot k
  unsigned long long* esum;    // sum of the errors in 2^-32 units, same slots
  float* lsum;                 // sum of the applied multipliers (null: not kept)
  const float4* xsave;         // position stars to restore when the warm start is rejected
  int strict;                  // reject the warm start when the max or the mean error grows
  int retry;                   // a rejected warm start is retried at kWarmRetryScale before the fallback to zero
};

// PBF_OPT_WARM_RETRY: the rejected warm start's multipliers are retried at this
// scale (exact in float) before the solve falls back to a zero warm start.
constexpr float kWarmRetryScale = 0.5f;

// The first density pass of a warm-started solve: the learned warm start's
// features, kept in device memory when feat is set, and with wlam set the
// network evaluated on them in the same pass:
// wlam[p] = lprev[p] + scale * net(features of p).
struct FeatCtl {
  float* feat;          // [n][8] (null: the features are not stored)
  const float* lprev;   // previous multiplier sums
  const short* w1;      // bf16 [64][16] first layer (features 8..15 zero)
  const short* w2;      // bf16 [64][64] second layer
  const float* mp;      // fp32 parameters (biases, output layer)
  float scale;
  float* wlam;          // null: no network in this pass
};

// Neighbour lists are tiled by groups of 64 particles: entry k of particle p
// sits at ((p / 64) * stride + k) * 64 + p % 64.  Entry k of a wavefront's 64
// particles is one 256-byte read, and a wavefront's whole list is one
// contiguous block of stride * 256 bytes.
__host__ __device__ __forceinline__ uint64_t list_at(uint32_t k, uint32_t p, uint32_t stride) {
  return ((uint64_t)(p >> 6) * stride + k) * 64u + (p & 63u);
}

struct KernelTimer {
  std::string name;
  uint64_t launches = 0;
  double total_ms = 0.0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> spare;  // events of folded launches, reused
};

}  // namespace pbf

struct pbf_context {
  pbf_params P;
  pbf::KParams K;
  hipStream_t stream = nullptr;
  int device = 0;
  uint64_t n = 0, capacity = 0;
  uint64_t dims[3] = {0, 0, 0};
  uint64_t ncells = 0;
  uint64_t grid_alloc = 0;  // cells allocated
  int cap = 60;
  int cap_alloc = 60;  // list rows allocated at pbf_create: the largest cap pbf_set_params accepts
  bool lists_valid = false;
  uint32_t last_iterations = 0;
  uint32_t dirty_stamp = 0;  // incremented per Jacobi iteration of the fused step
  bool last_was_solve = false;  // stats.iterations comes from the device counter
  bool track_lsum = false;      // lsum follows the particles through the reorder

  float4 *x = nullptr, *xs = nullptr, *v = nullptr;
  float4 *xa = nullptr, *xsa = nullptr, *va = nullptr;
  float4 *xt = nullptr, *scr = nullptr;
  float *rho = nullptr, *lam = nullptr, *rhoa = nullptr, *lama = nullptr;
  float4* scra = nullptr;
  uint32_t *oid = nullptr, *oida = nullptr, *key = nullptr, *keya = nullptr;
  uint32_t *slot = nullptr, *tmp = nullptr, *tmpo = nullptr;
  uint32_t *sk[2] = {nullptr, nullptr}, *sv[2] = {nullptr, nullptr};
  uint32_t *rhist = nullptr, *roffs = nullptr, *rpart = nullptr;
  uint32_t* perm = nullptr;  // sorted slot -> previous slot (points into sv or tmpo)
  uint32_t* heavy = nullptr;
  uint32_t* moved = nullptr;  // particles moved by the density fix in the current iteration
  uint32_t* dirty = nullptr;  // == dirty_stamp: the particle or one of its neighbours moved
  uint32_t* mflag = nullptr;  // == dirty_stamp: the density fix moved the particle in that iteration
  uint32_t* cell_start = nullptr;
  uint32_t* cbox = nullptr;  // [6][cells] bounding box of each cell's particles (encoded)
  float4* chbox = nullptr;   // [2][n/64] bounding box (min, max) of each aligned run of 64 slots
  uint32_t* scan_partials = nullptr;
  int32_t *nbr = nullptr, *nbr_count = nullptr;
  int32_t *cnb = nullptr, *cnb_count = nullptr;
  int32_t *snb = nullptr, *snb_count = nullptr;  // skin lists (entries within 1.1 h when built)
  float4* xs0 = nullptr;                          // position stars the lists were built from
  unsigned long long* skin_hint = nullptr;  // 1: the last step's moves kept the skin lists usable
  uint64_t* wmask = nullptr;  // [2][n] within-h masks of the last fused density pass
  unsigned long long* counters = nullptr;
  uint32_t* ehist = nullptr;
  unsigned long long* esum = nullptr;
  float* feat = nullptr;       // [n][8] features of the last warm-started solve
  bool feat_valid = false;
  short *mw1 = nullptr, *mw2 = nullptr;  // learned warm start: bf16 weights
  float* mp = nullptr;                   // and the fp32 parameters
  bool has_model = false;
  float warm_gain = 1.0f;  // learned warm start: Lambda_prev + warm_gain * S * net
  uint32_t nscenes = 1;
  std::vector<pbf::SceneGeo> scene_geo;   // host copy (batched contexts)
  pbf::SceneGeo* scenes = nullptr;        // device copy
  uint32_t *sid = nullptr, *sida = nullptr;  // scene of each slot (batched contexts)
  // captured pair of fused steps (pbf_step_graph)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_n = 0, graph_version = 0;
  std::vector<void*> graph_sig;  // buffer pointers when the graph was captured
  uint64_t version = 0;  // bumped by anything that changes the launches of a step
  bool capturing = false;
  // captured pair of iterate-to-tolerance steps (pbf_solve_graph) and what it was captured for
  hipGraph_t sgraph = nullptr;
  hipGraphExec_t sgraph_exec = nullptr;
  uint64_t sgraph_n = 0, sgraph_version = 0;
  float sgraph_tol = 0.f, sgraph_gain = 0.f;
  uint32_t sgraph_max_it = 0;
  int sgraph_warm = -1;
  std::vector<void*> sgraph_sig;
  // per-solve records: a ring of PBF_SOLVE_LOG_SIZE and the number written so far
  pbf_solve_record* slog = nullptr;
  unsigned long long* slog_count = nullptr;
  pbf_tuning tune{256u, 768u, 8192u};
  int coop_blocks[3] = {-1, -1, -1};  // blocks of a cooperative launch of the persistent Jacobi kernel: exact,
                                      // fast and static-range exact arithmetic (-1: not yet found)
  int coop_cus = 0;               // compute units of the device (found with coop_blocks)
  uint32_t coop_grid = 0;         // blocks of the last step's cooperative launch (0: none)
  unsigned long long* done_host = nullptr;  // pinned copy of C_DONE, polled between slot chunks
  float *lsum = nullptr, *lsuma = nullptr, *wlam = nullptr, *wgiven = nullptr;
  float4* xsave = nullptr;
  bool has_given = false;
  float* stage = nullptr;  // device staging for original-order transfers (4 floats / particle)
  uint64_t* stage64 = nullptr;

  hipEvent_t region_a = nullptr, region_b = nullptr;
  std::vector<pbf::KernelTimer> timers;
};

namespace pbf {

// error plumbing (pbf_api.cpp)
int set_error(int code, const std::string& msg);

// A roctx range over the host side of one entry point (shown by
// rocprofv3 --marker-trace; a push and a pop that return at once when no
// profiler is attached).
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};
#define PBF_TRACE(name) ::pbf::TraceRange pbf_trace_range_(name)
int check_hip(hipError_t e, const char* what);

// timing hooks around launches (pbf_api.cpp)
KernelTimer* timer_begin(pbf_context* c, const char* name);
void timer_end(pbf_context* c, KernelTimer* t);

// ---- launchers (pbf_grid.hip) ----
void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
void launch_fill_u64(unsigned long long* p, unsigned long long value, uint64_t n, hipStream_t s);
// exclusive scan of in[0..n) into out[0..n]; out[n] = total.  partials needs scan_partials_size(n).
uint64_t scan_partials_size(uint64_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partials, hipStream_t s);
void launch_bin(pbf_context* c, uint64_t bin_end);  // key, inverse permutation, keys by original index
void launch_sort(pbf_context* c);            // radix sort -> perm, cell_start
void launch_cell_box(pbf_context* c);        // bounding box of each cell's particles
// gather state by perm (without all_fields: densities, multipliers and the
// scratch vector stay where they are)
void launch_reorder(pbf_context* c, bool all_fields = true);
uint64_t radix_hist_size(uint64_t n);
int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint64_t n, uint32_t key_bits,
                     uint32_t* hist, uint32_t* offs, uint32_t* partials, hipStream_t s);
void launch_fill_neighbors(pbf_context* c, uint64_t fb, uint64_t fe, bool want_skin);
int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits, hipStream_t s);
void launch_bounds_counting(const uint32_t* sorted, uint64_t n, uint64_t b, uint64_t e, uint32_t* out,
                            int upper, hipStream_t s);

// ---- launchers (pbf_passes.hip) ----
// flags for the particle-local kernel
enum { F_FORCES = 1, F_PREDICT = 2, F_CLEAR = 4 };  // F_CLEAR: zero the counters first
void launch_forces_predict(pbf_context* c, int flags, uint64_t b, uint64_t e);
void launch_density(pbf_context* c, uint64_t b, uint64_t e);  // xs -> xt, rho
// ---- fused-step launchers (pbf_fused.hip): they build and use the compact lists ----
void launch_density_lambda_masked(pbf_context* c, int it, int eslot, const SolveCtl& S,
                                  const FeatCtl& F = FeatCtl{nullptr, nullptr});  // xs -> xt, rho, lam
void launch_mark_dirty(pbf_context* c, int it, const SolveCtl& S);     // particles next to a moved one
void launch_lambda_repair(pbf_context* c, int it, const SolveCtl& S);  // lam from xt for dirty particles
void launch_dp_update_masked(pbf_context* c, int it, int last, const SolveCtl& S);
void launch_warm_dp(pbf_context* c);      // xt, wlam -> xs (warm start), xsave = xt
void launch_add_warm(pbf_context* c);     // lsum += wlam when the warm start was kept
void launch_log_solve(pbf_context* c);    // one pbf_solve_record of the solve just queued
void launch_warm_retry(pbf_context* c);   // the damped warm start after a rejected one (a no-op otherwise)
// the Jacobi loop in one cooperative launch (PBF_OPT_PERSISTENT); advances dirty_stamp by iters
bool persistent_available(pbf_context* c);
uint32_t persistent_capacity(pbf_context* c);  // blocks of it the device holds at once (occupancy query)
bool persistent_fits(pbf_context* c);          // the particles fit one launch of it (one block per CU at most)
int launch_jacobi_persistent(pbf_context* c, uint32_t iters, bool solve, bool last_vel, const SolveCtl& S);
void launch_viscosity_first(pbf_context* c);                  // v -> va, compact lists of the final positions
void launch_vorticity_masked(pbf_context* c);                 // va -> scr
void launch_apply_vorticity_masked(pbf_context* c);           // va += ...
void launch_viscosity_last(pbf_context* c);                   // va -> v, x = x*
void launch_lambda(pbf_context* c, const float4* pos, uint64_t b, uint64_t e);
void launch_dp(pbf_context* c, uint64_t b, uint64_t e);                     // xs, lam -> scr
void launch_update_position_stars(pbf_context* c, uint64_t b, uint64_t e);
void launch_update_velocities(pbf_context* c, uint64_t b, uint64_t e);
void launch_update_positions(pbf_context* c, uint64_t b, uint64_t e);
void launch_viscosity(pbf_context* c, const float4* vin, float4* vout, int write_positions, uint64_t b,
                      uint64_t e);
void launch_vorticity(pbf_context* c, const float4* vin, uint64_t b, uint64_t e);
void launch_apply_vorticity(pbf_context* c, float4* vio, uint64_t b, uint64_t e);
void launch_surface_tension(pbf_context* c, uint64_t cb, uint64_t ce, uint64_t b, uint64_t e);
void launch_copy_f4(const float4* src, float4* dst, uint64_t n, hipStream_t s);
void launch_copy_f1(const float* src, float* dst, uint64_t n, hipStream_t s);
void launch_scale_f1(const float* src, float* dst, float scale, uint64_t n, hipStream_t s);
// ---- learned warm start (pbf_mlp.hip) ----
// original-order transfer helpers
void launch_gather_out_f4(pbf_context* c, const float4* src, float* dst3, int comps);
void launch_gather_out_f1(pbf_context* c, const float* src, float* dst);
void launch_gather_out_rows(pbf_context* c, const float* src, float* dst, int width);  // [n][width] rows
void launch_scatter_in_f4(pbf_context* c, const float* src3, float4* dst);
void launch_scatter_in_f1(pbf_context* c, const float* src, float* dst);

}  // namespace pbf
