// pbf_grid.hip -- cell binning, stable radix sort, cell bounds, reorder,
// neighbour lists and the sort / bound primitives.
//
// Reference: Simulation/Source/neighbors.h:108-248 (calculate_bin_id,
// calculate_bins, sort_bins, find_bin_bounds, calculate_neighbor_indices,
// fill_neighbors) and Common/sim_algorithms.h:57-99 (sort_by_key,
// lower_bound, upper_bound).
//
// Sort.  The reference's sort_by_key is Thrust's stable sort of (cell key,
// particle index).  Particles here live in cell order from the previous
// step, so a stable sort of the current slots would order by
// (key, previous key, index).  Instead the (key, slot) pairs are listed in
// original-index order (through the inverse permutation written by the
// binning pass) and sorted by a stable LSD radix sort on the key bits (8-bit
// digits, per-tile histograms, wavefront ballot ranking in LDS): the result is
// exactly the (key, original index) order the reference produces.
//
// Lists.  One lane per particle walks the 27 cells in the reference's order
// and stages its entries in LDS, then writes them tiled by groups of 64
// particles (list_at: entry k of a wavefront's 64 particles contiguous).  A
// particle with many candidates whose walk passes a budget is handed to a
// list of heavy particles that one wavefront each finishes with
// ballot-compacted appends, so that collapsed cells do not serialise a
// wavefront on one lane.
#include "pbf_device.h"

namespace pbf {

constexpr int kBlock = 256;

static inline uint32_t nblocks(uint64_t n, int per = kBlock) { return (uint32_t)((n + per - 1) / per); }

__global__ void k_fill_u32(uint32_t* p, uint32_t value, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long value, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

void launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(n)), dim3(kBlock), 0, s, p, value, n);
}
void launch_fill_u64(unsigned long long* p, unsigned long long value, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_u64, dim3(nblocks(n)), dim3(kBlock), 0, s, p, value, n);
}

// ---------------------------------------------------------------------------
// Exclusive scan (three passes: tile sums, scan of tile sums, tile scans).
// A tile is 256 threads x 8 items.  64-wide wavefront scans with __shfl_up.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread; returns the block total in *total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    const uint32_t s = wsum[i];
    if (i < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const uint32_t* in, uint64_t n, uint32_t* partials) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// one block scans all tile sums in place (exclusive), looping with a carry
__global__ void __launch_bounds__(kBlock) k_scan_partials(uint32_t* partials, uint32_t np) {
  uint32_t carry = 0;
  for (uint32_t base = 0; base < np; base += kBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < np ? partials[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan(v, &tot);
    if (i < np) partials[i] = carry + ex;
    carry += tot;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const uint32_t* in, uint32_t* out, uint64_t n,
                                                       const uint32_t* partials) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = (base + i < n) ? in[base + i] : 0u;
    s += vals[i];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot) + partials[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += vals[i];
    if (base + i + 1 == n) out[n] = run;  // total
  }
}

uint64_t scan_partials_size(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* partials, hipStream_t s) {
  if (n == 0) {
    launch_fill_u32(out, 0u, 1, s);
    return;
  }
  const uint32_t nt = (uint32_t)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, s, in, n, partials);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, s, partials, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kBlock), 0, s, in, out, n, (const uint32_t*)partials);
}

// ---------------------------------------------------------------------------
// calculate_bins (neighbors.h:120-125): cell key of every slot, and the
// inverse permutation (original index -> slot) so that the keys can be
// listed in original-index order for the stable sort.
// Also lists the (key, slot) pairs by original index: a stable sort of that
// sequence by key is the reference's stable sort of (key, index) pairs.
__global__ void __launch_bounds__(kBlock) k_bin(const float4* __restrict__ xs, uint32_t n, uint32_t bin_end,
                                                KParams K, const uint32_t* __restrict__ oid,
                                                uint32_t* __restrict__ key, uint32_t* __restrict__ inv,
                                                unsigned long long* __restrict__ counters,
                                                uint32_t* __restrict__ sk, uint32_t* __restrict__ sv) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {  // counted by the list build that follows
    counters[C_HEAVY] = 0ull;
    counters[C_TILE] = 0ull;
  }
  if (p >= n) return;
  const uint32_t o = oid[p];
  uint32_t k = K.ncells;  // particles past the bin span are placed in no cell
  if (o < bin_end) {
    const f3 ps = ld3(xs, p);
    k = cell_key(bin_coord(ps.x, K.s), bin_coord(ps.y, K.s), bin_coord(ps.z, K.s), grid_of(K, p), K);
    if (k == K.ncells) atomicAdd(&counters[C_OUTSIDE], 1ull);
  }
  key[p] = k;
  inv[o] = p;
  sk[o] = k;
  sv[o] = p;
}

void launch_bin(pbf_context* c, uint64_t bin_end) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "bin");
  hipLaunchKernelGGL(k_bin, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, (const float4*)c->xs, n,
                     (uint32_t)bin_end, c->K, (const uint32_t*)c->oid, c->key, c->slot, c->counters, c->sk[0],
                     c->sv[0]);
  timer_end(c, t);
}

// find_bin_bounds (neighbors.h:137-148): begin[c] = lower_bound(keys, c);
// end[c] = upper_bound(keys, c) = lower_bound(keys, c + 1) = begin[c + 1].
__global__ void k_cell_start(const uint32_t* __restrict__ sorted, uint32_t n, uint32_t ncells,
                             uint32_t* __restrict__ start) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells + 1) return;
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sorted[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  start[c] = lo;
}

void launch_sort(pbf_context* c) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "sort");
  uint32_t bits = 1;
  while (bits < 32 && (c->K.ncells >> bits) != 0) ++bits;  // keys go up to ncells (the outside bucket)
  const int which = radix_sort_pairs(c->sk[0], c->sv[0], c->sk[1], c->sv[1], n, bits, c->rhist, c->roffs,
                                     c->rpart, c->stream);
  c->perm = c->sv[which];
  hipLaunchKernelGGL(k_cell_start, dim3(nblocks((uint64_t)c->K.ncells + 2)), dim3(kBlock), 0, c->stream,
                     (const uint32_t*)c->sk[which], n, c->K.ncells, c->cell_start);
  timer_end(c, t);
}

// gather the persistent state into cell order
__global__ void __launch_bounds__(kBlock) k_reorder(uint32_t n, const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ x, const float4* __restrict__ xs,
                                                    const float4* __restrict__ v, const float* __restrict__ rho,
                                                    const float* __restrict__ lam, const float4* __restrict__ scr,
                                                    const uint32_t* __restrict__ oid,
                                                    const uint32_t* __restrict__ key,
                                                    const float* __restrict__ lsum,
                                                    const uint32_t* __restrict__ sid, float4* __restrict__ xo,
                                                    float4* __restrict__ xso, float4* __restrict__ vo,
                                                    float* __restrict__ rhoo, float* __restrict__ lamo,
                                                    float4* __restrict__ scro, uint32_t* __restrict__ oido,
                                                    uint32_t* __restrict__ keyo, float* __restrict__ lsumo,
                                                    uint32_t* __restrict__ sido) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t p = perm[s];
  if (lsum) lsumo[s] = lsum[p];
  if (sid) sido[s] = sid[p];
  xo[s] = x[p];
  xso[s] = xs[p];
  vo[s] = v[p];
  if (rho) {
    rhoo[s] = rho[p];
    lamo[s] = lam[p];
    scro[s] = scr[p];
  }
  oido[s] = oid[p];
  keyo[s] = key[p];
}

void launch_reorder(pbf_context* c, bool all_fields) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "reorder");
  hipLaunchKernelGGL(k_reorder, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, (const uint32_t*)c->perm,
                     (const float4*)c->x, (const float4*)c->xs, (const float4*)c->v,
                     all_fields ? (const float*)c->rho : nullptr,
                     (const float*)c->lam, (const float4*)c->scr, (const uint32_t*)c->oid,
                     (const uint32_t*)c->key, c->track_lsum ? (const float*)c->lsum : nullptr,
                     c->K.batched ? (const uint32_t*)c->sid : nullptr, c->xa, c->xsa, c->va, c->rhoa, c->lama,
                     c->scra, c->oida, c->keya, c->lsuma, c->sida);
  timer_end(c, t);
  if (c->track_lsum) std::swap(c->lsum, c->lsuma);
  if (c->K.batched) {
    std::swap(c->sid, c->sida);
    c->K.sid = c->sid;
  }
  std::swap(c->x, c->xa);
  std::swap(c->xs, c->xsa);
  std::swap(c->v, c->va);
  if (all_fields) {
    std::swap(c->rho, c->rhoa);
    std::swap(c->lam, c->lama);
    std::swap(c->scr, c->scra);
  }
  std::swap(c->oid, c->oida);
  std::swap(c->key, c->keya);
}

// ---------------------------------------------------------------------------
// Bounding box of every cell's particles, 6 order-preserving integer
// encodings per cell (min x y z, max x y z).  A neighbour walk skips a cell
// whose box is at least s away from the walking particle: float subtraction,
// squaring and addition are monotone, so every candidate's d^2 (computed with
// the same operations) is at least the box bound and none could be accepted.
// The lists are unchanged.  Only the wavefront builder tests cells (of more
// than kBoxMin particles): in the one-lane walk a skip saves time only when
// all 64 lanes skip, which is rare.
//
// The same pass writes the box of every aligned run of 64 slots (a chunk).
// Slots inside a cell are in original-index order, not spatial order, so a
// crowded cell (thousands of particles after a collapse) can be near a
// particle while most of its chunks are not; the wavefront builder tests 64
// chunk boxes at once and gathers only the chunks that can hold a neighbour.
constexpr uint32_t kBoxMin = 32;
constexpr uint32_t kChunkMin = 128;  // cells above this size are walked chunk by chunk

__device__ __forceinline__ uint32_t ord_enc(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ void k_box_init(uint32_t* __restrict__ box, uint64_t count, const unsigned long long* __restrict__ counters) {
  if (counters[C_HEAVY] == 0ull) return;  // the boxes serve only the wavefront builder
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) box[i] = (i % 6 < 3) ? 0xFFFFFFFFu : 0u;
}

// Slots are in cell order: each wavefront reduces its runs of equal keys and
// the first lane of a run updates the cell's box.
__global__ void __launch_bounds__(kBlock) k_cell_box(uint32_t n, uint32_t ncells, const uint32_t* __restrict__ key,
                                                     const float4* __restrict__ xs, uint32_t* __restrict__ box,
                                                     float4* __restrict__ chbox,
                                                     const unsigned long long* __restrict__ counters) {
  if (counters[C_HEAVY] == 0ull) return;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = p < n;
  const uint32_t k = valid ? key[p] : 0xFFFFFFFFu;
  const float4 t = valid ? xs[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t mn[3] = {ord_enc(t.x), ord_enc(t.y), ord_enc(t.z)};
  uint32_t mx[3] = {mn[0], mn[1], mn[2]};
  // chunk box over the wavefront's 64 slots (particles outside the grid are
  // never walked and left out)
  {
    const bool in = valid && k < ncells;
    uint32_t cmn[3], cmx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cmn[a] = in ? mn[a] : 0xFFFFFFFFu;
      cmx[a] = in ? mx[a] : 0u;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        cmn[a] = min(cmn[a], (uint32_t)__shfl_xor(cmn[a], d, 64));
        cmx[a] = max(cmx[a], (uint32_t)__shfl_xor(cmx[a], d, 64));
      }
    }
    if (lane == 0 && p < n) {
      const uint32_t ch = p >> 6;
      chbox[2 * ch] = make_float4(ord_dec(cmn[0]), ord_dec(cmn[1]), ord_dec(cmn[2]), 0.f);
      chbox[2 * ch + 1] = make_float4(ord_dec(cmx[0]), ord_dec(cmx[1]), ord_dec(cmx[2]), 0.f);
    }
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t kk = __shfl_down(k, d, 64);
    uint32_t omn[3], omx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      omn[a] = __shfl_down(mn[a], d, 64);
      omx[a] = __shfl_down(mx[a], d, 64);
    }
    if (lane + d < 64 && kk == k) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], omn[a]);
        mx[a] = max(mx[a], omx[a]);
      }
    }
  }
  // a wavefront inside one cell (all 64 keys equal) leaves the cell box to
  // k_cell_box_chunks, which merges such chunks before touching the cell
  const bool inside = __shfl(k, 0, 64) == __shfl(k, 63, 64) && __shfl(k, 63, 64) < ncells;
  const uint32_t prev = __shfl_up(k, 1, 64);
  if (!inside && valid && k < ncells && (lane == 0 || prev != k)) {
    uint32_t* b = box + 6ull * k;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&b[a], mn[a]);
      atomicMax(&b[3 + a], mx[a]);
    }
  }
}

// The chunks that lie inside one cell (a crowded cell holds thousands): one
// thread per chunk, runs of equal cells reduced per wavefront, one update of
// the cell box per run -- instead of one per chunk on the same six words.
__global__ void __launch_bounds__(kBlock) k_cell_box_chunks(uint32_t n, uint32_t ncells,
                                                            const uint32_t* __restrict__ key,
                                                            const float4* __restrict__ chbox,
                                                            uint32_t* __restrict__ box,
                                                            const unsigned long long* __restrict__ counters) {
  if (counters[C_HEAVY] == 0ull) return;
  const uint32_t ch = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint32_t nch = (n + 63) / 64;
  uint32_t k = 0xFFFFFFFFu;
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  if (ch < nch && 64ull * ch + 63 < n) {
    const uint32_t k0 = key[64 * ch], k1 = key[64 * ch + 63];
    if (k0 == k1 && k1 < ncells) {
      k = k0;
      const float4 lo = chbox[2 * ch], hi = chbox[2 * ch + 1];
      mn[0] = ord_enc(lo.x), mn[1] = ord_enc(lo.y), mn[2] = ord_enc(lo.z);
      mx[0] = ord_enc(hi.x), mx[1] = ord_enc(hi.y), mx[2] = ord_enc(hi.z);
    }
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t kk = __shfl_down(k, d, 64);
    uint32_t omn[3], omx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      omn[a] = __shfl_down(mn[a], d, 64);
      omx[a] = __shfl_down(mx[a], d, 64);
    }
    if (lane + d < 64 && kk == k) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], omn[a]);
        mx[a] = max(mx[a], omx[a]);
      }
    }
  }
  const uint32_t prev = __shfl_up(k, 1, 64);
  if (k < ncells && (lane == 0 || prev != k)) {
    uint32_t* b = box + 6ull * k;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&b[a], mn[a]);
      atomicMax(&b[3 + a], mx[a]);
    }
  }
}

void launch_cell_box(pbf_context* c) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  KernelTimer* t = timer_begin(c, "cell_box");
  const uint64_t count = 6ull * c->K.ncells;
  const unsigned long long* heavy = c->counters;
  hipLaunchKernelGGL(k_box_init, dim3(nblocks(count)), dim3(kBlock), 0, c->stream, c->cbox, count, heavy);
  hipLaunchKernelGGL(k_cell_box, dim3(nblocks(n)), dim3(kBlock), 0, c->stream, n, c->K.ncells,
                     (const uint32_t*)c->key, (const float4*)c->xs, c->cbox, c->chbox, heavy);
  hipLaunchKernelGGL(k_cell_box_chunks, dim3(nblocks((n + 63) / 64)), dim3(kBlock), 0, c->stream, n, c->K.ncells,
                     (const uint32_t*)c->key, (const float4*)c->chbox, c->cbox, heavy);
  timer_end(c, t);
}

// true when no point of the box [l, h] can lie within s of ps (see above)
__device__ __forceinline__ bool box_far(float lx, float ly, float lz, float hx, float hy, float hz, f3 ps, float ss) {
  const float dx = ps.x < lx ? lx - ps.x : (ps.x > hx ? ps.x - hx : 0.0f);
  const float dy = ps.y < ly ? ly - ps.y : (ps.y > hy ? ps.y - hy : 0.0f);
  const float dz = ps.z < lz ? lz - ps.z : (ps.z > hz ? ps.z - hz : 0.0f);
  return (dx * dx + dy * dy) + dz * dz >= ss;
}
__device__ __forceinline__ bool cell_far(const uint32_t* __restrict__ box, uint32_t cell, f3 ps, float ss) {
  const uint32_t* b = box + 6ull * cell;
  return box_far(ord_dec(b[0]), ord_dec(b[1]), ord_dec(b[2]), ord_dec(b[3]), ord_dec(b[4]), ord_dec(b[5]), ps, ss);
}

// ---------------------------------------------------------------------------
// fill_neighbors (neighbors.h:197-230).  A particle walks its 27 cells in the
// reference's order (x offset outer, z inner; cell ids from the FP
// coordinates x + i*s, neighbors.h:171-185), skips itself and accepts
// d^2 < s^2 while the list has room.  Lists hold sorted slots (list_at).
//
// Two tiers with identical results: one lane per particle, unless its 27
// cells hold more than twice K.heavy_budget candidates and its walk scans
// K.heavy_budget of them without filling the list (with the overflow count
// on: more than K.heavy_budget candidates); such a particle goes to a list
// that k_fill_heavy walks with a whole wavefront (64 candidates per step,
// accepted ones appended in order through a ballot).

struct WalkCells {
  float cx[3], cy[3], cz[3];
};

// a[i] for a loop index i without indexing the array at run time (which
// would put the array in scratch memory)
__device__ __forceinline__ float pick3(const float (&a)[3], int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

__device__ __forceinline__ WalkCells walk_cells(f3 ps, float s) {
  WalkCells w;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float off = (float)(i - 1) * s;
    w.cx[i] = bin_coord(ps.x + off, s);
    w.cy[i] = bin_coord(ps.y + off, s);
    w.cz[i] = bin_coord(ps.z + off, s);
  }
  return w;
}

// One wavefront per block; each lane stages its entries in its own LDS
// column (entry k of the 64 lanes is contiguous) and writes them out at the
// end, so that every global store of entry k covers the wavefront's 64
// consecutive slots instead of 64 scattered 4-byte writes.  A staged entry is
// 16 bits: the slot's offset in its cell (10 bits), the cell's place in the
// 27-cell walk (c, 5 bits) and whether the candidate lies within the skin
// radius (bit 15), which halves the staging and doubles the wavefronts the
// LDS holds.  The offset stays below kMaxCellOffset because a lane walks a
// cell in full only when its 27 cells hold at most that many candidates, and
// otherwise stops after a budget below it.
//
// Skin lists (with `skin`): the entries within 1.1 h (K.skin2) of the
// positions the lists are built from, in list order, each slot packed with
// its place in the full list (slot | place << 25; n < 2^25); x*_0 = x*.
// The Jacobi passes walk them instead of the full lists while no particle has
// moved far enough for an entry outside them to come within h
// (pbf_fused.hip).
constexpr int kFillBlock = 64;
constexpr uint32_t kMaxCellOffset = 1024u;
constexpr uint32_t kSkinPlaceShift = 25;

// The lane walk of particle p: writes its list and count and returns true,
// or returns false when the particle goes to the wavefront builder.  ovf
// receives the candidates dropped by the cap.
template <bool SKIN>
__device__ __forceinline__ bool walk_list(uint32_t p, uint32_t n, uint32_t lane, uint16_t* __restrict__ s_list,
                                          const float4* __restrict__ xs, const uint32_t* __restrict__ start,
                                          const KParams& K, int32_t* __restrict__ nbr,
                                          int32_t* __restrict__ nbr_count, uint32_t& ovf,
                                          int32_t* __restrict__ snb, int32_t* __restrict__ snb_count) {
  const f3 ps = ld3(xs, p);
  const WalkCells w = walk_cells(ps, K.s);
  const GridView G = grid_of(K, p);
  // the 27 cell ranges first, all loads in flight together (a cell outside
  // the grid is an empty range); the walk then waits only on candidates
  uint32_t cb[27], ce[27];
#pragma unroll
  for (int c = 0; c < 27; ++c) {
    const uint32_t cell = cell_key(pick3(w.cx, c / 9), pick3(w.cy, (c / 3) % 3), pick3(w.cz, c % 3), G, K);
    const bool in = cell < K.ncells;
    const uint32_t cc = in ? cell : 0u;
    const uint32_t b = start[cc], e = start[cc + 1];
    cb[c] = in ? b : 0u;
    ce[c] = in ? e : 0u;
  }
  uint32_t total = 0;
#pragma unroll
  for (int c = 0; c < 27; ++c) total += ce[c] - cb[c];
  // the lane's budget; below kMaxCellOffset - 8 so that a staged offset fits
  // in 10 bits (a batch of 8 may pass the budget before the test)
  const uint32_t hb = min(K.heavy_budget, kMaxCellOffset - 8u);
  // every candidate must be seen to count the overflow: route by the total
  if (K.count_overflow && total > hb) return false;
  // otherwise the walk ends when the list is full; a lane whose 27 cells hold
  // more than twice K.heavy_budget candidates and that has scanned
  // K.heavy_budget of them without filling its list hands the particle over
  // (the wavefront builder starts again and writes the same entries); lanes
  // with fewer candidates walk them all
  const uint32_t budget = total > min(2u * hb, kMaxCellOffset) ? hb : 0xFFFFFFFFu;
  uint32_t scanned = 0;
  int count = 0;
  const int cap = K.cap;
#pragma unroll
  for (int c = 0; c < 27; ++c) {
    const uint32_t b = cb[c], e = ce[c];
    // staged form of slot q: (c << 10) + (q - b), computed as q + eb in 16 bits
    const uint32_t eb = ((uint32_t)c << 10) - b;
    // candidates of a cell are contiguous: load 8 at a time, test in order
    // (position buffers carry 8 rows of padding, so the loads need no clamp)
    for (uint32_t q0 = b; q0 < e; q0 += 8) {
      scanned += 8;
      if (scanned > budget && !K.count_overflow) return false;
      float4 t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = xs[q0 + (uint32_t)u];
      // branch-free acceptance: every candidate is stored in row
      // min(count, cap) and only an accepted one advances count, so a
      // rejected one is overwritten by the next accepted one (rows >= the
      // final count, and the spare row cap, are never read)
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t q = q0 + (uint32_t)u;
        const float d2 = magsq3(sub3(ps, f3{t[u].x, t[u].y, t[u].z}));
        const bool acc = (q < e) & (q != p) & (d2 < K.ss);
        const uint32_t sb = SKIN ? (uint32_t)(d2 < K.skin2) << 15 : 0u;
        s_list[min(count, cap) * kFillBlock + lane] = (uint16_t)((q + eb) | sb);
        count += acc ? 1 : 0;
      }
      // a full list cannot change any more; only the overflow count could
      if (count >= cap && !K.count_overflow) goto done;
    }
  }
done:
  if (count > cap) {
    ovf = (uint32_t)(count - cap);
    count = cap;
  }
  nbr_count[p] = count;
  // staged entries come in walk order, so the cells' first slots are taken
  // from registers cell by cell
  int k = 0, j = 0;
#pragma unroll
  for (int c = 0; c < 27; ++c) {
    const uint32_t b = cb[c];
    for (; k < count; ++k) {
      const uint32_t e16 = s_list[k * kFillBlock + lane];
      if (((e16 >> 10) & 31u) != (uint32_t)c) break;
      const uint32_t slot = b + (e16 & (kMaxCellOffset - 1u));
      nbr[list_at(k, p, K.stride)] = (int32_t)slot;
      if (SKIN && (e16 >> 15)) snb[list_at(j++, p, K.stride)] = (int32_t)(slot | ((uint32_t)k << kSkinPlaceShift));
    }
  }
  if (SKIN) snb_count[p] = j;
  return true;
}

// ---------------------------------------------------------------------------
// The cell-tile walk: the default for particles of ordinary cells.  A
// wavefront's 64 consecutive slots fall into runs of one cell each
// (segments).  All particles of a segment walk the same candidate stream:
// the 27 cells around their cell in walk order, each cell's slots in order,
// i.e. the sequence the lane walk visits.  The stream (at most kTileMaxCand
// candidates) is held in registers, 64 candidates per round, one per lane;
// for each particle of the segment the wavefront tests every round's 64
// candidates at once, and the ballot of the accepted ones is that particle's
// next 64 acceptance bits, kept in the particle's own lane of two registers
// (the particle's position comes from its lane by readlane).
// Afterwards every lane stages its accepted stream positions in order up to
// the cap, then turns them into slots through the segment's table of cell
// starts.  The lists, counts, overflow counts and skin lists are the lane
// walk's, bit for bit: the same candidates in the same order under the same
// test.  Each candidate is tested by one lane for the whole segment instead
// of by every particle's lane for itself, with no divergence between lanes
// whose cells differ in size.
//
// A particle is not covered when its cell coordinates, computed in floating
// point like the walk's (x + i*s), are not exactly its cell's neighbours,
// when its cell is the outside bucket or one of its neighbour cells wraps
// onto its own, when its 27 cells hold more than kTileMaxCand candidates, or
// when its wavefront has more than kTileSegs segments; a wavefront with a
// particle that is not covered takes the lane walk for all its particles.
constexpr uint32_t kTileRounds = 8;
constexpr uint32_t kTileMaxCand = 64u * kTileRounds;
// (12 segment tables keep the wavefront's LDS under 10 KB, so that 16
// wavefronts share a CU, as many as the registers allow; 16 tables cover a few
// more particles at rest but fit only 14)
constexpr uint32_t kTileSegs = 12;
constexpr uint32_t kTileTab = 28;  // per segment: 27 cell starts; 28 stream offsets (the last is the total)

// v_writelane_b32: lane `lane` of v takes the uniform value `val`
__device__ __forceinline__ uint32_t write_lane(uint32_t v, uint32_t val, uint32_t lane) {
  asm("v_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(val), "{m0}"(lane));
  return v;
}
__device__ __forceinline__ float read_lane(float v, uint32_t lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

template <bool SKIN>
__device__ __forceinline__ bool tile_walk(uint32_t tb, uint32_t lane, uint32_t n, bool query, f3 ps,
                                          const uint32_t* __restrict__ key, const float4* __restrict__ xs,
                                          const uint32_t* __restrict__ start, const KParams& K,
                                          int32_t* __restrict__ nbr, int32_t* __restrict__ nbr_count,
                                          int32_t* __restrict__ snb, int32_t* __restrict__ snb_count,
                                          uint16_t* __restrict__ s_list, uint32_t* __restrict__ s_tab,
                                          uint16_t* __restrict__ s_off, uint64_t* __restrict__ s_qm,
                                          uint32_t& ovf) {
  const uint32_t p = tb + lane;
  const bool valid = p < n;
  const uint32_t kk = valid ? key[p] : 0xFFFFFFFFu;
  const WalkCells w = walk_cells(ps, K.s);
  const uint32_t prev = __shfl_up(kk, 1, 64);
  uint64_t heads = __builtin_amdgcn_ballot_w64(valid && (lane == 0 || kk != prev));
  const uint32_t last = min(64u, n - tb);
  // the current segment's stream (the slot of every stream position) lives in
  // the staging area, which the lists use only after the segments
  uint32_t* __restrict__ s_stream = reinterpret_cast<uint32_t*>(s_list);
  uint32_t wa[2 * kTileRounds], ws[2 * kTileRounds];
#pragma unroll
  for (uint32_t i = 0; i < 2 * kTileRounds; ++i) wa[i] = ws[i] = 0u;
  // pass 1: every segment's table of cells, and which particles it covers
  bool mine = false;
  uint32_t my_seg = 0, seg = 0;
  while (heads != 0ull) {
    const uint32_t s0 = (uint32_t)__builtin_ctzll(heads);
    heads &= heads - 1ull;
    const uint32_t s1 = heads ? (uint32_t)__builtin_ctzll(heads) : last;
    const uint32_t ck = (uint32_t)__builtin_amdgcn_readlane((int)kk, (int)s0);
    if (ck >= K.ncells) continue;  // the outside bucket (its particles are not covered)
    if (seg >= kTileSegs) continue;
    // the segment cell's coordinates in its grid (its scene's)
    const GridView G = grid_of(K, tb + s0);
    if (G.dxy > 0xFFFFFFFFull) continue;  // (a layer of 2^32 cells: coordinates need 64-bit division)
    const uint32_t local = ck - G.off;
    const uint32_t dxy = (uint32_t)G.dxy;
    const uint32_t bz = local / dxy, rem = local - bz * dxy;
    const uint32_t by = rem / G.dimx, bx = rem - by * G.dimx;
    const float fx = (float)bx, fy = (float)by, fz = (float)bz;
    // lanes 0..26: cell `lane` of the walk order (x outer, z inner)
    uint32_t cb = 0, cm = 0, cellk = K.ncells;
    if (lane < 27) {
      const float ox = (float)((int)(lane / 9u) - 1), oy = (float)((int)((lane / 3u) % 3u) - 1),
                  oz = (float)((int)(lane % 3u) - 1);
      cellk = cell_key(fx + ox, fy + oy, fz + oz, G, K);
      if (cellk < K.ncells) {
        cb = start[cellk];
        cm = start[cellk + 1] - cb;
      }
    }
    // exclusive scan of the cell sizes: lane c gets the stream offset of cell c, lane 27 the total
    uint32_t incl = cm;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if ((int)lane >= d) incl += t;
    }
    const uint32_t off = incl - cm;
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 26);
    // the segment's particles whose own walk cells are the segment's
    const bool canon = lane >= s0 && lane < s1 && query && w.cx[0] == fx - 1.0f && w.cx[1] == fx &&
                       w.cx[2] == fx + 1.0f && w.cy[0] == fy - 1.0f && w.cy[1] == fy && w.cy[2] == fy + 1.0f &&
                       w.cz[0] == fz - 1.0f && w.cz[1] == fz && w.cz[2] == fz + 1.0f;
    const uint64_t qm = __builtin_amdgcn_ballot_w64(canon);
    // a grid so narrow that a neighbour cell wraps onto the centre cell lists
    // a particle's own slot twice
    const bool dup = __builtin_amdgcn_ballot_w64(lane < 27 && lane != 13 && cellk == ck) != 0ull;
    if (total > kTileMaxCand || dup || qm == 0ull) continue;
    if (lane < 27) s_tab[seg * kTileTab + lane] = cb;
    if (lane < kTileTab) s_off[seg * kTileTab + lane] = (uint16_t)off;
    if (lane == 0) s_qm[seg] = qm;
    if (canon) {
      mine = true;
      my_seg = seg;
    }
    ++seg;
  }
  // a wavefront takes the lane walk throughout unless every particle that
  // needs a list is covered: the lane walk costs nearly as much for a few
  // lanes as for all of them (measured: leaving the uncovered lanes alone to
  // it was no faster at rest and 1-6% slower over the collapse)
  if (__builtin_amdgcn_ballot_w64(query && !mine) != 0ull) return false;
  __syncthreads();  // one wavefront per block
  // pass 2: the segments' streams
  for (uint32_t sg = 0; sg < seg; ++sg) {
    // (the same in every lane: read into scalar registers)
    const uint64_t qv = s_qm[sg];
    const uint64_t qm = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)qv) |
                        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(qv >> 32)) << 32);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_off[sg * kTileTab + 27]);
    uint32_t cb = 0, off = 0, cm = 0;
    if (lane < 27) {
      cb = s_tab[sg * kTileTab + lane];
      off = s_off[sg * kTileTab + lane];
      cm = (uint32_t)s_off[sg * kTileTab + lane + 1] - off;
    }
    // the candidates' slots through a table in the staging area (which the
    // lists use only afterwards), then their positions into registers: every
    // round's candidates stay there while the segment's particles go by
    for (uint32_t u = 0; u < cm; ++u) s_stream[off + u] = cb + u;
    __syncthreads();
    f3 t[kTileRounds];
#pragma unroll
    for (uint32_t r = 0; r < kTileRounds; ++r) {
      const uint32_t j = r * 64u + lane;
      t[r] = f3{__builtin_inff(), __builtin_inff(), __builtin_inff()};
      if (j < total) t[r] = ld3(xs, s_stream[j]);
    }
    // a particle's own stream position: its slot in the centre cell (place 13)
    const uint32_t jself0 = (uint32_t)__builtin_amdgcn_readlane((int)off, 13) + tb -
                            (uint32_t)__builtin_amdgcn_readlane((int)cb, 13);
    uint64_t qq = qm;
    while (qq != 0ull) {
      const uint32_t i = (uint32_t)__builtin_ctzll(qq);
      qq &= qq - 1ull;
      const f3 pi = {read_lane(ps.x, i), read_lane(ps.y, i), read_lane(ps.z, i)};
      const uint32_t js = jself0 + i;  // the particle itself is no candidate
#pragma unroll
      for (uint32_t r = 0; r < kTileRounds; ++r) {
        if (r * 64u >= total) break;
        const float d2 = magsq3(sub3(pi, t[r]));
        const uint64_t self = (js >> 6) == r ? 1ull << (js & 63u) : 0ull;
        const uint64_t am = __builtin_amdgcn_ballot_w64(d2 < K.ss) & ~self;
        wa[2 * r] = write_lane(wa[2 * r], (uint32_t)am, i);
        wa[2 * r + 1] = write_lane(wa[2 * r + 1], (uint32_t)(am >> 32), i);
        if (SKIN) {
          const uint64_t sm = __builtin_amdgcn_ballot_w64(d2 < K.skin2) & am;
          ws[2 * r] = write_lane(ws[2 * r], (uint32_t)sm, i);
          ws[2 * r + 1] = write_lane(ws[2 * r + 1], (uint32_t)(sm >> 32), i);
        }
      }
    }
    __syncthreads();  // the next segment rewrites the stream
  }
  if (!mine) return false;
  // stage the accepted stream positions (with the skin bit) in order, up to the cap
  const uint32_t base = my_seg * kTileTab;
  const uint32_t total = s_off[base + 27];
  const uint32_t cap = (uint32_t)K.cap;
  uint32_t k = 0, acc = 0;
#pragma unroll
  for (uint32_t wi = 0; wi < 2 * kTileRounds; ++wi) {
    if (32u * wi < total) {
      uint32_t bits = wa[wi];
      const uint32_t rest = total - 32u * wi;
      if (rest < 32u) bits &= (1u << rest) - 1u;
      acc += (uint32_t)__builtin_popcount(bits);
      const uint32_t sb = SKIN ? ws[wi] : 0u;
      while (bits != 0u && k < cap) {
        const uint32_t b = (uint32_t)__builtin_ctz(bits);
        bits &= bits - 1u;
        s_list[k * kFillBlock + lane] = (uint16_t)((32u * wi + b) | (((sb >> b) & 1u) << 15));
        ++k;
      }
    }
  }
  if (acc > k) ovf = acc - k;
  nbr_count[p] = (int32_t)k;
  // flush: stream positions to slots, cell by cell
  uint32_t c = 0, co = 0, nxt = s_off[base + 1], cs = s_tab[base];
  int js = 0;
  for (uint32_t e = 0; e < k; ++e) {
    const uint32_t e16 = s_list[e * kFillBlock + lane];
    const uint32_t j = e16 & 1023u;
    while (j >= nxt) {
      ++c;
      co = nxt;
      cs = s_tab[base + c];
      nxt = s_off[base + c + 1];
    }
    const uint32_t slot = cs + (j - co);
    nbr[list_at(e, p, K.stride)] = (int32_t)slot;
    if (SKIN && (e16 >> 15)) snb[list_at(js++, p, K.stride)] = (int32_t)(slot | (e << kSkinPlaceShift));
  }
  if (SKIN) snb_count[p] = js;
  return true;
}

// Hand-offs and overflow counts go to the shared counters once per
// wavefront (a same-address atomic per lane serialises in L2 when many
// particles of a collapsed region hand off together).
__global__ void __launch_bounds__(kFillBlock) k_fill_neighbors(uint32_t n, uint32_t fb, uint32_t fe,
                                                               const float4* __restrict__ xs,
                                                               const uint32_t* __restrict__ oid,
                                                               const uint32_t* __restrict__ key,
                                                               const uint32_t* __restrict__ start,
                                                               const uint32_t* __restrict__ box, KParams K,
                                                               int32_t* __restrict__ nbr,
                                                               int32_t* __restrict__ nbr_count,
                                                               uint32_t* __restrict__ heavy,
                                                               unsigned long long* __restrict__ counters,
                                                               int skin_mode,
                                                               const unsigned long long* __restrict__ hint,
                                                               int32_t* __restrict__ snb,
                                                               int32_t* __restrict__ snb_count,
                                                               float4* __restrict__ xs0) {
  extern __shared__ uint16_t s_list[];  // [cap + 1][kFillBlock], or the cell-tile walk's stream
  __shared__ uint32_t s_tab[kTileSegs * kTileTab];
  __shared__ uint16_t s_off[kTileSegs * kTileTab];
  __shared__ uint64_t s_qm[kTileSegs];
  const uint32_t lane = threadIdx.x;
  const uint32_t tb = xcd_block() * blockDim.x;
  const uint32_t p = tb + threadIdx.x;
  // skin lists only when the previous step kept them usable to its end
  const int skin = skin_mode && *hint != 0ull;
  if (skin && blockIdx.x == 0 && lane == 0) counters[C_SKIN] = 1ull;
  bool to_heavy = false;
  uint32_t ovf = 0;
  const uint32_t o = p < n ? oid[p] : 0u;
  const bool query = p < n && o >= fb && o < fe;
  bool built = false;
  if (K.tile_lists) {  // every lane of the wavefront takes part
    const f3 ps = p < n ? ld3(xs, p) : f3{0.f, 0.f, 0.f};
    built = skin ? tile_walk<true>(tb, lane, n, query, ps, key, xs, start, K, nbr, nbr_count, snb, snb_count,
                                   s_list, s_tab, s_off, s_qm, ovf)
                 : tile_walk<false>(tb, lane, n, query, ps, key, xs, start, K, nbr, nbr_count, snb, snb_count,
                                    s_list, s_tab, s_off, s_qm, ovf);
  }
  if (p < n) {
    if (skin_mode) xs0[p] = xs[p];
    if (built) {
      // listed by the cell-tile walk
    } else if (!query) {
      nbr_count[p] = 0;
      if (skin) snb_count[p] = 0;
    } else {
      to_heavy = skin ? !walk_list<true>(p, n, lane, s_list, xs, start, K, nbr, nbr_count, ovf, snb, snb_count)
                      : !walk_list<false>(p, n, lane, s_list, xs, start, K, nbr, nbr_count, ovf, snb, snb_count);
    }
  }
  // every lane of the wavefront (one per block) is here
  const uint64_t hm = __ballot(to_heavy);
  if (hm) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&counters[C_HEAVY], (unsigned long long)__popcll(hm));
    base = __shfl(base, 0, 64);
    if (to_heavy) heavy[base + (unsigned long long)__popcll(hm & ((1ull << lane) - 1ull))] = p;
  }
  if (K.count_overflow) {
    unsigned long long ov = ovf;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) ov += __shfl_xor(ov, d, 64);
    const unsigned long long tiles = (unsigned long long)__popcll(__ballot(built && p < n));
    if (lane == 0 && ov) atomicAdd(&counters[C_OVERFLOW], ov);
    if (lane == 0 && tiles) atomicAdd(&counters[C_TILE], tiles);
  }
}

// A group of 64 candidates of the wavefront builder, one per lane (valid:
// the lane holds a candidate other than p; t: its position): the accepted
// ones are appended in slot order through a ballot.  Returns true when the
// list is full and the walk may end.
__device__ __forceinline__ bool take_group(uint32_t q, bool valid, float4 t, f3 ps, const KParams& K, uint32_t n,
                                           uint32_t p, uint64_t below, uint32_t cap, int32_t* __restrict__ nbr,
                                           uint32_t& count, int skin, int32_t* __restrict__ snb, uint32_t& scount) {
  const float d2 = magsq3(sub3(ps, f3{t.x, t.y, t.z}));
  const bool acc = valid && d2 < K.ss;
  const uint64_t m = __ballot(acc);
  const uint32_t pos = count + (uint32_t)__popcll(m & below);
  if (acc && pos < cap) nbr[list_at(pos, p, K.stride)] = (int32_t)q;
  if (skin) {  // entries of the list (pos < cap) within the skin radius, in order
    const bool in = acc && pos < cap && d2 < K.skin2;
    const uint64_t ms = __ballot(in);
    if (in)
      snb[list_at(scount + (uint32_t)__popcll(ms & below), p, K.stride)] = (int32_t)(q | (pos << kSkinPlaceShift));
    scount += (uint32_t)__popcll(ms);
  }
  count += (uint32_t)__popcll(m);
  return count >= cap && !K.count_overflow;
}

// One wavefront per heavy particle; waves stride over the heavy list, whose
// length is read once per wave (every wave exits after the last entry).
__global__ void __launch_bounds__(kBlock) k_fill_heavy(uint32_t n, const float4* __restrict__ xs,
                                                       const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ box,
                                                       const float4* __restrict__ chbox, KParams K,
                                                       int32_t* __restrict__ nbr, int32_t* __restrict__ nbr_count,
                                                       const uint32_t* __restrict__ heavy,
                                                       unsigned long long* __restrict__ counters, int skin_mode,
                                                       const unsigned long long* __restrict__ hint,
                                                       int32_t* __restrict__ snb, int32_t* __restrict__ snb_count) {
  const int skin = skin_mode && *hint != 0ull;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nheavy = (uint32_t)counters[C_HEAVY];
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t cap = (uint32_t)K.cap;
  for (uint32_t h = wave; h < nheavy; h += nwaves) {
    const uint32_t p = heavy[h];
    const f3 ps = ld3(xs, p);
    const WalkCells w = walk_cells(ps, K.s);
    const GridView G = grid_of(K, p);
    uint32_t count = 0;   // accepted so far, uncapped (uniform across the wave)
    uint32_t scount = 0;  // skin entries so far
    // lane c < 27 fetches cell c's range and box test, all in flight at once;
    // the wavefront then walks the cells in order from registers
    uint32_t lb = 0, le = 0;
    if (lane < 27) {
      const uint32_t cell =
          cell_key(pick3(w.cx, (int)lane / 9), pick3(w.cy, ((int)lane / 3) % 3), pick3(w.cz, (int)lane % 3), G, K);
      if (cell < K.ncells) {
        lb = start[cell];
        le = start[cell + 1];
        if (le - lb > kBoxMin && cell_far(box, cell, ps, K.ss)) le = lb;  // far: walk nothing
      }
    }
    for (int c = 0; c < 27; ++c) {
      const uint32_t b = (uint32_t)__shfl((int)lb, c, 64), e = (uint32_t)__shfl((int)le, c, 64);
      if (b == e) continue;
      // candidates come in groups of 64 (one per lane); two groups are
      // loaded before the first is tested, so that two loads of a
      // wavefront are in flight instead of one
      if (e - b <= kChunkMin) {
        for (uint32_t base = b; base < e; base += 128) {
          const uint32_t qa = base + lane, qb = base + 64 + lane;
          const bool va = qa < e && qa != p, vb = qb < e && qb != p;
          float4 ta = make_float4(0.f, 0.f, 0.f, 0.f), tb = ta;
          if (va) ta = xs[qa];
          if (vb) tb = xs[qb];
          if (take_group(qa, va, ta, ps, K, n, p, below, cap, nbr, count, skin, snb, scount)) goto done;
          if (base + 64 < e && take_group(qb, vb, tb, ps, K, n, p, below, cap, nbr, count, skin, snb, scount)) goto done;
        }
        continue;
      }
      // crowded cell: each lane tests one chunk box, then the wavefront
      // gathers the chunks that may hold a neighbour, in slot order
      const uint32_t c_last = (e - 1) >> 6;
      for (uint32_t c0 = b >> 6; c0 <= c_last; c0 += 64) {
        const uint32_t ch = c0 + lane;
        bool near = false;
        if (ch <= c_last) {
          const float4 lo = chbox[2 * ch], hi = chbox[2 * ch + 1];
          near = !box_far(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, ps, K.ss);
        }
        uint64_t todo = __ballot(near);
        while (todo) {
          const uint32_t ca = (uint32_t)__builtin_ctzll(todo);
          todo &= todo - 1ull;
          const bool two = todo != 0ull;
          const uint32_t cb = two ? (uint32_t)__builtin_ctzll(todo) : ca;
          if (two) todo &= todo - 1ull;
          const uint32_t qa = ((c0 + ca) << 6) + lane, qb = ((c0 + cb) << 6) + lane;
          const bool va = qa >= b && qa < e && qa != p, vb = two && qb >= b && qb < e && qb != p;
          float4 ta = make_float4(0.f, 0.f, 0.f, 0.f), tb = ta;
          if (va) ta = xs[qa];
          if (vb) tb = xs[qb];
          if (take_group(qa, va, ta, ps, K, n, p, below, cap, nbr, count, skin, snb, scount)) goto done;
          if (two && take_group(qb, vb, tb, ps, K, n, p, below, cap, nbr, count, skin, snb, scount)) goto done;
        }
      }
    }
  done:
    if (lane == 0) {
      nbr_count[p] = (int32_t)(count < cap ? count : cap);
      if (skin) snb_count[p] = (int32_t)scount;
      if (count > cap && K.count_overflow) atomicAdd(&counters[C_OVERFLOW], (unsigned long long)(count - cap));
    }
  }
}

void launch_fill_neighbors(pbf_context* c, uint64_t fb, uint64_t fe, bool want_skin) {
  const uint32_t n = (uint32_t)c->n;
  if (!n) return;
  // skin lists for the fused step (their packed form needs slots below 2^25);
  // x*_0 is kept whenever they may be used, the lists themselves when the
  // previous step's moves stayed below their limit (c->skin_hint)
  const int skin = want_skin && c->K.skin_mode && c->n < (1ull << kSkinPlaceShift);
  KernelTimer* t = timer_begin(c, "fill_neighbors");
  // the staging area holds a wavefront's lists, or the cell-tile walk's stream of slots
  const uint32_t stage = std::max<uint32_t>(kFillBlock * (c->K.cap + 1) * sizeof(uint16_t),
                                            kTileMaxCand * sizeof(uint32_t));
  hipLaunchKernelGGL(k_fill_neighbors, dim3(nblocks(n, kFillBlock)), dim3(kFillBlock), stage, c->stream, n,
                     (uint32_t)fb,
                     (uint32_t)fe, (const float4*)c->xs, (const uint32_t*)c->oid, (const uint32_t*)c->key,
                     (const uint32_t*)c->cell_start,
                     (const uint32_t*)c->cbox, c->K, c->nbr, c->nbr_count, c->heavy, c->counters, skin,
                     (const unsigned long long*)c->skin_hint, c->snb, c->snb_count, c->xs0);
  timer_end(c, t);
  // the cell and chunk boxes of the wavefront builder, built only when some particle went to it
  launch_cell_box(c);
  // enough waves to cover the chip; each loops over the heavy list (usually empty)
  t = timer_begin(c, "fill_heavy");
  const uint32_t blocks = std::min<uint32_t>(nblocks((uint64_t)n * 64), c->tune.heavy_blocks);
  hipLaunchKernelGGL(k_fill_heavy, dim3(blocks), dim3(kBlock), 0, c->stream, n, (const float4*)c->xs,
                     (const uint32_t*)c->cell_start, (const uint32_t*)c->cbox, (const float4*)c->chbox, c->K,
                     c->nbr, c->nbr_count, (const uint32_t*)c->heavy, c->counters, skin,
                     (const unsigned long long*)c->skin_hint, c->snb, c->snb_count);
  timer_end(c, t);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (u32 key, u32 value) for the sort_by_key primitive.
// Digit = 8 bits.  Per pass: tile histograms (digit-major), scan, stable
// scatter.  Inside a tile, each element's rank among equal digits keeps
// input order: waves rank with ballots, waves are combined through LDS.
constexpr int kRadixItems = 4;                    // items per thread
constexpr int kRadixTile = kBlock * kRadixItems;  // 1024 keys per tile

__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint32_t* __restrict__ keys, uint64_t n, int shift,
                                                       uint32_t ntiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRadixTile;
  for (int i = 0; i < kRadixItems; ++i) {
    const uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter one tile.  Elements are processed in rounds of 256 (one per
// thread, in input order); within a round, rank = number of earlier lanes of
// the round with the same digit, found with 8 ballots per wave.
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals, uint64_t n, int shift,
                                                          uint32_t ntiles, const uint32_t* __restrict__ offs,
                                                          uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals) {
  __shared__ uint32_t base_d[256];       // running output offset per digit for this tile
  __shared__ uint32_t wave_cnt[4][256];  // per wave, count of each digit in this round
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  base_d[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
  const uint64_t tbase = (uint64_t)blockIdx.x * kRadixTile;
  for (int r = 0; r < kRadixItems; ++r) {
    for (int i = 0; i < 4; ++i) wave_cnt[i][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t idx = tbase + (uint64_t)r * kBlock + threadIdx.x;
    const bool valid = idx < n;
    const uint32_t k = valid ? keys[idx] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    // lanes of this wave with the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot(((d >> b) & 1u) != 0u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t rank_in_wave = (uint32_t)__popcll(peers & lower);
    const uint32_t wave_total = (uint32_t)__popcll(peers);
    if (valid && rank_in_wave == 0) wave_cnt[w][d] = wave_total;
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int i = 0; i < w; ++i) before += wave_cnt[i][d];
      const uint32_t pos = base_d[d] + before + rank_in_wave;
      okeys[pos] = k;
      ovals[pos] = vals[idx];
    }
    __syncthreads();
    // advance the digit offsets by this round's counts
    {
      const uint32_t dd = threadIdx.x;
      base_d[dd] += wave_cnt[0][dd] + wave_cnt[1][dd] + wave_cnt[2][dd] + wave_cnt[3][dd];
    }
    __syncthreads();
  }
}

uint64_t radix_hist_size(uint64_t n) { return 256ull * ((n + kRadixTile - 1) / kRadixTile); }

// Stable LSD sort of (keys, vals) over key_bits bits, ping-ponging between
// (k0, v0) and (k1, v1).  Returns 0 or 1: the buffer pair holding the result.
int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint64_t n, uint32_t key_bits,
                     uint32_t* hist, uint32_t* offs, uint32_t* partials, hipStream_t s) {
  if (n <= 1) return 0;
  const uint32_t ntiles = (uint32_t)((n + kRadixTile - 1) / kRadixTile);
  const uint64_t hn = 256ull * ntiles;
  uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
  int which = 0;
  for (uint32_t shift = 0; shift < key_bits; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, s, (const uint32_t*)ki, n, (int)shift, ntiles,
                       hist);
    launch_exclusive_scan(hist, offs, hn, partials, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kBlock), 0, s, (const uint32_t*)ki,
                       (const uint32_t*)vi, n, (int)shift, ntiles, (const uint32_t*)offs, ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
    which ^= 1;
  }
  return which;
}

int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t key_bits, hipStream_t s) {
  if (n <= 1) return PBF_OK;
  if (n > 0xFFFFFFFFull) return set_error(PBF_ERR_UNSUPPORTED, "sort_by_key: n must be < 2^32");
  if (key_bits == 0 || key_bits > 32) key_bits = 32;
  const uint64_t hn = radix_hist_size(n);
  uint32_t *k2 = nullptr, *v2 = nullptr, *hist = nullptr, *offs = nullptr, *part = nullptr;
  int rc = PBF_OK;
  if ((rc = check_hip(hipMallocAsync((void**)&k2, n * 4, s), "hipMallocAsync")) ||
      (rc = check_hip(hipMallocAsync((void**)&v2, n * 4, s), "hipMallocAsync")) ||
      (rc = check_hip(hipMallocAsync((void**)&hist, hn * 4, s), "hipMallocAsync")) ||
      (rc = check_hip(hipMallocAsync((void**)&offs, (hn + 1) * 4, s), "hipMallocAsync")) ||
      (rc = check_hip(hipMallocAsync((void**)&part, scan_partials_size(hn) * 4, s), "hipMallocAsync"))) {
    // the buffers allocated before the failure are freed below
  } else {
    if (radix_sort_pairs(keys, vals, k2, v2, n, key_bits, hist, offs, part, s)) {
      rc = check_hip(hipMemcpyAsync(keys, k2, n * 4, hipMemcpyDeviceToDevice, s), "sort_by_key copy");
      if (!rc) rc = check_hip(hipMemcpyAsync(vals, v2, n * 4, hipMemcpyDeviceToDevice, s), "sort_by_key copy");
    }
    if (!rc) rc = check_hip(hipGetLastError(), "radix sort launch");
  }
  for (uint32_t* b : {k2, v2, hist, offs, part})
    if (b) (void)hipFreeAsync(b, s);
  return rc;
}

// lower_bound / upper_bound of every value of a counting range
__global__ void k_bounds(const uint32_t* __restrict__ sorted, uint64_t n, uint64_t b, uint64_t cnt,
                         uint32_t* __restrict__ out, int upper) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint64_t val = b + i;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint64_t k = sorted[mid];
    if (upper ? (k <= val) : (k < val)) lo = mid + 1;
    else hi = mid;
  }
  out[i] = (uint32_t)lo;
}

void launch_bounds_counting(const uint32_t* sorted, uint64_t n, uint64_t b, uint64_t e, uint32_t* out, int upper,
                            hipStream_t s) {
  const uint64_t cnt = e - b;
  if (!cnt) return;
  hipLaunchKernelGGL(k_bounds, dim3(nblocks(cnt)), dim3(kBlock), 0, s, sorted, n, b, cnt, out, upper);
}

}  // namespace pbf
