// pbf_noalloc.hip -- sort_by_key for callers that cannot allocate device memory.
//
// pbf_sort_by_key_u32 takes its radix-sort work space (a second key/value
// pair of arrays, digit histograms, their scan and the scan's tile sums) from
// hipMallocAsync on every call.  pbf_sort_by_key_u32_noalloc runs the same
// kernels (radix_sort_pairs, pbf_grid.hip) on a fixed work space instead: the
// file-scope device array `scratch` below, one copy per GPU, which the GPU
// holds until the process ends.  It calls no hipMalloc, hipFree or allocating
// library routine.
//
// Every call on a GPU uses that GPU's one array, so calls that overlap on one
// GPU (two host threads, or two streams) race on it.  Nothing in it needs to
// start at zero: each pass writes every histogram entry, scan entry and tile
// sum it later reads, and the ping-pong arrays are written before they are read.
#include <mutex>
#include <vector>

#include "pbf_internal.h"

namespace pbf {

namespace {

// Largest n the allocation-free sort accepts.
constexpr uint64_t kMaxPairs = PBF_NOALLOC_MAX_PAIRS;
// Sizes in 32-bit words, as radix_hist_size and scan_partials_size give them
// for kMaxPairs (1,024 keys per radix tile, 2,048 values per scan tile).
constexpr uint64_t kHistWords = 256ull * ((kMaxPairs + 1023) / 1024);
constexpr uint64_t kOffsWords = kHistWords + 1;
constexpr uint64_t kPartialWords = (kHistWords + 2047) / 2048 + 1;
// each region starts on a 256-byte boundary
constexpr uint64_t align64(uint64_t w) { return (w + 63) & ~63ull; }
constexpr uint64_t kOffKeys = 0;
constexpr uint64_t kOffVals = kOffKeys + align64(kMaxPairs);
constexpr uint64_t kOffHist = kOffVals + align64(kMaxPairs);
constexpr uint64_t kOffOffs = kOffHist + align64(kHistWords);
constexpr uint64_t kOffPartials = kOffOffs + align64(kOffsWords);
constexpr uint64_t kScratchWords = kOffPartials + align64(kPartialWords);

}  // namespace

// The work space: 2 x 4 MiB of ping-pong keys and values, 1 MiB of digit
// histograms and 1 MiB of their scan at kMaxPairs = 2^20 (about 10 MiB).
// In host code this name is only a handle for hipGetSymbolAddress; kernels get
// the device address that call returns.
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

std::mutex g_scratch_mu;
std::vector<uint32_t*> g_scratch_ptr;  // device address of `scratch`, per device id

// The current device's copy of `scratch`, looked up once per device with that
// device current.
int scratch_for_current_device(uint32_t** out) {
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if ((rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))) return rc;
    if (!p) return set_error(PBF_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PBF_OK;
}

}  // namespace

}  // namespace pbf

using namespace pbf;

extern "C" int pbf_sort_by_key_u32_noalloc(uint32_t* d_keys, uint32_t* d_values, uint64_t n, uint32_t key_bits) {
  PBF_TRACE("pbf_sort_by_key_u32_noalloc");
  if (n && (!d_keys || !d_values)) return set_error(PBF_ERR_ARG, "null argument");
  if (n > kMaxPairs)
    return set_error(PBF_ERR_CAPACITY, "pbf_sort_by_key_u32_noalloc: n = " + std::to_string(n) +
                                      " exceeds its fixed work space of PBF_NOALLOC_MAX_PAIRS = " +
                                      std::to_string(kMaxPairs) + " pairs");
  if (n <= 1) return PBF_OK;
  if (key_bits == 0 || key_bits > 32) key_bits = 32;
  // the sizes the sort needs for n never exceed the regions laid out for kMaxPairs
  if (radix_hist_size(n) > kHistWords || scan_partials_size(radix_hist_size(n)) > kPartialWords)
    return set_error(PBF_ERR_UNSUPPORTED, "pbf_sort_by_key_u32_noalloc: work space layout");
  uint32_t* ws = nullptr;
  int rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  const hipStream_t s = 0;
  if (radix_sort_pairs(d_keys, d_values, ws + kOffKeys, ws + kOffVals, n, key_bits, ws + kOffHist, ws + kOffOffs,
                       ws + kOffPartials, s)) {
    rc = check_hip(hipMemcpyAsync(d_keys, ws + kOffKeys, n * 4, hipMemcpyDeviceToDevice, s), "sort_by_key copy");
    if (!rc)
      rc = check_hip(hipMemcpyAsync(d_values, ws + kOffVals, n * 4, hipMemcpyDeviceToDevice, s), "sort_by_key copy");
  }
  if (!rc) rc = check_hip(hipGetLastError(), "radix sort launch");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(s), "sort_by_key");
}
