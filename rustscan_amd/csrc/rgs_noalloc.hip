// Sort entry point that allocates nothing.  Its scratch memory (the sort's
// ping-pong key/value arrays and digit histograms) is a file-scope device
// array, so callers that must not allocate device memory (code under HIP
// stream capture, allocation-free serving loops) can still sort.
//
// The array's name is only a host-side stand-in: its device address differs
// per GPU and is looked up once per device at context creation with
// hipGetSymbolAddress; every later use goes through that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rgs.h"
#include "rgs_internal.h"

namespace rgs {
uint64_t radix_scratch_words(uint64_t cap);
int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch, const uint32_t* gsrc, uint32_t* gout, int count_ib);
}  // namespace rgs

namespace {

constexpr uint64_t kMaxPairs = RGS_SORT_NOALLOC_MAX_PAIRS;
// radix histograms: 256 digits x at most 2048 blocks, plus the digit totals
constexpr uint64_t kHistWords = 256ull * 2048ull + 512ull;
constexpr uint64_t kScratchWords = 2ull * kMaxPairs + kHistWords;  // 18.9 MB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

std::mutex g_mu;
uint32_t* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per GPU

}  // namespace

namespace rgs {

// Looks up this device's copy of the scratch array (the device must be
// current).  Called once per device, at context creation.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch_dev[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[device] = static_cast<uint32_t*>(p);
  return hipSuccess;
}

// Stable ascending sort of n <= kMaxPairs (key, value) pairs on `stream`
// with the device's scratch array; no allocation, no synchronisation.
int noalloc_sort_pairs(int device, hipStream_t stream, const uint32_t* keys, const uint32_t* values,
                       uint32_t* out_keys, uint32_t* out_values, uint32_t n, const char** err) {
  if (n > kMaxPairs) {
    *err = "n exceeds RGS_SORT_NOALLOC_MAX_PAIRS";
    return RGS_ERR_INVALID;
  }
  if (radix_scratch_words(n) > kHistWords) {
    *err = "sort histogram larger than the scratch array";
    return RGS_ERR_INVALID;
  }
  uint32_t* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
  }
  if (!base) {
    *err = "scratch array of this device was not set up";
    return RGS_ERR_STATE;
  }
  uint32_t* ping_k = base;
  uint32_t* ping_v = base + kMaxPairs;
  uint32_t* hist = base + 2 * kMaxPairs;
  // Four 8-bit passes: in -> ping -> out -> ping -> out, so the result lands
  // in the caller's arrays (which may alias the inputs: pass 0 reads the
  // inputs before pass 1 writes the outputs).
  const int where = radix_sort_pairs(stream, keys, values, ping_k, ping_v, out_keys, out_values, nullptr, n, n, 32,
                                     hist, nullptr, nullptr, 0);
  if (where != 2) {
    *err = "sort ended in the scratch array";
    return RGS_ERR_STATE;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return RGS_ERR_HIP;
  }
  return RGS_OK;
}

}  // namespace rgs
