// Allocation-free form of the pixel regrouping (snisr_regroup_pixels) for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops.  The regrouping's
// workspace (digit histograms, sorted keys, run heads, block offsets) lives in `scratch`, a
// file-scope device array of fixed size; each GPU holds its own copy, whose address is looked up
// once per device with hipGetSymbolAddress and kept.  Every counter in it is written by the
// regrouping's own kernels before it is read, so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
#include <mutex>

#include "snisr_common.hpp"

namespace snisr {

constexpr int64_t kStaticRegroupMaxPixels = int64_t(1) << 20;   // 1,048,576 pixels (a 1280x720 frame fits)
constexpr int64_t kStaticScratchBytes = 10 * 1024 * 1024;       // >= snisr_regroup_workspace_size(2^20)
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
static int scratch_for_current_device(void** out) {
  int dev = -1;
  SNISR_HIP(hipGetDevice(&dev));
  SNISR_REQUIRE(dev >= 0 && dev < kMaxDevices, SNISR_ERR_DEVICE, "regroup_pixels_static: device %d beyond %d", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SNISR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SNISR_REQUIRE(p, SNISR_ERR_DEVICE, "regroup_pixels_static: no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SNISR_OK;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_regroup_pixels_static_max(void) { return kStaticRegroupMaxPixels; }

int snisr_regroup_pixels_static(const int16_t* net_idx, int64_t n, int32_t pixels_per_block, int64_t max_blocks,
                                int32_t* query, int16_t* block_net, int32_t* block_starts, int32_t* n_blocks,
                                void* stream) {
  SNISR_REQUIRE(n >= 0, SNISR_ERR_ARG, "regroup_pixels_static: bad size");
  SNISR_REQUIRE(n <= kStaticRegroupMaxPixels, SNISR_ERR_CAPACITY,
                "regroup_pixels_static: %lld pixels above the static workspace's maximum of %lld (use "
                "snisr_regroup_pixels with a workspace)",
                (long long)n, (long long)kStaticRegroupMaxPixels);
  const int64_t need = snisr_regroup_workspace_size(n);
  SNISR_REQUIRE(need > 0 && need <= kStaticScratchBytes, SNISR_ERR_CAPACITY,
                "regroup_pixels_static: workspace of %lld bytes above the static %lld", (long long)need,
                (long long)kStaticScratchBytes);
  void* ws = nullptr;
  if (int e = scratch_for_current_device(&ws)) return e;
  return snisr_regroup_pixels(net_idx, n, pixels_per_block, max_blocks, query, block_net, block_starts, n_blocks, ws,
                              kStaticScratchBytes, stream);
}

}  // extern "C"
