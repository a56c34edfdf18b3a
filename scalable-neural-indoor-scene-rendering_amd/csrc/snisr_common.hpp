// Host-side helpers shared by the C-ABI translation units: error state and argument checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "snisr.h"

namespace snisr {

void set_error(const char* fmt, ...);

#define SNISR_REQUIRE(cond, code, ...)      \
  do {                                      \
    if (!(cond)) {                          \
      ::snisr::set_error(__VA_ARGS__);      \
      return (code);                        \
    }                                       \
  } while (0)

#define SNISR_HIP(call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      ::snisr::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
      return SNISR_ERR_DEVICE;                                                                 \
    }                                                                                          \
  } while (0)

#define SNISR_LAUNCHED()                                                                       \
  do {                                                                                         \
    hipError_t e_ = hipGetLastError();                                                         \
    if (e_ != hipSuccess) {                                                                    \
      ::snisr::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return SNISR_ERR_DEVICE;                                                                 \
    }                                                                                          \
  } while (0)

inline bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// Records stage event k of `out` (if the caller supplied that many) on `stream`.
inline hipError_t stage_event(const snisr_render_out* out, int k, hipStream_t stream) {
  if (!out || !out->stage_events || k >= out->n_stage_events || !out->stage_events[k]) return hipSuccess;
  return hipEventRecord((hipEvent_t)out->stage_events[k], stream);
}

// Checks every field of a scene descriptor that the kernels rely on.
int check_scene(const snisr_scene* s, bool need_feats);
int check_params(const snisr_march_params* p);
int check_rays(const snisr_rays* r, const snisr_march_params* p);

// f32 "transposed" MLP packing used by the FP32 kernels: per layer Wt[in][out4] then b[out4]
// (out4 = out rounded up to 4).
int64_t mlp_f32t_stride(int in, int width, int layers);

// MFMA weight packing (implemented with the MFMA kernels).
int64_t mfma_pack_bytes(const snisr_scene* scene, int dtype);
int mfma_pack(const snisr_scene* scene, int dtype, void* packed, void* stream);
int device_cus();  // compute units of the current device (render_mfma.hip)

}  // namespace snisr
