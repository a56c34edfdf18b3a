// One Adam step over a flat FP32 parameter tensor in a single pass (the optimizer of the training
// chain: torch.optim.Adam(params, weight_decay=1e-5) at training/scripts/tile/tileGroup.py:342,
// stepped at tile.py:780-785).  Per element, in the order of torch's single-tensor Adam:
//   g  = grad + weight_decay * param                     (when weight_decay != 0)
//   m  = m + (1 - beta1) * (g - m)                       (lerp with weight 1 - beta1 < 0.5)
//   v  = v * beta2 + (1 - beta2) * g * g
//   p  = p + (-step_size) * (m / (sqrt(v) / sqrt(bc2) + eps)),  step_size = lr / bc1,
// bc1 = 1 - beta1^step and bc2 = 1 - beta2^step computed on the host in double precision.  The
// parameter, gradient and both moments are read and written once (7 x 4 bytes per element).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "snisr_common.hpp"

namespace snisr {
namespace {

struct AdamScalars {
  float lr_bc1, bc2_sqrt, beta1c, beta2, beta2c, eps, wd;  // lr / bc1, sqrt(bc2), 1 - beta1, beta2, 1 - beta2
};

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, const AdamScalars& s) {
  if (s.wd != 0.0f) g = g + s.wd * p;
  m = m + s.beta1c * (g - m);
  v = v * s.beta2;
  v = v + s.beta2c * g * g;
  const float denom = __builtin_sqrtf(v) / s.bc2_sqrt + s.eps;
  p = p + -s.lr_bc1 * (m / denom);
}

// four elements per thread with 16-byte accesses when the pointers allow it, one otherwise
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, int64_t n, AdamScalars s, int vec4) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i >= n) return;
  if (vec4 && i + 4 <= n) {
    float4 P = *reinterpret_cast<const float4*>(p + i), G = *reinterpret_cast<const float4*>(g + i);
    float4 M = *reinterpret_cast<const float4*>(m + i), V = *reinterpret_cast<const float4*>(v + i);
    adam_one(P.x, G.x, M.x, V.x, s);
    adam_one(P.y, G.y, M.y, V.y, s);
    adam_one(P.z, G.z, M.z, V.z, s);
    adam_one(P.w, G.w, M.w, V.w, s);
    *reinterpret_cast<float4*>(p + i) = P;
    *reinterpret_cast<float4*>(m + i) = M;
    *reinterpret_cast<float4*>(v + i) = V;
  } else {
    for (int64_t k = i; k < i + 4 && k < n; ++k) adam_one(p[k], g[k], m[k], v[k], s);
  }
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" int snisr_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, double lr,
                               double beta1, double beta2, double eps, double weight_decay, int64_t step,
                               void* stream) {
  SNISR_REQUIRE(n >= 0, SNISR_ERR_ARG, "adam_step: n < 0");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(param && grad && exp_avg && exp_avg_sq, SNISR_ERR_ARG, "adam_step: NULL pointer");
  SNISR_REQUIRE(step >= 1, SNISR_ERR_ARG, "adam_step: step counts from 1 (got %lld)", (long long)step);
  SNISR_REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, SNISR_ERR_ARG,
                "adam_step: betas must lie in [0, 1)");
  // the scalars as torch forms them: in double precision on the host, rounded to FP32 once
  const double bc1 = 1.0 - std::pow(beta1, (double)step);
  const double bc2 = 1.0 - std::pow(beta2, (double)step);
  AdamScalars s;
  s.lr_bc1 = (float)(lr / bc1);
  s.bc2_sqrt = (float)std::sqrt(bc2);
  s.beta1c = (float)(1.0 - beta1);
  s.beta2 = (float)beta2;
  s.beta2c = (float)(1.0 - beta2);
  s.eps = (float)eps;
  s.wd = (float)weight_decay;
  const int vec4 = aligned(param, 16) && aligned(grad, 16) && aligned(exp_avg, 16) && aligned(exp_avg_sq, 16);
  const int64_t threads = (n + 3) / 4;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, (hipStream_t)stream, param, grad,
                     exp_avg, exp_avg_sq, n, s, vec4);
  SNISR_LAUNCHED();
  return SNISR_OK;
}
