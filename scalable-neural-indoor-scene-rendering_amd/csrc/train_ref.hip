// The reference's training-form compositing and specular head, forward and hand-written adjoint:
//   - utils.volume_rendering (training/scripts/utils.py:823-837): alpha = 1 - exp(-sigma * dist),
//     transparency = exclusive cumprod of (1 - alpha + 1e-10) started at init_T, w = alpha * T,
//     out = sum w rgb.  With init_T = the leftover transparency of the in-tile pass it is the
//     background pass chained behind it (training/scripts/tile/tile.py:703-704);
//   - the specular head of Tile.inference_specular (training/scripts/tile/tile.py:198-223): the MLP
//     output's 27 coefficients read as [basis][rgb], evaluated in real SH of degree <= 2 at the
//     normalised ray direction (training/scripts/sh.py:55-111) and passed through a sigmoid.
// One wavefront per ray (compositing, lane scans) or one thread per sample (SH).  FP32 throughout; the
// exponential is exp_cw.
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {
namespace {

constexpr float kTransEps = 1e-10f;  // the reference's 1 - alpha + 1e-10

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// One wavefront per ray, one lane per sample (64-sample chunks with the transparency carried):
// T_i = init_T * exclusive product of f_j = 1 - alpha_j + 1e-10 (a lane scan), out = sum alpha T rgb.
__global__ __launch_bounds__(256) void k_volume_render_fwd(const float* __restrict__ rgb, const float* __restrict__ sigma,
                                                           const float* __restrict__ dists,
                                                           const float* __restrict__ init_T, int64_t n_rays, int n,
                                                           float* __restrict__ out, float* __restrict__ trans,
                                                           float* __restrict__ alpha) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n_rays) return;
  const int lane = threadIdx.x & 63;
  float T0 = init_T ? init_T[r] : 1.0f;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const int64_t k = r * n + i;
    float a = 0.0f, f = 1.0f, x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
    if (i < n) {
      a = 1.0f - exp_cw(-(sigma[k] * dists[k]));
      f = 1.0f - a + kTransEps;
      x0 = rgb[3 * k];
      x1 = rgb[3 * k + 1];
      x2 = rgb[3 * k + 2];
    }
    float p = f;  // inclusive product over lanes 0..lane
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const float t = __shfl_up(p, o);
      if (lane >= o) p = p * t;
    }
    float ex = __shfl_up(p, 1);
    if (lane == 0) ex = 1.0f;
    const float T = T0 * ex, w = a * T;
    if (i < n) {
      trans[k] = T;
      alpha[k] = a;
    }
    c0 = c0 + wave_sum(w * x0);
    c1 = c1 + wave_sum(w * x1);
    c2 = c2 + wave_sum(w * x2);
    T0 = T0 * __shfl(p, 63);
  }
  if (lane == 0) {
    out[3 * r] = c0;
    out[3 * r + 1] = c1;
    out[3 * r + 2] = c2;
  }
}

// Reverse pass over the forward's saved transparency and alpha (T_i, alpha_i).  With
// e_i = exp(-sigma_i d_i), a_i = e_i + 1e-10, T_{i+1} = T_i a_i, w_i = alpha_i T_i and G_i =
// dL/dT_i (G_n = 0):
//   dL/dc_i = g w_i;  dL/dalpha_i = T_i (g . c_i) + ga_i;  G_i = alpha_i (g . c_i) + gt_i + a_i G_{i+1};
//   dL/de_i = -dL/dalpha_i + T_i G_{i+1};  dL/dsigma_i = -d_i e_i dL/de_i;  dL/dinit_T = G_0.
// e_i is recomputed as 1 - alpha_i, the value the forward used.  One wavefront per ray: the affine
// recurrence G_i = A_i + a_i G_{i+1} is a suffix scan of the maps x -> A_i + a_i x over the lanes,
// 64-sample chunks from the last, with G carried between chunks.
__global__ __launch_bounds__(256) void k_volume_render_bwd(const float* __restrict__ rgb, const float* __restrict__ dists,
                                                           const float* __restrict__ trans,
                                                           const float* __restrict__ alpha, int64_t n_rays, int n,
                                                           const float* __restrict__ g_out,
                                                           const float* __restrict__ g_trans,
                                                           const float* __restrict__ g_alpha,
                                                           float* __restrict__ g_rgb, float* __restrict__ g_sigma,
                                                           float* __restrict__ g_init) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n_rays) return;
  const int lane = threadIdx.x & 63;
  const float g0 = g_out[3 * r], g1 = g_out[3 * r + 1], g2 = g_out[3 * r + 2];
  float carry = 0.0f;  // G at the first sample after the chunk
  for (int i0 = ((n - 1) / 64) * 64; i0 >= 0; i0 -= 64) {
    const int i = i0 + lane;
    const int64_t k = r * n + i;
    const bool on = i < n;
    float T = 0.0f, a = 0.0f, gc = 0.0f;
    if (on) {
      T = trans[k];
      a = alpha[k];
      const float x0 = rgb[3 * k], x1 = rgb[3 * k + 1], x2 = rgb[3 * k + 2];
      gc = g0 * x0 + g1 * x1 + g2 * x2;
      const float w = a * T;
      g_rgb[3 * k] = g0 * w;
      g_rgb[3 * k + 1] = g1 * w;
      g_rgb[3 * k + 2] = g2 * w;
    }
    const float e = 1.0f - a;
    float pf = on ? e + kTransEps : 1.0f;                               // identity past the end
    float pa = on ? a * gc + (g_trans ? g_trans[k] : 0.0f) : 0.0f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {  // (pf, pa) := map_lane o ... o map_{lane + 2^s - 1}
      const float tf = __shfl_down(pf, o), ta = __shfl_down(pa, o);
      if (lane + o < 64) {
        pa = pa + pf * ta;
        pf = pf * tf;
      }
    }
    const float G = pa + pf * carry;     // G_i
    float Gn = __shfl_down(G, 1);        // G_{i+1}
    if (lane == 63) Gn = carry;
    if (on) {
      const float d_alpha = T * gc + (g_alpha ? g_alpha[k] : 0.0f);
      const float d_e = -d_alpha + T * Gn;
      g_sigma[k] = -dists[k] * e * d_e;
    }
    carry = __shfl(G, 0);
  }
  if (g_init && lane == 0) g_init[r] = n > 0 ? carry : 0.0f;
}

// real SH of degree <= 2 at unit direction (x, y, z), the reference's sign convention; unused
// entries of a lower degree are zero
__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float (&b)[9]) {
  constexpr float c0 = 0.28209479177387814f, c1 = 0.4886025119029199f, c2 = 1.0925484305920792f;
  constexpr float c22 = 0.31539156525252005f, c24 = 0.5462742152960396f;
  b[0] = c0;
#pragma unroll
  for (int j = 1; j < 9; ++j) b[j] = 0.0f;
  if (deg > 0) {
    b[1] = -c1 * y;
    b[2] = c1 * z;
    b[3] = -c1 * x;
  }
  if (deg > 1) {
    b[4] = c2 * (x * y);
    b[5] = -c2 * (y * z);
    b[6] = c22 * (2.0f * (z * z) - x * x - y * y);
    b[7] = -c2 * (x * z);
    b[8] = c24 * (x * x - y * y);
  }
}

__device__ __forceinline__ V3 unit_dir(const float* d) {
  const float inv = 1.0f / __builtin_sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  return v3(d[0] * inv, d[1] * inv, d[2] * inv);
}

// spec[s][c] = sigmoid(sum_b Y_b(dir) coeffs[s][b * 3 + c]), dir = dirs of ray s / spr
__global__ __launch_bounds__(256) void k_sh_sigmoid_fwd(const float* __restrict__ coeffs, int coeff_stride,
                                                        bool vec_rows, const float* __restrict__ dirs, int64_t n,
                                                        int spr, int deg, float* __restrict__ spec) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const V3 d = unit_dir(dirs + 3 * (s / spr));
  float b[9];
  sh_basis(deg, d.x, d.y, d.z, b);
  const int nb = (deg + 1) * (deg + 1);
  const float* q = coeffs + s * coeff_stride;
  // the 27 coefficients of the row: 16-byte loads when rows are 16-byte aligned (a row per lane,
  // so scalar loads would touch a cache line per lane 27 times)
  float v[28];
  if (vec_rows) {
#pragma unroll
    for (int t = 0; t < 7; ++t) {
      const float4 x = reinterpret_cast<const float4*>(q)[t];
      v[4 * t] = x.x;
      v[4 * t + 1] = x.y;
      v[4 * t + 2] = x.z;
      v[4 * t + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int t = 0; t < 27; ++t) v[t] = q[t];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < 9; ++j)  // fully unrolled: b and v stay in registers
      if (j < nb) acc = acc + b[j] * v[3 * j + c];
    spec[3 * s + c] = sigmoid_cw(acc);
  }
}

// dL/dcoeffs[s][b * 3 + c] = dL/dspec[s][c] spec (1 - spec) Y_b; coefficients of a higher degree
// than deg get zero
__global__ __launch_bounds__(256) void k_sh_sigmoid_bwd(const float* __restrict__ spec, const float* __restrict__ dirs,
                                                        int64_t n, int spr, int deg, const float* __restrict__ g_spec,
                                                        float* __restrict__ g_coeffs, int coeff_stride) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const V3 d = unit_dir(dirs + 3 * (s / spr));
  float b[9];
  sh_basis(deg, d.x, d.y, d.z, b);
  float gz[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float v = spec[3 * s + c];
    gz[c] = g_spec[3 * s + c] * (v * (1.0f - v));
  }
  float* q = g_coeffs + s * coeff_stride;
#pragma unroll
  for (int j = 0; j < 9; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) q[3 * j + c] = gz[c] * b[j];
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_volume_render_forward(const float* rgb, const float* sigma, const float* dists, const float* init_T,
                                int64_t n_rays, int32_t n, float* out, float* trans, float* alpha, void* stream) {
  SNISR_REQUIRE(n_rays >= 0 && n >= 0, SNISR_ERR_ARG, "volume_render_forward: negative size");
  if (n_rays == 0) return SNISR_OK;
  SNISR_REQUIRE(out && (n == 0 || (rgb && sigma && dists && trans && alpha)), SNISR_ERR_ARG,
                "volume_render_forward: NULL pointer");
  hipLaunchKernelGGL(k_volume_render_fwd, dim3((unsigned)((n_rays + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                     rgb, sigma, dists, init_T, n_rays, (int)n, out, trans, alpha);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_volume_render_backward(const float* rgb, const float* dists, const float* trans, const float* alpha,
                                 int64_t n_rays, int32_t n, const float* g_out, const float* g_trans,
                                 const float* g_alpha, float* g_rgb, float* g_sigma, float* g_init, void* stream) {
  SNISR_REQUIRE(n_rays >= 0 && n >= 0, SNISR_ERR_ARG, "volume_render_backward: negative size");
  if (n_rays == 0) return SNISR_OK;
  SNISR_REQUIRE(g_out && (n == 0 || (rgb && dists && trans && alpha && g_rgb && g_sigma)), SNISR_ERR_ARG,
                "volume_render_backward: NULL pointer");
  hipLaunchKernelGGL(k_volume_render_bwd, dim3((unsigned)((n_rays + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                     rgb, dists, trans, alpha, n_rays, (int)n, g_out, g_trans, g_alpha, g_rgb, g_sigma, g_init);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_sh_sigmoid_forward(const float* coeffs, int32_t coeff_stride, const float* dirs, int64_t n,
                             int32_t samples_per_ray, int32_t deg, float* spec, void* stream) {
  SNISR_REQUIRE(n >= 0 && samples_per_ray > 0 && deg >= 0 && deg <= 2, SNISR_ERR_ARG,
                "sh_sigmoid_forward: n >= 0, samples_per_ray > 0, deg in 0..2");
  SNISR_REQUIRE(coeff_stride >= 3 * (deg + 1) * (deg + 1), SNISR_ERR_ARG,
                "sh_sigmoid_forward: coeff_stride below 3 (deg + 1)^2");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(coeffs && dirs && spec, SNISR_ERR_ARG, "sh_sigmoid_forward: NULL pointer");
  const bool vec_rows = coeff_stride >= 28 && coeff_stride % 4 == 0 && aligned(coeffs, 16);
  hipLaunchKernelGGL(k_sh_sigmoid_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, coeffs,
                     (int)coeff_stride, vec_rows, dirs, n, (int)samples_per_ray, (int)deg, spec);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_sh_sigmoid_backward(const float* spec, const float* dirs, int64_t n, int32_t samples_per_ray, int32_t deg,
                              const float* g_spec, float* g_coeffs, int32_t coeff_stride, void* stream) {
  SNISR_REQUIRE(n >= 0 && samples_per_ray > 0 && deg >= 0 && deg <= 2, SNISR_ERR_ARG,
                "sh_sigmoid_backward: n >= 0, samples_per_ray > 0, deg in 0..2");
  SNISR_REQUIRE(coeff_stride >= 27, SNISR_ERR_ARG, "sh_sigmoid_backward: coeff_stride below 27");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(spec && dirs && g_spec && g_coeffs, SNISR_ERR_ARG, "sh_sigmoid_backward: NULL pointer");
  hipLaunchKernelGGL(k_sh_sigmoid_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, spec,
                     dirs, n, (int)samples_per_ray, (int)deg, g_spec, g_coeffs, (int)coeff_stride);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
