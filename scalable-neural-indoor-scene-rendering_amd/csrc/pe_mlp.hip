// networks.MLP of the reference's training (training/scripts/networks.py:4-41) in FP32, forward and
// hand-written adjoint: the positional encoding [x, sin(2^0 x), cos(2^0 x), ..., sin(2^(L-1) x),
// cos(2^(L-1) x)] of a 3-d point (Positional_Encoding, networks.py:4-15), then GeneralMLP
// (networks.py:17-32): depth Linear layers of width 64 with LeakyReLU(slope) between them (slope 0
// is ReLU) and n_out <= 32 outputs.  Parameters in nn.Linear order: per layer W[out][in], then
// b[out].  This is the surface MLP (depth 8, 4 outputs, ReLU) and the reflection MLP (depth 6, 28
// outputs, LeakyReLU 0.01) of the reference's training configuration (training/config/cfg.yaml:38-46),
// used for the compute_voxels bake, the background samples and the specular branch.
//
// One thread per sample with the 64 activations in registers; the weights, the same for every lane,
// are scalar loads.  The forward stores each layer's input row for the backward, which stores each
// layer's output-delta row; the weight gradients are the row products dW = sum delta (x) input,
// reduced on FP32 MFMA.
#include "snisr_common.hpp"
#include "snisr_device.hpp"
#include "train_dw.hpp"

namespace snisr {
namespace {

constexpr int kPW = 64;       // hidden width
constexpr int kPOut = 32;     // output rows computed (n_out <= 32, the rest zero weights)
constexpr int kPL = 10;       // frequencies of the positional encoding (63 inputs)

__host__ __device__ constexpr int pe_in(int L) { return 3 + 6 * L; }
__host__ __device__ inline int pe_out_l(int l, int depth, int n_out) { return l == depth - 1 ? n_out : kPW; }
__host__ __device__ inline int pe_in_l(int l, int L) { return l == 0 ? pe_in(L) : kPW; }
// padded row widths of the stored rows (x: layer inputs, d: layer output deltas)
__host__ __device__ inline int pe_xw(int l, int L) { return l == 0 ? ((pe_in(L) + 3) & ~3) : kPW; }
__host__ __device__ inline int pe_dw(int l, int depth, int n_out) { return l == depth - 1 ? ((n_out + 3) & ~3) : kPW; }
// float offset of layer l in the parameter vector
__host__ __device__ inline int64_t pe_param_off(int l, int L, int depth, int n_out) {
  int64_t s = 0;
  for (int m = 0; m < l; ++m) s += (int64_t)pe_out_l(m, depth, n_out) * pe_in_l(m, L) + pe_out_l(m, depth, n_out);
  return s;
}
// float offsets of layer l's x rows and d rows in the workspace (n rows each)
__host__ __device__ inline int64_t pe_x_off(int l, int L, int64_t n) {
  int64_t s = 0;
  for (int m = 0; m < l; ++m) s += pe_xw(m, L);
  return s * n;
}
__host__ __device__ inline int64_t pe_d_off(int l, int L, int depth, int n_out, int64_t n) {
  int64_t s = 0;
  for (int m = 0; m < depth; ++m) s += pe_xw(m, L);
  for (int m = 0; m < l; ++m) s += pe_dw(m, depth, n_out);
  return s * n;
}

__device__ __forceinline__ void store_row(float* dst, const float (&v)[kPW]) {
#pragma unroll
  for (int q = 0; q < kPW; q += 4) *reinterpret_cast<float4*>(dst + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
}

// y[o] = b[o] + sum_i W[o][i] x[i] for o < NO (the first `nout` real), W row-major with `nin`
// columns.  Every lane of a wavefront reads the same weights, so they are scalar loads (the
// weights never occupy vector registers or LDS).
template <int NO>
__device__ __forceinline__ void pe_layer(const float* __restrict__ W, int nin, int nout, const float (&x)[kPW],
                                         float (&y)[kPW]) {
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    float acc = 0.0f;
    if (o < nout) {
      const float* w = W + (int64_t)o * nin;
      acc = W[(int64_t)nout * nin + o];
#pragma unroll
      for (int i = 0; i < kPW; ++i)
        if (i < nin) acc = __builtin_fmaf(w[i], x[i], acc);
    }
    y[o] = acc;
  }
}

__global__ __launch_bounds__(256) void k_pemlp_fwd(const float* __restrict__ params, const float* __restrict__ pts,
                                                   int64_t n, int L, int depth, int n_out, float slope,
                                                   float* __restrict__ out, float* __restrict__ rows) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kPW];
  const float p[3] = {pts[3 * s], pts[3 * s + 1], pts[3 * s + 2]};
#pragma unroll
  for (int i = 0; i < kPW; ++i) x[i] = 0.0f;
  x[0] = p[0];
  x[1] = p[1];
  x[2] = p[2];
#pragma unroll
  for (int f = 0; f < kPL; ++f) {
    const float sc = (float)(1 << f);
#pragma unroll
    for (int j = 0; j < 3; ++j) sincosf(sc * p[j], &x[3 + 6 * f + j], &x[6 + 6 * f + j]);
  }
  if (rows) store_row(rows + pe_x_off(0, L, n) + s * kPW, x);   // pe_xw(0, 10) = 64
  for (int l = 0; l < depth - 1; ++l) {
    float y[kPW];
    pe_layer<kPW>(params + pe_param_off(l, L, depth, n_out), pe_in_l(l, L), kPW, x, y);
#pragma unroll
    for (int o = 0; o < kPW; ++o) x[o] = y[o] > 0.0f ? y[o] : slope * y[o];
    if (rows) store_row(rows + pe_x_off(l + 1, L, n) + s * kPW, x);
  }
  float y[kPW];
  pe_layer<kPOut>(params + pe_param_off(depth - 1, L, depth, n_out), kPW, n_out, x, y);
#pragma unroll
  for (int o = 0; o < kPOut; ++o)
    if (o < n_out) out[s * n_out + o] = y[o];
}

__global__ __launch_bounds__(256) void k_pemlp_bwd(const float* __restrict__ params, int64_t n, int L, int depth,
                                                   int n_out, float slope, const float* __restrict__ g_out,
                                                   float* __restrict__ rows) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float d[kPW];
#pragma unroll
  for (int o = 0; o < kPW; ++o) d[o] = (o < kPOut && o < n_out) ? g_out[s * n_out + (o < n_out ? o : 0)] : 0.0f;
  {
    const int dw = pe_dw(depth - 1, depth, n_out);
    float* dst = rows + pe_d_off(depth - 1, L, depth, n_out, n) + s * dw;
#pragma unroll
    for (int o = 0; o < kPOut; o += 4)
      if (o < dw) *reinterpret_cast<float4*>(dst + o) = make_float4(d[o], d[o + 1], d[o + 2], d[o + 3]);
  }
  for (int l = depth - 1; l >= 1; --l) {
    const int nout = pe_out_l(l, depth, n_out);
    const float* W = params + pe_param_off(l, L, depth, n_out);  // W[o][i], 64 inputs
    float acc[kPW];
#pragma unroll
    for (int i = 0; i < kPW; ++i) acc[i] = 0.0f;
#pragma unroll
    for (int o = 0; o < kPW; ++o)
      if (o < nout) {
        const float* w = W + (int64_t)o * kPW;
#pragma unroll
        for (int i = 0; i < kPW; ++i) acc[i] = __builtin_fmaf(w[i], d[o], acc[i]);
      }
    // the layer's input is the previous layer's activation: its sign is the pre-activation's
    const float4* xr = reinterpret_cast<const float4*>(rows + pe_x_off(l, L, n) + s * kPW);
#pragma unroll
    for (int q = 0; q < kPW / 4; ++q) {
      const float4 x4 = xr[q];
      d[4 * q] = acc[4 * q] * (x4.x > 0.0f ? 1.0f : slope);
      d[4 * q + 1] = acc[4 * q + 1] * (x4.y > 0.0f ? 1.0f : slope);
      d[4 * q + 2] = acc[4 * q + 2] * (x4.z > 0.0f ? 1.0f : slope);
      d[4 * q + 3] = acc[4 * q + 3] * (x4.w > 0.0f ? 1.0f : slope);
    }
    store_row(rows + pe_d_off(l - 1, L, depth, n_out, n) + s * kPW, d);
  }
}

// dW and db of layer blockIdx.y over split blockIdx.x of the samples (atomics into g_params)
template <int IP, int OP, int INL, int OUTL>
__device__ void pe_dw_layer(const float* rows, int64_t n, int L, int depth, int n_out, int l, float* g_params,
                            float* Ds, float* As) {
  const int64_t per = ((n + gridDim.x - 1) / gridDim.x + kDwRows - 1) / kDwRows * kDwRows;
  const int64_t jb = (int64_t)blockIdx.x * per;
  const int64_t je = jb + per < n ? jb + per : n;
  dw_layer<IP, OP, INL, OUTL>(rows + pe_x_off(l, L, n), rows + pe_d_off(l, L, depth, n_out, n), jb, je,
                              g_params + pe_param_off(l, L, depth, n_out), Ds, As);
}

template <int NOUT>
__global__ __launch_bounds__(256, 2) void k_pemlp_dw(const float* __restrict__ rows, int64_t n, int L, int depth,
                                                     float* __restrict__ g_params) {
  __shared__ __align__(16) float Ds[kDwRows * kDwPitch];
  __shared__ __align__(16) float As[kDwRows * kDwPitch];
  constexpr int NOP = (NOUT + 3) & ~3;
  const int l = blockIdx.y;
  if (l == 0) {
    pe_dw_layer<64, kPW, 63, kPW>(rows, n, L, depth, NOUT, l, g_params, Ds, As);
  } else if (l == depth - 1) {
    pe_dw_layer<kPW, NOP, kPW, NOUT>(rows, n, L, depth, NOUT, l, g_params, Ds, As);
  } else {
    pe_dw_layer<kPW, kPW, kPW, kPW>(rows, n, L, depth, NOUT, l, g_params, Ds, As);
  }
}

__global__ void k_pe_zero(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

int pe_check(int32_t L, int32_t depth, int32_t n_out) {
  SNISR_REQUIRE(L == 10, SNISR_ERR_SHAPE, "pemlp: the positional encoding has L = 10 frequencies (63 inputs)");
  SNISR_REQUIRE(depth >= 2 && depth <= 8, SNISR_ERR_SHAPE, "pemlp: depth in 2..8");
  SNISR_REQUIRE(n_out == 4 || n_out == 28, SNISR_ERR_SHAPE, "pemlp: n_out 4 (surface MLP) or 28 (reflection MLP)");
  return SNISR_OK;
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_pemlp_param_count(int32_t L, int32_t depth, int32_t n_out) {
  if (pe_check(L, depth, n_out)) return -1;
  return pe_param_off(depth, L, depth, n_out);
}

int64_t snisr_pemlp_workspace_size(int64_t n, int32_t L, int32_t depth, int32_t n_out) {
  if (n < 0 || pe_check(L, depth, n_out)) return -1;
  return 4 * pe_d_off(depth, L, depth, n_out, n) + 256;
}

int snisr_pemlp_forward(const float* params, const float* points, int64_t n, int32_t L, int32_t depth, int32_t n_out,
                        float slope, float* out, void* workspace, int64_t workspace_bytes, void* stream) {
  if (int e = pe_check(L, depth, n_out)) return e;
  SNISR_REQUIRE(n >= 0, SNISR_ERR_ARG, "pemlp_forward: negative size");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(params && points && out, SNISR_ERR_ARG, "pemlp_forward: NULL pointer");
  SNISR_REQUIRE(!workspace || workspace_bytes >= snisr_pemlp_workspace_size(n, L, depth, n_out), SNISR_ERR_CAPACITY,
                "pemlp_forward: workspace too small");
  SNISR_REQUIRE(!workspace || aligned(workspace, 16), SNISR_ERR_ALIGN, "pemlp_forward: workspace alignment");
  hipLaunchKernelGGL(k_pemlp_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                     points, n, (int)L, (int)depth, (int)n_out, slope, out, (float*)workspace);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_pemlp_backward(const float* params, int64_t n, int32_t L, int32_t depth, int32_t n_out, float slope,
                         const float* g_out, float* g_params, void* workspace, int64_t workspace_bytes, void* stream) {
  if (int e = pe_check(L, depth, n_out)) return e;
  SNISR_REQUIRE(n >= 0, SNISR_ERR_ARG, "pemlp_backward: negative size");
  SNISR_REQUIRE(params && g_params && workspace && (n == 0 || g_out), SNISR_ERR_ARG, "pemlp_backward: NULL pointer");
  SNISR_REQUIRE(workspace_bytes >= snisr_pemlp_workspace_size(n, L, depth, n_out), SNISR_ERR_CAPACITY,
                "pemlp_backward: workspace too small (the forward's rows)");
  hipStream_t st = (hipStream_t)stream;
  const int64_t np = pe_param_off(depth, L, depth, n_out);
  hipLaunchKernelGGL(k_pe_zero, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, g_params, np);
  SNISR_LAUNCHED();
  if (n == 0) return SNISR_OK;
  hipLaunchKernelGGL(k_pemlp_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, params, n, (int)L,
                     (int)depth, (int)n_out, slope, g_out, (float*)workspace);
  SNISR_LAUNCHED();
  int64_t splits = (n + 8191) / 8192;
  splits = splits < 1 ? 1 : (splits > 4 * (int64_t)device_cus() ? 4 * (int64_t)device_cus() : splits);
  const dim3 g((unsigned)splits, (unsigned)depth);
  if (n_out == 28)
    hipLaunchKernelGGL(k_pemlp_dw<28>, g, dim3(256), 0, st, (const float*)workspace, n, (int)L, (int)depth, g_params);
  else
    hipLaunchKernelGGL(k_pemlp_dw<4>, g, dim3(256), 0, st, (const float*)workspace, n, (int)L, (int)depth, g_params);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
