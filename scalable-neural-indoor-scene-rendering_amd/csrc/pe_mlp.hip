// networks.MLP of the reference's training (training/scripts/networks.py:4-41) in FP32, forward and
// hand-written adjoint: the positional encoding [x, sin(2^0 x), cos(2^0 x), ..., sin(2^(L-1) x),
// cos(2^(L-1) x)] of a 3-d point (Positional_Encoding, networks.py:4-15), then GeneralMLP
// (networks.py:17-32): depth Linear layers of width 64 with LeakyReLU(slope) between them (slope 0
// is ReLU) and n_out <= 32 outputs.  Parameters in nn.Linear order: per layer W[out][in], then
// b[out].  This is the surface MLP (depth 8, 4 outputs, ReLU) and the reflection MLP (depth 6, 28
// outputs, LeakyReLU 0.01) of the reference's training configuration (training/config/cfg.yaml:38-46),
// used for the compute_voxels bake, the background samples and the specular branch.
//
// FP32 MFMA (v_mfma_f32_32x32x2_f32) with the training chain's fragment layout: a wavefront takes 32
// samples as the 32 columns, computes their positional encoding cooperatively (30 sin/cos pairs per
// sample spread over the lanes, through LDS), and every layer's output (the D layout) is the next
// layer's B operand, so activations never leave registers.  The weights are packed once per call as
// forward and transposed fragments.  The forward writes only the outputs.  The backward recomputes
// the forward from the points, keeps the activation signs as bit masks, writes each layer's input
// row and output-delta row, and reduces the weight gradients dW = sum delta (x) input on FP32 MFMA.
#include "snisr_common.hpp"
#include "snisr_device.hpp"
#include "train_dw.hpp"
#include "train_mma.hpp"

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
// the last layer's delta rows: 4 floats for 4 outputs, 32 for 28 (the lane layout's pitches)
__host__ __device__ inline int pe_dw(int l, int depth, int n_out) { return l == depth - 1 ? (n_out <= 4 ? 4 : 32) : kPW; }
// rows stored per layer: the sample count rounded up to whole 32-sample blocks of the lane layout
__host__ __device__ inline int64_t pe_rows(int64_t n) { return (n + 31) & ~(int64_t)31; }
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

constexpr int kPIn = 63;                 // encoding width (3 + 6 L)
constexpr int kPNs0 = tr_ns0(kPIn);      // input steps of layer 0 (rows 0..63, row 63 zero)
constexpr int kPeLds = 65;               // LDS row pitch of the encoding (conflict-free column reads)

// float offsets of the packed fragments (one block of tr_stride(63, 64, depth, n_out) floats)
__host__ __device__ inline int64_t pe_frag_floats(int depth, int n_out) { return tr_stride(kPIn, kPW, depth, n_out); }

__device__ __forceinline__ float leaky_mask(uint64_t m, int r, float v, float slope) {
  return ((m >> r) & 1) ? v : slope * v;
}

// One wavefront = 32 samples.  Forward (BWD false): out[j][0..NOUT).  Backward: recomputes the
// forward, stores each layer's input row (pitch 64) and output-delta row (pitch 64, the last layer
// NOUT) for the weight-gradient reduction.
template <int NOUT, bool BWD>
__global__ __launch_bounds__(256, 2) void k_pemlp_mma(const float* __restrict__ P, const float* __restrict__ pts,
                                                      int64_t n, int depth, float slope, float* __restrict__ out,
                                                      const float* __restrict__ g_out, float* __restrict__ rows) {
  static_assert(NOUT % 4 == 0 && NOUT <= 32, "output rows in whole 4-row groups of one block");
  constexpr int NSBO = NOUT <= 8 ? 4 : 16;  // backward steps of the output layer
  __shared__ float pe_lds[4 * 32 * kPeLds];
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5, wv = threadIdx.x >> 6;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + wv) * 32;
  if (j0 >= n) return;  // whole wavefront past the end (only wavefront-level LDS sharing below)
  const int64_t j = j0 + col;
  const bool valid = j < n;
  float* E = pe_lds + wv * 32 * kPeLds;
  // ---- positional encoding of the wavefront's 32 samples: [x, sin(2^f x), cos(2^f x)] (f < 10)
  for (int q = lane; q < 32 * 30; q += 64) {
    const int smp = q / 30, t = q - 30 * smp, f = t / 3, c = t - 3 * f;
    const float v = j0 + smp < n ? pts[3 * (j0 + smp) + c] : 0.0f;
    float sn, cs;
    sincosf((float)(1 << f) * v, &sn, &cs);
    E[smp * kPeLds + 3 + 6 * f + c] = sn;
    E[smp * kPeLds + 6 + 6 * f + c] = cs;
  }
  for (int q = lane; q < 32 * 4; q += 64) {
    const int smp = q >> 2, c = q & 3;
    E[smp * kPeLds + (c < 3 ? c : kPIn)] = (c < 3 && j0 + smp < n) ? pts[3 * (j0 + smp) + c] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float x0[kPNs0];
#pragma unroll
  for (int s = 0; s < kPNs0; ++s) x0[s] = E[col * kPeLds + tr_row(s, h)];
  // rows in the lane layout of the weight-gradient reduction (32-sample blocks, one 1 KB store per
  // feature group and wavefront); the regions hold whole blocks
  const int64_t in_rows = pe_rows(n) * kPW;  // floats of one layer's input rows
  const int64_t blk = j0;                    // first sample of this wavefront's block
  if (BWD && valid) {
#pragma unroll
    for (int g = 0; g < kPNs0 / 4; ++g) store4(rows + blk * kPW + 256 * g + 4 * lane, x0 + 4 * g);
  }
  // ---- forward
  float a[32];
  uint64_t mask[8];
  mma_blocks<kPNs0, 2>(
      P + tr_fwd_off(kPIn, kPW, depth, 0), x0, lane,
      [&](int ob) { return bias_block(P + tr_bias_off(kPIn, kPW, depth, 0), ob, h); },
      [&](int ob, const f32x16& acc) {
#pragma unroll
        for (int r = 0; r < 16; ++r) a[ob * 16 + r] = acc[r] > 0.0f ? acc[r] : slope * acc[r];
      });
  uint64_t mk_last = 0;  // mask of the output layer's input
#pragma unroll
  for (int l = 1; l < 8; ++l) {
    mask[l] = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) mask[l] |= (uint64_t)(a[r] > 0.0f) << r;
    if (BWD && valid) {
      float* Al = rows + l * in_rows + blk * kPW + 4 * lane;
#pragma unroll
      for (int g = 0; g < 8; ++g) store4(Al + 256 * g, a + 4 * g);
    }
    if (l == depth - 1) {
      mk_last = mask[l];
      break;
    }
    float b[32];
    mma_blocks<32, 2>(
        P + tr_fwd_off(kPIn, kPW, depth, l), a, lane,
        [&](int ob) { return bias_block(P + tr_bias_off(kPIn, kPW, depth, l), ob, h); },
        [&](int ob, const f32x16& acc) {
#pragma unroll
          for (int r = 0; r < 16; ++r) b[ob * 16 + r] = acc[r] > 0.0f ? acc[r] : slope * acc[r];
        });
#pragma unroll
    for (int r = 0; r < 32; ++r) a[r] = b[r];
  }
  // ---- output layer: rows 8 (r / 4) + 4 h + r % 4 of column col
  if (!BWD) {
    f32x16 z = bias_block(P + tr_bias_off(kPIn, kPW, depth, depth - 1), 0, h);
    z = mma_f32_chain<32>(P + tr_fwd_off(kPIn, kPW, depth, depth - 1), a, z, lane);
    if (valid) {
#pragma unroll
      for (int g = 0; g < 4; ++g)
        if (8 * g + 4 * h < NOUT) {
          const float v[4] = {z[4 * g], z[4 * g + 1], z[4 * g + 2], z[4 * g + 3]};
          store4(out + j * NOUT + 8 * g + 4 * h, v);
        }
    }
    return;
  }
  // ---- backward: output delta (rows tr_row(s, h), s < NSBO), then d_{l-1} = (W_l^T d_l) * leaky'
  const int64_t d_base = (int64_t)depth * in_rows;  // delta rows follow the input rows
  float dout[NSBO];
#pragma unroll
  for (int g = 0; g < NSBO / 4; ++g) {
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid && 8 * g + 4 * h < NOUT) v = *reinterpret_cast<const float4*>(g_out + j * NOUT + 8 * g + 4 * h);
    dout[4 * g] = v.x;
    dout[4 * g + 1] = v.y;
    dout[4 * g + 2] = v.z;
    dout[4 * g + 3] = v.w;
  }
  if (valid) {
    constexpr int NOP = NOUT <= 4 ? 4 : 32;  // delta pitch of the output layer
    float* Dl = rows + d_base + (int64_t)(depth - 1) * in_rows + blk * NOP;
#pragma unroll
    for (int g = 0; g < NSBO / 4; ++g)  // rows NOUT .. NOP - 1 carry zero deltas
      if (8 * g + 4 * h < NOP) store4(Dl + lane_off(col, 8 * g + 4 * h), dout + 4 * g);
  }
  float d[32];
  mma_blocks<NSBO, 2>(
      P + tr_bwd_off(kPIn, kPW, depth, depth - 1, NOUT), dout, lane, [&](int) { return f32x16{}; },
      [&](int ib, const f32x16& acc) {
#pragma unroll
        for (int r = 0; r < 16; ++r) d[ib * 16 + r] = leaky_mask(mk_last, ib * 16 + r, acc[r], slope);
      });
#pragma unroll
  for (int l = 6; l >= 0; --l) {
    if (l > depth - 2) continue;
    if (valid) {  // d = delta of layer l's outputs
      float* Dl = rows + d_base + (int64_t)l * in_rows + blk * kPW + 4 * lane;
#pragma unroll
      for (int g = 0; g < 8; ++g) store4(Dl + 256 * g, d + 4 * g);
    }
    if (l == 0) break;
    float e[32];
    mma_blocks<32, 2>(
        P + tr_bwd_off(kPIn, kPW, depth, l, NOUT), d, lane, [&](int) { return f32x16{}; },
        [&](int ib, const f32x16& acc) {
#pragma unroll
          for (int r = 0; r < 16; ++r) e[ib * 16 + r] = leaky_mask(mask[l], ib * 16 + r, acc[r], slope);
        });
#pragma unroll
    for (int r = 0; r < 32; ++r) d[r] = e[r];
  }
}

// dW and db of layer blockIdx.y over split blockIdx.x of the samples (atomics into g_params)
template <int IP, int OP, int INL, int OUTL>
__device__ void pe_dw_layer(const float* rows, int64_t n, int L, int depth, int n_out, int l, float* g_params,
                            float* Ds, float* As) {
  const int64_t per = ((n + gridDim.x - 1) / gridDim.x + kDwRows - 1) / kDwRows * kDwRows;
  const int64_t jb = (int64_t)blockIdx.x * per;
  const int64_t je = jb + per < n ? jb + per : n;
  const int64_t nr = pe_rows(n);
  dw_layer<IP, OP, INL, OUTL, true>(rows + pe_x_off(l, L, nr), rows + pe_d_off(l, L, depth, n_out, nr), jb, je,
                                    g_params + pe_param_off(l, L, depth, n_out), Ds, As);
}

template <int NOUT>
__global__ __launch_bounds__(256, 2) void k_pemlp_dw(const float* __restrict__ rows, int64_t n, int L, int depth,
                                                     float* __restrict__ g_params) {
  __shared__ __align__(16) float Ds[kDwRows * kDwPitch];
  __shared__ __align__(16) float As[kDwRows * kDwPitch];
  constexpr int NOP = NOUT <= 4 ? 4 : 32;
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
  return 4 * (pe_frag_floats(depth, n_out) + pe_d_off(depth, L, depth, n_out, pe_rows(n))) + 256;
}

namespace {
int pe_pack(const float* params, int depth, int n_out, float* frag, hipStream_t st) {
  const int64_t tot = pe_frag_floats(depth, n_out);
  hipLaunchKernelGGL(k_pack_train_frag, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, params,
                     pe_param_off(depth, kPL, depth, n_out), frag, kPIn, kPW, depth, 1, n_out);
  SNISR_LAUNCHED();
  return SNISR_OK;
}
}  // namespace

int snisr_pemlp_forward(const float* params, const float* points, int64_t n, int32_t L, int32_t depth, int32_t n_out,
                        float slope, float* out, void* workspace, int64_t workspace_bytes, void* stream) {
  if (int e = pe_check(L, depth, n_out)) return e;
  SNISR_REQUIRE(n >= 0 && slope >= 0.0f, SNISR_ERR_ARG, "pemlp_forward: negative size or slope");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(params && points && out && workspace, SNISR_ERR_ARG, "pemlp_forward: NULL pointer");
  SNISR_REQUIRE(workspace_bytes >= snisr_pemlp_workspace_size(0, L, depth, n_out), SNISR_ERR_CAPACITY,
                "pemlp_forward: workspace too small (snisr_pemlp_workspace_size(0, ...) bytes)");
  SNISR_REQUIRE(aligned(workspace, 16) && aligned(out, 16), SNISR_ERR_ALIGN, "pemlp_forward: alignment");
  hipStream_t st = (hipStream_t)stream;
  float* frag = (float*)workspace;
  if (int e = pe_pack(params, depth, n_out, frag, st)) return e;
  const dim3 grid((unsigned)((n + 127) / 128));
  if (n_out == 28)
    hipLaunchKernelGGL((k_pemlp_mma<28, false>), grid, dim3(256), 0, st, frag, points, n, (int)depth, slope, out,
                       nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_pemlp_mma<4, false>), grid, dim3(256), 0, st, frag, points, n, (int)depth, slope, out,
                       nullptr, nullptr);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_pemlp_backward(const float* params, const float* points, int64_t n, int32_t L, int32_t depth, int32_t n_out,
                         float slope, const float* g_out, float* g_params, void* workspace, int64_t workspace_bytes,
                         void* stream) {
  if (int e = pe_check(L, depth, n_out)) return e;
  SNISR_REQUIRE(n >= 0 && slope >= 0.0f, SNISR_ERR_ARG, "pemlp_backward: negative size or slope");
  SNISR_REQUIRE(params && g_params && (n == 0 || (points && g_out && workspace)), SNISR_ERR_ARG,
                "pemlp_backward: NULL pointer");
  SNISR_REQUIRE(n == 0 || workspace_bytes >= snisr_pemlp_workspace_size(n, L, depth, n_out), SNISR_ERR_CAPACITY,
                "pemlp_backward: workspace too small (snisr_pemlp_workspace_size(n, ...) bytes)");
  SNISR_REQUIRE(n == 0 || (aligned(workspace, 16) && aligned(g_out, 16)), SNISR_ERR_ALIGN,
                "pemlp_backward: alignment");
  hipStream_t st = (hipStream_t)stream;
  const int64_t np = pe_param_off(depth, L, depth, n_out);
  hipLaunchKernelGGL(k_pe_zero, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, g_params, np);
  SNISR_LAUNCHED();
  if (n == 0) return SNISR_OK;
  float* frag = (float*)workspace;
  float* rows = frag + pe_frag_floats(depth, n_out);
  if (int e = pe_pack(params, depth, n_out, frag, st)) return e;
  const dim3 grid((unsigned)((n + 127) / 128));
  if (n_out == 28)
    hipLaunchKernelGGL((k_pemlp_mma<28, true>), grid, dim3(256), 0, st, (const float*)frag, points, n, (int)depth,
                       slope, nullptr, g_out, rows);
  else
    hipLaunchKernelGGL((k_pemlp_mma<4, true>), grid, dim3(256), 0, st, (const float*)frag, points, n, (int)depth,
                       slope, nullptr, g_out, rows);
  SNISR_LAUNCHED();
  // sample splits of the weight-gradient reduction: about four workgroups per CU over all layers,
  // at least 256 samples and at most 8192 per split
  const int64_t cus = device_cus();
  int64_t splits = (4 * cus + depth - 1) / depth;
  const int64_t lo = (n + 8191) / 8192, hi = (n + 255) / 256;
  splits = splits < lo ? lo : (splits > hi ? hi : splits);
  splits = splits < 1 ? 1 : splits;
  const dim3 g((unsigned)splits, (unsigned)depth);
  if (n_out == 28)
    hipLaunchKernelGGL(k_pemlp_dw<28>, g, dim3(256), 0, st, (const float*)rows, n, (int)L, (int)depth, g_params);
  else
    hipLaunchKernelGGL(k_pemlp_dw<4>, g, dim3(256), 0, st, (const float*)rows, n, (int)L, (int)depth, g_params);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
