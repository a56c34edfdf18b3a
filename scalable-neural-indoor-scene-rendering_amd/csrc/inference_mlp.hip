// A7 — the reference-format reflection MLP (inference_mlp) on v_mfma_f32_32x32x16_f16.
//
// Per sample: x = (p - group centre) * inverse_far, positional encoding with 10 bands (63 inputs),
// 63 -> 64 -> 4 x (64 -> 64) LeakyReLU(0.01) -> 28, sigma = max(out[27], 0), rgb = sigmoid of the
// degree-2 spherical-harmonics colour at the normalised ray direction with coefficient
// r = 3 * basis + channel (rendering/src/rayreflection_render.cu:111-300, inference_SH 61-108).
// The reference evaluates it with scalar FP16 FMAs and FP16 accumulation; here weights and
// activations are FP16 MFMA operands, accumulation, LeakyReLU, SH and sigmoid are FP32.
//
// Work split: the blocks of pixels (net_idx[b], pixel range [pixel_starts[b], pixel_starts[b+1]) of
// the sorted query list) are cut into equal contiguous spans, one per workgroup.  A workgroup stages
// its current group's packed weights (46,464 B) in LDS and reloads them only when the group changes
// (blocks arrive sorted by group).  Pixels whose transparency is below trans_stop are compacted away
// with a wave ballot, so every MFMA column is a live sample.  One wavefront evaluates 64 samples:
// two 32-column blocks; all six layers keep their activations in registers in the B-operand layout.
//
// Positional encoding: lane half h = 0 holds x, y and bands 0..4, h = 1 holds z, one zero and
// bands 5..9 (weights of layer 0 are packed in that order, so k slot s >= 2 is band slot s - 2 in
// both halves).  Each half evaluates the hardware sin / cos once per coordinate, at 2^(5h) x (the
// reference uses the same fast intrinsics, __sinf / __cosf), then four double-angle steps
// (sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a); the error stays near 1e-5, below the FP16
// rounding of the inputs.  LeakyReLU runs on the FP16 values, as in the reference (x < 0 -> 0.01 x
// in half precision), as packed-half multiply and max.
#include "mfma_ops.hpp"
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

constexpr int kRefIn = 63, kRefWidth = 64, kRefOut = 28, kRefLayers = 6;
constexpr int kRefParams = 22556;            // halves per group in the reference packing
constexpr int kRefHiddenBytes = 2 * (4 * 1024 + 128);
constexpr int kRefGroupBytes = (kRefLayers - 1) * kRefHiddenBytes + (4 * 1024 + 128);  // 46,464

__host__ __device__ constexpr int ref_in(int l) { return l == 0 ? kRefIn : kRefWidth; }
__host__ __device__ constexpr int ref_out(int l) { return l == kRefLayers - 1 ? kRefOut : kRefWidth; }
// halves before layer l in the reference packing (per layer: bias[out], then W[in][out])
__host__ __device__ constexpr int ref_src_offset(int l) {
  int o = 0;
  for (int m = 0; m < l; ++m) o += ref_out(m) + ref_in(m) * ref_out(m);
  return o;
}
static_assert(ref_src_offset(kRefLayers) == kRefParams, "reference parameter count");
static_assert(kRefGroupBytes % 16 == 0, "group bytes");

// layer-0 input index held in k slot s = 8 kc + e of lane half h (63 = zero padding)
__host__ __device__ constexpr int pe_input_index(int h, int s) {
  return h == 0 ? (s < 2 ? s : s + 1) : (s == 0 ? 2 : (s == 1 ? kRefIn : s + 31));
}

// Reference output held by MFMA row p of layer l (-1: a zero row).  The hidden layers keep their
// order.  The output layer's 32 rows are permuted so that slot i of both lane halves (row
// 8(i/4) + 4h + i%4) holds the same colour channel i % 3: half 0 holds the coefficients
// 3 b + c of bases b = 0..4 in slots 0..14 and sigma (output 27) in slot 15; half 1 holds bases
// 5..8 in slots 0..11 and zero rows in slots 12..15.  The SH sum then needs no per-half indexing.
__host__ __device__ constexpr int ref_out_row(int l, int p) {
  if (l != kRefLayers - 1) return p;
  const int h = (p / 4) % 2, i = 4 * (p / 8) + p % 4;
  return h == 0 ? (i < 15 ? i : kRefOut - 1) : (i < 12 ? 15 + i : -1);
}

// a / b for 0 <= a < 2^24, b > 0, from the FP32 quotient corrected by one step
__device__ __forceinline__ int div_small(int a, int b, float inv_b) {
  int q = (int)((float)a * inv_b);
  const int r = a - q * b;
  q += r < 0 ? -1 : (r >= b ? 1 : 0);
  return q;
}

__device__ __forceinline__ float half_bits_to_float(uint16_t b) {
  _Float16 v;
  __builtin_memcpy(&v, &b, 2);
  return (float)v;
}

// one thread per packed weight (16-bit, copied bit-exact) or bias (FP32) of one layer of one group
__global__ void k_pack_ref_layer(const uint16_t* __restrict__ src, uint8_t* __restrict__ dst, int l, int64_t n_groups) {
  const int in = ref_in(l), out = ref_out(l), nob = (out + 31) / 32;
  const int64_t nw = (int64_t)nob * 4 * 512, per = nw + nob * 32;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= per * n_groups) return;
  const int64_t grp = g / per, r = g - grp * per;
  const uint16_t* bias = src + grp * kRefParams + ref_src_offset(l);
  const uint16_t* Wt = bias + out;  // [in][out]
  uint8_t* D = dst + grp * kRefGroupBytes + (int64_t)l * kRefHiddenBytes;
  if (r < nw) {
    const int frag = (int)(r / 512), within = (int)(r % 512);
    const int ob = frag / 4, kc = frag % 4, lane = within / 8, e = within % 8;
    const int h = lane / 32, row = ref_out_row(l, 32 * ob + lane % 32);
    const int x = l == 0 ? pe_input_index(h, 8 * kc + e) : pack_input_index(l, kc, h, e);
    reinterpret_cast<uint16_t*>(D)[r] = (row >= 0 && row < out && x < in) ? Wt[x * out + row] : (uint16_t)0;
  } else {
    const int q = (int)(r - nw);
    const int ob = q / 32, h = (q / 16) % 2, i = q % 16;
    const int row = ref_out_row(l, 32 * ob + 8 * (i / 4) + 4 * h + (i % 4));
    reinterpret_cast<float*>(D + nob * 4 * 1024)[q] = row >= 0 && row < out ? half_bits_to_float(bias[row]) : 0.0f;
  }
}

// the 32 layer-0 inputs of lane half h for one sample, in k-slot order.  The hardware sin / cos
// take revolutions: rev = the argument times 1 / (2 pi), or times 32 / (2 pi) for half 1 (the
// power-of-two factor is exact, so this is the same product as __sinf(32 p) forms).
__device__ __forceinline__ void pe_half(const float (&p)[3], int h, float (&v)[32]) {
  constexpr float kRev = 0.15915494309189535f;  // 1 / (2 pi)
  const float rev_scale = h ? 32.0f * kRev : kRev;
  float sn[5][3], cs[5][3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float rev = p[r] * rev_scale;
    sn[0][r] = __builtin_amdgcn_sinf(rev);
    cs[0][r] = __builtin_amdgcn_cosf(rev);
#pragma unroll
    for (int k = 1; k < 5; ++k) {
      const float t = sn[k - 1][r] + sn[k - 1][r];
      sn[k][r] = t * cs[k - 1][r];                         // sin 2a = 2 sin a cos a
      cs[k][r] = __builtin_fmaf(-t, sn[k - 1][r], 1.0f);  // cos 2a = 1 - 2 sin^2 a
    }
  }
  v[0] = h ? p[2] : p[0];
  v[1] = h ? 0.0f : p[1];
#pragma unroll
  for (int s = 2; s < 32; ++s) {
    const int i = s - 2;
    v[s] = (i % 6) < 3 ? sn[i / 6][i % 6] : cs[i / 6][i % 6 - 3];
  }
}

// real SH of degree <= 2 (orthonormal on the sphere; the reference's sign convention)
__device__ __forceinline__ void sh2_basis(float x, float y, float z, float (&b)[9]) {
  constexpr float c0 = 0.28209479177387814f;   // 1 / (2 sqrt(pi))
  constexpr float c1 = 0.4886025119029199f;    // sqrt(3) c0
  constexpr float c2 = 1.0925484305920792f;    // sqrt(15) c0
  constexpr float c22 = 0.31539156525252005f;  // sqrt(5) c0 / 2
  constexpr float c24 = 0.5462742152960396f;   // sqrt(15) c0 / 2
  b[0] = c0;
  b[1] = -c1 * y; b[2] = c1 * z; b[3] = -c1 * x;
  b[4] = c2 * (x * y); b[5] = -c2 * (y * z); b[6] = c22 * (2.0f * (z * z) - x * x - y * y);
  b[7] = -c2 * (x * z); b[8] = c24 * (x * x - y * y);
}

// partial SH colour from the output rows this lane holds: slot i < 15 is coefficient
// 3 (5h + i/3) + i%3 (zero rows in slots 12..14 of half 1), w[j] the basis value of 5h + j
__device__ __forceinline__ void sh_partial(const f32x16& acc, const float (&w)[5], float (&pr)[3]) {
#pragma unroll
  for (int i = 0; i < 15; ++i) pr[i % 3] = __builtin_fmaf(w[i / 3], acc[i], pr[i % 3]);
}

// The inputs of one 64-column chunk as lane (h, s) needs them: the positions of columns s and
// 32 + s, and the record it finishes (column 32h + s): its task index (-1: no sample) and the ray
// direction of its pixel.  The loads are unconditional (a column past ncol reads sample 0 of
// pixel 0, which exists whenever the kernel runs), so the number of loads in flight is the same on
// every path and the wait for them can stay at their first use, one chunk later.
struct ChunkIn {
  float p[2][3];
  float d[3];
  int64_t task;
};

__device__ __forceinline__ void fetch_chunk(const snisr_mlp_query& Q, const int* act, int c0, int ncol, float inv_spr,
                                            float inv_w, int h, int s, ChunkIn& in) {
  const int spr = Q.samples_per_ray;
  int64_t tk = -1;
  int own = 0;
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int c = c0 + cb * 32 + s;
    const bool ok = c < ncol;
    const int cc = ok ? c : 0;
    const int q = div_small(cc, spr, inv_spr);  // < 64 for cc < ncol = 64 * samples_per_ray at most
    const int ori = ok ? act[q] : 0;
    const int64_t t = (int64_t)ori * spr + (cc - q * spr);
#pragma unroll
    for (int i = 0; i < 3; ++i) in.p[cb][i] = Q.samples[t * 3 + i];
    if (cb == h) {
      tk = ok ? t : -1;
      own = ori;
    }
  }
  in.task = tk;
  const int py = div_small(own, Q.W, inv_w), px = own - py * Q.W;
  const float* d = Q.rays_d + (int64_t)py * Q.d_row_stride + (int64_t)px * Q.d_col_stride;
#pragma unroll
  for (int i = 0; i < 3; ++i) in.d[i] = d[i];
}

// At least two waves per SIMD caps the kernel at 256 registers, so the compiler keeps the MFMA
// accumulators in VGPRs (with no cap it placed them in AGPRs and read each result back with a
// v_accvgpr_read: 352 extra instructions per 64 samples).
__global__ __launch_bounds__(256, 2) void k_inference_mlp(snisr_mlp_query Q) {
  __shared__ __align__(16) uint8_t lw[kRefGroupBytes];
  __shared__ int act[64];
  __shared__ float ctr[3];
  __shared__ int nact_s;
  using M = Mfma<SNISR_F16>;
  const int lane = threadIdx.x & 63, h = lane >> 5, s = lane & 31;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int spr = Q.samples_per_ray;
  const float inv_spr = 1.0f / (float)spr, inv_w = 1.0f / (float)Q.W;
  const int64_t npix = (int64_t)Q.H * Q.W;
  int64_t nb = Q.n_blocks;  // or the device-side count of a host-sync-free regrouping
  if (Q.n_blocks_device) {
    const int64_t d = *Q.n_blocks_device;
    nb = d < nb ? (d > 0 ? d : 0) : nb;
  }
  const int64_t per = (nb + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = (int64_t)blockIdx.x * per;
  const int64_t b1 = b0 + per < nb ? b0 + per : nb;
  const uint8_t* __restrict__ gw = reinterpret_cast<const uint8_t*>(Q.packed_params);
  int loaded = -1;
  for (int64_t b = b0; b < b1; ++b) {
    const int g = Q.net_idx[b];
    if (g < 0 || g >= Q.n_groups) continue;  // -1: pixels of no group
    if (g != loaded) {
      __syncthreads();
      const uint4* src = reinterpret_cast<const uint4*>(gw + (int64_t)g * kRefGroupBytes);
      uint4* dst = reinterpret_cast<uint4*>(lw);
      for (int k = threadIdx.x; k < kRefGroupBytes / 16; k += blockDim.x) dst[k] = src[k];
      if (threadIdx.x < 3) ctr[threadIdx.x] = Q.group_centers[3 * g + threadIdx.x];
      __syncthreads();
      loaded = g;
    }
    const int64_t p_end = Q.pixel_starts[b + 1] < Q.n_query ? Q.pixel_starts[b + 1] : Q.n_query;
    for (int64_t pb = Q.pixel_starts[b]; pb < p_end; pb += 64) {
      __syncthreads();  // act[] is free
      if (threadIdx.x < 64) {
        const int64_t p = pb + threadIdx.x;
        int ori = -1;
        bool ok = false;
        if (p >= 0 && p < p_end) {
          ori = Q.query_pixel_indices[p];
          ok = ori >= 0 && ori < npix && !(Q.transparency[ori] < Q.trans_stop);
        }
        const uint64_t m = __ballot(ok);
        if (ok) act[__popcll(m & ((1ull << threadIdx.x) - 1ull))] = ori;
        if (threadIdx.x == 0) nact_s = __popcll(m);
      }
      __syncthreads();
      const int ncol = nact_s * spr;
      // a chunk's inputs are loaded one chunk ahead, so their latency hides behind the layers
      ChunkIn cur, nxt;
      fetch_chunk(Q, act, wv * 64, ncol, inv_spr, inv_w, h, s, cur);
      for (int c0 = wv * 64; c0 < ncol; c0 += nwv * 64) {
        f16x8 xa[2][4], xb[2][4];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          float pt[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) pt[i] = (cur.p[cb][i] - ctr[i]) * Q.inverse_far;
          float v[32];
          pe_half(pt, h, v);
#pragma unroll
          for (int kc = 0; kc < 4; ++kc) xa[cb][kc] = M::cvt8(v + 8 * kc);
          __builtin_amdgcn_sched_barrier(0);
        }
        fetch_chunk(Q, act, c0 + nwv * 64, ncol, inv_spr, inv_w, h, s, nxt);
        // five hidden layers (A fragments from LDS), activations ping-pong xa <-> xb; the LeakyReLU
        // of one output block is issued under the MFMAs of the next
#pragma unroll
        for (int l = 0; l < kRefLayers - 1; ++l) {
          auto& src = (l & 1) ? xb : xa;
          auto& dst = (l & 1) ? xa : xb;
          const uint8_t* Lp = lw + l * kRefHiddenBytes;
          const float* bias = reinterpret_cast<const float*>(Lp + 2 * 4 * 1024);
          mma_layer<M, 4, 2, SNISR_MLP_PREFETCH, true>(Lp, bias, lane, h, src[0], src[1], dst[0], dst[1]);
        }
        // output layer: 28 rows in one 32-row block; the input is the output of layer 4 (in xb)
        const uint8_t* Lp = lw + (kRefLayers - 1) * kRefHiddenBytes;
        const float* bias = reinterpret_cast<const float*>(Lp + 4 * 1024);
        f32x16 acc[2];
        acc[0] = bias_block(bias, 0, h);
        acc[1] = acc[0];
        mma_block<M, 4>(Lp, lane, xb[0], xb[1], acc[0], acc[1]);
        // lane (h, s) finishes sample s of column block h: it evaluates that sample's SH basis, which
        // v_permlane32_swap hands to the other half, weights the rows of both blocks it holds, and one
        // swap per channel adds the two halves' partial sums of block h (sigma, slot 15 of lane half
        // 0, reaches the block-1 lanes the same way)
        const int64_t tk = cur.task;
        float bo[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (tk >= 0) {
          const V3 u = normalize_dir(v3(cur.d[0], cur.d[1], cur.d[2]));
          sh2_basis(u.x, u.y, u.z, bo);
        }
        float bs[2][9];
#pragma unroll
        for (int j = 0; j < 9; ++j) swap_halves(bo[j], bs[0][j], bs[1][j]);
        float pr[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          float w[5];
#pragma unroll
          for (int j = 0; j < 5; ++j) w[j] = h ? (j < 4 ? bs[cb][5 + j] : 0.0f) : bs[cb][j];
          sh_partial(acc[cb], w, pr[cb]);
        }
        float raw[3], sg, unused;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float lo, hi;
          swap_lanes(pr[0][k], pr[1][k], lo, hi);
          raw[k] = lo + hi;
        }
        swap_lanes(acc[0][15], acc[1][15], sg, unused);
        if (tk >= 0)
          reinterpret_cast<float4*>(Q.sample_results)[tk] =
              make_float4(sigmoid_hw(raw[0]), sigmoid_hw(raw[1]), sigmoid_hw(raw[2]), fmaxf(sg, 0.0f));
        cur = nxt;
      }
    }
  }
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_inference_mlp_packed_size(int64_t n_groups) { return n_groups < 0 ? -1 : n_groups * kRefGroupBytes; }

int snisr_pack_inference_mlp(const void* params_f16, int64_t n_groups, void* packed, void* stream) {
  SNISR_REQUIRE(n_groups >= 0, SNISR_ERR_ARG, "pack_inference_mlp: n_groups < 0");
  if (n_groups == 0) return SNISR_OK;
  SNISR_REQUIRE(params_f16 && packed, SNISR_ERR_ARG, "pack_inference_mlp: NULL pointer");
  SNISR_REQUIRE(aligned(packed, 16), SNISR_ERR_ALIGN, "pack_inference_mlp: packed must be 16-byte aligned");
  for (int l = 0; l < kRefLayers; ++l) {
    const int nob = (ref_out(l) + 31) / 32;
    const int64_t total = ((int64_t)nob * 4 * 512 + nob * 32) * n_groups;
    hipLaunchKernelGGL(k_pack_ref_layer, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t*)params_f16, (uint8_t*)packed, l, n_groups);
    SNISR_LAUNCHED();
  }
  return SNISR_OK;
}

int snisr_inference_mlp(const snisr_mlp_query* q, void* stream) {
  SNISR_REQUIRE(q, SNISR_ERR_ARG, "inference_mlp: query is NULL");
  SNISR_REQUIRE(q->samples_per_ray > 0 && q->n_blocks >= 0 && q->n_groups >= 0 && q->H >= 0 && q->W >= 0 &&
                    q->n_query >= 0,
                SNISR_ERR_ARG, "inference_mlp: bad size");
  if (q->n_blocks == 0 || q->n_groups == 0 || (int64_t)q->H * q->W == 0) return SNISR_OK;
  SNISR_REQUIRE(q->net_idx && q->pixel_starts && q->query_pixel_indices && q->transparency && q->group_centers &&
                    q->samples && q->rays_d && q->packed_params && q->sample_results,
                SNISR_ERR_ARG, "inference_mlp: NULL pointer");
  SNISR_REQUIRE(aligned(q->packed_params, 16) && aligned(q->sample_results, 16), SNISR_ERR_ALIGN,
                "inference_mlp: packed_params and sample_results must be 16-byte aligned");
  SNISR_REQUIRE((int64_t)q->H * q->W * q->samples_per_ray < (int64_t(1) << 31), SNISR_ERR_SHAPE,
                "inference_mlp: H*W*samples_per_ray must stay below 2^31");
  SNISR_REQUIRE((int64_t)q->H * q->W < (int64_t(1) << 24) && q->samples_per_ray <= (1 << 18), SNISR_ERR_SHAPE,
                "inference_mlp: H*W must stay below 2^24 and samples_per_ray at most 2^18");
  // about sixteen workgroups per CU (three are resident at a time), each with a short span of blocks,
  // so that blocks of unequal live-sample counts even out (720p graphed frame: 3 per CU 1.53 ms,
  // 6: 1.42, 12-24: 1.41, one block per workgroup: 1.42)
  const int64_t per_cu = 16;
  const int64_t grid = q->n_blocks < per_cu * (int64_t)device_cus() ? q->n_blocks : per_cu * (int64_t)device_cus();
  hipLaunchKernelGGL(k_inference_mlp, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, *q);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
