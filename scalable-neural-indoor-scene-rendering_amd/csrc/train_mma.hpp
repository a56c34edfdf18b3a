// FP32 MFMA helpers of the per-record training MLPs (the feature-scene training chain and the
// positional-encoding MLP): fragment layout and packing of the weights, and the product chains.
#pragma once

#include "mfma_ops.hpp"

namespace snisr {

// ---------------------------------------------------------------------------- FP32 MFMA training MLP
// The per-record MLP of training as v_mfma_f32_32x32x2_f32 products: one wavefront takes 32 sorted
// records (one tile: buckets are 64-aligned) as the 32 columns, so every layer is W (out x in) times
// a 32-column block of activations.  Lane layouts (h = lane / 32, c = lane % 32):
//   A:  lane holds A[row c][k = h]        B:  lane holds B[k = h][column c]
//   D:  lane holds rows 8(r/4) + 4h + (r%4), r = 0..15, of column c
// Step s of a product over an input block pairs row 8(r/4) + (r%4) (lane half 0) with row
// 8(r/4) + 4 + (r%4) (lane half 1), r = s % 16: the B operand of step s is then register r of the
// previous layer's D block s / 16, so activations and deltas never leave registers.  The weights are
// packed per tile in that step order: forward fragments A[o][i] = W[o][i] and, for the backward,
// transposed fragments A[i][o] = W[o][i]; a lane's values of 4 consecutive steps are contiguous
// (float offset 256 (s / 4) + 4 lane + s % 4 inside a block), so one 16-byte load per lane feeds
// four products and a wavefront reads 1 KB per load from global memory (L2-resident).  Layer 0's
// inputs (features, then the 15-value view encoding) follow the same row order; its last block
// only spans the steps that hold real inputs.
__host__ __device__ constexpr int tr_ns0(int in0) { return 16 * (in0 / 32) + 4 * ((in0 % 32 + 7) / 8); }
__host__ __device__ constexpr int tr_nob(int wd, int L, int l) { return l == L - 1 ? 1 : wd / 32; }
__host__ __device__ constexpr int tr_nsf(int in0, int wd, int l) { return l == 0 ? tr_ns0(in0) : wd / 2; }
__host__ __device__ constexpr int tr_nib(int wd, int l) { return l == 0 ? 1 : wd / 32; }
// backward steps of layer l: the output layer's transposed fragments span its nout rows (4 steps hold
// rows 0..7, 16 steps rows 0..31)
__host__ __device__ constexpr int tr_nsb(int wd, int L, int l, int nout = 3) { return l == L - 1 ? (nout <= 8 ? 4 : 16) : wd / 2; }
// float offsets inside one tile's fragment block: forward layer l (weights, then 32 bias floats per
// out-block in the D layout), then the transposed fragments of every layer
__host__ __device__ constexpr int64_t tr_fwd_off(int in0, int wd, int L, int l) {
  int64_t o = 0;
  for (int m = 0; m < l; ++m) o += (int64_t)tr_nob(wd, L, m) * (tr_nsf(in0, wd, m) * 64 + 32);
  return o;
}
__host__ __device__ constexpr int64_t tr_bias_off(int in0, int wd, int L, int l) {
  return tr_fwd_off(in0, wd, L, l) + (int64_t)tr_nob(wd, L, l) * tr_nsf(in0, wd, l) * 64;
}
__host__ __device__ constexpr int64_t tr_bwd_off(int in0, int wd, int L, int l, int nout = 3) {
  int64_t o = tr_fwd_off(in0, wd, L, L);
  for (int m = 0; m < l; ++m) o += (int64_t)tr_nib(wd, m) * tr_nsb(wd, L, m, nout) * 64;
  return o;
}
__host__ __device__ constexpr int64_t tr_stride(int in0, int wd, int L, int nout = 3) {
  return (tr_bwd_off(in0, wd, L, L, nout) + 63) / 64 * 64;
}
// input row held by lane half h at step s of an input block sequence
__host__ __device__ constexpr int tr_row(int s, int h) { return 32 * (s / 16) + 8 * ((s % 16) / 4) + 4 * h + (s % 4); }

namespace {
// one thread per packed float of one tile (mlp_params layout: per layer W[out][in], then b[out];
// the last layer has nout <= 32 rows)
__global__ void k_pack_train_frag(const float* __restrict__ src, int64_t src_stride, float* __restrict__ dst,
                                  int in0, int wd, int L, int n_tiles, int nout) {
  const int64_t per = tr_stride(in0, wd, L, nout);
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= per * n_tiles) return;
  const int64_t tile = g / per;
  const int64_t r = g - tile * per;
  float v = 0.0f;
  const float* P = src + tile * src_stride;
  int64_t poff = 0;
  for (int l = 0; l < L; ++l) {
    const int in = l == 0 ? in0 : wd, out = l == L - 1 ? nout : wd;
    const float* Wl = P + poff;
    const float* bl = Wl + (int64_t)out * in;
    const int64_t fo = tr_fwd_off(in0, wd, L, l), bo = tr_bias_off(in0, wd, L, l);
    const int nsf = tr_nsf(in0, wd, l), nob = tr_nob(wd, L, l);
    if (r >= fo && r < bo) {  // forward fragment (ob, s, lane)
      const int64_t q = r - fo, qq = q % (64 * nsf);
      const int lane = (int)((qq % 256) / 4), st = (int)(4 * (qq / 256) + qq % 4), ob = (int)(q / (64 * nsf));
      const int o = 32 * ob + lane % 32, i = tr_row(st, lane / 32);
      if (o < out && i < in) v = Wl[(int64_t)o * in + i];
    } else if (r >= bo && r < bo + 32 * nob) {  // bias in the D layout
      const int q = (int)(r - bo), ob = q / 32, h = (q / 16) % 2, rr = q % 16;
      const int o = 32 * ob + 8 * (rr / 4) + 4 * h + rr % 4;
      if (o < out) v = bl[o];
    }
    const int64_t wo = tr_bwd_off(in0, wd, L, l, nout), we = tr_bwd_off(in0, wd, L, l + 1, nout);
    if (r >= wo && r < we) {  // transposed fragment (ib, s, lane)
      const int nsb = tr_nsb(wd, L, l, nout);
      const int64_t q = r - wo, qq = q % (64 * nsb);
      const int lane = (int)((qq % 256) / 4), st = (int)(4 * (qq / 256) + qq % 4), ib = (int)(q / (64 * nsb));
      const int i = 32 * ib + lane % 32, o = tr_row(st, lane / 32);
      if (o < out && i < in) v = Wl[(int64_t)o * in + i];
    }
    poff += (int64_t)out * in + out;
  }
  dst[g] = v;
}
}  // namespace

// acc += sum_s A[s] x[s] over NS steps (NS % 4 == 0); each lane loads 4 steps' A values at once,
// PD such loads ahead of the products (a ring)
template <int NS, int PD = 6>
__device__ __forceinline__ f32x16 mma_f32_chain(const float* __restrict__ A, const float* x, f32x16 acc, int lane) {
  static_assert(NS % 4 == 0, "steps come in groups of 4");
  constexpr int N4 = NS / 4, D = PD < N4 ? PD : N4;
  const float4* p = reinterpret_cast<const float4*>(A) + lane;
  float4 ring[D];
#pragma unroll
  for (int q = 0; q < D; ++q) ring[q] = p[q * 64];
#pragma unroll
  for (int q = 0; q < N4; ++q) {
    const float4 a = ring[q % D];
    if (q + D < N4) ring[q % D] = p[(q + D) * 64];
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, x[4 * q], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, x[4 * q + 1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, x[4 * q + 2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, x[4 * q + 3], acc, 0, 0, 0);
  }
  return acc;
}

// two independent chains over the same B operands (two output blocks), interleaved so that
// consecutive products never depend on each other
template <int NS, int PD = 4>
__device__ __forceinline__ void mma_f32_chain2(const float* __restrict__ A0, const float* __restrict__ A1,
                                               const float* x, f32x16& acc0, f32x16& acc1, int lane) {
  static_assert(NS % 4 == 0, "steps come in groups of 4");
  constexpr int N4 = NS / 4, D = PD < N4 ? PD : N4;
  const float4* p0 = reinterpret_cast<const float4*>(A0) + lane;
  const float4* p1 = reinterpret_cast<const float4*>(A1) + lane;
  float4 r0[D], r1[D];
#pragma unroll
  for (int q = 0; q < D; ++q) {
    r0[q] = p0[q * 64];
    r1[q] = p1[q * 64];
  }
#pragma unroll
  for (int q = 0; q < N4; ++q) {
    const float4 a0 = r0[q % D], a1 = r1[q % D];
    if (q + D < N4) {
      r0[q % D] = p0[(q + D) * 64];
      r1[q % D] = p1[(q + D) * 64];
    }
    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, x[4 * q], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, x[4 * q], acc1, 0, 0, 0);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, x[4 * q + 1], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, x[4 * q + 1], acc1, 0, 0, 0);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, x[4 * q + 2], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, x[4 * q + 2], acc1, 0, 0, 0);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, x[4 * q + 3], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, x[4 * q + 3], acc1, 0, 0, 0);
  }
}

// out[blk * 16 + r] = f(block blk of sum_s A_blk[s] x[s] + init_blk) for NBLK blocks, two at a time
template <int NS, int NBLK, class Init, class Out>
__device__ __forceinline__ void mma_blocks(const float* __restrict__ A, const float* x, int lane, Init init, Out out) {
#pragma unroll
  for (int b = 0; b + 1 < NBLK; b += 2) {
    f32x16 a0 = init(b), a1 = init(b + 1);
    mma_f32_chain2<NS>(A + b * NS * 64, A + (b + 1) * NS * 64, x, a0, a1, lane);
    out(b, a0);
    out(b + 1, a1);
  }
  if (NBLK % 2) {
    f32x16 a0 = init(NBLK - 1);
    a0 = mma_f32_chain<NS>(A + (NBLK - 1) * NS * 64, x, a0, lane);
    out(NBLK - 1, a0);
  }
}

__device__ __forceinline__ void store4(float* p, const float* v) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

}  // namespace snisr
