// MFMA building blocks shared by the per-tile MLP kernels: v_mfma_f32_32x32x16_{f16,bf16}
// wrappers, fragment loads, the LDS-fed multiply of one 32-row output block, bias broadcast and the
// activation + conversion that turns an accumulator into the next layer's B operand.
//
// Lane layouts of v_mfma_f32_32x32x16 (64 lanes, h = lane / 32, s = lane % 32):
//   A (weights):     lane l holds W[row l%32][k slots 8h..8h+7]
//   B (activations): lane l holds X[k slots 8h..8h+7][column s]
//   D (result):      lane l holds rows 8(i/4) + 4h + (i%4), i = 0..15, of column s
// A layer's D for output block ob is the next layer's B for k-chunks 2ob and 2ob+1 once the next
// layer's weight columns are permuted by pack_input_index.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "snisr.h"

#ifndef SNISR_MLP_PREFETCH
#define SNISR_MLP_PREFETCH 2  // A-fragment LDS reads in flight ahead of the MFMA that uses them
#endif

namespace snisr {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

// input index held in k slot e (0..7) of lane half h for k-chunk kc of layer l
__host__ __device__ constexpr int pack_input_index(int l, int kc, int h, int e) {
  return l == 0 ? 16 * kc + 8 * h + e : 32 * (kc / 2) + 8 * (2 * (kc % 2) + e / 4) + 4 * h + (e % 4);
}

template <int DT>
struct Mfma;
template <>
struct Mfma<SNISR_F16> {
  typedef f16x8 vec;
  __device__ __forceinline__ static f32x16 mma(vec a, vec b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static vec cvt8(const float* v) {
    vec r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (_Float16)v[e];
    return r;
  }
};
template <>
struct Mfma<SNISR_BF16> {
  typedef bf16x8 vec;
  __device__ __forceinline__ static f32x16 mma(vec a, vec b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static vec cvt8(const float* v) {
    vec r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
    return r;
  }
};

template <class V>
__device__ __forceinline__ V load_frag(const uint8_t* __restrict__ p) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  return *reinterpret_cast<const V*>(&q);
}

// acc{0,1} += sum_kc A(ob, kc) * src{0,1}[kc] with the A fragments read from LDS PD steps ahead
// (a ring of PD + 1 fragments); the scheduling barrier keeps that ring the only live fragments.
template <class M, int NK, int PD = 1>
__device__ __forceinline__ void mma_block(const uint8_t* Ab, int lane, const typename M::vec (&s0)[NK],
                                          const typename M::vec (&s1)[NK], f32x16& acc0, f32x16& acc1) {
  using vec = typename M::vec;
  constexpr int R = PD + 1;
  vec ring[R];
#pragma unroll
  for (int p = 0; p < PD && p < NK; ++p) ring[p] = load_frag<vec>(Ab + p * 1024 + lane * 16);
#pragma unroll
  for (int kc = 0; kc < NK; ++kc) {
    if (kc + PD < NK) ring[(kc + PD) % R] = load_frag<vec>(Ab + (kc + PD) * 1024 + lane * 16);
    const vec a = ring[kc % R];
    acc0 = M::mma(a, s0[kc], acc0);
    acc1 = M::mma(a, s1[kc], acc1);
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ __forceinline__ f32x16 bias_block(const float* bias, int ob, int h) {
  f32x16 acc;
  const float4* bp = reinterpret_cast<const float4*>(bias + (ob * 2 + h) * 16);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 t4 = bp[q];
    acc[4 * q] = t4.x; acc[4 * q + 1] = t4.y; acc[4 * q + 2] = t4.z; acc[4 * q + 3] = t4.w;
  }
  return acc;
}

template <class M>
__device__ __forceinline__ void relu_cvt(const f32x16& acc, typename M::vec& lo, typename M::vec& hi) {
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = acc[i] > 0.0f ? acc[i] : 0.0f;
  lo = M::cvt8(v);
  hi = M::cvt8(v + 8);
}

// FP16 operands: convert first, then ReLU on packed halves (max(x, 0) commutes with rounding)
template <>
__device__ __forceinline__ void relu_cvt<Mfma<SNISR_F16>>(const f32x16& acc, f16x8& lo, f16x8& hi) {
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = acc[i];
  const f16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
  lo = __builtin_elementwise_max(Mfma<SNISR_F16>::cvt8(v), z);
  hi = __builtin_elementwise_max(Mfma<SNISR_F16>::cvt8(v + 8), z);
}

// BF16 operands: convert, then a packed 16-bit signed max with 0: a BF16 value whose sign bit is
// set is a negative 16-bit integer and becomes +0, every other value is unchanged, so one vector
// instruction applies the ReLU to two values.  max(x, 0) commutes with rounding, and a value that
// rounds to -0 becomes +0 either way.  A float max would first canonicalize its operand (IEEE mode):
// two vector instructions per value instead of one per two.
template <>
__device__ __forceinline__ void relu_cvt<Mfma<SNISR_BF16>>(const f32x16& acc, bf16x8& lo, bf16x8& hi) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  uint32_t w[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bf16x2 b = __builtin_convertvector((f32x2){acc[2 * q], acc[2 * q + 1]}, bf16x2);
    const s16x2 r = __builtin_elementwise_max(__builtin_bit_cast(s16x2, b), (s16x2){0, 0});
    w[q] = __builtin_bit_cast(uint32_t, r);
  }
  const uint4 l4 = make_uint4(w[0], w[1], w[2], w[3]), h4 = make_uint4(w[4], w[5], w[6], w[7]);
  lo = __builtin_bit_cast(bf16x8, l4);
  hi = __builtin_bit_cast(bf16x8, h4);
}

// LeakyReLU(0.01) on FP16 values: convert, then max(x, 0.01 x) on packed halves (x < 0 -> 0.01 x
// in half precision)
__device__ __forceinline__ void leaky_cvt_f16(const f32x16& acc, f16x8& lo, f16x8& hi) {
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = acc[i];
  const f16x8 slope = {0.01f16, 0.01f16, 0.01f16, 0.01f16, 0.01f16, 0.01f16, 0.01f16, 0.01f16};
  const f16x8 a = Mfma<SNISR_F16>::cvt8(v), b = Mfma<SNISR_F16>::cvt8(v + 8);
  lo = __builtin_elementwise_max(a, a * slope);
  hi = __builtin_elementwise_max(b, b * slope);
}

template <class M, bool LEAKY>
__device__ __forceinline__ void act_cvt(const f32x16& acc, typename M::vec& lo, typename M::vec& hi) {
  if constexpr (LEAKY) {
    static_assert(sizeof(typename M::vec) == sizeof(f16x8), "LeakyReLU on FP16 operands");
    leaky_cvt_f16(acc, lo, hi);
  } else {
    relu_cvt<M>(acc, lo, hi);
  }
}

// One layer over two 32-column blocks: d{0,1}[2ob], d{0,1}[2ob + 1] = act(W(ob) s{0,1} + b(ob))
// (ReLU, or LeakyReLU(0.01) in FP16 when LEAKY)
// for the NOB 32-row output blocks.  Software-pipelined over the output blocks: the ReLU and
// conversion of block ob - 1 are issued between the first MFMAs of block ob (vector work under
// the running matrix pipe instead of after it), and the bias of block ob + 1 is read from LDS
// before the last MFMA pair of block ob.  The layer's A fragments are consecutive 1 KB pieces
// of LDS, read PD steps ahead through a ring that runs across the output blocks.
template <class M, int NK, int NOB, int PD = SNISR_MLP_PREFETCH, bool LEAKY = false>
__device__ __forceinline__ void mma_layer(const uint8_t* A, const float* bias, int lane, int h,
                                          const typename M::vec (&s0)[NK], const typename M::vec (&s1)[NK],
                                          typename M::vec (&d0)[2 * NOB], typename M::vec (&d1)[2 * NOB]) {
  using vec = typename M::vec;
  static_assert(NK >= 2, "two k-steps at least");
  constexpr int K0 = 1, K1 = NK > 2 ? 2 : 1;  // k-steps that carry the pending activations
  constexpr int NG = NK * NOB, R = PD + 1;
  const uint8_t* Al = A + lane * 16;
  vec ring[R];
#pragma unroll
  for (int g = 0; g < PD && g < NG; ++g) ring[g] = load_frag<vec>(Al + g * 1024);
  f32x16 p0, p1;
  f32x16 bn = bias_block(bias, 0, h);
#pragma unroll
  for (int ob = 0; ob < NOB; ++ob) {
    f32x16 acc0 = bn, acc1 = bn;
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
      const int g = ob * NK + kc;
      if (g + PD < NG) ring[(g + PD) % R] = load_frag<vec>(Al + (g + PD) * 1024);
      if (kc == NK - 1 && ob + 1 < NOB) bn = bias_block(bias, ob + 1, h);
      const vec a = ring[g % R];
      acc0 = M::mma(a, s0[kc], acc0);
      acc1 = M::mma(a, s1[kc], acc1);
      if (ob > 0 && kc == K0) act_cvt<M, LEAKY>(p0, d0[2 * ob - 2], d0[2 * ob - 1]);
      if (ob > 0 && kc == K1) act_cvt<M, LEAKY>(p1, d1[2 * ob - 2], d1[2 * ob - 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
    p0 = acc0;
    p1 = acc1;
  }
  act_cvt<M, LEAKY>(p0, d0[2 * NOB - 2], d0[2 * NOB - 1]);
  act_cvt<M, LEAKY>(p1, d1[2 * NOB - 2], d1[2 * NOB - 1]);
}

// Sigmoid on the hardware exp2 and reciprocal (about 1 ulp each): the colour of the MFMA path, whose
// FP16 layers already set its error; the FP32 path keeps sigmoid_cw.  exp2 of a large argument
// gives +inf and 1 / inf = 0, so both tails saturate correctly.
__device__ __forceinline__ float sigmoid_hw(float z) {
  const float e = __builtin_amdgcn_exp2f(z * -1.44269504f);
  return __builtin_amdgcn_rcpf(1.0f + e);
}

// v_permlane32_swap on one value held by every lane: lo receives the value of lanes 0..31 and hi
// that of lanes 32..63, both in all 64 lanes (lane l and lane l ^ 32 end up with the same pair).
template <class T>
__device__ __forceinline__ void swap_halves(T own, T& lo, T& hi) {
  const uint32_t v = __builtin_bit_cast(uint32_t, own);
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  lo = __builtin_bit_cast(T, (uint32_t)r[0]);
  hi = __builtin_bit_cast(T, (uint32_t)r[1]);
}

// v_permlane32_swap of two values: lanes 32..63 of a are exchanged with lanes 0..31 of b, so
// a_out = (a of lanes 0..31, b of lanes 0..31) and b_out = (a of lanes 32..63, b of lanes 32..63).
__device__ __forceinline__ void swap_lanes(float a, float b, float& a_out, float& b_out) {
  const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                  false, false);
  a_out = __builtin_bit_cast(float, (uint32_t)r[0]);
  b_out = __builtin_bit_cast(float, (uint32_t)r[1]);
}

}  // namespace snisr
