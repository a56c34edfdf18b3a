// Weight-gradient reduction shared by the FP32 training kernels: dW += D^T A and db += sum D over
// a range of stored per-sample rows, on v_mfma_f32_32x32x2_f32.
#pragma once

#include "mfma_ops.hpp"

namespace snisr {

// dW_l[o][i] += sum_j D_l[j][o] A_l[j][i] and db_l[o] += sum_j D_l[j][o] over one split of one tile's
// sorted records, on v_mfma_f32_32x32x2_f32 with the records as the K dimension.  A workgroup owns
// every (out-block, in-block) pair of the layer (at most 16 pairs of 32 x 32; wave w takes pairs
// w, w + 4, ...) and walks its records in tiles of 32 staged through LDS: the A operand of step s
// is D rows 2s + h at columns ob*32 + c, the B operand A rows 2s + h at columns ib*32 + c.  Rows are
// padded to 160 floats so that the two lane halves read disjoint banks.  The next tile's rows are
// loaded into registers while the current tile's products run.
//
// Row layouts: with LANE false the rows are per record ([record][feature], pitch IP / OP); with LANE
// true each 32-record block (blocks start at multiples of 32) is stored in the order the training
// kernel's lanes hold it, feature group g (256 floats) = lane l's four features
// 32 (g / 4) + 8 (g % 4) + 4 (l / 32) .. + 3 of record l % 32, so that the kernel writes every
// group with one 1 KB wavefront store; a pitch of 4 is the same in both layouts.
constexpr int kDwRows = 32, kDwPitch = 160;
// float offset of (record, feature f, a multiple of 4) inside a 32-record block in the lane layout
__device__ __forceinline__ int lane_off(int rec, int f) {
  return 256 * (4 * (f >> 5) + ((f >> 3) & 3)) + 4 * (rec + 32 * ((f >> 2) & 1));
}
template <int IP, int OP, int INL, int OUTL, bool LANE = false>  // padded row widths, real in / out sizes
__device__ __forceinline__ void dw_layer(const float* __restrict__ A, const float* __restrict__ D, int64_t jb,
                                         int64_t je, float* __restrict__ G, float* Ds, float* As) {
  static_assert(!LANE || ((IP % 8 == 0 || IP == 4) && (OP % 8 == 0 || OP == 4)), "lane layout pitches");
  constexpr int IP4 = IP / 4, OP4 = OP / 4;
  constexpr int ND = (kDwRows * OP4 + 255) / 256, NA = (kDwRows * IP4 + 255) / 256;
  constexpr int NIB = (INL + 31) / 32, NOB = (OUTL + 31) / 32, NPAIR = NIB * NOB;
  static_assert(NPAIR <= 16, "at most 4 pairs per wave");
  static_assert(2 * OP <= 256, "two threads per output column");
  const int lane = threadIdx.x & 63, c = lane & 31, h = lane >> 5, wv = threadIdx.x >> 6;
  f32x16 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = f32x16{};
  float accb = 0.0f;
  float4 rd[ND], ra[NA];
  // thread q moves float4 (record q / P4, features 4 (q % P4) ..) of the 32-record tile, so that the
  // LDS stores below are conflict-free; in the lane layout its global load then reads 16 bytes of
  // a line that the workgroup's other loads of the tile complete
  auto fetch = [&](int64_t r0) {
#pragma unroll
    for (int u = 0; u < ND; ++u) {
      const int q = threadIdx.x + 256 * u, rr = q / OP4, cc = 4 * (q % OP4);
      rd[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (q < kDwRows * OP4 && r0 + rr < je)
        rd[u] = *reinterpret_cast<const float4*>(D + r0 * OP + (LANE ? lane_off(rr, cc) : rr * OP + cc));
    }
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int q = threadIdx.x + 256 * u, rr = q / IP4, cc = 4 * (q % IP4);
      ra[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (q < kDwRows * IP4 && r0 + rr < je)
        ra[u] = *reinterpret_cast<const float4*>(A + r0 * IP + (LANE ? lane_off(rr, cc) : rr * IP + cc));
    }
  };
  if (jb < je) fetch(jb);
  for (int64_t r0 = jb; r0 < je; r0 += kDwRows) {
    __syncthreads();  // the previous tile's reads of Ds / As are done
#pragma unroll
    for (int u = 0; u < ND; ++u) {
      const int q = threadIdx.x + 256 * u;
      if (q < kDwRows * OP4) *reinterpret_cast<float4*>(Ds + (q / OP4) * kDwPitch + 4 * (q % OP4)) = rd[u];
    }
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int q = threadIdx.x + 256 * u;
      if (q < kDwRows * IP4) *reinterpret_cast<float4*>(As + (q / IP4) * kDwPitch + 4 * (q % IP4)) = ra[u];
    }
    __syncthreads();
    if (r0 + kDwRows < je) fetch(r0 + kDwRows);
    // db: thread t < 2 OP sums rows (t / OP) * 16 .. + 15 of column t % OP, four independent
    // partial sums so that the LDS reads do not wait on each other
    if (threadIdx.x < 2 * OP) {
      const int col = threadIdx.x % OP, r0h = (threadIdx.x / OP) * (kDwRows / 2);
      float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int rr = 0; rr < kDwRows / 2; ++rr) part[rr & 3] += Ds[(r0h + rr) * kDwPitch + col];
      accb += (part[0] + part[1]) + (part[2] + part[3]);
    }
    // the 16 steps' operands of the wave's next pair are read from LDS while the current pair's
    // products run (a ring of two operand sets), so no product waits on its own LDS read
    float xs[2][kDwRows / 2], ys[2][kDwRows / 2];
    auto load_pair = [&](int p, float (&x)[kDwRows / 2], float (&y)[kDwRows / 2]) {
      const int ob = p / NIB, ib = p % NIB;
#pragma unroll
      for (int st = 0; st < kDwRows / 2; ++st) {
        x[st] = Ds[(2 * st + h) * kDwPitch + ob * 32 + c];
        y[st] = As[(2 * st + h) * kDwPitch + ib * 32 + c];
      }
    };
    if (wv < NPAIR) load_pair(wv, xs[0], ys[0]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int p = wv + 4 * q;
      if (p < NPAIR) {  // wave-uniform
        if (q + 1 < 4 && p + 4 < NPAIR) load_pair(p + 4, xs[(q + 1) & 1], ys[(q + 1) & 1]);
        f32x16 a = acc[q];
#pragma unroll
        for (int st = 0; st < kDwRows / 2; ++st) a = __builtin_amdgcn_mfma_f32_32x32x2f32(xs[q & 1][st], ys[q & 1][st], a, 0, 0, 0);
        acc[q] = a;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int p = wv + 4 * q;
    if (p >= NPAIR) break;
    const int ob = p / NIB, ib = p % NIB;
    const int i = ib * 32 + c;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = ob * 32 + 8 * (r / 4) + 4 * h + (r % 4);
      if (o < OUTL && i < INL && acc[q][r] != 0.0f) atomicAdd(&G[(int64_t)o * INL + i], acc[q][r]);
    }
  }
  if (threadIdx.x < 2 * OP && threadIdx.x % OP < OUTL && accb != 0.0f)
    atomicAdd(&G[(int64_t)OUTL * INL + threadIdx.x % OP], accb);
}

}  // namespace snisr
