// The reference's screen-space refinement U-Net (rendering/scripts/networks.py:5-73: ten 3x3
// convolutions with ReLU, 2x2 average pooling, bilinear x2 upsampling with align_corners and two
// channel concatenations; applied to the composited frame at rendering/src/__init__.py:336-339) on
// FP16 MFMA with FP32 accumulation.
//
// Activations are NHWC FP16.  A convolution is an implicit GEMM: output channels are the rows
// (32-row blocks, the A operand of v_mfma_f32_32x32x16_f16), pixels the columns (a wavefront takes
// 64 consecutive pixels of one row as two 32-column blocks) and K runs over the 9 taps x 16-channel
// chunks.  Lane (h, s) loads the B operand of one step as one 16-byte piece: channels 16kc + 8h ..
// +7 of pixel s shifted by the tap.  Pixels outside the image read a zeroed row of the workspace
// instead (zero padding without branches).  A concatenated input is two NHWC sources, chosen per
// 16-channel chunk, so the concatenation is never materialised.  Each workgroup stages the A
// fragments of its output blocks (all taps and chunks, packed by k_unet_pack_layer) in LDS once and
// loops over pixel tiles; the bias and ReLU are applied to the accumulators before the FP16 store.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mfma_ops.hpp"
#include "snisr_common.hpp"

namespace snisr {
namespace {

struct UnetLayer {
  int cin, cin_pad, cout, nob;
};
// conv1 .. conv10 of Unet (networks.py:8-15); cin_pad: channels rounded up to 16, nob: 32-row blocks
constexpr UnetLayer kUnet[10] = {{5, 16, 64, 2},    {64, 64, 32, 1},   {32, 32, 64, 2},  {64, 64, 64, 2},
                                 {64, 64, 128, 4},  {128, 128, 128, 4}, {192, 192, 64, 2}, {64, 64, 64, 2},
                                 {96, 96, 32, 1},   {32, 32, 3, 1}};

// packed layer: per output block ob, per step g = tap * (cin_pad / 16) + kc, a 1 KB A fragment
// (lane l = (h, s): W[32 ob + s][16 kc + 8 h + e][tap], e = 0..7, as FP16); then per output block
// the bias in accumulator order (32 floats)
__host__ __device__ constexpr int64_t unet_frag_bytes(int l) {
  return (int64_t)kUnet[l].nob * 9 * (kUnet[l].cin_pad / 16) * 1024;
}
__host__ __device__ constexpr int64_t unet_layer_bytes(int l) { return unet_frag_bytes(l) + kUnet[l].nob * 128; }
__host__ __device__ constexpr int64_t unet_layer_off(int l) {
  int64_t s = 0;
  for (int m = 0; m < l; ++m) s += unet_layer_bytes(m);
  return s;
}

__global__ void k_unet_pack_layer(const float* __restrict__ w, const float* __restrict__ b, uint8_t* __restrict__ dst,
                                  int l) {
  const UnetLayer L = kUnet[l];
  const int nkc = L.cin_pad / 16;
  const int64_t nfrag = (int64_t)L.nob * 9 * nkc;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nfrag * 64) {
    const int lane = (int)(i & 63), h = lane >> 5, s = lane & 31;
    const int64_t f = i >> 6;
    const int ob = (int)(f / (9 * nkc)), g = (int)(f % (9 * nkc)), tap = g / nkc, kc = g % nkc;
    const int co = 32 * ob + s;
    _Float16 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int ci = 16 * kc + 8 * h + e;
      v[e] = (co < L.cout && ci < L.cin) ? (_Float16)w[((int64_t)co * L.cin + ci) * 9 + tap] : (_Float16)0.0f;
    }
    *reinterpret_cast<uint4*>(dst + f * 1024 + lane * 16) = *reinterpret_cast<const uint4*>(v);
  } else if (i < nfrag * 64 + L.nob * 32) {
    const int j = (int)(i - nfrag * 64), ob = j >> 5, h = (j >> 4) & 1, r = j & 15;
    const int co = 32 * ob + 8 * (r >> 2) + 4 * h + (r & 3);  // accumulator row of element r in lane half h
    reinterpret_cast<float*>(dst + unet_frag_bytes(l))[j] = co < L.cout ? b[co] : 0.0f;
  }
}

struct ConvArgs {
  const _Float16* src0;  // NHWC, C0 channels: input channels [0, C0)
  const _Float16* src1;  // NHWC, the remaining input channels (unused when src0 holds them all)
  int H, W;
  const _Float16* zero;  // a zeroed row of at least 192 halves (the padding pixel)
  const uint8_t* layer;  // packed layer
  _Float16* dst;         // MODE 0: NHWC output with cout channels
  float* out;            // MODE 1: [H, W, 3] float, min(ReLU(x), 1)
  int cout;
};

constexpr int kConvThreads = 512;

// NKC 16-channel input chunks, the first C0 / 16 of them from src0; OBW output blocks per workgroup;
// MODE 0: FP16 NHWC output with ReLU, 1: the final three channels as min(ReLU(x), 1) in FP32
template <int NKC, int C0, int OBW, int MODE>
__global__ __launch_bounds__(kConvThreads) void k_unet_conv(ConvArgs a) {
  using M = Mfma<SNISR_F16>;
  using vec = f16x8;
  constexpr int NS = 9 * NKC;  // steps of one output block: tap-major, then 16-channel chunk
  constexpr int PD = 4;        // steps whose B operands are loaded ahead
  constexpr int R = PD + 1;
  extern __shared__ __align__(16) uint8_t lds[];
  const int nob_total = (int)gridDim.y * OBW;
  const int ob0 = blockIdx.y * OBW;
  {
    const uint4* src = reinterpret_cast<const uint4*>(a.layer + (int64_t)ob0 * NS * 1024);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int k = threadIdx.x; k < OBW * NS * 64; k += blockDim.x) dst[k] = src[k];
    const float4* bs = reinterpret_cast<const float4*>(a.layer + (int64_t)nob_total * NS * 1024) + ob0 * 8;
    float4* bd = reinterpret_cast<float4*>(lds + OBW * NS * 1024);
    for (int k = threadIdx.x; k < OBW * 8; k += blockDim.x) bd[k] = bs[k];
  }
  __syncthreads();
  const float* bias = reinterpret_cast<const float*>(lds + OBW * NS * 1024);
  const int lane = threadIdx.x & 63, h = lane >> 5, s = lane & 31;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int H = a.H, W = a.W;
  const int tpr = (W + 63) >> 6;
  const int64_t ntiles = (int64_t)H * tpr;
  for (int64_t tile = (int64_t)blockIdx.x * nwv + wv; tile < ntiles; tile += (int64_t)gridDim.x * nwv) {
    const int y = (int)(tile / tpr), x0 = (int)(tile % tpr) * 64;
    // per tap and column block: the row of the shifted pixel in each source (or the zero row, whose
    // 192 halves cover every chunk offset); chunk kc then sits 16 kc halves further
    const auto row = [&](int tap, int cb, int src) -> const _Float16* {
      const int yy = y + tap / 3 - 1, x = x0 + 32 * cb + s, xx = x + tap % 3 - 1;
      const bool ok = x < W && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
      const int64_t pix = ok ? (int64_t)yy * W + xx : 0;
      if (src == 0) return ok ? a.src0 + pix * C0 + 8 * h : a.zero + 8 * h;
      return ok ? a.src1 + pix * (NKC * 16 - C0) + 8 * h - C0 : a.zero + 8 * h - C0;
    };
    // B operand of step g for column block cb: 16 bytes of the shifted pixel (or of the zero row)
    const auto load_b = [&](int g, int cb) -> vec {
      const int tap = g / NKC, kc = g % NKC;
      const _Float16* p = row(tap, cb, 16 * kc < C0 ? 0 : 1) + 16 * kc;
      const uint4 q = *reinterpret_cast<const uint4*>(p);
      return *reinterpret_cast<const vec*>(&q);
    };
    f32x16 acc[OBW][2];
#pragma unroll
    for (int ob = 0; ob < OBW; ++ob) {
      acc[ob][0] = bias_block(bias, ob, h);
      acc[ob][1] = acc[ob][0];
    }
    vec ring[R][2];
#pragma unroll
    for (int g = 0; g < PD; ++g) {
      ring[g][0] = load_b(g, 0);
      ring[g][1] = load_b(g, 1);
    }
#pragma unroll
    for (int g = 0; g < NS; ++g) {
      if (g + PD < NS) {
        ring[(g + PD) % R][0] = load_b(g + PD, 0);
        ring[(g + PD) % R][1] = load_b(g + PD, 1);
      }
#pragma unroll
      for (int ob = 0; ob < OBW; ++ob) {
        const vec af = load_frag<vec>(lds + (ob * NS + g) * 1024 + lane * 16);
        acc[ob][0] = M::mma(af, ring[g % R][0], acc[ob][0]);
        acc[ob][1] = M::mma(af, ring[g % R][1], acc[ob][1]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    // epilogue: lane (h, s) holds rows 8j + 4h + (0..3) of pixel column s of each block
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int x = x0 + 32 * cb + s;
      if (x >= W) continue;
      const int64_t pix = (int64_t)y * W + x;
      if constexpr (MODE == 0) {
#pragma unroll
        for (int ob = 0; ob < OBW; ++ob)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            _Float16 v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float t = acc[ob][cb][4 * j + r];
              v[r] = (_Float16)(t > 0.0f ? t : 0.0f);
            }
            const int co = 32 * (ob0 + ob) + 8 * j + 4 * h;
            *reinterpret_cast<uint2*>(a.dst + pix * a.cout + co) = *reinterpret_cast<const uint2*>(v);
          }
      } else {
        if (h == 0) {
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float t = acc[0][cb][r];
            a.out[pix * 3 + r] = t > 0.0f ? (t < 1.0f ? t : 1.0f) : 0.0f;
          }
        }
      }
    }
  }
}

// Layers whose workgroups hold a single 32-channel output block are bound by their B loads when
// every tap reloads its pixels (one 16-byte load per MFMA).  Used for conv2 (64 -> 32 at full
// resolution: 162 -> 119 us at 720p); conv9 (96 -> 32) fits only 4 rows next to its weights and
// is faster on k_unet_conv (220 vs 259 us).  This variant stages the ROWS + 2 input
// rows of a 66-pixel span (a tile of ROWS x 64 output pixels and its halo; zero outside the image,
// read from the zero row) in LDS once per tile, and the nine taps of its ROWS wavefronts (one
// output row each) take their B operands from there.  The next tile's halo is loaded into
// registers while the current tile is computed, and written to LDS after it.
template <int NKC, int C0, int ROWS>
__global__ __launch_bounds__(64 * ROWS) void k_unet_conv_rows(ConvArgs a) {
  using M = Mfma<SNISR_F16>;
  using vec = f16x8;
  constexpr int NS = 9 * NKC;
  constexpr int SPAN = 66;                   // pixels per staged row
  constexpr int CHK = 2 * NKC;               // 16-byte chunks per pixel
  constexpr int PITCH = NKC * 32 + 16;       // bytes per staged pixel (the 16 extra bytes spread the banks)
  constexpr int TOTAL = (ROWS + 2) * SPAN * CHK;
  constexpr int PER = (TOTAL + 64 * ROWS - 1) / (64 * ROWS);
  extern __shared__ __align__(16) uint8_t lds[];
  uint8_t* halo = lds + NS * 1024 + 128;
  const int ob = blockIdx.y;
  {
    const uint4* src = reinterpret_cast<const uint4*>(a.layer + (int64_t)ob * NS * 1024);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int k = threadIdx.x; k < NS * 64; k += blockDim.x) dst[k] = src[k];
    const float4* bs = reinterpret_cast<const float4*>(a.layer + (int64_t)gridDim.y * NS * 1024) + ob * 8;
    float4* bd = reinterpret_cast<float4*>(lds + NS * 1024);
    if (threadIdx.x < 8) bd[threadIdx.x] = bs[threadIdx.x];
  }
  const float* bias = reinterpret_cast<const float*>(lds + NS * 1024);
  const int lane = threadIdx.x & 63, h = lane >> 5, s = lane & 31, wv = threadIdx.x >> 6;
  const int H = a.H, W = a.W;
  const int tpr = (W + 63) >> 6;
  const int64_t ntiles = (int64_t)((H + ROWS - 1) / ROWS) * tpr;
  // this thread's chunks of a tile's halo, loaded into registers
  uint4 hv[PER];
  const auto fetch = [&](int64_t tile) {
    const int y0 = (int)(tile / tpr) * ROWS, x0 = (int)(tile % tpr) * 64;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int k = threadIdx.x + j * 64 * ROWS;
      const int c = k % CHK, px = (k / CHK) % SPAN, r = k / (CHK * SPAN);
      const int yy = y0 + r - 1, xx = x0 + px - 1, ch = 8 * c;
      const bool ok = k < TOTAL && (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
      const int64_t pix = ok ? (int64_t)yy * W + xx : 0;
      const _Float16* p = ch < C0 ? a.src0 + pix * C0 + ch : a.src1 + pix * (NKC * 16 - C0) + (ch - C0);
      hv[j] = *reinterpret_cast<const uint4*>(ok ? p : a.zero + 8 * (c & 1));
    }
  };
  int64_t tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  for (; tile < ntiles; tile += gridDim.x) {
    __syncthreads();  // the weights are staged; the previous tile's halo reads are done
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int k = threadIdx.x + j * 64 * ROWS;
      if (k < TOTAL) {
        const int c = k % CHK, px = (k / CHK) % SPAN, r = k / (CHK * SPAN);
        *reinterpret_cast<uint4*>(halo + (r * SPAN + px) * PITCH + c * 16) = hv[j];
      }
    }
    __syncthreads();
    if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);  // in flight during the MFMAs below
    const int y = (int)(tile / tpr) * ROWS + wv, x0 = (int)(tile % tpr) * 64;
    if (y >= H) continue;
    f32x16 acc[2];
    acc[0] = bias_block(bias, 0, h);
    acc[1] = acc[0];
#pragma unroll
    for (int g = 0; g < NS; ++g) {
      const int tap = g / NKC, kc = g % NKC, dy = tap / 3, dx = tap % 3;
      const vec af = load_frag<vec>(lds + g * 1024 + lane * 16);
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const vec bf = load_frag<vec>(halo + ((wv + dy) * SPAN + 32 * cb + s + dx) * PITCH + (16 * kc + 8 * h) * 2);
        acc[cb] = M::mma(af, bf, acc[cb]);
      }
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int x = x0 + 32 * cb + s;
      if (x >= W) continue;
      const int64_t pix = (int64_t)y * W + x;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        _Float16 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float t = acc[cb][4 * j + r];
          v[r] = (_Float16)(t > 0.0f ? t : 0.0f);
        }
        const int co = 32 * ob + 8 * j + 4 * h;
        *reinterpret_cast<uint2*>(a.dst + pix * a.cout + co) = *reinterpret_cast<const uint2*>(v);
      }
    }
  }
}

// 2x2 average pooling of an NHWC FP16 tensor (nn.AvgPool2d(2, 2)), FP32 sum of the window in row order
__global__ void k_unet_pool(const _Float16* __restrict__ in, int H, int W, int C, _Float16* __restrict__ out) {
  const int G = C / 8, Ho = H / 2, Wo = W / 2;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)Ho * Wo * G) return;
  const int g = (int)(i % G);
  const int64_t p = i / G;
  const int yo = (int)(p / Wo), xo = (int)(p % Wo);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
      const uint4 q = *reinterpret_cast<const uint4*>(in + ((int64_t)(2 * yo + dy) * W + 2 * xo + dx) * C + 8 * g);
      const _Float16* hv = reinterpret_cast<const _Float16*>(&q);
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[e] = acc[e] + (float)hv[e];
    }
  _Float16 r[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = (_Float16)(acc[e] / 4.0f);
  *reinterpret_cast<uint4*>(out + p * C + 8 * g) = *reinterpret_cast<const uint4*>(r);
}

// bilinear x2 upsampling with align_corners (nn.Upsample(scale_factor=2, mode="bilinear",
// align_corners=True)): source coordinate (in - 1) / (out - 1) * dst in FP32, the four neighbours
// weighted in the row-then-column order of the usual align-corners kernel
__global__ void k_unet_up(const _Float16* __restrict__ in, int Hi, int Wi, int C, _Float16* __restrict__ out) {
  const int G = C / 8, Ho = 2 * Hi, Wo = 2 * Wi;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)Ho * Wo * G) return;
  const int g = (int)(i % G);
  const int64_t p = i / G;
  const int yo = (int)(p / Wo), xo = (int)(p % Wo);
  const float sh = Ho > 1 ? (float)(Hi - 1) / (float)(Ho - 1) : 0.0f;
  const float sw = Wo > 1 ? (float)(Wi - 1) / (float)(Wo - 1) : 0.0f;
  const float hr = sh * (float)yo, wr = sw * (float)xo;
  const int h1 = (int)hr, w1 = (int)wr;
  const int hp = h1 < Hi - 1 ? 1 : 0, wp = w1 < Wi - 1 ? 1 : 0;
  const float l1 = hr - (float)h1, l0 = 1.0f - l1, m1 = wr - (float)w1, m0 = 1.0f - m1;
  const auto px = [&](int yy, int xx) {
    const uint4 q = *reinterpret_cast<const uint4*>(in + ((int64_t)yy * Wi + xx) * C + 8 * g);
    return q;
  };
  const uint4 q00 = px(h1, w1), q01 = px(h1, w1 + wp), q10 = px(h1 + hp, w1), q11 = px(h1 + hp, w1 + wp);
  const _Float16 *a = reinterpret_cast<const _Float16*>(&q00), *b = reinterpret_cast<const _Float16*>(&q01),
                 *c = reinterpret_cast<const _Float16*>(&q10), *d = reinterpret_cast<const _Float16*>(&q11);
  _Float16 r[8];
#pragma unroll
  for (int e = 0; e < 8; ++e)
    r[e] = (_Float16)(l0 * (m0 * (float)a[e] + m1 * (float)b[e]) + l1 * (m0 * (float)c[e] + m1 * (float)d[e]));
  *reinterpret_cast<uint4*>(out + p * C + 8 * g) = *reinterpret_cast<const uint4*>(r);
}

// the 5-channel input (rgb, inverse depth, boundary flag; __init__.py:336-338) as NHWC16 FP16, and
// the zero row used as the padding pixel
__global__ void k_unet_input(const float* __restrict__ frame, const float* __restrict__ inv_near,
                             const uint8_t* __restrict__ boundary, int64_t n, _Float16* __restrict__ x16,
                             _Float16* __restrict__ zero) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 192) zero[i] = (_Float16)0.0f;
  if (i >= n) return;
  _Float16 v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) v[e] = (_Float16)0.0f;
  v[0] = (_Float16)frame[3 * i];
  v[1] = (_Float16)frame[3 * i + 1];
  v[2] = (_Float16)frame[3 * i + 2];
  v[3] = (_Float16)inv_near[i];
  v[4] = (_Float16)(boundary[i] ? 1.0f : 0.0f);
  uint4* d = reinterpret_cast<uint4*>(x16 + i * 16);
  d[0] = reinterpret_cast<const uint4*>(v)[0];
  d[1] = reinterpret_cast<const uint4*>(v)[1];
}

// workspace: the FP16 NHWC activations of every stage, 256-byte aligned
struct UnetWs {
  _Float16 *x16, *a1, *o2, *p1, *a3, *o5, *p2, *a5, *e7, *u7, *a7, *e4, *u9, *a9, *zero;
};

int64_t unet_ws(int64_t H, int64_t W, uint8_t* base, UnetWs* w) {
  const int64_t P = H * W, P2 = P / 4, P4 = P / 16;
  const int64_t halves[15] = {P * 16, P * 64, P * 32, P2 * 32, P2 * 64, P2 * 64, P4 * 64, P4 * 128,
                              P4 * 128, P2 * 128, P2 * 64, P2 * 64, P * 64, P * 32, 256};
  _Float16** slots[15] = {&w->x16, &w->a1, &w->o2, &w->p1, &w->a3, &w->o5, &w->p2, &w->a5,
                          &w->e7,  &w->u7, &w->a7, &w->e4, &w->u9, &w->a9, &w->zero};
  int64_t off = 0;
  for (int k = 0; k < 15; ++k) {
    if (base) *slots[k] = reinterpret_cast<_Float16*>(base + off);
    off += (halves[k] * 2 + 255) & ~(int64_t)255;
  }
  return off;
}

template <int NKC, int C0, int OBW, int MODE>
int conv_launch(int l, const _Float16* s0, const _Float16* s1, int H, int W, const UnetWs& ws, const uint8_t* packed,
                _Float16* dst, float* out, hipStream_t st) {
  static_assert(NKC * 16 <= 192 && C0 % 16 == 0 && C0 <= NKC * 16, "chunks within the zero row and the sources");
  const UnetLayer L = kUnet[l];
  const size_t lds = (size_t)OBW * 9 * NKC * 1024 + (size_t)OBW * 128;
  auto kern = k_unet_conv<NKC, C0, OBW, MODE>;
  static bool lds_set = false;
  if (!lds_set) {
    SNISR_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    lds_set = true;
  }
  ConvArgs a;
  a.src0 = s0;
  a.src1 = s1;
  a.H = H;
  a.W = W;
  a.zero = ws.zero;
  a.layer = packed + unet_layer_off(l);
  a.dst = dst;
  a.out = out;
  a.cout = L.cout;
  const int gy = L.nob / OBW;
  const int64_t tiles = (int64_t)H * ((W + 63) / 64);
  const int per_cu = lds > 80 * 1024 ? 1 : 2;
  int64_t gx = ((int64_t)per_cu * device_cus() + gy - 1) / gy;
  const int64_t need = (tiles + kConvThreads / 64 - 1) / (kConvThreads / 64);
  if (gx > need) gx = need;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(kern, dim3((unsigned)gx, (unsigned)gy), dim3(kConvThreads), lds, st, a);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

template <int NKC, int C0, int ROWS>
int conv_rows_launch(int l, const _Float16* s0, const _Float16* s1, int H, int W, const UnetWs& ws,
                     const uint8_t* packed, _Float16* dst, hipStream_t st) {
  static_assert(NKC * 16 <= 192 && C0 % 16 == 0 && C0 <= NKC * 16, "chunks within the zero row and the sources");
  const UnetLayer L = kUnet[l];
  const size_t lds = (size_t)9 * NKC * 1024 + 128 + (size_t)(ROWS + 2) * 66 * (NKC * 32 + 16);
  auto kern = k_unet_conv_rows<NKC, C0, ROWS>;
  static bool lds_set = false;
  if (!lds_set) {
    SNISR_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    lds_set = true;
  }
  ConvArgs a;
  a.src0 = s0;
  a.src1 = s1;
  a.H = H;
  a.W = W;
  a.zero = ws.zero;
  a.layer = packed + unet_layer_off(l);
  a.dst = dst;
  a.out = nullptr;
  a.cout = L.cout;
  const int64_t tiles = (int64_t)((H + ROWS - 1) / ROWS) * ((W + 63) / 64);
  const int per_cu = lds > 80 * 1024 ? 1 : 2;
  int64_t gx = ((int64_t)per_cu * device_cus() + L.nob - 1) / L.nob;
  if (gx > tiles) gx = tiles;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(kern, dim3((unsigned)gx, (unsigned)L.nob), dim3(64 * ROWS), lds, st, a);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int pool_launch(const _Float16* in, int H, int W, int C, _Float16* out, hipStream_t st) {
  const int64_t n = (int64_t)(H / 2) * (W / 2) * (C / 8);
  hipLaunchKernelGGL(k_unet_pool, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, H, W, C, out);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int up_launch(const _Float16* in, int Hi, int Wi, int C, _Float16* out, hipStream_t st) {
  const int64_t n = (int64_t)(2 * Hi) * (2 * Wi) * (C / 8);
  hipLaunchKernelGGL(k_unet_up, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, Hi, Wi, C, out);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_unet_packed_size(void) { return unet_layer_off(10); }

int snisr_unet_pack(const float* const* weights, const float* const* biases, void* packed, void* stream) {
  SNISR_REQUIRE(weights && biases && packed, SNISR_ERR_ARG, "unet_pack: NULL pointer");
  SNISR_REQUIRE(aligned(packed, 256), SNISR_ERR_ALIGN, "unet_pack: packed must be 256-byte aligned");
  for (int l = 0; l < 10; ++l) {
    SNISR_REQUIRE(weights[l] && biases[l], SNISR_ERR_ARG, "unet_pack: layer %d has a NULL tensor", l);
    const int64_t n = unet_frag_bytes(l) / 16 + kUnet[l].nob * 32;
    hipLaunchKernelGGL(k_unet_pack_layer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       weights[l], biases[l], (uint8_t*)packed + unet_layer_off(l), l);
    SNISR_LAUNCHED();
  }
  return SNISR_OK;
}

int64_t snisr_unet_workspace_size(int32_t H, int32_t W) {
  if (H <= 0 || W <= 0 || H % 4 || W % 4) {
    set_error("unet_workspace_size: H and W must be positive multiples of 4 (got %d x %d)", H, W);
    return -1;
  }
  return unet_ws(H, W, nullptr, nullptr);
}

int snisr_unet_refine(const void* packed, const float* frame, const float* inverse_near, const uint8_t* boundary,
                      int32_t H, int32_t W, void* workspace, int64_t workspace_bytes, float* out, void* stream) {
  SNISR_REQUIRE(H > 0 && W > 0 && H % 4 == 0 && W % 4 == 0, SNISR_ERR_SHAPE,
                "unet_refine: H and W must be positive multiples of 4 (got %d x %d)", H, W);
  SNISR_REQUIRE(packed && frame && inverse_near && boundary && workspace && out, SNISR_ERR_ARG,
                "unet_refine: NULL pointer");
  SNISR_REQUIRE(aligned(packed, 256) && aligned(workspace, 256), SNISR_ERR_ALIGN,
                "unet_refine: packed and workspace must be 256-byte aligned");
  SNISR_REQUIRE(workspace_bytes >= unet_ws(H, W, nullptr, nullptr), SNISR_ERR_CAPACITY,
                "unet_refine: workspace too small (snisr_unet_workspace_size)");
  UnetWs w;
  unet_ws(H, W, (uint8_t*)workspace, &w);
  const hipStream_t st = (hipStream_t)stream;
  const uint8_t* pk = (const uint8_t*)packed;
  const int64_t P = (int64_t)H * W;
  const int H2 = H / 2, W2 = W / 2, H4 = H / 4, W4 = W / 4;
  hipLaunchKernelGGL(k_unet_input, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, frame, inverse_near, boundary,
                     P, w.x16, w.zero);
  SNISR_LAUNCHED();
  int e;
  // o2 = relu(conv2(relu(conv1(x))))
  if ((e = conv_launch<1, 16, 2, 0>(0, w.x16, nullptr, H, W, w, pk, w.a1, nullptr, st))) return e;
  if ((e = conv_rows_launch<4, 64, 8>(1, w.a1, nullptr, H, W, w, pk, w.o2, st))) return e;
  // o5 = relu(conv4(relu(conv3(down(o2)))))
  if ((e = pool_launch(w.o2, H, W, 32, w.p1, st))) return e;
  if ((e = conv_launch<2, 32, 2, 0>(2, w.p1, nullptr, H2, W2, w, pk, w.a3, nullptr, st))) return e;
  if ((e = conv_launch<4, 64, 2, 0>(3, w.a3, nullptr, H2, W2, w, pk, w.o5, nullptr, st))) return e;
  // e7 = relu(conv6(relu(conv5(down(o5)))))
  if ((e = pool_launch(w.o5, H2, W2, 64, w.p2, st))) return e;
  if ((e = conv_launch<4, 64, 4, 0>(4, w.p2, nullptr, H4, W4, w, pk, w.a5, nullptr, st))) return e;
  if ((e = conv_launch<8, 128, 2, 0>(5, w.a5, nullptr, H4, W4, w, pk, w.e7, nullptr, st))) return e;
  // e4 = relu(conv8(relu(conv7(cat(up(e7), o5)))))
  if ((e = up_launch(w.e7, H4, W4, 128, w.u7, st))) return e;
  if ((e = conv_launch<12, 128, 1, 0>(6, w.u7, w.o5, H2, W2, w, pk, w.a7, nullptr, st))) return e;
  if ((e = conv_launch<4, 64, 2, 0>(7, w.a7, nullptr, H2, W2, w, pk, w.e4, nullptr, st))) return e;
  // out = min(relu(conv10(relu(conv9(cat(up(e4), o2))))), 1)
  if ((e = up_launch(w.e4, H2, W2, 64, w.u9, st))) return e;
  if ((e = conv_launch<6, 64, 1, 0>(8, w.u9, w.o2, H, W, w, pk, w.a9, nullptr, st))) return e;
  if ((e = conv_launch<2, 32, 1, 1>(9, w.a9, nullptr, H, W, w, pk, nullptr, out, st))) return e;
  return SNISR_OK;
}

}  // extern "C"
