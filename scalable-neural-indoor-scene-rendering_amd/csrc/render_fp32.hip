// A1 camera rays, A2 tile DDA, and the one-thread-per-ray FP32 renderer (the plain path).
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

// ---------------------------------------------------------------------------- A1
// d = R [(px-cx)*invf, (py-cy)*invf, 1]   (rendering/src/compute_rays.cu:7-44)
__global__ __launch_bounds__(256) void k_compute_rays(const float* __restrict__ intr, const float* __restrict__ c2w,
                                                      int H, int W, float* __restrict__ rays_d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)H * W) return;
  const int py = (int)(i / W), px = (int)(i - (int64_t)py * W);
  const float x = ((float)px - intr[0]) * intr[2];
  const float y = ((float)py - intr[1]) * intr[2];
#pragma unroll
  for (int r = 0; r < 3; ++r) rays_d[i * 3 + r] = c2w[3 * r] * x + c2w[3 * r + 1] * y + c2w[3 * r + 2];
}

// ---------------------------------------------------------------------------- A2
// One wave per 64 rays.  With STAGE, each lane writes its K entries into an LDS row (odd stride,
// so the 64 rows hit distinct banks) and the wave then stores the block's 64*K contiguous entries
// with coalesced writes; without it (K > kTraceStageMaxK) every lane writes its row directly.
constexpr int kTraceStageMaxK = 127;
template <bool STAGE>
__global__ __launch_bounds__(64) void k_trace_tiles(snisr_scene S, const float* __restrict__ origin,
                                                    const float* __restrict__ rays_d, int64_t n, int K,
                                                    int32_t* __restrict__ visited) {
  extern __shared__ int32_t trace_lds[];
  const int64_t base = (int64_t)blockIdx.x * 64;
  const int64_t i = base + threadIdx.x;
  const int KP = K | 1;
  if (i < n) {
    const V3 o = v3(origin[0], origin[1], origin[2]);
    const V3 d = v3(rays_d[3 * i], rays_d[3 * i + 1], rays_d[3 * i + 2]);
    TileWalker w;
    w.init(S, o, d, K);
    if (STAGE) {
      int32_t* row = trace_lds + threadIdx.x * KP;
      for (int j = 0; j < K; ++j) row[j] = w.next();
    } else {
      int32_t* out = visited + i * K;
      for (int j = 0; j < K; ++j) out[j] = w.next();
    }
  }
  if (!STAGE) return;
  __syncthreads();
  const int64_t rows = n - base < 64 ? n - base : 64;
  const int total = (int)rows * K;
  int32_t* out = visited + base * K;
  for (int q = threadIdx.x; q < total; q += 64) {
    const int r = q / K;
    out[q] = trace_lds[r * KP + (q - r * K)];
  }
}

// ---------------------------------------------------------------------------- MLP packing (f32t)
// mlp_params per tile: per layer W[out][in] then b[out]  ->  Wt[in][out4] then b[out4]
__global__ void k_pack_f32t(const float* __restrict__ src, int64_t src_stride, float* __restrict__ dst,
                            int64_t dst_stride, int in0, int width, int layers, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t tile = g / dst_stride;
  int64_t r = g - tile * dst_stride;
  const float* s = src + tile * src_stride;
  int in = in0;
  for (int l = 0; l < layers; ++l) {
    const int out = (l == layers - 1) ? 3 : width;
    const int o4 = (out + 3) & ~3;
    const int64_t wsz = (int64_t)in * o4;
    if (r < wsz) {
      const int i = (int)(r / o4), j = (int)(r - (int64_t)(r / o4) * o4);
      dst[g] = j < out ? s[(int64_t)j * in + i] : 0.0f;
      return;
    }
    r -= wsz;
    if (r < o4) {
      dst[g] = r < out ? s[(int64_t)out * in + r] : 0.0f;
      return;
    }
    r -= o4;
    s += (int64_t)out * in + out;
    in = out;
  }
}

// ---------------------------------------------------------------------------- FP32 render
template <int F, int WD, int L, bool HALF>
__global__ __launch_bounds__(64) void k_render_fp32(snisr_scene S, snisr_march_params P, snisr_rays R,
                                                    snisr_render_out O, int packet_w) {
  constexpr int IN = F + 15;
  constexpr int XD = IN > WD ? IN : WD;
  __shared__ float lds[XD * 64];
  float* xs = lds + threadIdx.x;
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (tid >= R.n) return;
  const int64_t i = ray_of_thread(tid, R.n, packet_w);
  const float* og = R.origins + i * R.origin_stride;
  const V3 origin = v3(og[0], og[1], og[2]);
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  const float t_hit = R.t_hit ? R.t_hit[i] : __builtin_inff();
  float enc[15];
  view_encoding(d, enc);
  constexpr int64_t wstride = f32t_stride(IN, WD, L);
  float col[3] = {0.0f, 0.0f, 0.0f};
  int32_t* bricks = O.bricks ? O.bricks + i * O.brick_cap : nullptr;
  MarchResult res;
  march_ray(S, P, origin, d, t_hit, R.visited ? R.visited + i * P.max_tiles : nullptr, bricks, O.brick_cap,
            [&](int tile, const V3&, const Corners& c, float w, float, float, float) {
              gather_features<F, HALF>(S.feats, c, xs);
#pragma unroll
              for (int k = 0; k < 15; ++k) xs[(F + k) * 64] = enc[k];
              float rgb[3];
              const int t0 = __builtin_amdgcn_readfirstlane(tile);
              if (__all(tile == t0)) mlp_f32<IN, WD, L>(S.mlp_f32t + (int64_t)t0 * wstride, xs, rgb);
              else mlp_f32<IN, WD, L>(S.mlp_f32t + (int64_t)tile * wstride, xs, rgb);
#pragma unroll
              for (int ch = 0; ch < 3; ++ch) col[ch] = col[ch] + w * rgb[ch];
            },
            res);
  if (bricks)
    for (int k = res.n_bricks; k < O.brick_cap; ++k) bricks[k] = -1;
  if (O.rgb) {
    O.rgb[3 * i] = col[0];
    O.rgb[3 * i + 1] = col[1];
    O.rgb[3 * i + 2] = col[2];
  }
  if (O.T) O.T[i] = res.T;
  if (O.n_samples) O.n_samples[i] = res.n_samples;
  if (O.depth) O.depth[i] = res.depth;
}

// sigma-only march (no MLP): counts, T, depth, bricks; used when rgb is not requested.
__global__ __launch_bounds__(64) void k_march_sigma(snisr_scene S, snisr_march_params P, snisr_rays R,
                                                    snisr_render_out O, int packet_w) {
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (tid >= R.n) return;
  const int64_t i = ray_of_thread(tid, R.n, packet_w);
  const float* og = R.origins + i * R.origin_stride;
  const V3 origin = v3(og[0], og[1], og[2]);
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  const float t_hit = R.t_hit ? R.t_hit[i] : __builtin_inff();
  int32_t* bricks = O.bricks ? O.bricks + i * O.brick_cap : nullptr;
  MarchResult res;
  march_ray(S, P, origin, d, t_hit, R.visited ? R.visited + i * P.max_tiles : nullptr, bricks, O.brick_cap,
            [&](int, const V3&, const Corners&, float, float, float, float) {}, res);
  if (bricks)
    for (int k = res.n_bricks; k < O.brick_cap; ++k) bricks[k] = -1;
  if (O.T) O.T[i] = res.T;
  if (O.n_samples) O.n_samples[i] = res.n_samples;
  if (O.depth) O.depth[i] = res.depth;
}

int packet_width(const snisr_rays& R) {
  if (R.image_width <= 0 || R.image_width % 8 != 0) return 0;
  if (R.n % R.image_width != 0) return 0;
  const int64_t rows = R.n / R.image_width;
  return rows % 8 == 0 ? R.image_width : 0;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_compute_rays(const float* intrinsic, const float* c2w, int32_t H, int32_t W, float* rays_d,
                       void* stream) {
  SNISR_REQUIRE(intrinsic && c2w && rays_d, SNISR_ERR_ARG, "compute_rays: NULL pointer");
  SNISR_REQUIRE(H >= 0 && W >= 0, SNISR_ERR_ARG, "compute_rays: negative size");
  const int64_t n = (int64_t)H * W;
  if (n == 0) return SNISR_OK;
  hipLaunchKernelGGL(k_compute_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     intrinsic, c2w, H, W, rays_d);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_trace_tiles(const snisr_scene* scene, const float* origin, const float* rays_d, int64_t n, int32_t K,
                      int32_t* visited, void* stream) {
  // only the tile grid is read: tile_shape, scene_min, tile_size, tile_index_map
  SNISR_REQUIRE(scene, SNISR_ERR_ARG, "scene descriptor is NULL");
  SNISR_REQUIRE(scene->tile_shape[0] > 0 && scene->tile_shape[1] > 0 && scene->tile_shape[2] > 0, SNISR_ERR_ARG,
                "tile_shape must be positive");
  SNISR_REQUIRE(scene->tile_size > 0.0f, SNISR_ERR_ARG, "tile_size must be positive");
  SNISR_REQUIRE(scene->tile_index_map, SNISR_ERR_ARG, "tile_index_map is NULL");
  SNISR_REQUIRE(origin && (n == 0 || (rays_d && visited)), SNISR_ERR_ARG, "trace_tiles: NULL pointer");
  SNISR_REQUIRE(n >= 0 && K >= 0, SNISR_ERR_ARG, "trace_tiles: negative size");
  if (n == 0 || K == 0) return SNISR_OK;
  const dim3 grid((unsigned)((n + 63) / 64));
  if (K <= kTraceStageMaxK)
    hipLaunchKernelGGL(k_trace_tiles<true>, grid, dim3(64), (size_t)64 * (K | 1) * sizeof(int32_t),
                       (hipStream_t)stream, *scene, origin, rays_d, n, K, visited);
  else
    hipLaunchKernelGGL(k_trace_tiles<false>, grid, dim3(64), 0, (hipStream_t)stream, *scene, origin, rays_d, n, K,
                       visited);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_render_fp32(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                      const snisr_render_out* out, void* stream) {
  SNISR_REQUIRE(out, SNISR_ERR_ARG, "render_fp32: out is NULL");
  const bool want_rgb = out->rgb != nullptr;
  if (int e = check_scene(scene, want_rgb)) return e;
  if (int e = check_params(params)) return e;
  if (int e = check_rays(rays, params)) return e;
  SNISR_REQUIRE(!out->bricks || out->brick_cap > 0, SNISR_ERR_ARG, "render_fp32: bricks needs brick_cap > 0");
  if (rays->n == 0) return SNISR_OK;
  const int pw = packet_width(*rays);
  const dim3 grid((unsigned)((rays->n + 63) / 64));
  hipStream_t st = (hipStream_t)stream;
  if (!want_rgb) {
    SNISR_HIP(stage_event(out, 0, st));
    hipLaunchKernelGGL(k_march_sigma, grid, dim3(64), 0, st, *scene, *params, *rays, *out, pw);
    SNISR_LAUNCHED();
    SNISR_HIP(stage_event(out, 1, st));
    return SNISR_OK;
  }
  SNISR_REQUIRE(scene->mlp_f32t, SNISR_ERR_ARG, "render_fp32: scene.mlp_f32t is NULL (snisr_pack_mlp(SNISR_F32))");
  SNISR_REQUIRE(aligned(scene->mlp_f32t, 16), SNISR_ERR_ALIGN, "render_fp32: mlp_f32t must be 16-byte aligned");
  const int F = scene->feat_dim, WD = scene->mlp_width, L = scene->mlp_layers;
  const bool half = scene->feat_dtype == SNISR_F16;
#define SNISR_FP32_CASE(F_, W_, L_, H_)                                                                   \
  if (F == F_ && WD == W_ && L == L_ && half == H_) {                                                     \
    SNISR_HIP(stage_event(out, 0, st));                                                                   \
    hipLaunchKernelGGL((k_render_fp32<F_, W_, L_, H_>), grid, dim3(64), 0, st, *scene, *params, *rays, *out, pw); \
    SNISR_LAUNCHED();                                                                                     \
    SNISR_HIP(stage_event(out, 1, st));                                                                   \
    return SNISR_OK;                                                                                      \
  }
  SNISR_FP32_CASE(16, 64, 3, false)
  SNISR_FP32_CASE(16, 64, 3, true)
  SNISR_FP32_CASE(32, 128, 4, true)
  SNISR_FP32_CASE(32, 128, 4, false)
  SNISR_FP32_CASE(16, 32, 3, false)
#undef SNISR_FP32_CASE
  set_error("render_fp32: unsupported shape F=%d width=%d layers=%d feat_dtype=%d", F, WD, L, scene->feat_dtype);
  return SNISR_ERR_SHAPE;
}

int64_t snisr_pack_mlp_size(const snisr_scene* scene, int32_t dtype) {
  if (!scene) return -1;
  if (dtype == SNISR_F32)
    return (int64_t)scene->n_tiles * mlp_f32t_stride(scene->mlp_in, scene->mlp_width, scene->mlp_layers) * 4;
  return mfma_pack_bytes(scene, dtype);
}

int snisr_pack_mlp(const snisr_scene* scene, int32_t dtype, void* packed, void* stream) {
  SNISR_REQUIRE(scene && packed && scene->mlp_params, SNISR_ERR_ARG, "pack_mlp: NULL pointer");
  SNISR_REQUIRE(scene->n_tiles > 0 && scene->mlp_layers >= 2, SNISR_ERR_ARG, "pack_mlp: bad MLP shape");
  if (dtype == SNISR_F32) {
    const int64_t ds = mlp_f32t_stride(scene->mlp_in, scene->mlp_width, scene->mlp_layers);
    const int64_t total = ds * scene->n_tiles;
    hipLaunchKernelGGL(k_pack_f32t, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       scene->mlp_params, scene->mlp_stride, (float*)packed, ds, scene->mlp_in, scene->mlp_width,
                       scene->mlp_layers, total);
    SNISR_LAUNCHED();
    return SNISR_OK;
  }
  return mfma_pack(scene, dtype, packed, stream);
}

}  // extern "C"
