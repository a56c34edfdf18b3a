// Training-side samplers of the reference (CUDA_EXT Sample_sparse / Sample_reflection / Sample_uniform /
// Sample_bg, training/src/Sample_kernel.cu): per ray and tile, the sample depths that feed the
// tile's training batch.  One thread per ray (the reference gives each thread a contiguous run of
// rays; the per-ray results do not depend on the split).
//
//  - sparse: a voxel DDA over the tile's num_voxel^3 occupancy flags measures the total length of
//    the ray inside flagged voxels, then a second DDA gives every run of consecutive flagged voxels
//    ceil((N - 8) * run / total) uniform samples, with 4 samples in the voxel before the first run
//    and 4 in the voxel after the last (Sample_kernel.cu:181-395);
//  - reflection: the same DDA accumulates transmittance through the flagged voxels (density from an
//    overlap-trilinear lookup at each voxel's entry point) and, once it drops to T_th, places N
//    inverse-depth samples between that voxel's exit and `far` (16-137);
//  - uniform: N samples over the tile's chord (436-479); bg: N samples around a background depth,
//    pushed behind the tile's exit (481-542).
// Arithmetic follows the written operation order (a float3 divided by a scalar is a multiply by the
// reciprocal, make_int3 truncates), so depths and sample counts are bit-identical to the FP32 oracle;
// the reflection sampler's exp is exp_cw.  Guards the reference lacks: with rows shorter than 8
// samples the sparse sampler's run counts are clamped at 0 and its 4 leading samples at the row end
// (the reference then writes outside the row), and the overlap lookup
// clamps its corner indices into the grid (the reference reads past it for points on the upper face).
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

struct SmpTile {
  float c[3], ts, hs, vs, mn[3];
  int nv;
};

__host__ __device__ inline SmpTile smp_tile(float cx, float cy, float cz, float tile_size, int num_voxel) {
  SmpTile t;
  t.c[0] = cx; t.c[1] = cy; t.c[2] = cz;
  t.ts = tile_size;
  t.hs = tile_size * 0.5f;
  t.nv = num_voxel;
  t.vs = tile_size / (float)num_voxel;
  for (int a = 0; a < 3; ++a) t.mn[a] = t.c[a] - t.hs;
  return t;
}

struct SmpDda {
  int cur[3], stp[3];
  float tmax[3], tdel[3];
};

__device__ __forceinline__ void smp_dda_init(const SmpTile& T, V3 o, V3 d, float near, SmpDda& D) {
  const float inv = 1.0f / T.vs;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float ol = clampx(get(o, a) + near * get(d, a) - T.mn[a], 0.0f, T.ts - 0.000001f);
    D.cur[a] = min(max((int)(ol * inv), 0), T.nv - 1);
    D.stp[a] = get(d, a) >= 0.0f ? 1 : -1;
    float nb = (float)(D.cur[a] + D.stp[a]) * T.vs;
    if (D.stp[a] < 0) nb += T.vs;
    D.tmax[a] = safe_div(nb - ol, get(d, a));
    D.tdel[a] = safe_div((float)D.stp[a] * T.vs, get(d, a));
  }
}

__device__ __forceinline__ bool smp_dda_step(const SmpTile& T, SmpDda& D) {
  if (D.tmax[0] < D.tmax[1]) {
    if (D.tmax[0] < D.tmax[2]) { D.cur[0] += D.stp[0]; D.tmax[0] += D.tdel[0]; }
    else { D.cur[2] += D.stp[2]; D.tmax[2] += D.tdel[2]; }
  } else {
    if (D.tmax[1] < D.tmax[2]) { D.cur[1] += D.stp[1]; D.tmax[1] += D.tdel[1]; }
    else { D.cur[2] += D.stp[2]; D.tmax[2] += D.tdel[2]; }
  }
  return D.cur[0] >= 0 && D.cur[1] >= 0 && D.cur[2] >= 0 && D.cur[0] < T.nv && D.cur[1] < T.nv && D.cur[2] < T.nv;
}

__device__ __forceinline__ void smp_voxel_bound(const SmpTile& T, V3 o, V3 d, const int (&cur)[3], float& lo,
                                                float& hi) {
  const V3 vc = v3(T.mn[0] + (float)cur[0] * T.vs + 0.5f * T.vs, T.mn[1] + (float)cur[1] * T.vs + 0.5f * T.vs,
                   T.mn[2] + (float)cur[2] * T.vs + 0.5f * T.vs);
  const float h = 0.5f * T.vs;
  ray_aabb(o, d, vc, v3(h, h, h), lo, hi);
}

// uniform_sample_bound_v3 (training/src/include/cuda_utils.h:101-113), at most `room` entries
__device__ __forceinline__ void smp_uniform_v3(float* z, float* dist, float near, float far, int n, int room) {
  const float interval = (far - near) / (float)n;
  for (int i = 0; i < n && i < room; ++i) {
    z[i] = near + (float)i * interval;
    dist[i] = interval;
  }
}

// uniform_sample_bound (cuda_utils.h:77-88)
__device__ __forceinline__ void smp_uniform(float* z, float near, float far, int n) {
  const float interval = (far - near) / (float)(n - 1);
  for (int i = 0; i < n; ++i) z[i] = near + (float)i * interval;
}

__device__ __forceinline__ V3 ray_of(const float* p, int64_t r) { return v3(p[3 * r], p[3 * r + 1], p[3 * r + 2]); }

// The sparse sampler's occupancy test (flag == 1) from a bit per voxel in LDS when the tile has at
// most kSmpBitsMax voxels (NUM_VOXEL <= 80), else from the int16 flags.  Each DDA step waits for its
// flag before it branches, so the walk is a chain of dependent loads: an LDS read in place of an
// L2 read shortens every link.
constexpr int kSmpBitsMax = 64 * 1024 * 8;
constexpr int kSmpBoundCache = 32;  // voxel bounds kept per lane between the sparse sampler's passes
struct SmpFlags {
  const int16_t* g;
  const uint32_t* bits;  // LDS bitmap or nullptr
  __device__ __forceinline__ bool on(int v) const {
    return bits ? ((bits[v >> 5] >> (v & 31)) & 1u) != 0u : g[v] == 1;
  }
};

__global__ __launch_bounds__(64) void k_sample_sparse(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                                                      int64_t n, const int16_t* __restrict__ flags, SmpTile T,
                                                      int nsamp, float* __restrict__ z_vals,
                                                      float* __restrict__ dists) {
  __shared__ uint32_t fbits[kSmpBitsMax / 32];
  // the voxel bounds of a lane's first kSmpBoundCache occupied voxels, from the first pass, so that the
  // second pass does not evaluate the same ray-box slabs (and their divisions) again
  __shared__ float2 bcache[kSmpBoundCache * 64];
  const int64_t nvox = (int64_t)T.nv * T.nv * T.nv;
  const bool lds = nvox <= kSmpBitsMax;
  if (lds) {  // word w = flags 32 w .. 32 w + 31, read as four 16-byte loads when aligned
    const int64_t nw = (nvox + 31) / 32;
    const bool vec = (reinterpret_cast<uintptr_t>(flags) & 15) == 0 && (nvox & 31) == 0;
    for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) {
      uint32_t m = 0;
      if (vec) {
        const uint4* q = reinterpret_cast<const uint4*>(flags + 32 * w);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint4 x = q[u];
          const uint32_t h[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            m |= (uint32_t)((int16_t)(h[t] & 0xffffu) == 1) << (8 * u + 2 * t);
            m |= (uint32_t)((int16_t)(h[t] >> 16) == 1) << (8 * u + 2 * t + 1);
          }
        }
      } else {
        for (int t = 0; t < 32; ++t)
          if (32 * w + t < nvox && flags[32 * w + t] == 1) m |= 1u << t;
      }
      fbits[w] = m;
    }
    __syncthreads();
  }
  const SmpFlags F{flags, lds ? fbits : nullptr};
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const V3 o = ray_of(rays_o, r), d = ray_of(rays_d, r);
  float* z = z_vals + r * nsamp;
  float* ds = dists + r * nsamp;
  float near, far;
  ray_aabb(o, d, v3(T.c[0], T.c[1], T.c[2]), v3(T.hs, T.hs, T.hs), near, far);
  if (near == -1.0f && far == -1.0f) return;
  SmpDda D;
  smp_dda_init(T, o, d, near, D);
  const SmpDda D0 = D;
  float total = 0.0f;
  int m = 0;  // occupied voxels so far
  float2* cache = bcache + threadIdx.x;
  while (true) {
    if (F.on((D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2])) {
      float lo, hi;
      smp_voxel_bound(T, o, d, D.cur, lo, hi);
      total = total + (hi - lo);
      if (m < kSmpBoundCache) cache[64 * m] = make_float2(lo, hi);
      ++m;
    }
    if (!smp_dda_step(T, D)) break;
  }
  if (total == 0.0f) return;
  D = D0;
  m = 0;
  int count = 0;
  bool acc = false, first = false;
  const int bsmp = 4, ins = nsamp - 2 * bsmp;
  while (true) {
    if (F.on((D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2])) {
      float lo, hi;
      if (m < kSmpBoundCache) {
        const float2 b = cache[64 * m];
        lo = b.x;
        hi = b.y;
      } else {
        smp_voxel_bound(T, o, d, D.cur, lo, hi);
      }
      ++m;
      if (lo != -1.0f && hi != -1.0f) {
        if (acc) {
          far = hi;
        } else {
          near = lo;
          far = hi;
          acc = true;
          if (!first) {
            first = true;
            smp_uniform_v3(z + count, ds + count, near - T.vs, near - 0.000001f, bsmp, nsamp - count);
            count = min(count + bsmp, nsamp);
          }
        }
      }
    } else if (acc) {
      int k = (int)ceilf((float)ins * (far - near) / total);
      k = max(min(nsamp - count, k), 0);
      smp_uniform_v3(z + count, ds + count, near, far, k, nsamp - count);
      count = count + k;
      acc = false;
    }
    if (!smp_dda_step(T, D)) break;
  }
  if (acc) {
    int k = (int)ceilf((float)ins * (far - near) / total);
    k = max(min(nsamp - count, k), 0);
    smp_uniform_v3(z + count, ds + count, near, far, k, nsamp - count);
    count = count + k;
  }
  const int tail = min(nsamp - count, bsmp);
  smp_uniform_v3(z + count, ds + count, far + 0.000001f, far + T.vs, tail, nsamp - count);
}

__global__ __launch_bounds__(256) void k_sample_uniform(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                                                        int64_t n, float cx, float cy, float cz, float tile_size,
                                                        int nsamp, float* __restrict__ z_vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float hs = tile_size / 2.0f;
  float near, far;
  ray_aabb(ray_of(rays_o, r), ray_of(rays_d, r), v3(cx, cy, cz), v3(hs, hs, hs), near, far);
  smp_uniform(z_vals + r * nsamp, near, far, nsamp);
}

__global__ __launch_bounds__(256) void k_sample_bg(const float* __restrict__ rays_o, const float* __restrict__ rays_d,
                                                   const float* __restrict__ bg_zval, int64_t n,
                                                   const float* __restrict__ tile_center, float tile_size,
                                                   float sample_range, int nsamp, float* __restrict__ z_vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const float bg = bg_zval[r];
  if (bg == -1.0f) return;
  const float half_range = sample_range / 2.0f;
  const float hs = tile_size / 2.0f;
  float lo, hi;
  ray_aabb(ray_of(rays_o, r), ray_of(rays_d, r), v3(tile_center[0], tile_center[1], tile_center[2]), v3(hs, hs, hs),
           lo, hi);
  float near = bg - half_range, far = bg + half_range;
  if (near <= hi) {
    near = hi + 0.00001f;
    far = near + sample_range;
  }
  smp_uniform(z_vals + r * nsamp, near, far, nsamp);
}

// overlap_interpolation + trilinear_base (training/src/include/interpolation.h:43-134): density only
__device__ __forceinline__ float smp_overlap_sigma(const float4* __restrict__ vox, const SmpTile& T, V3 p) {
  const float inv = 1.0f / T.vs;
  const int M = T.nv;
  float w[3];
  int i0[3], i1[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float corner = T.c[a] - T.ts / 2.0f;
    const float f = (get(p, a) - corner) * inv + 0.5f;
    i0[a] = (int)f;
    i1[a] = i0[a] + 1;
    const float t = f - (float)i0[a];
    w[a] = t < 0.0f ? 0.0f : t;
  }
  auto V = [&](int x, int y, int z) {
    x = min(max(x, 0), M - 1);
    y = min(max(y, 0), M - 1);
    z = min(max(z, 0), M - 1);
    return vox[((int64_t)x * M + y) * M + z].w;
  };
  auto lerp = [](float a, float b, float t) { return a * (1.0f - t) + b * t; };
  const bool ex = i1[0] >= M, ey = i1[1] >= M, ez = i1[2] >= M;
  const float v000 = V(i0[0], i0[1], i0[2]);
  if (ex) {
    if (ey) return ez ? v000 : lerp(v000, V(i0[0], i0[1], i1[2]), w[2]);
    if (ez) return lerp(v000, V(i0[0], i1[1], i0[2]), w[1]);
    return lerp(lerp(v000, V(i0[0], i1[1], i0[2]), w[1]), lerp(V(i0[0], i0[1], i1[2]), V(i0[0], i1[1], i1[2]), w[1]),
                w[2]);
  }
  if (ey) {
    if (ez) return lerp(v000, V(i1[0], i0[1], i0[2]), w[0]);
    return lerp(lerp(v000, V(i1[0], i0[1], i0[2]), w[0]), lerp(V(i0[0], i0[1], i1[2]), V(i1[0], i0[1], i1[2]), w[0]),
                w[2]);
  }
  if (ez)
    return lerp(lerp(v000, V(i1[0], i0[1], i0[2]), w[0]), lerp(V(i0[0], i1[1], i0[2]), V(i1[0], i1[1], i0[2]), w[0]),
                w[1]);
  const float s =
      lerp(lerp(v000, V(i1[0], i0[1], i0[2]), w[0]), lerp(V(i0[0], i1[1], i0[2]), V(i1[0], i1[1], i0[2]), w[0]), w[1]);
  const float t = lerp(lerp(V(i0[0], i0[1], i1[2]), V(i1[0], i0[1], i1[2]), w[0]),
                       lerp(V(i0[0], i1[1], i1[2]), V(i1[0], i1[1], i1[2]), w[0]), w[1]);
  return lerp(s, t, w[2]);
}

__global__ __launch_bounds__(256) void k_sample_reflection(const float* __restrict__ rays_o,
                                                           const float* __restrict__ rays_d, int64_t n,
                                                           const int16_t* __restrict__ flags,
                                                           const float4* __restrict__ voxels, SmpTile T, float far,
                                                           float T_th, int nsamp, float* __restrict__ z_vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const V3 o = ray_of(rays_o, r), d = ray_of(rays_d, r);
  float bl, bh;
  ray_aabb(o, d, v3(T.c[0], T.c[1], T.c[2]), v3(T.hs, T.hs, T.hs), bl, bh);
  if (bl == -1.0f && bh == -1.0f) return;
  SmpDda D;
  smp_dda_init(T, o, d, bl, D);
  float trans = 1.0f, near = -1.0f;
  while (true) {
    if (flags[(D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2]] == 1) {
      float lo, hi;
      smp_voxel_bound(T, o, d, D.cur, lo, hi);
      if (lo != -1.0f && hi != -1.0f) {
        const V3 p = v3(o.x + lo * d.x, o.y + lo * d.y, o.z + lo * d.z);
        const float sigma = smp_overlap_sigma(voxels, T, p);
        const float dist = hi - lo;
        const float alpha = 1.0f - exp_cw(-1.0f * sigma * dist);
        trans = trans * (1.0f - alpha);
        if (trans <= T_th) {
          near = hi;
          break;
        }
      }
    }
    if (!smp_dda_step(T, D)) break;
  }
  if (near == -1.0f) return;
  float* z = z_vals + r * nsamp;
  const float inv_near = 1.0f / near, inv_far = 1.0f / far, inv_bound = inv_far - inv_near;
  const float step = 1.0f / (float)(nsamp - 1);
  for (int i = 0; i < nsamp; ++i) z[i] = 1.0f / (step * (float)i * inv_bound + inv_near);
}

static dim3 ray_grid(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_sample_sparse(const float* rays_o, const float* rays_d, int64_t n, const int16_t* node_flags, float cx,
                        float cy, float cz, float tile_size, int32_t num_voxel, int32_t n_samples, float* z_vals,
                        float* dists, void* stream) {
  SNISR_REQUIRE(n >= 0 && n_samples >= 0, SNISR_ERR_ARG, "sample_sparse: negative size");
  if (n == 0 || n_samples == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && node_flags && z_vals && dists, SNISR_ERR_ARG, "sample_sparse: NULL pointer");
  SNISR_REQUIRE(num_voxel > 0 && tile_size > 0.0f, SNISR_ERR_ARG, "sample_sparse: num_voxel and tile_size must be positive");
  SNISR_REQUIRE(n < (int64_t(1) << 40), SNISR_ERR_SHAPE, "sample_sparse: too many rays");
  hipLaunchKernelGGL(k_sample_sparse, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, (hipStream_t)stream, rays_o, rays_d,
                     n, node_flags,
                     smp_tile(cx, cy, cz, tile_size, num_voxel), n_samples, z_vals, dists);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_sample_uniform(const float* rays_o, const float* rays_d, int64_t n, float cx, float cy, float cz,
                         float tile_size, int32_t n_samples, float* z_vals, void* stream) {
  SNISR_REQUIRE(n >= 0 && n_samples >= 0, SNISR_ERR_ARG, "sample_uniform: negative size");
  if (n == 0 || n_samples == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && z_vals, SNISR_ERR_ARG, "sample_uniform: NULL pointer");
  hipLaunchKernelGGL(k_sample_uniform, ray_grid(n), dim3(256), 0, (hipStream_t)stream, rays_o, rays_d, n, cx, cy, cz,
                     tile_size, n_samples, z_vals);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_sample_bg(const float* rays_o, const float* rays_d, const float* bg_zval, int64_t n, const float* tile_center,
                    float tile_size, float sample_range, int32_t n_samples, float* z_vals, void* stream) {
  SNISR_REQUIRE(n >= 0 && n_samples >= 0, SNISR_ERR_ARG, "sample_bg: negative size");
  if (n == 0 || n_samples == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && bg_zval && tile_center && z_vals, SNISR_ERR_ARG, "sample_bg: NULL pointer");
  hipLaunchKernelGGL(k_sample_bg, ray_grid(n), dim3(256), 0, (hipStream_t)stream, rays_o, rays_d, bg_zval, n,
                     tile_center, tile_size, sample_range, n_samples, z_vals);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_sample_reflection(const float* rays_o, const float* rays_d, int64_t n, const int16_t* node_flags,
                            const float* voxels, float cx, float cy, float cz, float tile_size, float far,
                            int32_t num_voxel, float T_th, int32_t n_samples, float* z_vals, void* stream) {
  SNISR_REQUIRE(n >= 0 && n_samples >= 0, SNISR_ERR_ARG, "sample_reflection: negative size");
  if (n == 0 || n_samples == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && node_flags && voxels && z_vals, SNISR_ERR_ARG, "sample_reflection: NULL pointer");
  SNISR_REQUIRE(num_voxel > 0 && tile_size > 0.0f, SNISR_ERR_ARG,
                "sample_reflection: num_voxel and tile_size must be positive");
  SNISR_REQUIRE(aligned(voxels, 16), SNISR_ERR_ALIGN, "sample_reflection: voxels must be 16-byte aligned");
  hipLaunchKernelGGL(k_sample_reflection, ray_grid(n), dim3(256), 0, (hipStream_t)stream, rays_o, rays_d, n,
                     node_flags, (const float4*)voxels, smp_tile(cx, cy, cz, tile_size, num_voxel), far, T_th,
                     n_samples, z_vals);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
