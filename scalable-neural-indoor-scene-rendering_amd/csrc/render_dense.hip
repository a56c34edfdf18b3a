// A10: the reference's plain FP32 marcher over dense per-tile grids, render_diffuse_kernel
// (training/src/cuda_func.cu:1687-1802).  A group's tiles each hold a dilated grid of nv^3 cells
// (the training tile's NUM_VOXEL + 2, one voxel of overlap on every side) of float4 (r, g, b, sigma)
// and an int16 cell flag; only cells flagged 1 are sampled.  Per ray, over its visited tiles in
// order: the tile's ray-AABB chord is marched from its entry in steps of sample_step; a sample in a
// flagged cell is interpolated with the OVERLAP rule (interpolation.h:120-134: trilinear at the cell
// coordinate + 0.5 on the dilated grid, degrading to lower-order interpolation at the upper grid
// edge, interpolation.h:41-116) and composited; the depth is the first sample reached with
// T <= 0.5; a tile outside the group (sparse_to_group -1) ends the ray as invalid.  One thread per
// ray, FP32, the reference's operation order (a float3 divided by a scalar is a multiply by the
// FP32 reciprocal, cutil_math.h:396-400; make_int3 truncates).
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

struct F4 {
  float x, y, z, w;
};

__device__ __forceinline__ F4 lerp4(const F4& a, const F4& b, float t) {   // interpolate1D
  const float u = 1.0f - t;
  return F4{a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t, a.w * u + b.w * t};
}

__device__ __forceinline__ F4 ld4(const float4* __restrict__ v, int64_t i) {
  const float4 q = v[i];
  return F4{q.x, q.y, q.z, q.w};
}

// trilinear_base (interpolation.h:41-116) on an M^3 grid with the upper-edge degradation.  Corner
// indices are clamped into the grid (the reference reads whatever lies there); for points inside
// the tile every index is in range and the clamp changes nothing.
__device__ __forceinline__ F4 dense_trilinear(const float4* __restrict__ vox, int M, float px, float py, float pz) {
  const int x0 = (int)px, y0 = (int)py, z0 = (int)pz;
  const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  float x = px - (float)x0, y = py - (float)y0, z = pz - (float)z0;
  if (x < 0.0f) x = 0.0f;
  if (y < 0.0f) y = 0.0f;
  if (z < 0.0f) z = 0.0f;
  const int lo = 0, hi = M - 1;
  const int cx0 = min(max(x0, lo), hi), cy0 = min(max(y0, lo), hi), cz0 = min(max(z0, lo), hi);
  const int cx1 = min(max(x1, lo), hi), cy1 = min(max(y1, lo), hi), cz1 = min(max(z1, lo), hi);
  auto at = [&](int a, int b, int c) { return ld4(vox, ((int64_t)a * M + b) * M + c); };
  const bool ex = x1 >= M, ey = y1 >= M, ez = z1 >= M;
  if (ex) {
    if (ey) return ez ? at(cx0, cy0, cz0) : lerp4(at(cx0, cy0, cz0), at(cx0, cy0, cz1), z);
    if (ez) return lerp4(at(cx0, cy0, cz0), at(cx0, cy1, cz0), y);
    return lerp4(lerp4(at(cx0, cy0, cz0), at(cx0, cy1, cz0), y), lerp4(at(cx0, cy0, cz1), at(cx0, cy1, cz1), y), z);
  }
  if (ey) {
    if (ez) return lerp4(at(cx0, cy0, cz0), at(cx1, cy0, cz0), x);
    return lerp4(lerp4(at(cx0, cy0, cz0), at(cx1, cy0, cz0), x), lerp4(at(cx0, cy0, cz1), at(cx1, cy0, cz1), x), z);
  }
  if (ez)
    return lerp4(lerp4(at(cx0, cy0, cz0), at(cx1, cy0, cz0), x), lerp4(at(cx0, cy1, cz0), at(cx1, cy1, cz0), x), y);
  const F4 s = lerp4(lerp4(at(cx0, cy0, cz0), at(cx1, cy0, cz0), x), lerp4(at(cx0, cy1, cz0), at(cx1, cy1, cz0), x), y);
  const F4 t = lerp4(lerp4(at(cx0, cy0, cz1), at(cx1, cy0, cz1), x), lerp4(at(cx0, cy1, cz1), at(cx1, cy1, cz1), x), y);
  return lerp4(s, t, z);
}

__global__ __launch_bounds__(256) void k_render_dense(snisr_dense_grid G, const float* __restrict__ rays_o,
                                                      const float* __restrict__ rays_d,
                                                      const int32_t* __restrict__ visited, int64_t n, int K,
                                                      float* __restrict__ diffuse, float* __restrict__ depth,
                                                      int32_t* __restrict__ valid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float early = 0.00001f;
  const int M = G.num_voxel;
  const int64_t cells = (int64_t)M * M * M;
  const V3 o = v3(rays_o[3 * i], rays_o[3 * i + 1], rays_o[3 * i + 2]);
  const V3 d = v3(rays_d[3 * i], rays_d[3 * i + 1], rays_d[3 * i + 2]);
  const int32_t* vis = visited + i * K;
  const float ts = G.tile_size, vs = G.voxel_size, step = G.sample_step;
  const float half = ts / 2.0f;
  const float inv_vs = 1.0f / vs;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, T = 1.0f;
  int first_tile = -1;
  float dep = 0.0f;
  for (int j = 0; j < K; ++j) {
    const int tile = vis[j];
    if (tile == -1) break;
    if (T < early) break;
    const int slot = G.sparse_to_group[tile];
    if (slot == -1) {
      first_tile = -1;
      break;
    }
    const float4* vox = reinterpret_cast<const float4*>(G.voxels) + (int64_t)slot * cells;
    const int16_t* nodes = G.nodes + (int64_t)slot * cells;
    const V3 tc = v3(G.centers[3 * tile], G.centers[3 * tile + 1], G.centers[3 * tile + 2]);
    float tl, th;
    ray_aabb(o, d, tc, v3(half, half, half), tl, th);
    const V3 dc = v3(tc.x - half - vs, tc.y - half - vs, tc.z - half - vs);   // dilated grid corner
    const V3 corner = v3(tc.x - half, tc.y - half, tc.z - half);
    float z = tl;
    while (z < th) {
      const V3 p = v3(o.x + z * d.x, o.y + z * d.y, o.z + z * d.z);
      const int px = (int)((p.x - dc.x) * inv_vs), py = (int)((p.y - dc.y) * inv_vs), pz = (int)((p.z - dc.z) * inv_vs);
      const bool in = px >= 0 && py >= 0 && pz >= 0 && px < M && py < M && pz < M;
      if (!in || nodes[((int64_t)px * M + py) * M + pz] != 1) {
        z = z + step;
        continue;
      }
      if (first_tile == -1 && T <= 0.5f) {
        dep = z;
        first_tile = tile;
      }
      const F4 v = dense_trilinear(vox, M, (p.x - corner.x) * inv_vs + 0.5f, (p.y - corner.y) * inv_vs + 0.5f,
                                   (p.z - corner.z) * inv_vs + 0.5f);
      const float interval = fminx(step, th - z);
      const float alpha = 1.0f - exp_cw(-1.0f * v.w * interval);
      const float wgt = alpha * T;
      c0 = c0 + wgt * v.x;
      c1 = c1 + wgt * v.y;
      c2 = c2 + wgt * v.z;
      T = T * (1.0f - alpha);
      if (T < early) break;
      z = z + step;
    }
  }
  if (first_tile != -1) {
    valid[i] = 1;
    depth[i] = dep;
  } else {
    valid[i] = 0;
    depth[i] = -1.0f;
  }
  diffuse[3 * i] = c0;
  diffuse[3 * i + 1] = c1;
  diffuse[3 * i + 2] = c2;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_render_dense(const snisr_dense_grid* grid, const float* rays_o, const float* rays_d, const int32_t* visited,
                       int64_t n, int32_t K, float* diffuse, float* depth, int32_t* valid, void* stream) {
  SNISR_REQUIRE(grid, SNISR_ERR_ARG, "render_dense: grid descriptor is NULL");
  SNISR_REQUIRE(n >= 0 && K >= 0, SNISR_ERR_ARG, "render_dense: negative size");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && diffuse && depth && valid && (K == 0 || visited), SNISR_ERR_ARG,
                "render_dense: NULL pointer");
  SNISR_REQUIRE(grid->voxels && grid->nodes && grid->sparse_to_group && grid->centers, SNISR_ERR_ARG,
                "render_dense: grid tables must be non-NULL");
  SNISR_REQUIRE(grid->num_voxel >= 2 && grid->num_voxel <= 1024, SNISR_ERR_ARG, "render_dense: num_voxel %d",
                grid->num_voxel);
  SNISR_REQUIRE(grid->sample_step > 0.0f && grid->voxel_size > 0.0f && grid->tile_size > 0.0f, SNISR_ERR_ARG,
                "render_dense: sample_step, voxel_size and tile_size must be positive");
  SNISR_REQUIRE(aligned(grid->voxels, 16), SNISR_ERR_ALIGN, "render_dense: voxels must be 16-byte aligned");
  SNISR_REQUIRE(n < (int64_t(1) << 40), SNISR_ERR_SHAPE, "render_dense: too many rays");
  hipLaunchKernelGGL(k_render_dense, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *grid,
                     rays_o, rays_d, visited, n, K, diffuse, depth, valid);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
