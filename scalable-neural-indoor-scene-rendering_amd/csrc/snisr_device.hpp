// Device building blocks of the tiled sparse-brick ray marcher (gfx950).
//
// Compiled with -ffp-contract=off: every float expression is evaluated in the written order
// with separately rounded operations, FMAs only where __builtin_fmaf is written.  Together
// with IEEE division/sqrt (hipcc's default for f32) and the exp below, the traversal, the
// sample positions, sigma, alpha and T are bit-identical to a host build of the same
// operations, which is what makes visit lists and sample counts reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "snisr.h"

namespace snisr {

constexpr float kSafeDivBig = 100000000.0f;  // safe_divide's 1e8 (cutil_math.h:867-869)
#ifndef SNISR_TILE_QUORUM
#define SNISR_TILE_QUORUM 4
#endif
// the flat march runs a tile step once 1/kTileQuorum of a wavefront's live lanes wait for it
constexpr int kTileQuorum = SNISR_TILE_QUORUM;

// Floats per tile of the FP32 "transposed" MLP packing: per layer Wt[in][out4] then b[out4].
__host__ __device__ constexpr int64_t f32t_stride(int in, int width, int layers) {
  int64_t s = 0;
  for (int l = 0; l < layers; ++l) {
    const int o = (l == layers - 1) ? 3 : width;
    const int o4 = (o + 3) & ~3;
    s += (int64_t)in * o4 + o4;
    in = o;
  }
  return s;
}

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ float get(const V3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ float fminx(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmaxx(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float clampx(float v, float lo, float hi) { return fmaxx(lo, fminx(v, hi)); }
__device__ __forceinline__ float safe_div(float a, float b) { return b != 0.0f ? a / b : kSafeDivBig; }

// e^x for x in [ln(FLT_MIN), ln(FLT_MAX)] (the reduction and polynomial of exp_cw)
__device__ __forceinline__ float exp_cw_in_range(float x) {
  const float k = __builtin_rintf(x * 0x1.715476p+0f);
  float r = __builtin_fmaf(-k, 0x1.62e400p-1f, x);
  r = __builtin_fmaf(-k, 0x1.7f7d1cp-20f, r);
  float p = 0x1.a01a02p-13f;
  p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
  p = __builtin_fmaf(p, r, 0x1.111112p-7f);
  p = __builtin_fmaf(p, r, 0x1.555556p-5f);
  p = __builtin_fmaf(p, r, 0x1.555556p-3f);
  p = __builtin_fmaf(p, r, 0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}

// e^x by Cody-Waite reduction (x = k ln2 + r with ln2 split into a 16-bit head and a tail)
// and the degree-7 Taylor polynomial of e^r evaluated by Horner's rule with FMAs, scaled by
// 2^k.  Inputs below ln(FLT_MIN) give 0, above ln(FLT_MAX) give +inf, NaN gives itself.  Branch
// free: the polynomial runs on the input clamped to that range and the ends are selected after.
__device__ __forceinline__ float exp_cw(float x) {
  const float lo = -87.33654475f, hi = 88.72283905f;
  const float xc = x < lo ? lo : (x > hi ? hi : x);
  const float e = exp_cw_in_range(xc);
  return !(x == x) ? x : (x < lo ? 0.0f : (x > hi ? __builtin_inff() : e));
}

__device__ __forceinline__ float sigmoid_cw(float z) { return 1.0f / (1.0f + exp_cw(-z)); }

__device__ __forceinline__ V3 normalize_dir(V3 d) {
  const float inv = 1.0f / __builtin_sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  return v3(inv * d.x, inv * d.y, inv * d.z);
}

// Slab test of RayAABBIntersection (rendering/src/include/cuda_utils.h:553-598); (-1,-1) on a miss.
__device__ __forceinline__ void ray_aabb(V3 o, V3 d, V3 c, V3 h, float& tlo, float& thi) {
  float lo = 0.0f, hi = 100000.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float inv = safe_div(1.0f, get(d, a));
    float l = (get(c, a) - get(h, a) - get(o, a)) * inv;
    float u = (get(c, a) + get(h, a) - get(o, a)) * inv;
    if (u < l) { const float t = l; l = u; u = t; }
    if (u < lo || l > hi) { tlo = thi = -1.0f; return; }
    lo = l > lo ? l : lo;
    hi = u < hi ? u : hi;
    if (lo > hi) { tlo = thi = -1.0f; return; }
  }
  tlo = lo;
  thi = hi;
}

// Amanatides-Woo walk over the tile grid yielding non-empty tile ids lazily, capped at K
// (voxel_traversal, rendering/src/include/cuda_utils.h:158-238, with the scene-entry clamp of
// tracing_tiles_kernel, rendering/src/raytile_tracing.cu:30-47).
struct TileWalker {
  int cur[3], step[3], M[3];
  float tmax[3], tdel[3];
  int emitted, K;
  bool alive, first;
  const int32_t* map;

  __device__ __forceinline__ void init(const snisr_scene& S, V3 origin, V3 d, int Kcap) {
    K = Kcap;
    emitted = 0;
    first = true;
    map = S.tile_index_map;
    M[0] = S.tile_shape[0]; M[1] = S.tile_shape[1]; M[2] = S.tile_shape[2];
    const float ts = S.tile_size;
    const V3 size = v3((float)M[0] * ts, (float)M[1] * ts, (float)M[2] * ts);
    V3 o = v3(origin.x - S.scene_min[0], origin.y - S.scene_min[1], origin.z - S.scene_min[2]);
    if (o.x < 0.0f || o.x >= size.x || o.y < 0.0f || o.y >= size.y || o.z < 0.0f || o.z >= size.z) {
      const V3 half = v3(size.x / 2.0f, size.y / 2.0f, size.z / 2.0f);
      const V3 ctr = v3(S.scene_min[0] + half.x, S.scene_min[1] + half.y, S.scene_min[2] + half.z);
      float tl, th;
      ray_aabb(origin, d, ctr, half, tl, th);
      o = v3(o.x + tl * d.x, o.y + tl * d.y, o.z + tl * d.z);
      o = v3(clampx(o.x, 0.0f, size.x - 0.000001f), clampx(o.y, 0.0f, size.y - 0.000001f),
             clampx(o.z, 0.0f, size.z - 0.000001f));
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float oa = get(o, a), da = get(d, a);
      cur[a] = (int)__builtin_floorf(oa / ts);
      step[a] = da >= 0.0f ? 1 : -1;
      float nb = (float)(cur[a] + step[a]) * ts;
      if (step[a] < 0) nb += ts;
      tmax[a] = da != 0.0f ? (nb - oa) / da : kSafeDivBig;
      tdel[a] = da != 0.0f ? (ts / da) * (float)step[a] : kSafeDivBig;
    }
    alive = inside() && K > 0;
  }
  __device__ __forceinline__ bool inside() const {
    return cur[0] >= 0 && cur[1] >= 0 && cur[2] >= 0 && cur[0] < M[0] && cur[1] < M[1] && cur[2] < M[2];
  }
  __device__ __forceinline__ void advance() {
    int a;
    if (tmax[0] < tmax[1]) a = (tmax[0] < tmax[2]) ? 0 : 2;
    else a = (tmax[1] < tmax[2]) ? 1 : 2;
    // select by index without dynamic register indexing
    if (a == 0) { cur[0] += step[0]; tmax[0] += tdel[0]; }
    else if (a == 1) { cur[1] += step[1]; tmax[1] += tdel[1]; }
    else { cur[2] += step[2]; tmax[2] += tdel[2]; }
  }
  // next non-empty tile id, or -1 once K ids were produced or the walk left the grid
  __device__ __forceinline__ int next() {
    while (alive && emitted < K) {
      if (!first) {
        advance();
        if (!inside()) { alive = false; break; }
      }
      first = false;
      const int id = map[(cur[0] * M[1] + cur[1]) * M[2] + cur[2]];
      if (id != -1) { ++emitted; return id; }
    }
    alive = false;
    return -1;
  }
};

// The eight trilinear corners of continuous voxel coordinate f (= p/voxel - 0.5) in one tile:
// brick-resolved slots (-1 = empty/outside, reads as zero) and the fractional weights.
struct Corners {
  int32_t slot[8];  // order x0y0z0, x1y0z0, x0y1z0, x1y1z0, x0y0z1, x1y0z1, x0y1z1, x1y1z1
  float wx, wy, wz;
};

// Branch-free: all eight brick-table loads are issued together (a corner outside the tile reads
// the tile's first entry, always present since n >= 8, and is then discarded), so one memory
// round trip resolves the sample instead of eight dependent ones.  The brick index and the
// in-brick offset of a corner are sums of per-axis terms, formed once per axis in 32-bit integer
// arithmetic; the indices are non-negative, so the loads use zero-extended offsets.
__device__ __forceinline__ void find_corners(const int32_t* __restrict__ bmap, int n, V3 f, Corners& c) {
  const float fx = __builtin_floorf(f.x), fy = __builtin_floorf(f.y), fz = __builtin_floorf(f.z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  c.wx = f.x - fx;
  c.wy = f.y - fy;
  c.wz = f.z - fz;
  const int nb = n >> 3;
  const int i0[3] = {ix, iy, iz};
  const int scale[3] = {nb * nb, nb, 1};
  const int lscale[3] = {64, 8, 1};
  bool in[3][2];
  uint32_t bt[3][2], lt[3][2];  // brick-index and in-brick-offset terms per axis and corner side
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int x = i0[a] + s;
      in[a][s] = (unsigned)x < (unsigned)n;
      bt[a][s] = (uint32_t)((x >> 3) * scale[a]);
      lt[a][s] = (uint32_t)((x & 7) * lscale[a]);
    }
  bool inside[8];
  int32_t b[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = k & 1, sy = (k >> 1) & 1, sz = k >> 2;
    inside[k] = in[0][sx] && in[1][sy] && in[2][sz];
    const uint32_t idx = inside[k] ? bt[0][sx] + bt[1][sy] + bt[2][sz] : 0u;
    b[k] = bmap[idx];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int sx = k & 1, sy = (k >> 1) & 1, sz = k >> 2;
    const int32_t local = (int32_t)(lt[0][sx] + lt[1][sy] + lt[2][sz]);
    c.slot[k] = inside[k] && b[k] >= 0 ? b[k] * 512 + local : -1;
  }
}

__device__ __forceinline__ float lerpx(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// interpolate3D nesting (rendering/src/include/interpolation.h:13-38): x, then y, then z.
__device__ __forceinline__ float tri_combine(const float v[8], float wx, float wy, float wz) {
  const float s = lerpx(lerpx(v[0], v[1], wx), lerpx(v[2], v[3], wx), wy);
  const float t = lerpx(lerpx(v[4], v[5], wx), lerpx(v[6], v[7], wx), wy);
  return lerpx(s, t, wz);
}

// All ones for an occupied corner, zero for an empty one.  The empty asm statement hides the
// select from the optimizer: ANDed into the raw bits of an unconditional load, it zeroes the empty
// corners without the load being moved into a branch of its own (where every corner would wait
// for its own memory round trip).
__device__ __forceinline__ uint32_t corner_mask(bool occ) {
  uint32_t m = occ ? 0xffffffffu : 0u;
  asm volatile("" : "+v"(m));
  return m;
}

// Only called for samples inside an occupied brick, so the scene holds a brick and slot 0 (read
// for the empty corners, then discarded) exists; the eight loads are unconditional.
__device__ __forceinline__ float tri_sigma(const float* __restrict__ sigma, const Corners& c) {
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float s = sigma[c.slot[k] >= 0 ? (uint32_t)c.slot[k] : 0u];
    v[k] = c.slot[k] >= 0 ? s : 0.0f;
  }
  return tri_combine(v, c.wx, c.wy, c.wz);
}

struct MarchResult {
  float T;
  int n_samples;
  float depth;
  int n_bricks;
};

// Marches one ray (unit direction d) through its non-empty tiles.  Per tile: RayAABB entry,
// local origin clamped to [0, tile - 1e-5] (raydiffuse_render_octree_fp16.cu:202-213); the
// brick DDA of block_traversal_rendering (raydiffuse_render_octree_fp16.cu:39-138) over 8^3-voxel
// bricks; in an occupied brick samples z_k = fma(k, step, near) for z_k < far, interval
// min(step, far - z_k); alpha = 1 - exp(-sigma*interval), T *= 1 - alpha
// (rendering/src/include/cuda_utils.h:13-20); stop once T < early_stop or at t >= t_hit.
// `on_sample(tile, f, corners, w, sigma, interval, T_after)` receives every composited sample:
// f = voxel coordinate (p/voxel - 0.5 in the tile), its trilinear corners, w = alpha*T, the
// clamped density, the interval and the transmittance after the sample.
template <class OnSample>
__device__ __forceinline__ void march_ray(const snisr_scene& S, const snisr_march_params& P, V3 origin, V3 d,
                                          float t_hit, const int32_t* __restrict__ vis, int32_t* bricks,
                                          int brick_cap, OnSample&& on_sample, MarchResult& res) {
  TileWalker walker;
  if (!vis) walker.init(S, origin, d, P.max_tiles);
  float T = 1.0f;
  int nsamp = 0, nbr = 0;
  float depth = -1.0f;
  bool first_hit = false, done = false;
  const float ts = S.tile_size;
  const float step = P.sample_step;
  const float hs = ts / 2.0f;
  for (int vi = 0; vi < P.max_tiles && !done; ++vi) {
    const int tile = vis ? vis[vi] : walker.next();
    if (tile == -1) break;
    if (T < P.early_stop) break;
    const V3 tc = v3(S.tile_centers[3 * tile], S.tile_centers[3 * tile + 1], S.tile_centers[3 * tile + 2]);
    float tl, th;
    ray_aabb(origin, d, tc, v3(hs, hs, hs), tl, th);
    if (tl == -1.0f) continue;
    const V3 corner = v3(tc.x - hs, tc.y - hs, tc.z - hs);
    const float hi = ts - 0.00001f;
    V3 o = v3(origin.x + tl * d.x - corner.x, origin.y + tl * d.y - corner.y, origin.z + tl * d.z - corner.z);
    o = v3(clampx(o.x, 0.0f, hi), clampx(o.y, 0.0f, hi), clampx(o.z, 0.0f, hi));
    const int n = S.tile_res[tile];
    const int nb = n >> 3;
    const float bs = ts / (float)nb;
    const float inv_vs = 1.0f / (ts / (float)n);
    const int32_t* __restrict__ bmap = S.brick_map + S.tile_brick_base[tile];
    int cur[3], stp[3];
    float tmax[3], tdel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float oa = get(o, a), da = get(d, a);
      cur[a] = (int)(oa / bs);
      stp[a] = da >= 0.0f ? 1 : -1;
      float nbd = (float)(cur[a] + stp[a]) * bs;
      if (stp[a] < 0) nbd += bs;
      tmax[a] = safe_div(nbd - oa, da);
      tdel[a] = safe_div(bs * (float)stp[a], da);
    }
    float near_ = 0.0f, far_ = fminx(fminx(tmax[0], tmax[1]), tmax[2]);
    while (true) {
      if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= nb || cur[1] >= nb || cur[2] >= nb) break;
      const int b = bmap[(cur[0] * nb + cur[1]) * nb + cur[2]];
      if (b != -1) {
        if (bricks && nbr < brick_cap) bricks[nbr] = b;
        ++nbr;
        for (int k = 0;; ++k) {
          const float z = __builtin_fmaf((float)k, step, near_);
          if (!(z < far_)) break;
          const float tg = tl + z;
          if (tg >= t_hit) { done = true; break; }
          const V3 p = v3(o.x + z * d.x, o.y + z * d.y, o.z + z * d.z);
          const V3 f = v3(p.x * inv_vs - 0.5f, p.y * inv_vs - 0.5f, p.z * inv_vs - 0.5f);
          Corners c;
          find_corners(bmap, n, f, c);
          float sig = tri_sigma(S.sigma, c);
          sig = sig > 0.0f ? sig : 0.0f;
          if (!first_hit && T < P.trans_th) { first_hit = true; depth = tg; }
          const float interval = fminx(step, far_ - z);
          const float alpha = 1.0f - exp_cw(-(sig * interval));
          const float w = alpha * T;
          const float T_next = T * (1.0f - alpha);
          on_sample(tile, f, c, w, sig, interval, T_next);
          T = T_next;
          ++nsamp;
          if (T < P.early_stop) { done = true; break; }
        }
        if (done) break;
      }
      int a;
      if (tmax[0] < tmax[1]) a = (tmax[0] < tmax[2]) ? 0 : 2;
      else a = (tmax[1] < tmax[2]) ? 1 : 2;
      if (a == 0) { cur[0] += stp[0]; tmax[0] += tdel[0]; }
      else if (a == 1) { cur[1] += stp[1]; tmax[1] += tdel[1]; }
      else { cur[2] += stp[2]; tmax[2] += tdel[2]; }
      near_ = far_;
      far_ = fminx(fminx(tmax[0], tmax[1]), tmax[2]);
    }
  }
  res.T = T;
  res.n_samples = nsamp;
  res.depth = depth;
  res.n_bricks = nbr;
}

// march_ray with its three nested loops (tiles, bricks of a tile, samples of a brick) flattened into
// a resumable state machine whose every step does one unit of work: set up the next tile, test /
// step one brick cell, or take one sample.  Lanes of a wavefront that are in different phases then
// share iterations instead of waiting for each other's inner loops, and a lane whose ray ended can
// take a new ray while the others continue.  The arithmetic and the order of the tiles, cells,
// samples and callbacks are those of march_ray, so the results are identical.
struct FlatMarch {
  enum { kTile, kCell, kSample, kDone };
  TileWalker walker;
  V3 origin, d, o;
  float t_hit, T, depth, tl, inv_vs, near_, far_;
  const int32_t* vis;
  const int32_t* bmap;
  int nsamp, nbr, vi, tile, n, nb, k, phase;
  bool first_hit;
  int cur[3], stp[3];
  float tmax[3], tdel[3];

  __device__ __forceinline__ void init(const snisr_scene& S, const snisr_march_params& P, V3 origin_, V3 d_,
                                       float t_hit_, const int32_t* __restrict__ vis_) {
    origin = origin_;
    d = d_;
    t_hit = t_hit_;
    vis = vis_;
    if (!vis) walker.init(S, origin, d, P.max_tiles);
    T = 1.0f;
    nsamp = 0;
    nbr = 0;
    depth = -1.0f;
    first_hit = false;
    vi = 0;
    tile = -1;
    n = nb = k = 0;
    tl = inv_vs = near_ = far_ = 0.0f;
    o = v3(0.0f, 0.0f, 0.0f);
    bmap = nullptr;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cur[a] = 0;
      stp[a] = 1;
      tmax[a] = tdel[a] = 0.0f;
    }
    phase = kTile;
  }
  __device__ __forceinline__ void dda_step() {
    int a;
    if (tmax[0] < tmax[1]) a = (tmax[0] < tmax[2]) ? 0 : 2;
    else a = (tmax[1] < tmax[2]) ? 1 : 2;
    if (a == 0) { cur[0] += stp[0]; tmax[0] += tdel[0]; }
    else if (a == 1) { cur[1] += stp[1]; tmax[1] += tdel[1]; }
    else { cur[2] += stp[2]; tmax[2] += tdel[2]; }
    near_ = far_;
    far_ = fminx(fminx(tmax[0], tmax[1]), tmax[2]);
  }
  // One unit of work.  A lane waiting for a tile step does nothing unless `tile_turn` (see
  // march_ray_flat).
  template <class OnSample>
  __device__ __forceinline__ void step(const snisr_scene& S, const snisr_march_params& P, bool tile_turn,
                                       int32_t* bricks, int brick_cap, OnSample&& on_sample) {
    const float ts = S.tile_size;
    const float hs = ts / 2.0f;
    if (phase == kTile) {
      if (!tile_turn) return;
      if (vi >= P.max_tiles) { phase = kDone; return; }
      const int t = vis ? vis[vi] : walker.next();
      ++vi;
      if (t == -1 || T < P.early_stop) { phase = kDone; return; }
      const V3 tc = v3(S.tile_centers[3 * t], S.tile_centers[3 * t + 1], S.tile_centers[3 * t + 2]);
      float th;
      ray_aabb(origin, d, tc, v3(hs, hs, hs), tl, th);
      if (tl == -1.0f) return;
      tile = t;
      const V3 corner = v3(tc.x - hs, tc.y - hs, tc.z - hs);
      const float hi = ts - 0.00001f;
      o = v3(origin.x + tl * d.x - corner.x, origin.y + tl * d.y - corner.y, origin.z + tl * d.z - corner.z);
      o = v3(clampx(o.x, 0.0f, hi), clampx(o.y, 0.0f, hi), clampx(o.z, 0.0f, hi));
      n = S.tile_res[tile];
      nb = n >> 3;
      const float bs = ts / (float)nb;
      inv_vs = 1.0f / (ts / (float)n);
      bmap = S.brick_map + S.tile_brick_base[tile];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float oa = get(o, a), da = get(d, a);
        cur[a] = (int)(oa / bs);
        stp[a] = da >= 0.0f ? 1 : -1;
        float nbd = (float)(cur[a] + stp[a]) * bs;
        if (stp[a] < 0) nbd += bs;
        tmax[a] = safe_div(nbd - oa, da);
        tdel[a] = safe_div(bs * (float)stp[a], da);
      }
      near_ = 0.0f;
      far_ = fminx(fminx(tmax[0], tmax[1]), tmax[2]);
      phase = kCell;
    } else if (phase == kCell) {
      if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= nb || cur[1] >= nb || cur[2] >= nb) {
        phase = kTile;
        return;
      }
      const int b = bmap[(cur[0] * nb + cur[1]) * nb + cur[2]];
      if (b != -1) {
        if (bricks && nbr < brick_cap) bricks[nbr] = b;
        ++nbr;
        k = 0;
        phase = kSample;
      } else {
        dda_step();
      }
    } else if (phase == kSample) {
      const float step = P.sample_step;
      const float z = __builtin_fmaf((float)k, step, near_);
      if (!(z < far_)) {
        dda_step();
        phase = kCell;
        return;
      }
      const float tg = tl + z;
      if (tg >= t_hit) { phase = kDone; return; }
      const V3 p = v3(o.x + z * d.x, o.y + z * d.y, o.z + z * d.z);
      const V3 f = v3(p.x * inv_vs - 0.5f, p.y * inv_vs - 0.5f, p.z * inv_vs - 0.5f);
      Corners c;
      find_corners(bmap, n, f, c);
      float sig = tri_sigma(S.sigma, c);
      sig = sig > 0.0f ? sig : 0.0f;
      if (!first_hit && T < P.trans_th) { first_hit = true; depth = tg; }
      const float interval = fminx(step, far_ - z);
      const float alpha = 1.0f - exp_cw(-(sig * interval));
      const float w = alpha * T;
      const float T_next = T * (1.0f - alpha);
      on_sample(tile, f, c, w, sig, interval, T_next);
      T = T_next;
      ++nsamp;
      ++k;
      if (T < P.early_stop) phase = kDone;
    }
  }
  __device__ __forceinline__ void result(MarchResult& res) const {
    res.T = T;
    res.n_samples = nsamp;
    res.depth = depth;
    res.n_bricks = nbr;
  }
};

// Whether the lanes waiting for a tile step run it in this iteration: a wavefront of diverged lanes
// issues every phase that any of its lanes is in, each iteration.  The tile step (tile walk, slab
// test, twelve divisions) costs more vector instructions than the cell and sample steps together
// and each lane needs it only a few times, so a lane waiting for it idles until a quorum of the live
// lanes waits (or nothing else is left): the wavefront then issues it once for all of them.  Each
// lane's arithmetic is unchanged, only the interleaving of lanes.  Called by every live lane.
__device__ __forceinline__ bool tile_quorum(bool waiting) {
  const uint64_t live = __builtin_amdgcn_ballot_w64(true);
  const uint64_t wait = __builtin_amdgcn_ballot_w64(waiting);
  return wait == live || kTileQuorum * __builtin_popcountll(wait) >= __builtin_popcountll(live);
}

template <class OnSample>
__device__ __forceinline__ void march_ray_flat(const snisr_scene& S, const snisr_march_params& P, V3 origin, V3 d,
                                               float t_hit, const int32_t* __restrict__ vis, int32_t* bricks,
                                               int brick_cap, OnSample&& on_sample, MarchResult& res) {
  FlatMarch m;
  m.init(S, P, origin, d, t_hit, vis);
  // the ballots at the head of the single loop also keep the compiler from turning the phase
  // machine back into nested loops
  while (m.phase != FlatMarch::kDone) {
    const bool turn = tile_quorum(m.phase == FlatMarch::kTile);
    m.step(S, P, turn, bricks, brick_cap, on_sample);
  }
  m.result(res);
}

// Camera packets (8x8 pixels per wavefront) walk nearly the same cells in the same order, so the
// nested loops of march_ray keep their lanes together; arbitrary rays (training batches, batched
// tiles) diverge, and the flattened loop is faster for them (measured 2.5-2.8x on such batches,
// 0.55x on camera packets).
template <class OnSample>
__device__ __forceinline__ void march_ray_any(bool packets, const snisr_scene& S, const snisr_march_params& P,
                                              V3 origin, V3 d, float t_hit, const int32_t* __restrict__ vis,
                                              int32_t* bricks, int brick_cap, OnSample&& on_sample, MarchResult& res) {
  if (packets) march_ray(S, P, origin, d, t_hit, vis, bricks, brick_cap, on_sample, res);
  else march_ray_flat(S, P, origin, d, t_hit, vis, bricks, brick_cap, on_sample, res);
}

// 15-d view encoding Positional_Encoding(L=2) (training/scripts/networks.py:8-13) of a unit dir.
__device__ __forceinline__ void view_encoding(V3 d, float e[15]) {
  e[0] = d.x; e[1] = d.y; e[2] = d.z;
#pragma unroll
  for (int l = 0; l < 2; ++l) {
    const float s = (float)(1 << l);
    e[3 + 6 * l + 0] = sinf(s * d.x);
    e[3 + 6 * l + 1] = sinf(s * d.y);
    e[3 + 6 * l + 2] = sinf(s * d.z);
    e[3 + 6 * l + 3] = cosf(s * d.x);
    e[3 + 6 * l + 4] = cosf(s * d.y);
    e[3 + 6 * l + 5] = cosf(s * d.z);
  }
}

// ---------------------------------------------------------------------------- per-lane FP32 MLP
// Activations live in a lane-private LDS column (entry k of lane t at xs[k * 64], wavefront-sized
// blocks), weights in the f32t layout Wt[in][out4] then b[out4].
// One dense layer for one lane: y = b + x Wt, x read from this lane's LDS column (stride 64),
// Wt rows read as float4 (all lanes of a wave in one tile read the same address).
template <int IN, int OUT>
__device__ __forceinline__ void dense_f32(const float* __restrict__ Wt, const float* __restrict__ b,
                                          const float* xs, float (&y)[OUT]) {
  static_assert(OUT % 4 == 0, "padded width");
#pragma unroll
  for (int j = 0; j < OUT; ++j) y[j] = b[j];
#pragma unroll 1
  for (int i = 0; i < IN; ++i) {
    const float xi = xs[i * 64];
    const float4* w = reinterpret_cast<const float4*>(Wt + i * OUT);
#pragma unroll
    for (int j = 0; j < OUT / 4; ++j) {
      const float4 q = w[j];
      y[4 * j + 0] = __builtin_fmaf(xi, q.x, y[4 * j + 0]);
      y[4 * j + 1] = __builtin_fmaf(xi, q.y, y[4 * j + 1]);
      y[4 * j + 2] = __builtin_fmaf(xi, q.z, y[4 * j + 2]);
      y[4 * j + 3] = __builtin_fmaf(xi, q.w, y[4 * j + 3]);
    }
  }
}

// Per-sample MLP on one lane: input in xs[0..IN), hidden layers ReLU, sigmoid RGB.
template <int IN, int WD, int L>
__device__ __forceinline__ void mlp_f32(const float* __restrict__ P, float* xs, float rgb[3]) {
  float y[WD];
  dense_f32<IN, WD>(P, P + IN * WD, xs, y);
  P += IN * WD + WD;
#pragma unroll
  for (int j = 0; j < WD; ++j) xs[j * 64] = y[j] > 0.0f ? y[j] : 0.0f;
#pragma unroll 1
  for (int l = 1; l < L - 1; ++l) {
    dense_f32<WD, WD>(P, P + WD * WD, xs, y);
    P += WD * WD + WD;
#pragma unroll
    for (int j = 0; j < WD; ++j) xs[j * 64] = y[j] > 0.0f ? y[j] : 0.0f;
  }
  float z[4];
  dense_f32<WD, 4>(P, P + WD * 4, xs, z);
#pragma unroll
  for (int c = 0; c < 3; ++c) rgb[c] = sigmoid_cw(z[c]);
}

template <bool HALF>
struct FeatLoad;
template <>
struct FeatLoad<true> {  // 8 halves per 16-byte load, raw bits ANDed with m
  static constexpr int kVec = 8;
  __device__ __forceinline__ static void load(const void* base, int64_t off, uint32_t m, float* v) {
    uint4 q = *reinterpret_cast<const uint4*>(reinterpret_cast<const _Float16*>(base) + off);
    q.x &= m; q.y &= m; q.z &= m; q.w &= m;
    const _Float16* h = reinterpret_cast<const _Float16*>(&q);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (float)h[k];
  }
};
template <>
struct FeatLoad<false> {  // 4 floats per 16-byte load, raw bits ANDed with m
  static constexpr int kVec = 4;
  __device__ __forceinline__ static void load(const void* base, int64_t off, uint32_t m, float* v) {
    const uint4 q = *reinterpret_cast<const uint4*>(reinterpret_cast<const float*>(base) + off);
    v[0] = __uint_as_float(q.x & m); v[1] = __uint_as_float(q.y & m);
    v[2] = __uint_as_float(q.z & m); v[3] = __uint_as_float(q.w & m);
  }
};

// Trilinear features of one sample into this lane's LDS column xs[0..F).
template <int F, bool HALF>
__device__ __forceinline__ void gather_features(const void* __restrict__ feats, const Corners& c, float* xs) {
  using LD = FeatLoad<HALF>;
  constexpr int V = LD::kVec;
  static_assert(F % V == 0, "feature dim must fill whole 16-byte vectors");
  uint32_t m[8];  // computed ahead of the loads: the masking statements bound scheduling regions
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = corner_mask(c.slot[k] >= 0);
#pragma unroll
  for (int ch = 0; ch < F / V; ++ch) {
    float v[8][V];
#pragma unroll
    for (int k = 0; k < 8; ++k)  // unconditional loads (slot 0 for empty corners), masked to zero
      LD::load(feats, (int64_t)(c.slot[k] >= 0 ? c.slot[k] : 0) * F + ch * V, m[k], v[k]);
#pragma unroll
    for (int e = 0; e < V; ++e) {
      const float q[8] = {v[0][e], v[1][e], v[2][e], v[3][e], v[4][e], v[5][e], v[6][e], v[7][e]};
      xs[(ch * V + e) * 64] = tri_combine(q, c.wx, c.wy, c.wz);
    }
  }
}

// Ray index of lane `tid` of a launch over n rays.  With image_width W > 0 (and W % 8 == 0,
// rows % 8 == 0) each 64-lane wavefront takes an 8x8 pixel packet, so its rays stay coherent.
__device__ __forceinline__ int64_t ray_of_thread(int64_t tid, int64_t n, int W) {
  if (W <= 0) return tid;
  const int64_t packet = tid >> 6;
  const int lane = (int)(tid & 63);
  const int64_t pw = W >> 3;
  const int64_t py = packet / pw, px = packet - py * pw;
  return (py * 8 + (lane >> 3)) * W + px * 8 + (lane & 7);
}

}  // namespace snisr
