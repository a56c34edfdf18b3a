// ORACLE — test infrastructure only.
//
// A host C++/OpenMP restatement of the reference renderer's per-ray logic, templated
// on the arithmetic type (float = parity oracle, double = accuracy oracle).  It is
// used by tests/, by the smoke check and by bench.py's cpu_baseline leg, never by
// the product path.  Every function cites the reference file:line it restates
// (paths relative to the reference checkout).
//
// Arithmetic contract shared with the GPU kernels (built with -ffp-contract=off):
//   * no FMA contraction except where fma() is written explicitly;
//   * float division and sqrt are IEEE correctly rounded;
//   * exp() in float is the Cody-Waite / Taylor-Horner routine orc_expf below
//     (the GPU path implements the same published method independently);
//   * the DDA tie rules, floor/truncation and safe_divide's 1e8 follow the reference.
//
// Pinning (what each function is checked against; see tests/test_oracle_*.py):
//   ray_aabb            slab closed form on random boxes / rays
//   trace_tiles         brute-force enumeration of grid cells hit by the ray segment
//   brick DDA + march   brute-force sample enumeration on tiny scenes; Python loop
//   trilinear           torch.nn.functional.grid_sample(align_corners=True, zeros)
//   composite           constant-sigma closed form T = exp(-sigma L); sum w + T = 1
//   mlp                 numpy matmul
//   mesh hit            analytic hits on axis-aligned boxes; numpy vectorised MT
//   inverse_z / integrate / assign   closed forms and numpy restatements
//   train adjoint       FP64 central finite differences
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ------------------------------------------------------------------ math helpers
// exp in float: Cody-Waite range reduction x = k*ln2 + r (ln2 split in a 16-bit
// head and a tail, Cody & Waite 1980), then the degree-7 Taylor polynomial of e^r
// in Horner form with fused multiply-adds, scaled by 2^k with ldexp.  Arguments
// below ln(FLT_MIN) return 0 so that no denormal is ever produced.
static inline float orc_expf(float x) {
  if (!(x == x)) return x;
  if (x < -87.33654475f) return 0.0f;
  if (x > 88.72283905f) return std::numeric_limits<float>::infinity();
  const float k = rintf(x * 0x1.715476p+0f);
  float r = fmaf(-k, 0x1.62e400p-1f, x);
  r = fmaf(-k, 0x1.7f7d1cp-20f, r);
  float p = 0x1.a01a02p-13f;          // 1/5040
  p = fmaf(p, r, 0x1.6c16c2p-10f);    // 1/720
  p = fmaf(p, r, 0x1.111112p-7f);     // 1/120
  p = fmaf(p, r, 0x1.555556p-5f);     // 1/24
  p = fmaf(p, r, 0x1.555556p-3f);     // 1/6
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}
template <class R> static inline R rexp(R x);
template <> inline float rexp<float>(float x) { return orc_expf(x); }
template <> inline double rexp<double>(double x) { return std::exp(x); }

template <class R> struct V3 { R x, y, z; };
template <class R> static inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> static inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> static inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> static inline R at(const V3<R>& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class R> static inline void set(V3<R>& a, int i, R v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }

// safe_divide (rendering/src/include/cutil_math.h:867-869): b != 0 ? a/b : 1e8
template <class R> static inline R safe_div(R a, R b) { return b != R(0) ? a / b : R(100000000.0); }
// signf (cutil_math.h:856-859): a >= 0 ? 1 : -1  (so -0.0 gives +1)
template <class R> static inline int signi(R a) { return a >= R(0) ? 1 : -1; }
// clamp(v, lo, hi) = max(lo, min(v, hi)) per component (cutil_math.h clamp)
template <class R> static inline R clampr(R v, R lo, R hi) { return std::max(lo, std::min(v, hi)); }

// normalised direction: d * (1 / sqrt(dot(d, d)))   (training path normalises,
// training/src/cuda_func.cu:99; the render entry point applies it to every ray)
template <class R> static inline V3<R> normalize3(V3<R> d) {
  R inv = R(1) / std::sqrt(dot(d, d));
  return inv * d;
}

// RayAABBIntersection (rendering/src/include/cuda_utils.h:553-598), per-axis half sizes.
template <class R>
static inline void ray_aabb(V3<R> o, V3<R> d, V3<R> c, V3<R> h, R& tlo, R& thi) {
  R lo = 0, hi = R(100000.0);
  for (int a = 0; a < 3; ++a) {
    R inv = safe_div(R(1), at(d, a));
    R l = (at(c, a) - at(h, a) - at(o, a)) * inv;
    R u = (at(c, a) + at(h, a) - at(o, a)) * inv;
    if (u < l) std::swap(l, u);
    if (u < lo || l > hi) { tlo = thi = R(-1); return; }
    lo = l > lo ? l : lo;
    hi = u < hi ? u : hi;
    if (lo > hi) { tlo = thi = R(-1); return; }
  }
  tlo = lo; thi = hi;
}

// ------------------------------------------------------------------ scene
struct Scene {
  int tile_shape[3];
  float scene_min[3];
  float tile_size;
  const int* tile_index_map;   // dense tile cell -> tile id or -1   (x*Ty*Tz + y*Tz + z)
  int n_tiles;
  const float* tile_centers;   // [T,3]
  const int* tile_res;         // [T] voxels per side (multiple of 8)
  const int* tile_brick_base;  // [T] offset of the tile's brick table in brick_map
  const int* brick_map;        // bricks of 8^3 voxels -> brick id or -1
  const float* sigma;          // [NB, 512]
  const void* feats;           // [NB, 512, F]
  int feat_dim;
  int feat_half;               // 1: feats are IEEE binary16, 0: float32
  int mlp_in, mlp_width, mlp_layers;  // out = 3
  const float* mlp_params;     // [T, P]  per layer: W[out][in] then b[out]
  long long mlp_stride;        // P
};

struct Params {
  int max_tiles;        // K, tiles recorded per ray (base.yaml:8)
  float sample_step;    // world units along the unit direction
  float early_stop;     // stop the ray once T < early_stop
  float trans_th;       // first-hit threshold on T (__init__.py:37, trans_decay)
  int brick_debug;      // capacity of the visited-brick debug list (0 = off)
};

static inline float half_to_float(uint16_t h) {
  uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
  uint32_t bits;
  if (e == 0) {
    if (m == 0) bits = s;
    else {  // subnormal half -> normal float
      int ee = -1;
      do { m <<= 1; ++ee; } while (!(m & 0x400));
      bits = s | ((uint32_t)(127 - 15 - ee) << 23) | ((m & 0x3ff) << 13);
    }
  } else if (e == 31) bits = s | 0x7f800000u | (m << 13);
  else bits = s | ((e + 112) << 23) | (m << 13);
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// one voxel corner lookup with zero padding outside the tile / in empty bricks
static inline long long voxel_slot(const Scene& S, int tile, int ix, int iy, int iz) {
  const int n = S.tile_res[tile];
  if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n) return -1;
  const int nb = n >> 3;
  const int b = S.brick_map[S.tile_brick_base[tile] + ((ix >> 3) * nb + (iy >> 3)) * nb + (iz >> 3)];
  if (b < 0) return -1;
  return (long long)b * 512 + (((ix & 7) * 8 + (iy & 7)) * 8 + (iz & 7));
}

template <class R> static inline R lerp(R a, R b, R t) { return a * (R(1) - t) + b * t; }

// Trilinear interpolation at continuous voxel coordinate f (= p/voxel - 0.5):
// corner order and the x, then y, then z lerp nesting follow interpolate3D
// (rendering/src/include/interpolation.h:13-38); corner weights from floor(f),
// which is the truncation of trilinear_base (interpolation.h:49-73) on the
// non-negative coordinates it sees; empty voxels read as zero (tile.py:232
// grid_sample padding_mode="zeros").  Writes sigma into *sig and F features.
template <class R>
static void trilinear(const Scene& S, int tile, V3<R> f, R* sig, R* feat) {
  const R fx = std::floor(f.x), fy = std::floor(f.y), fz = std::floor(f.z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const R wx = f.x - fx, wy = f.y - fy, wz = f.z - fz;
  long long slot[8];
  // v1..v8 of interpolate3D: x0y0z0, x1y0z0, x0y1z0, x1y1z0, x0y0z1, x1y0z1, x0y1z1, x1y1z1
  for (int c = 0; c < 8; ++c) slot[c] = voxel_slot(S, tile, ix + (c & 1), iy + ((c >> 1) & 1), iz + (c >> 2));
  auto sv = [&](int c) -> R { return slot[c] < 0 ? R(0) : R(S.sigma[slot[c]]); };
  {
    R s = lerp(lerp(sv(0), sv(1), wx), lerp(sv(2), sv(3), wx), wy);
    R t = lerp(lerp(sv(4), sv(5), wx), lerp(sv(6), sv(7), wx), wy);
    *sig = lerp(s, t, wz);
  }
  if (!feat) return;
  const int F = S.feat_dim;
  for (int k = 0; k < F; ++k) {
    auto fv = [&](int c) -> R {
      if (slot[c] < 0) return R(0);
      long long i = slot[c] * F + k;
      return S.feat_half ? R(half_to_float(((const uint16_t*)S.feats)[i])) : R(((const float*)S.feats)[i]);
    };
    R s = lerp(lerp(fv(0), fv(1), wx), lerp(fv(2), fv(3), wx), wy);
    R t = lerp(lerp(fv(4), fv(5), wx), lerp(fv(6), fv(7), wx), wy);
    feat[k] = lerp(s, t, wz);
  }
}

// Per-sample MLP: x = [features(F), PE_2(dir)(15)] -> Linear/ReLU ... -> 3 -> sigmoid.
// Linear layout W[out][in] then b[out] per layer (networks.py:17-31 nn.Linear order);
// the view encoding is Positional_Encoding(L=2) (networks.py:8-13): [d, sin d, cos d,
// sin 2d, cos 2d]; ReLU hidden activations, sigmoid RGB (tile.py:194).
template <class R>
static void mlp_eval(const Scene& S, int tile, const R* x, R* rgb_out, std::vector<R>& bufa, std::vector<R>& bufb) {
  const float* P = S.mlp_params + (long long)tile * S.mlp_stride;
  const int L = S.mlp_layers, Wd = S.mlp_width;
  int in = S.mlp_in;
  bufa.assign(x, x + in);
  for (int l = 0; l < L; ++l) {
    const int out = (l == L - 1) ? 3 : Wd;
    const float* W = P;
    const float* b = P + (long long)out * in;
    bufb.resize(out);
    for (int j = 0; j < out; ++j) {
      R acc = R(b[j]);
      const float* w = W + (long long)j * in;
      for (int i = 0; i < in; ++i) acc += R(w[i]) * bufa[i];
      bufb[j] = (l == L - 1) ? acc : (acc > R(0) ? acc : R(0));
    }
    P = b + out;
    in = out;
    std::swap(bufa, bufb);
  }
  for (int c = 0; c < 3; ++c) rgb_out[c] = R(1) / (R(1) + rexp(-bufa[c]));
}

template <class R>
static void view_encoding(V3<R> d, R* e) {
  e[0] = d.x; e[1] = d.y; e[2] = d.z;
  for (int l = 0; l < 2; ++l) {
    const R s = R(1 << l);
    e[3 + 6 * l + 0] = std::sin(s * d.x);
    e[3 + 6 * l + 1] = std::sin(s * d.y);
    e[3 + 6 * l + 2] = std::sin(s * d.z);
    e[3 + 6 * l + 3] = std::cos(s * d.x);
    e[3 + 6 * l + 4] = std::cos(s * d.y);
    e[3 + 6 * l + 5] = std::cos(s * d.z);
  }
}

// ------------------------------------------------------------------ A1 camera rays
// compute_rays_kernel / get_direction (rendering/src/compute_rays.cu:7-44):
// x = (px - cx) * invf, y = (py - cy) * invf, d = R [x, y, 1]  (not normalised).
template <class R>
static void compute_rays(int H, int W, const float* intr, const float* c2w, R* out) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)H * W; ++i) {
    const int py = (int)(i / W), px = (int)(i - (long long)py * W);
    const R x = (R(px) - R(intr[0])) * R(intr[2]);
    const R y = (R(py) - R(intr[1])) * R(intr[2]);
    for (int r = 0; r < 3; ++r) out[i * 3 + r] = R(c2w[3 * r]) * x + R(c2w[3 * r + 1]) * y + R(c2w[3 * r + 2]);
  }
}

// ------------------------------------------------------------------ A2 tile DDA
// voxel_traversal (rendering/src/include/cuda_utils.h:158-238) over the tile grid with
// the entry clamp of tracing_tiles_kernel (rendering/src/raytile_tracing.cu:30-47).
// Deviation: a start cell outside the grid (possible when scene_size - 1e-6 rounds
// up to scene_size) records nothing instead of reading IndexMap out of bounds.
template <class R>
static int trace_tiles_one(const Scene& S, V3<R> origin, V3<R> d, int K, int* visited) {
  for (int j = 0; j < K; ++j) visited[j] = -1;
  if (K <= 0) return 0;
  const V3<R> smin = v3<R>(S.scene_min[0], S.scene_min[1], S.scene_min[2]);
  const R ts = R(S.tile_size);
  const V3<R> size = v3<R>(R(S.tile_shape[0]) * ts, R(S.tile_shape[1]) * ts, R(S.tile_shape[2]) * ts);
  V3<R> o = origin - smin;
  if (o.x < 0 || o.x >= size.x || o.y < 0 || o.y >= size.y || o.z < 0 || o.z >= size.z) {
    const V3<R> half = v3<R>(size.x / R(2), size.y / R(2), size.z / R(2));
    R tl, th;
    ray_aabb(origin, d, smin + half, half, tl, th);
    o = o + tl * d;
    const R eps = R(0.000001f);
    o = v3<R>(clampr(o.x, R(0), size.x - eps), clampr(o.y, R(0), size.y - eps), clampr(o.z, R(0), size.z - eps));
  }
  const int M[3] = {S.tile_shape[0], S.tile_shape[1], S.tile_shape[2]};
  int cur[3], step[3];
  R tmax[3], tdel[3];
  for (int a = 0; a < 3; ++a) {
    cur[a] = (int)std::floor(at(o, a) / ts);
    step[a] = at(d, a) >= 0 ? 1 : -1;
    R nb = R(cur[a] + step[a]) * ts;
    if (step[a] < 0) nb += ts;
    tmax[a] = at(d, a) != 0 ? (nb - at(o, a)) / at(d, a) : R(100000000);
    tdel[a] = at(d, a) != 0 ? (ts / at(d, a)) * R(step[a]) : R(100000000);
  }
  auto inside = [&]() {
    return cur[0] >= 0 && cur[1] >= 0 && cur[2] >= 0 && cur[0] < M[0] && cur[1] < M[1] && cur[2] < M[2];
  };
  if (!inside()) return 0;
  int idx = 0;
  int id = S.tile_index_map[(cur[0] * M[1] + cur[1]) * M[2] + cur[2]];
  if (id != -1) visited[idx++] = id;
  while (idx < K) {
    int a;
    if (tmax[0] < tmax[1]) a = (tmax[0] < tmax[2]) ? 0 : 2;
    else a = (tmax[1] < tmax[2]) ? 1 : 2;
    cur[a] += step[a];
    tmax[a] += tdel[a];
    if (!inside()) break;
    id = S.tile_index_map[(cur[0] * M[1] + cur[1]) * M[2] + cur[2]];
    if (id != -1) visited[idx++] = id;
  }
  return idx;
}

// ------------------------------------------------------------------ A3/A4/A5/A7/A10 march
template <class R>
struct SampleRec {
  int tile;
  V3<R> f;       // voxel coordinate p/voxel - 0.5 in the tile
  R sig;         // max(trilinear sigma, 0)
  R interval;    // min(step, far - z)
  R w;           // alpha * T
  R T_after;     // T * (1 - alpha)
};

struct RayOut {
  float rgb[3];
  float T;
  int n_samples;
  float depth;  // t of the first sample reached with T < trans_th, -1 if none
  int n_bricks;
};

// One ray through the visited tiles.  Per tile: RayAABB entry (raydiffuse_render_octree_fp16.cu:202-213,
// local origin clamped to [0, tile_size - 1e-5]), then the block DDA of block_traversal_rendering
// (raydiffuse_render_octree_fp16.cu:39-138) over 8^3-voxel bricks; inside an occupied brick the
// samples restart at the brick entry "near" and are z_k = fma(k, step, near) while z_k < far
// (the reference accumulates zval += step, :71-105; the fma form is the stated deviation),
// interval = min(step, far - z_k) (:101); compositing alpha = 1 - exp(-sigma*interval),
// C += alpha*T*rgb, T *= 1 - alpha (rendering/src/include/cuda_utils.h:13-20); the ray stops once
// T < early_stop after a sample (training/src/cuda_func.cu:1778-1791) or at the first sample whose
// global t reaches the proxy-mesh hit t_hit.  sigma = max(trilinear sigma, 0); rgb = MLP(features).
template <class R>
static void march_ray(const Scene& S, const Params& P, V3<R> origin, V3<R> d_raw, R t_hit,
                      const int* visited_in, RayOut& out, int* brick_dbg, bool with_mlp,
                      std::vector<SampleRec<R>>* recs = nullptr) {
  const V3<R> d = normalize3(d_raw);
  std::vector<int> vis(std::max(P.max_tiles, 1));
  int nvis;
  if (visited_in) {
    nvis = 0;
    while (nvis < P.max_tiles && visited_in[nvis] != -1) { vis[nvis] = visited_in[nvis]; ++nvis; }
  } else {
    nvis = trace_tiles_one(S, origin, d, P.max_tiles, vis.data());
  }
  R color[3] = {0, 0, 0};
  R T = 1;
  int nsamp = 0, nbr = 0;
  R depth = -1;
  bool first_hit = false, done = false;
  const R ts = R(S.tile_size);
  const R step = R(P.sample_step);
  const int F = S.feat_dim;
  std::vector<R> x(S.mlp_in), ba, bb;
  R enc[15];
  if (with_mlp) view_encoding(d, enc);
  for (int vi = 0; vi < nvis && !done; ++vi) {
    const int tile = vis[vi];
    if (T < R(P.early_stop)) break;
    const V3<R> tc = v3<R>(S.tile_centers[3 * tile], S.tile_centers[3 * tile + 1], S.tile_centers[3 * tile + 2]);
    const R hs = ts / R(2);
    R tl, th;
    ray_aabb(origin, d, tc, v3<R>(hs, hs, hs), tl, th);
    if (tl == R(-1)) continue;
    const V3<R> corner = v3<R>(tc.x - hs, tc.y - hs, tc.z - hs);
    V3<R> o = (origin + tl * d) - corner;
    const R hi = ts - R(0.00001f);
    o = v3<R>(clampr(o.x, R(0), hi), clampr(o.y, R(0), hi), clampr(o.z, R(0), hi));
    const int n = S.tile_res[tile];
    const int nb = n >> 3;
    const R bs = ts / R(nb);
    const R inv_vs = R(1) / (ts / R(n));
    const int* bmap = S.brick_map + S.tile_brick_base[tile];
    int cur[3], stp[3];
    R tmax[3], tdel[3];
    for (int a = 0; a < 3; ++a) {
      cur[a] = (int)(at(o, a) / bs);  // make_int3(float3) truncates; o >= 0
      stp[a] = signi(at(d, a));
      R nbd = R(cur[a] + stp[a]) * bs;
      if (stp[a] < 0) nbd += bs;
      tmax[a] = safe_div(nbd - at(o, a), at(d, a));
      tdel[a] = safe_div(bs * R(stp[a]), at(d, a));
    }
    R near_ = 0, far_ = std::min(std::min(tmax[0], tmax[1]), tmax[2]);
    while (true) {
      if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= nb || cur[1] >= nb || cur[2] >= nb) break;
      const int b = bmap[(cur[0] * nb + cur[1]) * nb + cur[2]];
      if (b != -1) {
        if (brick_dbg && nbr < P.brick_debug) brick_dbg[nbr] = b;
        ++nbr;
        for (int k = 0;; ++k) {
          const R z = std::fma(R(k), step, near_);
          if (!(z < far_)) break;
          const R tg = tl + z;
          if (tg >= t_hit) { done = true; break; }
          const V3<R> p = o + z * d;
          const V3<R> f = v3<R>(p.x * inv_vs - R(0.5), p.y * inv_vs - R(0.5), p.z * inv_vs - R(0.5));
          R sig;
          trilinear(S, tile, f, &sig, with_mlp ? x.data() : nullptr);
          sig = sig > R(0) ? sig : R(0);
          if (!first_hit && T < R(P.trans_th)) { first_hit = true; depth = tg; }
          const R interval = std::min(step, far_ - z);
          const R alpha = R(1) - rexp(-(sig * interval));
          const R w = alpha * T;
          if (with_mlp) {
            for (int i = 0; i < 15; ++i) x[F + i] = enc[i];
            R rgb[3];
            mlp_eval(S, tile, x.data(), rgb, ba, bb);
            for (int c = 0; c < 3; ++c) color[c] = color[c] + w * rgb[c];
          }
          T = T * (R(1) - alpha);
          if (recs) recs->push_back(SampleRec<R>{tile, f, sig, interval, w, T});
          ++nsamp;
          if (T < R(P.early_stop)) { done = true; break; }
        }
        if (done) break;
      }
      int a;
      if (tmax[0] < tmax[1]) a = (tmax[0] < tmax[2]) ? 0 : 2;
      else a = (tmax[1] < tmax[2]) ? 1 : 2;
      cur[a] += stp[a];
      tmax[a] += tdel[a];
      near_ = far_;
      far_ = std::min(std::min(tmax[0], tmax[1]), tmax[2]);
    }
  }
  for (int c = 0; c < 3; ++c) out.rgb[c] = float(color[c]);
  out.T = float(T);
  out.n_samples = nsamp;
  out.depth = float(depth);
  out.n_bricks = nbr;
}

// ------------------------------------------------------------------ A11 training adjoint
// Gradient of L = sum_rays <g_ray, C_ray> for the render of march_ray (features, density and the
// per-tile MLP are the parameters; sample positions depend on geometry only).  Restates the
// derivative of the compositing of utils.volume_rendering (training/scripts/utils.py:823-837,
// here with the marcher's per-sample T) and of the MLP/trilinear chain that torch autograd
// differentiates in Tile.train_batch (training/scripts/tile/tile.py:631-820):
//   dL/dc_i     = w_i g
//   dL/dsigma_i = interval_i (T_{i+1} <c_i, g> - sum_{k>i} w_k <c_k, g>)   (0 where sigma was clamped)
//   MLP and trilinear by the chain rule; the trilinear weight of corner q is the product of the
//   per-axis lerp weights.
template <class R>
static void accum(R* p, R v) {
#pragma omp atomic
  *p += v;
}

template <class R>
static void train_ray(const Scene& S, const Params& P, V3<R> origin, V3<R> d_raw, R t_hit, const R* g, R* C_out,
                      R* dsigma, R* dfeat, R* dmlp) {
  std::vector<SampleRec<R>> recs;
  RayOut ro;
  march_ray<R>(S, P, origin, d_raw, t_hit, nullptr, ro, nullptr, false, &recs);
  const V3<R> d = normalize3(d_raw);
  R enc[15];
  view_encoding(d, enc);
  const int F = S.feat_dim, L = S.mlp_layers, Wd = S.mlp_width, IN = S.mlp_in;
  const int n = (int)recs.size();
  // forward per sample, keeping every layer's activations
  std::vector<std::vector<std::vector<R>>> act(n);  // act[i][l] = input of layer l (l = 0..L-1)
  std::vector<std::array<R, 3>> col(n);
  R C[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    std::vector<R> x(IN);
    R sg;
    trilinear(S, recs[i].tile, recs[i].f, &sg, x.data());
    for (int k = 0; k < 15; ++k) x[F + k] = enc[k];
    act[i].push_back(x);
    const float* Pm = S.mlp_params + (long long)recs[i].tile * S.mlp_stride;
    int in = IN;
    for (int l = 0; l < L; ++l) {
      const int out = (l == L - 1) ? 3 : Wd;
      const float* Wt = Pm;
      const float* b = Pm + (long long)out * in;
      std::vector<R> y(out);
      for (int j = 0; j < out; ++j) {
        R acc = R(b[j]);
        for (int q = 0; q < in; ++q) acc += R(Wt[(long long)j * in + q]) * act[i][l][q];
        y[j] = (l == L - 1) ? acc : (acc > R(0) ? acc : R(0));
      }
      Pm = b + out;
      in = out;
      if (l < L - 1) act[i].push_back(y);
      else for (int c = 0; c < 3; ++c) col[i][c] = R(1) / (R(1) + rexp(-y[c]));
    }
    for (int c = 0; c < 3; ++c) C[c] += recs[i].w * col[i][c];
  }
  for (int c = 0; c < 3; ++c) C_out[c] = C[c];
  // reverse pass over the samples
  R suffix = 0;
  for (int i = n - 1; i >= 0; --i) {
    const SampleRec<R>& r = recs[i];
    const R cg = col[i][0] * g[0] + col[i][1] * g[1] + col[i][2] * g[2];
    R dsig = r.interval * (r.T_after * cg - suffix);
    suffix += r.w * cg;
    if (!(r.sig > R(0))) dsig = 0;
    // MLP backward
    std::vector<R> delta(3);
    for (int c = 0; c < 3; ++c) delta[c] = r.w * g[c] * col[i][c] * (R(1) - col[i][c]);
    // parameter offsets of each layer
    std::vector<long long> woff(L), boff(L);
    std::vector<int> lin(L), lout(L);
    long long off = 0;
    int in = IN;
    for (int l = 0; l < L; ++l) {
      lin[l] = in;
      lout[l] = (l == L - 1) ? 3 : Wd;
      woff[l] = off;
      boff[l] = off + (long long)lout[l] * in;
      off = boff[l] + lout[l];
      in = lout[l];
    }
    R* gm = dmlp + (long long)r.tile * S.mlp_stride;
    const float* Pm = S.mlp_params + (long long)r.tile * S.mlp_stride;
    std::vector<R> dx;
    for (int l = L - 1; l >= 0; --l) {
      const std::vector<R>& a = act[i][l];
      for (int o = 0; o < lout[l]; ++o) {
        if (delta[o] == R(0)) continue;
        accum(&gm[boff[l] + o], delta[o]);
        for (int q = 0; q < lin[l]; ++q) accum(&gm[woff[l] + (long long)o * lin[l] + q], delta[o] * a[q]);
      }
      std::vector<R> prev(lin[l], R(0));
      for (int q = 0; q < lin[l]; ++q) {
        R acc = 0;
        for (int o = 0; o < lout[l]; ++o) acc += R(Pm[woff[l] + (long long)o * lin[l] + q]) * delta[o];
        prev[q] = acc;
      }
      if (l > 0)
        for (int q = 0; q < lin[l]; ++q) prev[q] = a[q] > R(0) ? prev[q] : R(0);
      delta.swap(prev);
      if (l == 0) dx = delta;
    }
    // trilinear adjoint: corner weights are products of the per-axis lerp weights
    const R fx = std::floor(r.f.x), fy = std::floor(r.f.y), fz = std::floor(r.f.z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const R wx = r.f.x - fx, wy = r.f.y - fy, wz = r.f.z - fz;
    for (int q = 0; q < 8; ++q) {
      const long long slot = voxel_slot(S, r.tile, ix + (q & 1), iy + ((q >> 1) & 1), iz + (q >> 2));
      if (slot < 0) continue;
      const R wq = ((q & 1) ? wx : R(1) - wx) * (((q >> 1) & 1) ? wy : R(1) - wy) * ((q >> 2) ? wz : R(1) - wz);
      if (dsig != R(0)) accum(&dsigma[slot], wq * dsig);
      for (int e = 0; e < F; ++e) accum(&dfeat[slot * F + e], wq * dx[e]);
    }
  }
}

// ------------------------------------------------------------------ A9 mesh hit
// RayTriangleIntersection (rendering/src/include/cuda_utils.h:403-436, blur = 0) with the
// two-sided vertex flip of ray_firstHit_inTile_kernel (training/src/cuda_func.cu:215-231).
// Deviation: the flip tests the sign of dot(d, cross(v1-v0, v2-v1)) without normalising the
// normal (same sign).  Returns -1e8 on a miss.
template <class R>
static inline R ray_tri(V3<R> o, V3<R> d, V3<R> a, V3<R> b, V3<R> c) {
  const V3<R> e1 = b - a, e2 = c - b;
  const R sim = dot(d, cross(e1, e2));
  V3<R> v1 = b, v2 = c;
  if (sim > 0) { v1 = c; v2 = b; }
  const V3<R> v0v1 = v1 - a, v0v2 = v2 - a, v0o = o - a;
  const V3<R> pv = cross(d, v0v2);
  R det = dot(v0v1, pv);
  if (det < R(0.00000001f)) return R(-100000000);
  det = R(1) / det;
  const R u = dot(v0o, pv) * det;
  if (u < R(0) || u > R(1)) return R(-100000000);
  const V3<R> qv = cross(v0o, v0v1);
  const R v = dot(d, qv) * det;
  if (v < R(0) || v > R(1)) return R(-100000000);
  if ((u + v) < R(0) || (u + v) > R(1)) return R(-100000000);
  return dot(v0v2, qv) * det;
}

// nearest hit: smallest t > 0, ties to the smaller triangle id (brute force over all triangles)
template <class R>
static void mesh_hit_one(const float* V, const int* Fc, int nf, V3<R> o, V3<R> d, R* t_out, int* id_out) {
  R best = std::numeric_limits<R>::infinity();
  int bid = -1;
  for (int f = 0; f < nf; ++f) {
    const int* ix = Fc + 3 * f;
    const V3<R> a = v3<R>(V[3 * ix[0]], V[3 * ix[0] + 1], V[3 * ix[0] + 2]);
    const V3<R> b = v3<R>(V[3 * ix[1]], V[3 * ix[1] + 1], V[3 * ix[1] + 2]);
    const V3<R> c = v3<R>(V[3 * ix[2]], V[3 * ix[2] + 1], V[3 * ix[2] + 2]);
    const R t = ray_tri(o, d, a, b, c);
    if (t > R(0) && t < best) { best = t; bid = f; }
  }
  *t_out = best;
  *id_out = bid;
}

// A7: inference_mlp_kernel (rendering/src/rayreflection_render.cu:111-300) with inference_SH
// (61-108), leakyRelu (38-49) and Sigmoid (51-58), in FP64.  Per sample: x = (p - centre) *
// inverse_far in FP32 as the reference computes it, positional encoding of L = 10 bands
// (x, sin 2^k x, cos 2^k x; 63 inputs), 63 -> 64 -> 4 x (64 -> 64) with LeakyReLU(0.01) -> 28;
// sigma = max(out[27], 0); rgb = sigmoid(SH degree 2 at the normalised ray direction with
// coefficient r = 3 * basis + channel).  Parameters per group: 22,556 halves, each layer's bias
// (out) followed by W as [in][out].  quant16 != 0 rounds the 63 inputs and every hidden activation
// to FP16 (round to nearest even) where the MFMA kernel's operands are FP16; 0 is exact math on
// the FP16 parameters.
static inline double round_half(double v) {
  if (!(std::fabs(v) < 65520.0)) return std::isnan(v) ? v : std::copysign(INFINITY, v);
  int e;
  std::frexp(v, &e);                        // |v| = m 2^e, m in [0.5, 1)
  const int q = std::max(e - 11, -24);      // 11 significant bits; subnormal quantum 2^-24
  return std::ldexp(std::nearbyint(std::ldexp(v, -q)), q);
}

constexpr int kRefIn = 63, kRefWidth = 64, kRefOut = 28, kRefLayers = 6, kRefParams = 22556;

static void sh2_basis(const double* d, double* b) {
  const double pi = 3.14159265358979323846;
  const double c0 = 0.5 / std::sqrt(pi), c1 = std::sqrt(3.0) * c0, c2 = std::sqrt(15.0) * c0,
               c22 = std::sqrt(5.0) * 0.5 * c0, c24 = std::sqrt(15.0) * 0.5 * c0;
  const double x = d[0], y = d[1], z = d[2];
  b[0] = c0;
  b[1] = -c1 * y; b[2] = c1 * z; b[3] = -c1 * x;
  b[4] = c2 * x * y; b[5] = -c2 * y * z; b[6] = c22 * (2.0 * z * z - x * x - y * y);
  b[7] = -c2 * x * z; b[8] = c24 * (x * x - y * y);
}

static void ref_mlp_sample(const uint16_t* P, const float* point, const double* dir_unit, int quant16, float* out4) {
  double a[kRefWidth], b[kRefWidth];
  a[0] = point[0]; a[1] = point[1]; a[2] = point[2];
  for (int k = 0; k < 10; ++k)
    for (int r = 0; r < 3; ++r) {
      const double t = std::ldexp((double)point[r], k);
      a[3 + 6 * k + r] = std::sin(t);
      a[6 + 6 * k + r] = std::cos(t);
    }
  if (quant16)
    for (int i = 0; i < kRefIn; ++i) a[i] = round_half(a[i]);
  int off = 0, in = kRefIn;
  double* x = a;
  double* y = b;
  for (int l = 0; l < kRefLayers; ++l) {
    const int out = l == kRefLayers - 1 ? kRefOut : kRefWidth;
    const uint16_t* bias = P + off;
    const uint16_t* Wt = bias + out;  // [in][out]
    for (int j = 0; j < out; ++j) {
      double acc = half_to_float(bias[j]);
      for (int i = 0; i < in; ++i) acc += x[i] * (double)half_to_float(Wt[i * out + j]);
      y[j] = acc;
    }
    if (l < kRefLayers - 1)
      for (int j = 0; j < out; ++j) {
        y[j] = y[j] < 0.0 ? 0.01 * y[j] : y[j];
        if (quant16) y[j] = round_half(y[j]);
      }
    off += out + in * out;
    in = out;
    std::swap(x, y);
  }
  double basis[9];
  sh2_basis(dir_unit, basis);
  for (int c = 0; c < 3; ++c) {
    double raw = 0.0;
    for (int k = 0; k < 9; ++k) raw += basis[k] * x[3 * k + c];
    out4[c] = (float)(1.0 / (1.0 + std::exp(-raw)));
  }
  out4[3] = (float)std::max(x[kRefOut - 1], 0.0);
}

}  // namespace orc
extern "C" {

void orc_inference_mlp_f64(int spr, long long nblocks, float inverse_far, float trans_stop, const int16_t* net,
                           const float* T, const int* qpi, const int* pixel_starts, const float* centers,
                           int n_groups, const float* samples, const float* rays_d, long long d_row, long long d_col,
                           int W, long long npix, const uint16_t* params, int quant16, float* results) {
#pragma omp parallel for schedule(dynamic, 4)
  for (long long b = 0; b < nblocks; ++b) {
    const int g = net[b];
    if (g < 0 || g >= n_groups) continue;
    const uint16_t* P = params + (long long)g * orc::kRefParams;
    for (long long t = (long long)pixel_starts[b] * spr; t < (long long)pixel_starts[b + 1] * spr; ++t) {
      const long long ori = qpi[t / spr];
      if (ori < 0 || ori >= npix) continue;
      if (T[ori] < trans_stop) continue;
      const long long task = ori * spr + t % spr;
      float p[3];
      for (int i = 0; i < 3; ++i) p[i] = (samples[task * 3 + i] - centers[g * 3 + i]) * inverse_far;
      const float* d = rays_d + (ori / W) * d_row + (ori % W) * d_col;
      const double n = std::sqrt((double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2]);
      const double du[3] = {d[0] / n, d[1] / n, d[2] / n};
      orc::ref_mlp_sample(P, p, du, quant16, results + task * 4);
    }
  }
}

void orc_sh2_basis_f64(const double* dirs, long long n, double* out) {
  for (long long i = 0; i < n; ++i) orc::sh2_basis(dirs + 3 * i, out + 9 * i);
}

double orc_round_half_f64(double v) { return orc::round_half(v); }

}  // extern "C"
namespace orc {
}  // namespace orc

using namespace orc;

// =================================================================== C entry points
extern "C" {

struct orc_scene_c {
  int tile_shape[3];
  float scene_min[3];
  float tile_size;
  const int* tile_index_map;
  int n_tiles;
  const float* tile_centers;
  const int* tile_res;
  const int* tile_brick_base;
  const int* brick_map;
  const float* sigma;
  const void* feats;
  int feat_dim, feat_half;
  int mlp_in, mlp_width, mlp_layers;
  const float* mlp_params;
  long long mlp_stride;
};

struct orc_params_c {
  int max_tiles;
  float sample_step, early_stop, trans_th;
  int brick_debug;
};

static Scene to_scene(const orc_scene_c* c) {
  Scene S;
  std::memcpy(S.tile_shape, c->tile_shape, sizeof S.tile_shape);
  std::memcpy(S.scene_min, c->scene_min, sizeof S.scene_min);
  S.tile_size = c->tile_size;
  S.tile_index_map = c->tile_index_map;
  S.n_tiles = c->n_tiles;
  S.tile_centers = c->tile_centers;
  S.tile_res = c->tile_res;
  S.tile_brick_base = c->tile_brick_base;
  S.brick_map = c->brick_map;
  S.sigma = c->sigma;
  S.feats = c->feats;
  S.feat_dim = c->feat_dim;
  S.feat_half = c->feat_half;
  S.mlp_in = c->mlp_in;
  S.mlp_width = c->mlp_width;
  S.mlp_layers = c->mlp_layers;
  S.mlp_params = c->mlp_params;
  S.mlp_stride = c->mlp_stride;
  return S;
}
static Params to_params(const orc_params_c* p) {
  return Params{p->max_tiles, p->sample_step, p->early_stop, p->trans_th, p->brick_debug};
}

float orc_expf_c(float x) { return orc_expf(x); }
float orc_half_to_float_c(uint16_t h) { return half_to_float(h); }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_compute_rays_f32(int H, int W, const float* intr, const float* c2w, float* out) {
  compute_rays<float>(H, W, intr, c2w, out);
}
void orc_compute_rays_f64(int H, int W, const float* intr, const float* c2w, double* out) {
  compute_rays<double>(H, W, intr, c2w, out);
}

void orc_ray_aabb_f32(const float* o, const float* d, const float* c, const float* h, float* out) {
  float tl, th;
  ray_aabb<float>(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), v3(c[0], c[1], c[2]), v3(h[0], h[1], h[2]), tl, th);
  out[0] = tl; out[1] = th;
}

// A2 over N rays sharing one origin (tracing_tiles_kernel, rendering/src/raytile_tracing.cu:9-60)
void orc_trace_tiles_f32(const orc_scene_c* sc, const float* origin, const float* dirs, long long N, int K,
                         int* visited) {
  Scene S = to_scene(sc);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long i = 0; i < N; ++i)
    trace_tiles_one<float>(S, v3(origin[0], origin[1], origin[2]),
                           v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), K, visited + i * K);
}

}  // extern "C"
template <class R>
static void render_impl(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                        const float* dirs, const float* t_hit, const int* visited, long long N, const long long* ray_index,
                        float* rgb, float* T, int* nsamp, float* depth, int* brick_dbg, int with_mlp) {
  Scene S = to_scene(sc);
  Params P = to_params(pc);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long j = 0; j < N; ++j) {
    const long long i = ray_index ? ray_index[j] : j;
    RayOut o;
    const R th = t_hit ? R(t_hit[i]) : std::numeric_limits<R>::infinity();
    const float* og = origin + i * ostride;
    march_ray<R>(S, P, v3<R>(og[0], og[1], og[2]), v3<R>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]),
                 th, visited ? visited + i * P.max_tiles : nullptr, o,
                 brick_dbg ? brick_dbg + j * P.brick_debug : nullptr, with_mlp != 0);
    if (rgb) for (int c = 0; c < 3; ++c) rgb[3 * j + c] = o.rgb[c];
    if (T) T[j] = o.T;
    if (nsamp) nsamp[j] = o.n_samples;
    if (depth) depth[j] = o.depth;
  }
}

extern "C" {
// Full per-ray render of N rays (or of the subset ray_index[0..N) when ray_index != NULL;
// outputs are then indexed by j).  brick_dbg, when non-NULL, receives per ray the first
// brick_debug occupied brick ids marched, -1 padded (caller pre-fills with -1).
void orc_render_f32(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                    const float* dirs, const float* t_hit, const int* visited, long long N, const long long* ray_index, float* rgb,
                    float* T, int* nsamp, float* depth, int* brick_dbg, int with_mlp) {
  render_impl<float>(sc, pc, origin, ostride, dirs, t_hit, visited, N, ray_index, rgb, T, nsamp, depth, brick_dbg, with_mlp);
}
void orc_render_f64(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                    const float* dirs, const float* t_hit, const int* visited, long long N, const long long* ray_index, float* rgb,
                    float* T, int* nsamp, float* depth, int* brick_dbg, int with_mlp) {
  render_impl<double>(sc, pc, origin, ostride, dirs, t_hit, visited, N, ray_index, rgb, T, nsamp, depth, brick_dbg, with_mlp);
}

}  // extern "C"
template <class R>
static void train_impl(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                       const float* dirs, const float* t_hit, long long N, const R* g, R* C, R* dsigma, R* dfeat,
                       R* dmlp) {
  Scene S = to_scene(sc);
  Params P = to_params(pc);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < N; ++i) {
    const float* og = origin + i * ostride;
    const R th = t_hit ? R(t_hit[i]) : std::numeric_limits<R>::infinity();
    train_ray<R>(S, P, v3<R>(og[0], og[1], og[2]), v3<R>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), th,
                 g + 3 * i, C + 3 * i, dsigma, dfeat, dmlp);
  }
}
extern "C" {
// Training forward colours C[N,3] and the gradients of L = sum <g_i, C_i> with respect to sigma
// [NB,512], features [NB,512,F] and the MLP parameters [T,P] (accumulated into zeroed buffers).
void orc_train_grad_f64(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                        const float* dirs, const float* t_hit, long long N, const double* g, double* C,
                        double* dsigma, double* dfeat, double* dmlp) {
  train_impl<double>(sc, pc, origin, ostride, dirs, t_hit, N, g, C, dsigma, dfeat, dmlp);
}
void orc_train_grad_f32(const orc_scene_c* sc, const orc_params_c* pc, const float* origin, int ostride,
                        const float* dirs, const float* t_hit, long long N, const float* g, float* C, float* dsigma,
                        float* dfeat, float* dmlp) {
  train_impl<float>(sc, pc, origin, ostride, dirs, t_hit, N, g, C, dsigma, dfeat, dmlp);
}

// trilinear at M continuous voxel coordinates of one tile (sigma + features), for unit tests
void orc_trilinear_f32(const orc_scene_c* sc, int tile, const float* f, long long M, float* sig, float* feat) {
  Scene S = to_scene(sc);
  for (long long i = 0; i < M; ++i)
    trilinear<float>(S, tile, v3(f[3 * i], f[3 * i + 1], f[3 * i + 2]), sig + i, feat ? feat + i * S.feat_dim : nullptr);
}

// MLP on M input rows x[M, mlp_in] of one tile -> rgb[M,3]
void orc_mlp_f64(const orc_scene_c* sc, int tile, const double* x, long long M, double* rgb) {
  Scene S = to_scene(sc);
  std::vector<double> a, b;
  for (long long i = 0; i < M; ++i) mlp_eval<double>(S, tile, x + i * S.mlp_in, rgb + 3 * i, a, b);
}

void orc_view_encoding_f64(const double* d, long long M, double* e) {
  for (long long i = 0; i < M; ++i) view_encoding<double>(v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), e + 15 * i);
}

// A9: nearest triangle hit per ray (brute force), rays with per-ray origins
void orc_mesh_hit_f32(const float* V, const int* Fc, int nf, const float* orig, const float* dirs, long long N,
                      float* t, int* id) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < N; ++i) {
    float tt;
    int ii;
    mesh_hit_one<float>(V, Fc, nf, v3(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]),
                        v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), &tt, &ii);
    t[i] = tt;
    id[i] = ii;
  }
}

void orc_normalize_f32(const float* d, long long N, float* out) {
  for (long long i = 0; i < N; ++i) {
    V3<float> n = normalize3(v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
    out[3 * i] = n.x; out[3 * i + 1] = n.y; out[3 * i + 2] = n.z;
  }
}

// A6: inverse_z_sampling_kernel (rendering/src/sampling.cu:5-46).  For pixel p and point j:
// z = 1/((offset+j)*ib[p] + in[p]), next = 1/((offset+j+1)*ib[p] + in[p]), step = max(next - z, 0),
// point = origin + z * d[p].
void orc_inverse_z_sampling_f32(const float* origin, const float* dirs, const float* inv_near,
                                const float* inv_bound, long long npix, int offset, int spr, float* points,
                                float* steps) {
  for (long long t = 0; t < npix * spr; ++t) {
    const long long p = t / spr;
    const int j = (int)(t % spr);
    const float z = 1.0f / ((float)(offset + j) * inv_bound[p] + inv_near[p]);
    const float nz = 1.0f / ((float)(offset + j + 1) * inv_bound[p] + inv_near[p]);
    steps[t] = std::max(nz - z, 0.0f);
    for (int c = 0; c < 3; ++c) points[t * 3 + c] = origin[c] + z * dirs[p * 3 + c];
  }
}

// A5: integrate_points_kernel (rendering/src/volume_rendering.cu:5-58), 16 samples per pixel,
// stateful T and colour across calls; skips pixels with T < early_stop.
void orc_integrate_points_f32(const float* res, const float* steps, long long npix, float early, float* T,
                              float* color) {
  for (long long p = 0; p < npix; ++p) {
    float t = T[p];
    if (t < early) continue;
    float c[3] = {0, 0, 0};
    for (int i = 0; i < 16; ++i) {
      const float* v = res + (p * 16 + i) * 4;
      const float alpha = 1.0f - orc_expf(-1.0f * v[3] * steps[p * 16 + i]);
      for (int k = 0; k < 3; ++k) c[k] = c[k] + alpha * t * v[k];
      t = t * (1 - alpha);
    }
    T[p] = t;
    for (int k = 0; k < 3; ++k) color[p * 3 + k] += c[k];
  }
}

// A8: assign_pixels_kernel (rendering/src/reorder.cu:25-67)
void orc_assign_pixels(const int16_t* net, const int* pixel_starts, const int* block_starts, int nnet,
                       int16_t* new_net, int* new_starts) {
  for (int i = 0; i < nnet; ++i) {
    const int start_loc = block_starts[i];
    int ps = pixel_starts[i];
    const int npx = pixel_starts[i + 1] - pixel_starts[i];
    const int nb = block_starts[i + 1] - start_loc;
    if (nb <= 0) continue;
    const int base = npx / nb, add = npx - base * nb;
    for (int b = 0; b < nb; ++b) {
      new_net[start_loc + b] = net[i];
      new_starts[start_loc + b] = ps;
      ps += (b < add) ? base + 1 : base;
    }
  }
}

}  // extern "C"

// ------------------------------------------------------------------ A3: reference octree format
// rendering_diffuse_octree_fp16_kernel (rendering/src/raydiffuse_render_octree_fp16.cu:141-235) with
// block_traversal_rendering (20-139) and the FP16 trilinear_base (rendering/src/include/half_ope.h:
// 246-284), restated in FP32 with the reference's operation order: float3 / s is a multiply by the
// float reciprocal 1/s (cutil_math.h:397-401), make_int3 truncates, the march advances by
// zval += sample_step.  Arithmetic the reference does in FP16 (interpolation, alpha, weights,
// colour, transmittance) is FP32 here, exp is orc_expf.  Three guards the reference does not have:
// nodes_IndexMap entries of -1 read as an all-zero voxel, voxel / corner coordinates are clamped
// into the dilated node grid, and one block's march takes at most 2^16 steps.
namespace orc {

struct OctreeC {
  const int32_t* group_map;       // [n_tiles]
  const uint16_t* data_voxels;    // half4 [*]
  const int32_t* block_index_map; // [n_tiles * 64]
  const int16_t* nodes_index_map; // [n_nodes * nv^3]
  const uint8_t* nodes_sample_flag; // [n_nodes * nv^3]
  const int64_t* voxels_start;    // [n_nodes]
  const float* centers;           // [n_tiles * 3]
  float tile_size, voxel_size, sample_step;
  int num_voxel;
};

static inline void octree_voxel(const OctreeC& S, int64_t start, const int16_t* nodes, int idx, float* v) {
  const int k = nodes[idx];
  if (k < 0) { v[0] = v[1] = v[2] = v[3] = 0.0f; return; }
  const uint16_t* p = S.data_voxels + 4 * (start + k);
  for (int c = 0; c < 4; ++c) v[c] = half_to_float(p[c]);
}

static inline float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static void octree_trilinear(const OctreeC& S, int64_t start, const int16_t* nodes, const float* fidx, float* out) {
  const int M = S.num_voxel, MYZ = M * M;
  int i0[3];
  float w[3];
  for (int a = 0; a < 3; ++a) {
    i0[a] = std::min(std::max((int)fidx[a], 0), M - 2);
    w[a] = fidx[a] - (float)(int)fidx[a];
    if (w[a] < 0.0f) w[a] = 0.0f;
  }
  float c[8][4];
  for (int q = 0; q < 8; ++q) {
    const int x = i0[0] + (q & 1), y = i0[1] + ((q >> 1) & 1), z = i0[2] + (q >> 2);
    octree_voxel(S, start, nodes, x * MYZ + y * M + z, c[q]);
  }
  for (int ch = 0; ch < 4; ++ch) {
    const float s = lerpf(lerpf(c[0][ch], c[1][ch], w[0]), lerpf(c[2][ch], c[3][ch], w[0]), w[1]);
    const float t = lerpf(lerpf(c[4][ch], c[5][ch], w[0]), lerpf(c[6][ch], c[7][ch], w[0]), w[1]);
    out[ch] = lerpf(s, t, w[2]);
  }
}

struct OctreeRay {
  float color[3];
  float T;
  bool first_hit;
  int32_t n_samples;
};

static void octree_blocks(const OctreeC& S, const float* o, const float* d, int tile, float tile_near, int group,
                          float trans_th, OctreeRay& R, float* inv_near, int16_t* net) {
  constexpr int NB = 4;
  const int M = S.num_voxel, total = M * M * M;
  const float bs = S.tile_size / (float)NB;
  const float inv_bs = 1.0f / bs, inv_vs = 1.0f / S.voxel_size;
  int cur[3], step[3];
  float tmax[3], tdel[3];
  for (int a = 0; a < 3; ++a) {
    cur[a] = (int)(o[a] * inv_bs);
    step[a] = d[a] >= 0.0f ? 1 : -1;
    float nb = (float)(cur[a] + step[a]) * bs;
    if (step[a] < 0) nb += bs;
    tmax[a] = safe_div(nb - o[a], d[a]);
    tdel[a] = safe_div(bs * (float)step[a], d[a]);
  }
  float near = 0.0f, far = std::fmin(std::fmin(tmax[0], tmax[1]), tmax[2]);
  while (true) {
    const int bidx = cur[0] * NB * NB + cur[1] * NB + cur[2];
    const int nidx = S.block_index_map[tile * NB * NB * NB + bidx];
    if (nidx != -1) {
      const int16_t* nodes = S.nodes_index_map + (int64_t)nidx * total;
      const uint8_t* flags = S.nodes_sample_flag + (int64_t)nidx * total;
      const int64_t start = S.voxels_start[nidx];
      float corner[3];
      for (int a = 0; a < 3; ++a) corner[a] = (float)cur[a] * bs - S.voxel_size;
      float zval = near;
      for (int guard = 0; zval < far && guard < (1 << 16); ++guard) {  // step cap of the GPU kernel
        float rel[3];
        int vl[3];
        for (int a = 0; a < 3; ++a) {
          const float p = o[a] + zval * d[a];
          rel[a] = (p - corner[a]) * inv_vs;
          vl[a] = std::min(std::max((int)rel[a], 0), M - 1);
        }
        const int vidx = vl[0] * M * M + vl[1] * M + vl[2];
        if (!flags[vidx]) {
          zval += S.sample_step;
          continue;
        }
        if (!R.first_hit && R.T < trans_th) {
          R.first_hit = true;
          *inv_near = 1.0f / (zval + tile_near);
          *net = (int16_t)group;
        }
        float fidx[3] = {rel[0] - 0.5f, rel[1] - 0.5f, rel[2] - 0.5f};
        float v[4];
        octree_trilinear(S, start, nodes, fidx, v);
        const float interval = std::fmin(S.sample_step, far - zval);
        const float alpha = 1.0f - orc_expf(-1.0f * v[3] * interval);
        const float wgt = alpha * R.T;
        for (int c = 0; c < 3; ++c) R.color[c] = R.color[c] + wgt * v[c];
        R.T = R.T * (1.0f - alpha);
        ++R.n_samples;
        zval += S.sample_step;
      }
    }
    if (tmax[0] < tmax[1]) {
      if (tmax[0] < tmax[2]) { cur[0] += step[0]; tmax[0] += tdel[0]; }
      else { cur[2] += step[2]; tmax[2] += tdel[2]; }
    } else {
      if (tmax[1] < tmax[2]) { cur[1] += step[1]; tmax[1] += tdel[1]; }
      else { cur[2] += step[2]; tmax[2] += tdel[2]; }
    }
    near = far;
    far = std::fmin(std::fmin(tmax[0], tmax[1]), tmax[2]);
    if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= NB || cur[1] >= NB || cur[2] >= NB) break;
  }
}

}  // namespace orc
extern "C" {

void orc_render_octree_f32(const int32_t* group_map, const uint16_t* data_voxels, const int32_t* block_index_map,
                           const int16_t* nodes_index_map, const uint8_t* nodes_sample_flag,
                           const int64_t* voxels_start, const float* centers, float tile_size, float voxel_size,
                           float sample_step, int num_voxel, const float* origin, const float* rays_d,
                           const int32_t* visited, long long n, int K, float trans_th, float early_stop,
                           int soft_hit, float* frame, float* inverse_near, int16_t* net_idx, int32_t* n_samples) {
  const orc::OctreeC S{group_map, data_voxels, block_index_map, nodes_index_map, nodes_sample_flag, voxels_start,
                       centers, tile_size, voxel_size, sample_step, num_voxel};
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < n; ++r) {
    orc::OctreeRay R{{0.0f, 0.0f, 0.0f}, 1.0f, false, 0};
    const float d[3] = {rays_d[3 * r], rays_d[3 * r + 1], rays_d[3 * r + 2]};
    const orc::V3<float> dv{d[0], d[1], d[2]}, ov{origin[0], origin[1], origin[2]};
    int group = -1;
    float bl = -1.0f, bh = -1.0f;
    for (int i = 0; i < K; ++i) {
      const int tile = visited[r * K + i];
      if (tile == -1) break;
      if (R.T < early_stop) break;
      group = group_map[tile];
      const orc::V3<float> c{centers[3 * tile], centers[3 * tile + 1], centers[3 * tile + 2]};
      const float hs = tile_size / 2.0f;
      orc::ray_aabb(ov, dv, c, orc::V3<float>{hs, hs, hs}, bl, bh);
      if (bl == -1.0f) continue;
      float ol[3];
      const float cc[3] = {c.x, c.y, c.z};
      for (int a = 0; a < 3; ++a) {
        const float corner = cc[a] - hs;
        ol[a] = orc::clampr(origin[a] + bl * d[a] - corner, 0.0f, tile_size - 0.00001f);
      }
      orc::octree_blocks(S, ol, d, tile, bl, group, trans_th, R, inverse_near + r, net_idx + r);
    }
    if (soft_hit && !R.first_hit && group != -1) {
      inverse_near[r] = 1.0f / bh;
      net_idx[r] = (int16_t)group;
    }
    for (int c = 0; c < 3; ++c) frame[3 * r + c] = R.color[c];
    if (n_samples) n_samples[r] = R.n_samples;
  }
}

}  // extern "C"

// ------------------------------------------------------------------ training samplers
// Restatement of training/src/Sample_kernel.cu (Sample_sparse 181-395, Sample_reflection 16-137,
// Sample_uniform 436-479, Sample_bg 481-542) and the sample helpers of
// training/src/include/cuda_utils.h (inverse_z_sample_bound 61-75, uniform_sample_bound 77-88,
// uniform_sample_bound_v3 101-113), FP32, one ray at a time.  Operation order as written there: a
// float3 divided by a scalar is a multiply by the FP32 reciprocal, make_int3 truncates.
namespace orc {
struct SmpTile {
  float c[3], ts, hs, vs, mn[3];
  int nv;
};
static inline SmpTile smp_tile(float cx, float cy, float cz, float tile_size, int num_voxel) {
  SmpTile t;
  t.c[0] = cx; t.c[1] = cy; t.c[2] = cz;
  t.ts = tile_size;
  t.hs = tile_size * 0.5f;
  t.nv = num_voxel;
  t.vs = tile_size / (float)num_voxel;
  for (int a = 0; a < 3; ++a) t.mn[a] = t.c[a] - t.hs;
  return t;
}
static inline void smp_aabb(const float* o, const float* d, const float* c, float h, float& lo, float& hi) {
  ray_aabb(V3<float>{o[0], o[1], o[2]}, V3<float>{d[0], d[1], d[2]}, V3<float>{c[0], c[1], c[2]},
           V3<float>{h, h, h}, lo, hi);
}
// voxel DDA state over the tile's num_voxel^3 grid from the entry point at t = near
struct SmpDda {
  int cur[3], stp[3];
  float nb[3], tmax[3], tdel[3], ol[3];
};
static inline void smp_dda_init(const SmpTile& T, const float* o, const float* d, float near, SmpDda& D) {
  const float inv = 1.0f / T.vs;
  for (int a = 0; a < 3; ++a) {
    D.ol[a] = clampr(o[a] + near * d[a] - T.mn[a], 0.0f, T.ts - 0.000001f);
    D.cur[a] = std::min(std::max((int)(D.ol[a] * inv), 0), T.nv - 1);
    D.stp[a] = d[a] >= 0.0f ? 1 : -1;
    D.nb[a] = (float)(D.cur[a] + D.stp[a]) * T.vs;
    if (D.stp[a] < 0) D.nb[a] += T.vs;
    D.tmax[a] = safe_div(D.nb[a] - D.ol[a], d[a]);
    D.tdel[a] = safe_div((float)D.stp[a] * T.vs, d[a]);
  }
}
static inline bool smp_dda_step(const SmpTile& T, SmpDda& D) {
  int a;
  if (D.tmax[0] < D.tmax[1]) a = D.tmax[0] < D.tmax[2] ? 0 : 2;
  else a = D.tmax[1] < D.tmax[2] ? 1 : 2;
  D.cur[a] += D.stp[a];
  D.tmax[a] += D.tdel[a];
  return D.cur[0] >= 0 && D.cur[1] >= 0 && D.cur[2] >= 0 && D.cur[0] < T.nv && D.cur[1] < T.nv && D.cur[2] < T.nv;
}
static inline void smp_voxel_bound(const SmpTile& T, const float* o, const float* d, const int* cur, float& lo,
                                   float& hi) {
  float vc[3];
  for (int a = 0; a < 3; ++a) vc[a] = T.mn[a] + (float)cur[a] * T.vs + 0.5f * T.vs;
  smp_aabb(o, d, vc, 0.5f * T.vs, lo, hi);
}
// uniform_sample_bound_v3 writing at most `room` entries (the reference writes n; n <= room there
// except for rows shorter than 8 samples)
static inline void smp_uniform_v3(float* z, float* dist, float near, float far, int n, int room) {
  const float interval = (far - near) / (float)n;
  for (int i = 0; i < n && i < room; ++i) {
    z[i] = near + (float)i * interval;
    dist[i] = interval;
  }
}
static inline void smp_uniform(float* z, float near, float far, int n) {
  const float interval = (far - near) / (float)(n - 1);
  for (int i = 0; i < n; ++i) z[i] = near + (float)i * interval;
}
// overlap_interpolation + trilinear_base (training/src/include/interpolation.h:43-134) over a dense
// num_voxel^3 float4 grid; corner indices are clamped into the grid (the reference reads out of
// range for points on the upper face)
static inline void smp_overlap(const float* vox, const SmpTile& T, const float* p, float* out) {
  const float inv = 1.0f / T.vs;
  const int M = T.nv;
  float f[3], w[3];
  int i0[3], i1[3];
  for (int a = 0; a < 3; ++a) {
    const float corner = T.c[a] - T.ts / 2.0f;
    f[a] = (p[a] - corner) * inv + 0.5f;
    i0[a] = (int)f[a];
    i1[a] = i0[a] + 1;
    w[a] = f[a] - (float)i0[a];
    if (w[a] < 0.0f) w[a] = 0.0f;
  }
  auto V = [&](int x, int y, int z, int c) {
    x = std::min(std::max(x, 0), M - 1); y = std::min(std::max(y, 0), M - 1); z = std::min(std::max(z, 0), M - 1);
    return vox[(((int64_t)x * M + y) * M + z) * 4 + c];
  };
  for (int c = 0; c < 4; ++c) {
    const bool ex = i1[0] >= M, ey = i1[1] >= M, ez = i1[2] >= M;
    const float v000 = V(i0[0], i0[1], i0[2], c);
    float r;
    if (ex) {
      if (ey) r = ez ? v000 : lerp(v000, V(i0[0], i0[1], i1[2], c), w[2]);
      else if (ez) r = lerp(v000, V(i0[0], i1[1], i0[2], c), w[1]);
      else r = lerp(lerp(v000, V(i0[0], i1[1], i0[2], c), w[1]), lerp(V(i0[0], i0[1], i1[2], c), V(i0[0], i1[1], i1[2], c), w[1]), w[2]);
    } else if (ey) {
      if (ez) r = lerp(v000, V(i1[0], i0[1], i0[2], c), w[0]);
      else r = lerp(lerp(v000, V(i1[0], i0[1], i0[2], c), w[0]), lerp(V(i0[0], i0[1], i1[2], c), V(i1[0], i0[1], i1[2], c), w[0]), w[2]);
    } else if (ez) {
      r = lerp(lerp(v000, V(i1[0], i0[1], i0[2], c), w[0]), lerp(V(i0[0], i1[1], i0[2], c), V(i1[0], i1[1], i0[2], c), w[0]), w[1]);
    } else {
      const float s = lerp(lerp(v000, V(i1[0], i0[1], i0[2], c), w[0]), lerp(V(i0[0], i1[1], i0[2], c), V(i1[0], i1[1], i0[2], c), w[0]), w[1]);
      const float t = lerp(lerp(V(i0[0], i0[1], i1[2], c), V(i1[0], i0[1], i1[2], c), w[0]), lerp(V(i0[0], i1[1], i1[2], c), V(i1[0], i1[1], i1[2], c), w[0]), w[1]);
      r = lerp(s, t, w[2]);
    }
    out[c] = r;
  }
}
}  // namespace orc

extern "C" {

void orc_sample_sparse(const float* rays_o, const float* rays_d, long long n, const int16_t* flags, float cx, float cy,
                       float cz, float tile_size, int num_voxel, int nsamp, float* z_vals, float* dists) {
  using namespace orc;
  const SmpTile T = smp_tile(cx, cy, cz, tile_size, num_voxel);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < n; ++r) {
    const float* o = rays_o + 3 * r;
    const float* d = rays_d + 3 * r;
    float* z = z_vals + r * nsamp;
    float* ds = dists + r * nsamp;
    float near, far;
    smp_aabb(o, d, T.c, T.hs, near, far);
    if (near == -1.0f && far == -1.0f) continue;
    SmpDda D;
    smp_dda_init(T, o, d, near, D);
    const SmpDda D0 = D;
    float total = 0.0f;
    while (true) {
      const int v = (D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2];
      if (flags[v] == 1) {
        float lo, hi;
        smp_voxel_bound(T, o, d, D.cur, lo, hi);
        total = total + (hi - lo);
      }
      if (!smp_dda_step(T, D)) break;
    }
    if (total == 0.0f) continue;
    D = D0;
    int count = 0;
    bool acc = false, first = false;
    const int bsmp = 4, ins = nsamp - 2 * bsmp;
    while (true) {
      const int v = (D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2];
      if (flags[v] == 1) {
        float lo, hi;
        smp_voxel_bound(T, o, d, D.cur, lo, hi);
        if (lo != -1.0f && hi != -1.0f) {
          if (acc) far = hi;
          else {
            near = lo;
            far = hi;
            acc = true;
            if (!first) {
              first = true;
              smp_uniform_v3(z + count, ds + count, near - T.vs, near - 0.000001f, bsmp, nsamp - count);
              count = std::min(count + bsmp, nsamp);
            }
          }
        }
      } else if (acc) {
        int k = (int)std::ceil((float)ins * (far - near) / total);
        k = std::max(std::min(nsamp - count, k), 0);
        smp_uniform_v3(z + count, ds + count, near, far, k, nsamp - count);
        count = count + k;
        acc = false;
      }
      if (!smp_dda_step(T, D)) break;
    }
    if (acc) {
      int k = (int)std::ceil((float)ins * (far - near) / total);
      k = std::max(std::min(nsamp - count, k), 0);
      smp_uniform_v3(z + count, ds + count, near, far, k, nsamp - count);
      count = count + k;
    }
    const int tail = std::min(nsamp - count, bsmp);
    smp_uniform_v3(z + count, ds + count, far + 0.000001f, far + T.vs, tail, nsamp - count);
  }
}

void orc_sample_uniform(const float* rays_o, const float* rays_d, long long n, float cx, float cy, float cz,
                        float tile_size, int nsamp, float* z_vals) {
  using namespace orc;
  const float c[3] = {cx, cy, cz};
  const float hs = tile_size / 2.0f;
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < n; ++r) {
    float near, far;
    smp_aabb(rays_o + 3 * r, rays_d + 3 * r, c, hs, near, far);
    smp_uniform(z_vals + r * nsamp, near, far, nsamp);
  }
}

void orc_sample_bg(const float* rays_o, const float* rays_d, const float* bg_zval, long long n, const float* tile_center,
                   float tile_size, float sample_range, int nsamp, float* z_vals) {
  using namespace orc;
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < n; ++r) {
    const float bg = bg_zval[r];
    if (bg == -1.0f) continue;
    const float half_range = sample_range / 2.0f;
    float lo, hi;
    smp_aabb(rays_o + 3 * r, rays_d + 3 * r, tile_center, tile_size / 2.0f, lo, hi);
    float near = bg - half_range, far = bg + half_range;
    if (near <= hi) {
      near = hi + 0.00001f;
      far = near + sample_range;
    }
    smp_uniform(z_vals + r * nsamp, near, far, nsamp);
  }
}

void orc_sample_reflection(const float* rays_o, const float* rays_d, long long n, const int16_t* flags,
                           const float* voxels, float cx, float cy, float cz, float tile_size, float far, int num_voxel,
                           float T_th, int nsamp, float* z_vals) {
  using namespace orc;
  const SmpTile T = smp_tile(cx, cy, cz, tile_size, num_voxel);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < n; ++r) {
    const float* o = rays_o + 3 * r;
    const float* d = rays_d + 3 * r;
    float bl, bh;
    smp_aabb(o, d, T.c, T.hs, bl, bh);
    if (bl == -1.0f && bh == -1.0f) continue;
    SmpDda D;
    smp_dda_init(T, o, d, bl, D);
    float trans = 1.0f, near = -1.0f;
    while (true) {
      const int v = (D.cur[0] * T.nv + D.cur[1]) * T.nv + D.cur[2];
      if (flags[v] == 1) {
        float lo, hi;
        smp_voxel_bound(T, o, d, D.cur, lo, hi);
        if (lo != -1.0f && hi != -1.0f) {
          float p[3], rgba[4];
          for (int a = 0; a < 3; ++a) p[a] = o[a] + lo * d[a];
          smp_overlap(voxels, T, p, rgba);
          const float dist = hi - lo;
          const float alpha = 1.0f - orc_expf(-1.0f * rgba[3] * dist);
          trans = trans * (1.0f - alpha);
          if (trans <= T_th) {
            near = hi;
            break;
          }
        }
      }
      if (!smp_dda_step(T, D)) break;
    }
    if (near == -1.0f) continue;
    float* z = z_vals + r * nsamp;
    const float inv_near = 1.0f / near, inv_far = 1.0f / far, inv_bound = inv_far - inv_near;
    const float step = 1.0f / (float)(nsamp - 1);
    for (int i = 0; i < nsamp; ++i) z[i] = 1.0f / (step * (float)i * inv_bound + inv_near);
  }
}

}  // extern "C"

extern "C" {
// compute_group_boundary_kernel (rendering/src/compute_group_boundary.cu:4-50), one pixel at a time
void orc_group_boundary(const int16_t* net, int H, int W, int ksize, uint8_t* boundary) {
  const int r = ksize / 2;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)H * W; ++i) {
    const int py = (int)(i / W), px = (int)(i - (long long)py * W);
    const int16_t c = net[i];
    const int x0 = std::max(px - r, 0), x1 = std::min(px + r, W - 1);
    const int y0 = std::max(py - r, 0), y1 = std::min(py + r, H - 1);
    for (int y = y0; y <= y1 && !boundary[i]; ++y)
      for (int x = x0; x <= x1; ++x)
        if (net[(long long)y * W + x] != c) { boundary[i] = 1; break; }
  }
}
}  // extern "C"

// ------------------------------------------------------------------ A10 dense FP32 marcher
// Restatement of render_diffuse_kernel (training/src/cuda_func.cu:1687-1802) with the OVERLAP
// interpolation (training/src/include/interpolation.h:120-134) and trilinear_base's upper-edge
// degradation (interpolation.h:41-116).  Grids: nv^3 cells of float4 (r, g, b, sigma) per group slot,
// cell (x*nv + y)*nv + z; flags int16 per cell, only 1 is sampled.  Out-of-grid cells (not reachable
// from inside a tile) read as unflagged / clamped, the device's guard.
namespace orc {

struct RGBA {
  float v[4];
};

static inline RGBA lerp_rgba(const RGBA& a, const RGBA& b, float t) {   // interpolate1D
  RGBA r;
  for (int c = 0; c < 4; ++c) r.v[c] = a.v[c] * (1.0f - t) + b.v[c] * t;
  return r;
}

static RGBA dense_overlap(const float* vox, int M, float px, float py, float pz) {
  int i0[3] = {(int)px, (int)py, (int)pz};
  float f[3] = {px - (float)i0[0], py - (float)i0[1], pz - (float)i0[2]};
  bool edge[3];
  int c0[3], c1[3];
  for (int a = 0; a < 3; ++a) {
    if (f[a] < 0.0f) f[a] = 0.0f;
    edge[a] = i0[a] + 1 >= M;
    c0[a] = std::min(std::max(i0[a], 0), M - 1);
    c1[a] = std::min(std::max(i0[a] + 1, 0), M - 1);
  }
  auto cell = [&](int x, int y, int z) {
    RGBA r;
    const float* p = vox + 4 * (((long long)x * M + y) * M + z);
    for (int c = 0; c < 4; ++c) r.v[c] = p[c];
    return r;
  };
  // axes that stay: interpolate along x, then y, then z over the non-degenerate axes only
  const int xs[2] = {c0[0], c1[0]}, ys[2] = {c0[1], c1[1]}, zs[2] = {c0[2], c1[2]};
  const int nx = edge[0] ? 1 : 2, ny = edge[1] ? 1 : 2, nz = edge[2] ? 1 : 2;
  RGBA plane[2][2];
  for (int zi = 0; zi < nz; ++zi)
    for (int yi = 0; yi < ny; ++yi)
      plane[zi][yi] = nx == 2 ? lerp_rgba(cell(xs[0], ys[yi], zs[zi]), cell(xs[1], ys[yi], zs[zi]), f[0])
                              : cell(xs[0], ys[yi], zs[zi]);
  RGBA line[2];
  for (int zi = 0; zi < nz; ++zi) line[zi] = ny == 2 ? lerp_rgba(plane[zi][0], plane[zi][1], f[1]) : plane[zi][0];
  return nz == 2 ? lerp_rgba(line[0], line[1], f[2]) : line[0];
}

}  // namespace orc

extern "C" {

// OVERLAP interpolation at world points p (tile centre c, edge tile_size, cell size vs)
void orc_dense_overlap_f32(const float* vox, int M, const float* c, float tile_size, float vs, const float* p,
                           long long n, float* out) {
  const float half = tile_size / 2.0f, inv = 1.0f / vs;
  for (long long i = 0; i < n; ++i) {
    float q[3];
    for (int a = 0; a < 3; ++a) q[a] = (p[3 * i + a] - (c[a] - half)) * inv + 0.5f;
    const orc::RGBA r = orc::dense_overlap(vox, M, q[0], q[1], q[2]);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = r.v[k];
  }
}

void orc_render_dense_f32(const float* voxels, const int16_t* nodes, const int32_t* sparse_to_group,
                          const float* centers, float tile_size, float voxel_size, float sample_step, int M,
                          const float* rays_o, const float* rays_d, const int32_t* visited, long long n, int K,
                          float* diffuse, float* depth, int32_t* valid) {
  const long long cells = (long long)M * M * M;
#pragma omp parallel for schedule(dynamic, 64)
  for (long long r = 0; r < n; ++r) {
    const orc::V3<float> o{rays_o[3 * r], rays_o[3 * r + 1], rays_o[3 * r + 2]};
    const orc::V3<float> d{rays_d[3 * r], rays_d[3 * r + 1], rays_d[3 * r + 2]};
    float col[3] = {0.0f, 0.0f, 0.0f}, T = 1.0f, dep = 0.0f;
    int hit_tile = -1;
    const float half = tile_size / 2.0f, inv = 1.0f / voxel_size;
    for (int j = 0; j < K; ++j) {
      const int tile = visited[r * K + j];
      if (tile == -1 || T < 0.00001f) break;
      const int slot = sparse_to_group[tile];
      if (slot == -1) {
        hit_tile = -1;
        break;
      }
      const float* vox = voxels + 4 * cells * slot;
      const int16_t* fl = nodes + cells * slot;
      const float cc[3] = {centers[3 * tile], centers[3 * tile + 1], centers[3 * tile + 2]};
      float near, far;
      orc::ray_aabb(o, d, orc::V3<float>{cc[0], cc[1], cc[2]}, orc::V3<float>{half, half, half}, near, far);
      float z = near;
      while (z < far) {
        const float p[3] = {o.x + z * d.x, o.y + z * d.y, o.z + z * d.z};
        int idx[3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
          idx[a] = (int)((p[a] - (cc[a] - half - voxel_size)) * inv);
          inside = inside && idx[a] >= 0 && idx[a] < M;
        }
        if (!inside || fl[((long long)idx[0] * M + idx[1]) * M + idx[2]] != 1) {
          z = z + sample_step;
          continue;
        }
        if (hit_tile == -1 && T <= 0.5f) {
          dep = z;
          hit_tile = tile;
        }
        float q[3];
        for (int a = 0; a < 3; ++a) q[a] = (p[a] - (cc[a] - half)) * inv + 0.5f;
        const orc::RGBA v = orc::dense_overlap(vox, M, q[0], q[1], q[2]);
        const float interval = std::min(sample_step, far - z);
        const float alpha = 1.0f - orc_expf(-1.0f * v.v[3] * interval);
        const float w = alpha * T;
        for (int c = 0; c < 3; ++c) col[c] = col[c] + w * v.v[c];
        T = T * (1.0f - alpha);
        if (T < 0.00001f) break;
        z = z + sample_step;
      }
    }
    valid[r] = hit_tile != -1 ? 1 : 0;
    depth[r] = hit_tile != -1 ? dep : -1.0f;
    for (int c = 0; c < 3; ++c) diffuse[3 * r + c] = col[c];
  }
}

}  // extern "C"

// ------------------------------------------------------------------ A9 per-tile first / double hit
// Restatement of ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel (training/src/cuda_func.cu:
// 130-444) without face lists: in every block of the walk, every face whose bounding box comes
// within 2e-5 of the block is tested (a face farther away cannot produce a hit point within the
// kernel's 1e-5 acceptance band), in face order, with the reference's acceptance and update rules.
extern "C" {

void orc_mesh_hit_tiles_f32(int mode, const float* V, const int* Fc, int nf, const float* centers, float tile_size,
                            int nb, const float* rays_o, const float* rays_d, long long n_rays, int K,
                            const int32_t* visited, float* hits, int32_t* hit_face) {
  std::vector<float> fb(6 * (size_t)nf);   // face bounding boxes
  for (int f = 0; f < nf; ++f)
    for (int a = 0; a < 3; ++a) {
      float lo = V[3 * Fc[3 * f] + a], hi = lo;
      for (int k = 1; k < 3; ++k) {
        lo = std::min(lo, V[3 * Fc[3 * f + k] + a]);
        hi = std::max(hi, V[3 * Fc[3 * f + k] + a]);
      }
      fb[6 * f + a] = lo;
      fb[6 * f + 3 + a] = hi;
    }
  const float BIG = 100000000.0f;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long it = 0; it < n_rays * K; ++it) {
    if (mode == 1) {
      hits[2 * it] = -1.0f;
      hits[2 * it + 1] = -1.0f;
    } else {
      hit_face[it] = -1;
    }
    const int tile = visited[it];
    if (tile == -1) continue;
    const long long r = it / K;
    const orc::V3<float> o{rays_o[3 * r], rays_o[3 * r + 1], rays_o[3 * r + 2]};
    const orc::V3<float> d{rays_d[3 * r], rays_d[3 * r + 1], rays_d[3 * r + 2]};
    const float bs = tile_size / (float)nb, half = tile_size / 2.0f, hb = bs / 2.0f;
    const float c[3] = {centers[3 * tile], centers[3 * tile + 1], centers[3 * tile + 2]};
    const float mc[3] = {c[0] - half, c[1] - half, c[2] - half};
    float bl, bh;
    orc::ray_aabb(o, d, orc::V3<float>{c[0], c[1], c[2]}, orc::V3<float>{half, half, half}, bl, bh);
    const float od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    int cur[3], step[3];
    float tmax[3], tdel[3];
    const float inv_bs = 1.0f / bs;
    for (int a = 0; a < 3; ++a) {
      const float no = orc::clampr(od[a] + bl * dd[a] - mc[a], 0.0f, tile_size - 0.000001f);
      cur[a] = std::min((int)(no * inv_bs), nb - 1);
      step[a] = dd[a] >= 0.0f ? 1 : -1;
      float nbd = (float)(cur[a] + step[a]) * bs;
      if (step[a] < 0) nbd += bs;
      tmax[a] = orc::safe_div(nbd - no, dd[a]);
      tdel[a] = orc::safe_div((float)step[a] * bs, dd[a]);
    }
    float near = BIG, second = BIG;
    int face = -1;
    for (;;) {
      float bc[3];
      for (int a = 0; a < 3; ++a) bc[a] = mc[a] + (float)cur[a] * bs + hb;
      for (int f = 0; f < nf; ++f) {
        bool close = true;
        for (int a = 0; a < 3; ++a)
          close = close && fb[6 * f + a] <= bc[a] + hb + 2e-5f && fb[6 * f + 3 + a] >= bc[a] - hb - 2e-5f;
        if (!close) continue;
        const int* ix = Fc + 3 * f;
        const orc::V3<float> A{V[3 * ix[0]], V[3 * ix[0] + 1], V[3 * ix[0] + 2]};
        const orc::V3<float> B{V[3 * ix[1]], V[3 * ix[1] + 1], V[3 * ix[1] + 2]};
        const orc::V3<float> Cc{V[3 * ix[2]], V[3 * ix[2] + 1], V[3 * ix[2] + 2]};
        const float t = orc::ray_tri(o, d, A, B, Cc);
        if (!(t > 0.0f && t < (mode == 1 ? second : near))) continue;
        bool inside = true;
        for (int a = 0; a < 3; ++a) inside = inside && std::fabs(od[a] + t * dd[a] - bc[a]) - hb <= 0.00001f;
        if (!inside) continue;
        if (mode == 1) {
          if (t < near) {
            second = near;
            near = t;
          } else {
            second = t;
          }
        } else {
          near = t;
          face = f;
        }
      }
      if (mode == 1 ? second != BIG : near != BIG) break;
      const int ax = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
      cur[ax] += step[ax];
      tmax[ax] += tdel[ax];
      if (cur[ax] < 0 || cur[ax] >= nb) break;
    }
    if (mode == 1) {
      if (near != BIG) {
        hits[2 * it] = near;
        if (second != BIG) hits[2 * it + 1] = second;
      }
    } else if (near != BIG) {
      hits[it] = near;
      hit_face[it] = face;
    }
  }
}

}  // extern "C"

// ------------------------------------------------------------------ A5 / A7 training forms
// utils.volume_rendering (training/scripts/utils.py:823-837) in FP64 and its adjoint, per ray over
// n samples: alpha = 1 - exp(-sigma d), trans = init * exclusive cumprod(1 - alpha + 1e-10),
// out = sum alpha trans rgb.  The adjoint is the reverse recursion of the cumprod (the gradient of
// L = <g_out, out> + <g_trans, trans> + <g_alpha, alpha>).
extern "C" {

void orc_volume_render_f64(const double* rgb, const double* sigma, const double* dists, const double* init_T,
                           long long n_rays, int n, const double* g_out, const double* g_trans,
                           const double* g_alpha, double* out, double* trans, double* alpha, double* g_rgb,
                           double* g_sigma, double* g_init) {
  for (long long r = 0; r < n_rays; ++r) {
    double T = init_T ? init_T[r] : 1.0;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) {
      const long long k = r * n + i;
      const double a = 1.0 - std::exp(-(sigma[k] * dists[k]));
      for (int ch = 0; ch < 3; ++ch) c[ch] += a * T * rgb[3 * k + ch];
      trans[k] = T;
      alpha[k] = a;
      T = T * (1.0 - a + 1e-10);
    }
    for (int ch = 0; ch < 3; ++ch) out[3 * r + ch] = c[ch];
    if (!g_out) continue;
    double G = 0.0;  // dL/dT_{i+1}
    for (int i = n - 1; i >= 0; --i) {
      const long long k = r * n + i;
      const double Ti = trans[k], a = alpha[k], e = std::exp(-(sigma[k] * dists[k]));
      double gc = 0.0;
      for (int ch = 0; ch < 3; ++ch) {
        gc += g_out[3 * r + ch] * rgb[3 * k + ch];
        g_rgb[3 * k + ch] = g_out[3 * r + ch] * a * Ti;
      }
      const double d_alpha = Ti * gc + (g_alpha ? g_alpha[k] : 0.0);
      const double d_e = -d_alpha + Ti * G;
      g_sigma[k] = -dists[k] * e * d_e;
      G = a * gc + (g_trans ? g_trans[k] : 0.0) + (e + 1e-10) * G;
    }
    if (g_init) g_init[r] = G;
  }
}

// specular head of Tile.inference_specular (training/scripts/tile/tile.py:212-219, sh.py:55-111):
// spec = sigmoid(sum_b Y_b(d) coeffs[3 b + c]) at the normalised direction of each sample's ray, and
// the adjoint dL/dcoeffs[3 b + c] = g_c spec_c (1 - spec_c) Y_b (degree <= 2)
void orc_sh_sigmoid_f64(const double* coeffs, int stride, const double* dirs, long long n, int spr, int deg,
                        const double* g_spec, double* spec, double* g_coeffs) {
  for (long long s = 0; s < n; ++s) {
    const double* d = dirs + 3 * (s / spr);
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double u[3] = {d[0] / len, d[1] / len, d[2] / len};
    double b[9];
    orc::sh2_basis(u, b);
    const int nb = (deg + 1) * (deg + 1);
    for (int j = nb; j < 9; ++j) b[j] = 0.0;
    for (int c = 0; c < 3; ++c) {
      double z = 0.0;
      for (int j = 0; j < nb; ++j) z += b[j] * coeffs[s * stride + 3 * j + c];
      const double v = 1.0 / (1.0 + std::exp(-z));
      spec[3 * s + c] = v;
      if (g_spec)
        for (int j = 0; j < 9; ++j) g_coeffs[s * stride + 3 * j + c] = g_spec[3 * s + c] * v * (1.0 - v) * b[j];
    }
  }
}

}  // extern "C"

// networks.MLP (training/scripts/networks.py:4-41) in FP64 with its adjoint: Positional_Encoding
// with L = 10 ([x, sin(2^0 x), cos(2^0 x), ..., cos(2^9 x)]), depth Linear layers of width 64 with
// LeakyReLU(slope) between them and n_out outputs; params in nn.Linear order (W[out][in], b[out]).
// With g_out, g_params receives the gradient of <g_out, out> (accumulated over the samples).
extern "C" {

void orc_pemlp_f64(const double* params, const double* pts, long long n, int depth, int n_out, double slope,
                   const double* g_out, double* out, double* g_params) {
  const int W = 64, L = 10, in0 = 3 + 6 * L;
  std::vector<long long> off(depth + 1, 0);
  for (int l = 0; l < depth; ++l) {
    const int ni = l == 0 ? in0 : W, no = l == depth - 1 ? n_out : W;
    off[l + 1] = off[l] + (long long)no * ni + no;
  }
  if (g_params) std::fill(g_params, g_params + off[depth], 0.0);
  std::vector<std::vector<double>> act(depth + 1);
  for (long long s = 0; s < n; ++s) {
    std::vector<double>& x0 = act[0];
    x0.assign(in0, 0.0);
    for (int j = 0; j < 3; ++j) x0[j] = pts[3 * s + j];
    for (int f = 0; f < L; ++f)
      for (int j = 0; j < 3; ++j) {
        const double a = std::ldexp(pts[3 * s + j], f);
        x0[3 + 6 * f + j] = std::sin(a);
        x0[6 + 6 * f + j] = std::cos(a);
      }
    for (int l = 0; l < depth; ++l) {
      const int ni = l == 0 ? in0 : W, no = l == depth - 1 ? n_out : W;
      const double* Wl = params + off[l];
      std::vector<double>& y = act[l + 1];
      y.assign(no, 0.0);
      for (int o = 0; o < no; ++o) {
        double acc = Wl[(long long)no * ni + o];
        for (int i = 0; i < ni; ++i) acc += Wl[(long long)o * ni + i] * act[l][i];
        y[o] = (l == depth - 1) ? acc : (acc > 0.0 ? acc : slope * acc);
      }
    }
    for (int o = 0; o < n_out; ++o) out[s * n_out + o] = act[depth][o];
    if (!g_out) continue;
    std::vector<double> d(g_out + s * n_out, g_out + (s + 1) * n_out);
    for (int l = depth - 1; l >= 0; --l) {
      const int ni = l == 0 ? in0 : W, no = l == depth - 1 ? n_out : W;
      const double* Wl = params + off[l];
      double* Gl = g_params + off[l];
      for (int o = 0; o < no; ++o) {
        for (int i = 0; i < ni; ++i) Gl[(long long)o * ni + i] += d[o] * act[l][i];
        Gl[(long long)no * ni + o] += d[o];
      }
      if (l == 0) break;
      std::vector<double> dp(ni, 0.0);
      for (int i = 0; i < ni; ++i) {
        double acc = 0.0;
        for (int o = 0; o < no; ++o) acc += Wl[(long long)o * ni + i] * d[o];
        dp[i] = acc * (act[l][i] > 0.0 ? 1.0 : slope);   // act[l] = leaky(z): same sign as z
      }
      d.swap(dp);
    }
  }
}

}  // extern "C"

// Tile.query_voxel (training/scripts/tile/tile.py:225-236) in FP64: F.grid_sample (trilinear,
// align_corners=True, zero padding) of a channel-last grid vox[D][D][D][4] at world points, through
// the reference's coordinate chain.  With g_rgb / g_sigma, g_vox (zeroed here) receives the gradient.
extern "C" {

void orc_voxel_query_f64(const double* vox, int D, const double* pts, long long n, const double* mc, double vs,
                         const double* g_rgb, const double* g_sigma, double* rgb, double* sigma, double* g_vox) {
  const double dm1 = D - 1;
  if (g_vox) std::fill(g_vox, g_vox + 4LL * D * D * D, 0.0);
  for (long long s = 0; s < n; ++s) {
    double u[3];
    long long c0[3];
    for (int a = 0; a < 3; ++a) {
      const double idx = (pts[3 * s + a] - mc[a]) / vs - 0.5;
      const double g = idx / dm1 * 2.0 - 1.0;
      u[a] = (g + 1.0) / 2.0 * dm1;
      c0[a] = (long long)std::floor(u[a]);
    }
    double acc[4] = {0, 0, 0, 0};
    const double gg[4] = {g_rgb ? g_rgb[3 * s] : 0.0, g_rgb ? g_rgb[3 * s + 1] : 0.0, g_rgb ? g_rgb[3 * s + 2] : 0.0,
                          g_sigma ? g_sigma[s] : 0.0};
    for (int j = 0; j < 8; ++j) {
      const long long x = c0[0] + (j & 1), y = c0[1] + ((j >> 1) & 1), z = c0[2] + (j >> 2);
      if (x < 0 || x >= D || y < 0 || y >= D || z < 0 || z >= D) continue;
      const double w = (1.0 - std::fabs(u[0] - x)) * (1.0 - std::fabs(u[1] - y)) * (1.0 - std::fabs(u[2] - z));
      const long long cell = (z * D + y) * D + x;
      for (int c = 0; c < 4; ++c) {
        acc[c] += w * vox[4 * cell + c];
        if (g_vox) g_vox[4 * cell + c] += w * gg[c];
      }
    }
    for (int c = 0; c < 3; ++c) rgb[3 * s + c] = acc[c];
    sigma[s] = acc[3];
  }
}

}  // extern "C"
