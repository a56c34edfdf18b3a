// A9 — nearest proxy-mesh hit per ray over the BVH of bvh_build.cpp.
//
// Triangle test: RayTriangleIntersection (rendering/src/include/cuda_utils.h:403-436, blur 0)
// with the two-sided vertex flip of ray_firstHit_inTile_kernel (training/src/cuda_func.cu:215-231);
// the flip uses the sign of dot(d, cross(v1 - v0, v2 - v1)) (the reference normalises that normal
// first, which does not change its sign).  A hit needs t > 0; the nearest t wins, ties go to the
// smaller triangle id, so the result does not depend on the traversal order and equals a brute
// force over all triangles.  Each 64-byte node holds both child boxes, so one node fetch (four
// 16-byte loads of adjacent addresses) tests two children.  A triangle is 64 bytes: its first vertex
// and id, its two edges from that vertex and the edge-cross normal that decides the vertex flip.
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

constexpr int kBvhStack = 64;

__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// t of the hit, or -1e8 on a miss, from the triangle's first vertex a, its edges e1 = b - a and
// e2 = c - a and n = (b - a) x (c - b), the terms that do not depend on the ray (the BVH stores them
// per triangle): the same operations as from the three vertices, so the same t.
__device__ __forceinline__ float ray_triangle_pre(V3 o, V3 d, V3 a, V3 e1, V3 e2, V3 n) {
  const bool flip = dot3(d, n) > 0.0f;
  const V3 v0v1 = flip ? e2 : e1, v0v2 = flip ? e1 : e2, v0o = sub3(o, a);
  const V3 pv = cross3(d, v0v2);
  float det = dot3(v0v1, pv);
  if (det < 0.00000001f) return -100000000.0f;
  det = 1.0f / det;
  const float u = dot3(v0o, pv) * det;
  if (u < 0.0f || u > 1.0f) return -100000000.0f;
  const V3 qv = cross3(v0o, v0v1);
  const float v = dot3(d, qv) * det;
  if (v < 0.0f || v > 1.0f) return -100000000.0f;
  if ((u + v) < 0.0f || (u + v) > 1.0f) return -100000000.0f;
  return dot3(v0v2, qv) * det;
}
__device__ __forceinline__ float ray_triangle(V3 o, V3 d, V3 a, V3 b, V3 c) {
  return ray_triangle_pre(o, d, a, sub3(b, a), sub3(c, a), cross3(sub3(b, a), sub3(c, b)));
}

// triangle j of the BVH's leaf-ordered array (four float4: a + id, e1, e2, n)
__device__ __forceinline__ float bvh_triangle(const float4* __restrict__ tris, int j, V3 o, V3 d, int& id) {
  const float4 t0 = tris[4 * j], t1 = tris[4 * j + 1], t2 = tris[4 * j + 2], t3 = tris[4 * j + 3];
  id = __float_as_int(t0.w);
  return ray_triangle_pre(o, d, v3(t0.x, t0.y, t0.z), v3(t1.x, t1.y, t1.z), v3(t2.x, t2.y, t2.z), v3(t3.x, t3.y, t3.z));
}

// slab test against a (widened) box: entry/exit t; a direction component of 0 uses +-1e30.  Each
// plane's t is one FMA, plane * inv - o * inv with o * inv formed once per ray: it differs from
// (plane - o) * inv by a few ulps of the products, far inside the boxes' widening (a relative 1e-6 of
// the scene extent), so no box whose triangles the exact test would hit is rejected.
__device__ __forceinline__ void slab(const float* b, V3 oi, V3 inv, float& tn, float& tf) {
  const float x0 = __builtin_fmaf(b[0], inv.x, -oi.x), x1 = __builtin_fmaf(b[3], inv.x, -oi.x);
  const float y0 = __builtin_fmaf(b[1], inv.y, -oi.y), y1 = __builtin_fmaf(b[4], inv.y, -oi.y);
  const float z0 = __builtin_fmaf(b[2], inv.z, -oi.z), z1 = __builtin_fmaf(b[5], inv.z, -oi.z);
  tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
  tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
}

__global__ __launch_bounds__(64) void k_mesh_first_hit(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                                       snisr_rays R, float* __restrict__ t_out,
                                                       int32_t* __restrict__ id_out, int packet_w) {
  __shared__ int stack_mem[kBvhStack * 64];
  int* stack = stack_mem + threadIdx.x;  // lane-strided: entry k at stack[k * 64]
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (tid >= R.n) return;
  const int64_t i = ray_of_thread(tid, R.n, packet_w);
  const float* og = R.origins + i * R.origin_stride;
  const V3 o = v3(og[0], og[1], og[2]);
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  const V3 inv = v3(d.x != 0.0f ? 1.0f / d.x : __builtin_copysignf(1e30f, d.x),
                    d.y != 0.0f ? 1.0f / d.y : __builtin_copysignf(1e30f, d.y),
                    d.z != 0.0f ? 1.0f / d.z : __builtin_copysignf(1e30f, d.z));
  const V3 oi = v3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  float best = __builtin_inff();
  int best_id = -1;
  int sp = 0;
  int node = 0;
  while (true) {
    const float4 q0 = nodes[4 * node], q1 = nodes[4 * node + 1], q2 = nodes[4 * node + 2], q3 = nodes[4 * node + 3];
    const float b0[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
    const float b1[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    const int c[2] = {__float_as_int(q3.x), __float_as_int(q3.y)};
    float tn[2], tf[2];
    slab(b0, oi, inv, tn[0], tf[0]);
    slab(b1, oi, inv, tn[1], tf[1]);
    bool hit[2];
    int inner[2], ni = 0;
    float inner_t[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      hit[k] = tn[k] <= tf[k] && tf[k] >= 0.0f && tn[k] <= best;
      if (!hit[k]) continue;
      if (c[k] >= 0) {
        inner[ni] = c[k];
        inner_t[ni] = tn[k];
        ++ni;
      } else {
        const int code = ~c[k];
        const int first = code >> 3, cnt = code & 7;
        for (int j = 0; j < cnt; ++j) {
          int id;
          const float t = bvh_triangle(tris, first + j, o, d, id);
          if (t > 0.0f && (t < best || (t == best && id < best_id))) {
            best = t;
            best_id = id;
          }
        }
      }
    }
    if (ni == 2) {
      const bool swap = inner_t[1] < inner_t[0];
      const int near_n = swap ? inner[1] : inner[0], far_n = swap ? inner[0] : inner[1];
      stack[sp * 64] = far_n;  // sp < kBvhStack: the tree depth is below the stack size
      ++sp;
      node = near_n;
    } else if (ni == 1) {
      node = inner[0];
    } else {
      if (sp == 0) break;
      --sp;
      node = stack[sp * 64];
    }
  }
  t_out[i] = best;
  id_out[i] = best_id;
}

// Packet traversal for coherent rays (camera 8x8 pixel packets): the wavefront walks one shared
// traversal order, the union of the nodes its 64 rays need.  Each lane tests its own ray against
// both child boxes of the current node with its own nearest t; a child is visited when any lane
// hits it, the nearer one first by a lane vote; a leaf's triangles are tested by every lane once any
// lane hits the leaf box.  The stack lives in LDS, one per wavefront, and every lane holds the same
// stack pointer, so node fetches are wave-uniform (one 64-byte line for the whole wavefront).  The
// nearest hit of each ray is the same as in the per-lane traversal: the order of visits never
// changes the minimum t, and ties go to the smaller triangle id.
__global__ __launch_bounds__(64) void k_mesh_first_hit_packet(const float4* __restrict__ nodes,
                                                              const float4* __restrict__ tris, snisr_rays R,
                                                              float* __restrict__ t_out, int32_t* __restrict__ id_out,
                                                              int packet_w) {
  __shared__ int stack[kBvhStack];
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool live = tid < R.n;
  const int64_t i = live ? ray_of_thread(tid, R.n, packet_w) : 0;
  V3 o = v3(0.0f, 0.0f, 0.0f), d = v3(0.0f, 0.0f, 1.0f);
  if (live) {
    const float* og = R.origins + i * R.origin_stride;
    o = v3(og[0], og[1], og[2]);
    d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  }
  const V3 inv = v3(d.x != 0.0f ? 1.0f / d.x : __builtin_copysignf(1e30f, d.x),
                    d.y != 0.0f ? 1.0f / d.y : __builtin_copysignf(1e30f, d.y),
                    d.z != 0.0f ? 1.0f / d.z : __builtin_copysignf(1e30f, d.z));
  const V3 oi = v3(o.x * inv.x, o.y * inv.y, o.z * inv.z);
  float best = live ? __builtin_inff() : -1.0f;  // a dead lane never wants a node
  int best_id = -1;
  int sp = 0;
  int node = 0;
  while (true) {
    const float4 q0 = nodes[4 * node], q1 = nodes[4 * node + 1], q2 = nodes[4 * node + 2], q3 = nodes[4 * node + 3];
    const float b0[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
    const float b1[6] = {q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
    const int c[2] = {__float_as_int(q3.x), __float_as_int(q3.y)};
    float tn[2], tf[2];
    slab(b0, oi, inv, tn[0], tf[0]);
    slab(b1, oi, inv, tn[1], tf[1]);
    bool any[2];
    bool mine[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      mine[k] = tn[k] <= tf[k] && tf[k] >= 0.0f && tn[k] <= best;
      any[k] = __ballot(mine[k]) != 0;
    }
    int inner[2], ni = 0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!any[k]) continue;
      if (c[k] >= 0) {
        inner[ni++] = c[k];
      } else {
        const int code = ~c[k];
        const int first = code >> 3, cnt = code & 7;
        for (int j = 0; j < cnt; ++j) {
          int id;
          const float t = bvh_triangle(tris, first + j, o, d, id);
          if (live && t > 0.0f && (t < best || (t == best && id < best_id))) {
            best = t;
            best_id = id;
          }
        }
      }
    }
    if (ni == 2) {
      // nearer child first, by the vote of the lanes that hit both
      const bool both = mine[0] && mine[1];
      const int n_both = __popcll(__ballot(both));
      const int n_first1 = __popcll(__ballot(both && tn[1] < tn[0]));
      const bool swap = 2 * n_first1 > n_both;
      const int near_n = swap ? inner[1] : inner[0], far_n = swap ? inner[0] : inner[1];
      if (threadIdx.x == 0) stack[sp] = far_n;  // sp < kBvhStack: the tree depth is below the stack size
      ++sp;
      node = near_n;
    } else if (ni == 1) {
      node = inner[0];
    } else {
      if (sp == 0) break;
      --sp;
      __syncthreads();  // one wavefront per workgroup, uniform branch: orders the stack write before the read
      node = stack[sp];
    }
  }
  if (live) {
    t_out[i] = best;
    id_out[i] = best_id;
  }
}

// ---------------------------------------------------------------------------- per-tile form
// ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel (training/src/cuda_func.cu:130-444): one
// thread per (ray, visited tile).  The ray enters the tile's num_block^3 block grid at its AABB entry
// (clamped into the tile) and walks it with the Amanatides-Woo tie rules; in each block it tests the
// block's face list (face_num[block] = (start, count) into face_idx) and accepts hits with t > 0 whose
// point lies inside the block within 1e-5.  First-hit mode stops at the first block with a hit and
// reports the nearest t and its face; double-hit mode keeps the two nearest accepted t and stops at
// the first block after which both are known.
template <bool DOUBLE>
__global__ __launch_bounds__(256) void k_mesh_hit_tiles(snisr_tile_mesh Mh, const float* __restrict__ rays_o,
                                                        const float* __restrict__ rays_d,
                                                        const int32_t* __restrict__ visited, int64_t n_items, int K,
                                                        float* __restrict__ hits, int32_t* __restrict__ hit_face) {
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= n_items) return;
  if (DOUBLE) {
    hits[2 * item] = -1.0f;
    hits[2 * item + 1] = -1.0f;
  } else {
    hit_face[item] = -1;
  }
  const int tile = visited[item];
  if (tile == -1) return;
  const int64_t ray = item / K;
  const int nb = Mh.num_block;
  const int2* fnum = reinterpret_cast<const int2*>(Mh.face_num) + (int64_t)tile * nb * nb * nb;
  const V3 o = v3(rays_o[3 * ray], rays_o[3 * ray + 1], rays_o[3 * ray + 2]);
  const V3 d = v3(rays_d[3 * ray], rays_d[3 * ray + 1], rays_d[3 * ray + 2]);
  const float ts = Mh.tile_size;
  const float bs = ts / (float)nb;
  const float half = ts / 2.0f;
  const V3 tc = v3(Mh.centers[3 * tile], Mh.centers[3 * tile + 1], Mh.centers[3 * tile + 2]);
  const V3 mc = v3(tc.x - half, tc.y - half, tc.z - half);
  float bl, bh;
  ray_aabb(o, d, tc, v3(half, half, half), bl, bh);
  V3 no = v3(o.x + bl * d.x - mc.x, o.y + bl * d.y - mc.y, o.z + bl * d.z - mc.z);
  no = v3(clampx(no.x, 0.0f, ts - 0.000001f), clampx(no.y, 0.0f, ts - 0.000001f), clampx(no.z, 0.0f, ts - 0.000001f));
  const float inv_bs = 1.0f / bs;
  // entry block; the clamp into the grid only matters where rounding of no / bs could reach num_block
  int cur[3] = {min((int)(no.x * inv_bs), nb - 1), min((int)(no.y * inv_bs), nb - 1), min((int)(no.z * inv_bs), nb - 1)};
  int step[3];
  float tmax[3], tdel[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float da = get(d, a);
    step[a] = da >= 0.0f ? 1 : -1;
    float nbd = (float)(cur[a] + step[a]) * bs;
    if (step[a] < 0) nbd += bs;
    tmax[a] = safe_div(nbd - get(no, a), da);
    tdel[a] = safe_div((float)step[a] * bs, da);
  }
  const float hb = bs / 2.0f;
  float near = 100000000.0f, second = 100000000.0f;
  int face = -1;
  while (true) {
    // cur stays inside the block grid: the walk starts inside and ends as soon as it leaves
    const int bidx = cur[2] + cur[1] * nb + cur[0] * nb * nb;
    const V3 bc = v3(mc.x + (float)cur[0] * bs + hb, mc.y + (float)cur[1] * bs + hb, mc.z + (float)cur[2] * bs + hb);
    const int2 fl = fnum[bidx];
    for (int j = 0; j < fl.y; ++j) {
      const int f = Mh.face_idx[fl.x + j];
      const int32_t* fv = Mh.faces + 3 * (int64_t)f;
      const float* a = Mh.vertices + 3 * (int64_t)fv[0];
      const float* b = Mh.vertices + 3 * (int64_t)fv[1];
      const float* c = Mh.vertices + 3 * (int64_t)fv[2];
      const float t = ray_triangle(o, d, v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2]));
      if (t > 0.0f && t < (DOUBLE ? second : near)) {
        const float dx = fabsf(o.x + t * d.x - bc.x) - hb;
        const float dy = fabsf(o.y + t * d.y - bc.y) - hb;
        const float dz = fabsf(o.z + t * d.z - bc.z) - hb;
        if (dx <= 0.00001f && dy <= 0.00001f && dz <= 0.00001f) {
          if (DOUBLE) {
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
      }
    }
    if (DOUBLE ? second != 100000000.0f : near != 100000000.0f) break;
    int ax;
    if (tmax[0] < tmax[1]) ax = tmax[0] < tmax[2] ? 0 : 2;
    else ax = tmax[1] < tmax[2] ? 1 : 2;
    cur[ax] += step[ax];
    tmax[ax] += tdel[ax];
    if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= nb || cur[1] >= nb || cur[2] >= nb) break;
  }
  if (DOUBLE) {
    if (near != 100000000.0f) {
      hits[2 * item] = near;
      if (second != 100000000.0f) hits[2 * item + 1] = second;
    }
  } else if (near != 100000000.0f) {
    hits[item] = near;
    hit_face[item] = face;
  }
}

int packet_width(const snisr_rays& R);  // render_fp32.hip

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_mesh_first_hit(const void* nodes, int32_t n_nodes, const void* tris, int64_t n_tris, int32_t depth,
                         const snisr_rays* rays, float* t_out, int32_t* id_out, void* stream) {
  SNISR_REQUIRE(nodes && tris && rays && t_out && id_out, SNISR_ERR_ARG, "mesh_first_hit: NULL pointer");
  SNISR_REQUIRE(n_nodes >= 1 && n_tris >= 0, SNISR_ERR_ARG, "mesh_first_hit: empty BVH");
  SNISR_REQUIRE(depth >= 0 && depth < kBvhStack, SNISR_ERR_SHAPE, "mesh_first_hit: BVH depth %d >= %d", depth,
                kBvhStack);
  SNISR_REQUIRE(aligned(nodes, 16) && aligned(tris, 16), SNISR_ERR_ALIGN, "mesh_first_hit: 16-byte alignment");
  if (int e = check_rays(rays, nullptr)) return e;
  if (rays->n == 0) return SNISR_OK;
  // camera packets (8x8 pixels per wavefront) share one traversal; other rays walk one per lane
  const int pw = packet_width(*rays);
  if (pw > 0)
    hipLaunchKernelGGL(k_mesh_first_hit_packet, dim3((unsigned)((rays->n + 63) / 64)), dim3(64), 0,
                       (hipStream_t)stream, (const float4*)nodes, (const float4*)tris, *rays, t_out, id_out, pw);
  else
    hipLaunchKernelGGL(k_mesh_first_hit, dim3((unsigned)((rays->n + 63) / 64)), dim3(64), 0, (hipStream_t)stream,
                       (const float4*)nodes, (const float4*)tris, *rays, t_out, id_out, pw);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_mesh_hit_tiles(const snisr_tile_mesh* mesh, const float* rays_o, const float* rays_d, int64_t n_rays,
                         int32_t K, const int32_t* visited, int32_t mode, float* hits, int32_t* hit_face,
                         void* stream) {
  SNISR_REQUIRE(mesh, SNISR_ERR_ARG, "mesh_hit_tiles: mesh descriptor is NULL");
  SNISR_REQUIRE(mode == 0 || mode == 1, SNISR_ERR_ARG, "mesh_hit_tiles: mode must be 0 (first) or 1 (double)");
  SNISR_REQUIRE(n_rays >= 0 && K >= 0, SNISR_ERR_ARG, "mesh_hit_tiles: negative size");
  const int64_t items = n_rays * K;
  if (items == 0) return SNISR_OK;
  SNISR_REQUIRE(rays_o && rays_d && visited && hits && (mode == 1 || hit_face), SNISR_ERR_ARG,
                "mesh_hit_tiles: NULL pointer");
  SNISR_REQUIRE(mesh->centers && mesh->face_num && (mesh->n_face_idx == 0 || mesh->face_idx) &&
                    (mesh->n_faces == 0 || (mesh->faces && mesh->vertices)),
                SNISR_ERR_ARG, "mesh_hit_tiles: mesh tables must be non-NULL");
  SNISR_REQUIRE(mesh->num_block >= 1 && mesh->num_block <= 1024 && mesh->tile_size > 0.0f, SNISR_ERR_ARG,
                "mesh_hit_tiles: num_block in [1, 1024] and a positive tile_size");
  SNISR_REQUIRE(aligned(mesh->face_num, 8), SNISR_ERR_ALIGN, "mesh_hit_tiles: face_num must be 8-byte aligned");
  const dim3 g((unsigned)((items + 255) / 256)), b(256);
  if (mode == 1)
    hipLaunchKernelGGL(k_mesh_hit_tiles<true>, g, b, 0, (hipStream_t)stream, *mesh, rays_o, rays_d, visited, items, K,
                       hits, hit_face);
  else
    hipLaunchKernelGGL(k_mesh_hit_tiles<false>, g, b, 0, (hipStream_t)stream, *mesh, rays_o, rays_d, visited, items,
                       K, hits, hit_face);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
