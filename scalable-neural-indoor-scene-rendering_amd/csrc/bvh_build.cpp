// Host BVH builder for the proxy mesh: binned SAH (16 bins per axis), leaves of <= 4 triangles,
// index-median splits below depth 32, so the depth stays below 32 + log2(2^28 / 4) < 64 (the
// traversal stack size).
//
// Output layout (consumed by the traversal kernel in mesh_hit.hip):
//   node  = 64 bytes: child 0 box (lo.xyz, hi.xyz), child 1 box, child 0 code, child 1 code;
//           code >= 0: internal node index; code < 0: leaf ~(first << 3 | count), count <= 7
//           (count 0 = empty child); node 0 is the root.
//   tri   = 64 bytes: v0.xyz + original id (as int bits), e1 = v1 - v0 + 0, e2 = v2 - v0 + 0,
//           n = (v1 - v0) x (v2 - v1) + 0, in leaf order (the ray-independent terms of the triangle
//           test, in FP32 with the device's operation order).
// Boxes are widened outward by a relative 1e-6 of the scene extent (plus one float ulp) so the
// traversal's float slab test never rejects a box whose triangles the exact test would hit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "snisr_common.hpp"

namespace snisr {
namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) { lo[a] = std::numeric_limits<float>::infinity(); hi[a] = -lo[a]; }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  void grow(const float* p) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  float area() const {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

struct Prim {
  Box box;
  float c[3];
  int id;
};

struct GpuNode {
  float b0[6];
  float b1[6];
  int c0, c1;
  int pad[2];
};
static_assert(sizeof(GpuNode) == 64, "node size");

struct Builder {
  std::vector<Prim> prims;
  std::vector<GpuNode> nodes;
  float pad_abs = 0.0f;
  int max_depth = 0;

  Box bounds(int b, int e) const {
    Box x;
    x.reset();
    for (int i = b; i < e; ++i) x.grow(prims[i].box);
    return x;
  }
  void widen(const Box& in, float* out) const {
    for (int a = 0; a < 3; ++a) {
      const float l = in.lo[a] - pad_abs - std::fabs(in.lo[a]) * 1e-6f;
      const float h = in.hi[a] + pad_abs + std::fabs(in.hi[a]) * 1e-6f;
      out[a] = std::nextafter(l, -std::numeric_limits<float>::infinity());
      out[3 + a] = std::nextafter(h, std::numeric_limits<float>::infinity());
    }
  }
  static int leaf_code(int first, int count) { return ~((first << 3) | count); }

  // returns the child code for prims [b, e)
  int build(int b, int e, int depth) {
    max_depth = std::max(max_depth, depth);
    const int n = e - b;
    if (n <= 4) return leaf_code(b, n);
    Box cb;
    cb.reset();
    for (int i = b; i < e; ++i) cb.grow(prims[i].c);
    int mid = -1;
    if (depth < 32) {
      constexpr int NB = 16;
      float best = std::numeric_limits<float>::infinity();
      int best_axis = -1, best_split = -1;
      for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        if (!(ext > 0)) continue;
        Box bb[NB];
        int cnt[NB] = {0};
        for (int k = 0; k < NB; ++k) bb[k].reset();
        const float scale = NB / ext;
        for (int i = b; i < e; ++i) {
          int k = (int)((prims[i].c[a] - cb.lo[a]) * scale);
          k = std::min(NB - 1, std::max(0, k));
          ++cnt[k];
          bb[k].grow(prims[i].box);
        }
        float left_area[NB];
        int left_cnt[NB];
        Box acc;
        acc.reset();
        int c = 0;
        for (int k = 0; k < NB - 1; ++k) {
          acc.grow(bb[k]);
          c += cnt[k];
          left_area[k] = acc.area();
          left_cnt[k] = c;
        }
        acc.reset();
        c = 0;
        for (int k = NB - 1; k >= 1; --k) {
          acc.grow(bb[k]);
          c += cnt[k];
          const int lc = left_cnt[k - 1];
          if (lc == 0 || c == 0) continue;
          const float cost = left_area[k - 1] * lc + acc.area() * c;
          if (cost < best) { best = cost; best_axis = a; best_split = k; }
        }
      }
      if (best_axis >= 0) {
        const float ext = cb.hi[best_axis] - cb.lo[best_axis];
        const float scale = 16 / ext;
        Prim* p = std::partition(prims.data() + b, prims.data() + e, [&](const Prim& q) {
          int k = (int)((q.c[best_axis] - cb.lo[best_axis]) * scale);
          k = std::min(15, std::max(0, k));
          return k < best_split;
        });
        mid = (int)(p - prims.data());
        if (mid <= b || mid >= e) mid = -1;
      }
    }
    if (mid < 0) {  // index median on the widest centroid axis
      int a = 0;
      for (int q = 1; q < 3; ++q)
        if (cb.hi[q] - cb.lo[q] > cb.hi[a] - cb.lo[a]) a = q;
      mid = b + n / 2;
      std::nth_element(prims.data() + b, prims.data() + mid, prims.data() + e,
                       [a](const Prim& x, const Prim& y) { return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.id < y.id); });
    }
    const int idx = (int)nodes.size();
    nodes.emplace_back();
    const Box lb = bounds(b, mid), rb = bounds(mid, e);
    const int c0 = build(b, mid, depth + 1);
    const int c1 = build(mid, e, depth + 1);
    GpuNode& nd = nodes[idx];
    widen(lb, nd.b0);
    widen(rb, nd.b1);
    nd.c0 = c0;
    nd.c1 = c1;
    nd.pad[0] = nd.pad[1] = 0;
    return idx;
  }
};

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_bvh_build_bounds(int64_t n_faces, int64_t* node_bytes, int64_t* tri_bytes) {
  SNISR_REQUIRE(n_faces >= 0 && n_faces < (int64_t(1) << 28) && node_bytes && tri_bytes, SNISR_ERR_ARG,
                "bvh_build_bounds: bad arguments");
  *node_bytes = 64 * (n_faces + 2);
  *tri_bytes = 64 * (n_faces + 1);
  return SNISR_OK;
}

int snisr_bvh_build(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces, void* nodes_out,
                    int64_t node_bytes, void* tris_out, int64_t tri_bytes, int32_t* n_nodes, int32_t* depth) {
  SNISR_REQUIRE(vertices && faces && nodes_out && tris_out && n_nodes, SNISR_ERR_ARG, "bvh_build: NULL pointer");
  SNISR_REQUIRE(n_faces >= 0 && n_faces < (int64_t(1) << 28), SNISR_ERR_ARG, "bvh_build: too many faces");
  SNISR_REQUIRE(node_bytes >= 64 * (n_faces + 2) && tri_bytes >= 64 * n_faces, SNISR_ERR_CAPACITY,
                "bvh_build: output buffers too small");
  Builder B;
  B.prims.resize(n_faces);
  Box scene;
  scene.reset();
  for (int64_t f = 0; f < n_faces; ++f) {
    Prim& p = B.prims[f];
    p.box.reset();
    for (int k = 0; k < 3; ++k) {
      const int32_t v = faces[3 * f + k];
      SNISR_REQUIRE(v >= 0 && v < n_vertices, SNISR_ERR_ARG, "bvh_build: face %lld has vertex %d out of range",
                    (long long)f, v);
      p.box.grow(vertices + 3 * (int64_t)v);
    }
    for (int a = 0; a < 3; ++a) p.c[a] = 0.5f * (p.box.lo[a] + p.box.hi[a]);
    p.id = (int)f;
    scene.grow(p.box);
  }
  float ext = 0.0f;
  for (int a = 0; a < 3; ++a) ext = std::max(ext, scene.hi[a] - scene.lo[a]);
  B.pad_abs = n_faces ? ext * 1e-6f : 0.0f;
  B.nodes.reserve(n_faces / 2 + 2);
  B.nodes.emplace_back();  // root placeholder
  int root_code;
  if (n_faces <= 4) {
    root_code = Builder::leaf_code(0, (int)n_faces);
    GpuNode& r = B.nodes[0];
    Box bb = B.bounds(0, (int)n_faces);
    if (n_faces == 0) {
      for (int a = 0; a < 3; ++a) { bb.lo[a] = 1.0f; bb.hi[a] = -1.0f; }
      for (int a = 0; a < 3; ++a) { r.b0[a] = bb.lo[a]; r.b0[3 + a] = bb.hi[a]; }
    } else {
      B.widen(bb, r.b0);
    }
    for (int a = 0; a < 3; ++a) { r.b1[a] = 1.0f; r.b1[3 + a] = -1.0f; }  // empty box
    r.c0 = root_code;
    r.c1 = Builder::leaf_code(0, 0);
    r.pad[0] = r.pad[1] = 0;
  } else {
    // build the subtree under index 0: let build() create nodes after the placeholder, then move
    const int top = B.build(0, (int)n_faces, 0);
    // top is an internal node index >= 1: copy it to the root slot and keep the copy (unused)
    B.nodes[0] = B.nodes[top];
  }
  std::memcpy(nodes_out, B.nodes.data(), B.nodes.size() * sizeof(GpuNode));
  float* T = reinterpret_cast<float*>(tris_out);
  for (int64_t i = 0; i < n_faces; ++i) {
    const Prim& p = B.prims[i];
    const float* v[3];
    for (int k = 0; k < 3; ++k) v[k] = vertices + 3 * (int64_t)faces[3 * (int64_t)p.id + k];
    float e1[3], e2[3], f[3], n[3];
    for (int a = 0; a < 3; ++a) {
      e1[a] = v[1][a] - v[0][a];
      e2[a] = v[2][a] - v[0][a];
      f[a] = v[2][a] - v[1][a];
    }
    // cross((v1 - v0), (v2 - v1)) component by component, each a difference of two products
    n[0] = e1[1] * f[2] - e1[2] * f[1];
    n[1] = e1[2] * f[0] - e1[0] * f[2];
    n[2] = e1[0] * f[1] - e1[1] * f[0];
    float* t = T + 16 * i;
    for (int a = 0; a < 3; ++a) {
      t[a] = v[0][a];
      t[4 + a] = e1[a];
      t[8 + a] = e2[a];
      t[12 + a] = n[a];
    }
    const int32_t tag = p.id, zero = 0;
    std::memcpy(&t[3], &tag, 4);
    std::memcpy(&t[7], &zero, 4);
    std::memcpy(&t[11], &zero, 4);
    std::memcpy(&t[15], &zero, 4);
  }
  *n_nodes = (int32_t)B.nodes.size();
  if (depth) *depth = B.max_depth;
  return SNISR_OK;
}

}  // extern "C"
