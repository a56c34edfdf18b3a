// Host builder of the per-tile per-block face lists consumed by snisr_mesh_hit_tiles (the reference
// loads them as BConFaceIdx / BConFaceNum, training/scripts/tools.py:364-379; its builder belongs to
// the data preparation).  A face is listed in every block of every tile whose box, widened by
// max(2e-5, 1e-4 * block edge), it intersects.  The widening covers the kernel's 1e-5 acceptance
// band around a block and the rounding of its hit points, so every hit the kernel can accept is
// listed.  Triangle-box overlap by the separating-axis test (Akenine-Moeller, "Fast 3D triangle-box
// overlap testing", 2001): the three box normals, the triangle normal and the nine edge cross
// products, evaluated in double precision.
#include <algorithm>
#include <cmath>
#include <vector>

#include "snisr_common.hpp"

namespace snisr {
namespace {

struct D3 {
  double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 crs(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// overlap of the triangle (a, b, c) with the box centred at ctr with half extents h
bool tri_box(D3 ctr, D3 h, D3 a, D3 b, D3 c) {
  const D3 v0 = sub(a, ctr), v1 = sub(b, ctr), v2 = sub(c, ctr);
  const D3 e[3] = {sub(v1, v0), sub(v2, v1), sub(v0, v2)};
  const D3 ax[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto separated = [&](D3 n) {
    const double p0 = dt(v0, n), p1 = dt(v1, n), p2 = dt(v2, n);
    const double r = h.x * std::fabs(n.x) + h.y * std::fabs(n.y) + h.z * std::fabs(n.z);
    return std::min({p0, p1, p2}) > r || std::max({p0, p1, p2}) < -r;
  };
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (separated(crs(ax[j], e[i]))) return false;
  for (int j = 0; j < 3; ++j)
    if (separated(ax[j])) return false;
  return !separated(crs(e[0], e[1]));
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_tile_face_lists(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces,
                          const float* centers, int32_t n_tiles, float tile_size, int32_t num_block, int32_t* face_num,
                          int32_t* face_idx, int64_t capacity, int64_t* total) {
  SNISR_REQUIRE(total && face_num && centers, SNISR_ERR_ARG, "tile_face_lists: NULL pointer");
  SNISR_REQUIRE(n_tiles >= 0 && n_faces >= 0 && n_vertices >= 0, SNISR_ERR_ARG, "tile_face_lists: negative size");
  SNISR_REQUIRE(n_faces == 0 || (vertices && faces), SNISR_ERR_ARG, "tile_face_lists: NULL mesh");
  SNISR_REQUIRE(num_block >= 1 && num_block <= 1024 && tile_size > 0.0f, SNISR_ERR_ARG,
                "tile_face_lists: num_block in [1, 1024] and a positive tile_size");
  for (int64_t f = 0; f < 3 * n_faces; ++f)
    SNISR_REQUIRE(faces[f] >= 0 && faces[f] < n_vertices, SNISR_ERR_ARG, "tile_face_lists: face %lld refers to "
                  "vertex %d of %lld", (long long)(f / 3), faces[f], (long long)n_vertices);
  const int nb = num_block;
  const int64_t per_tile = (int64_t)nb * nb * nb;
  const double ts = tile_size, bs = (double)(tile_size / (float)nb);
  const double margin = std::max(2e-5, 1e-4 * bs);
  const D3 h{bs / 2 + margin, bs / 2 + margin, bs / 2 + margin};
  // per block its faces, tile by tile, in face order
  std::vector<std::vector<int32_t>> lists((size_t)per_tile);
  int64_t count = 0;
  for (int32_t t = 0; t < n_tiles; ++t) {
    for (auto& l : lists) l.clear();
    const D3 lo{centers[3 * t] - ts / 2, centers[3 * t + 1] - ts / 2, centers[3 * t + 2] - ts / 2};
    for (int64_t f = 0; f < n_faces; ++f) {
      D3 v[3];
      double mn[3], mx[3];
      for (int k = 0; k < 3; ++k) {
        const float* p = vertices + 3 * (int64_t)faces[3 * f + k];
        v[k] = {p[0], p[1], p[2]};
      }
      const double c[3][3] = {{v[0].x, v[1].x, v[2].x}, {v[0].y, v[1].y, v[2].y}, {v[0].z, v[1].z, v[2].z}};
      const double o[3] = {lo.x, lo.y, lo.z};
      int b0[3], b1[3];
      bool out = false;
      for (int a = 0; a < 3; ++a) {
        mn[a] = std::min({c[a][0], c[a][1], c[a][2]});
        mx[a] = std::max({c[a][0], c[a][1], c[a][2]});
        b0[a] = (int)std::floor((mn[a] - margin - o[a]) / bs);
        b1[a] = (int)std::floor((mx[a] + margin - o[a]) / bs);
        b0[a] = std::max(b0[a], 0);
        b1[a] = std::min(b1[a], nb - 1);
        if (b0[a] > b1[a]) out = true;
      }
      if (out) continue;
      for (int x = b0[0]; x <= b1[0]; ++x)
        for (int y = b0[1]; y <= b1[1]; ++y)
          for (int z = b0[2]; z <= b1[2]; ++z) {
            const D3 ctr{lo.x + (x + 0.5) * bs, lo.y + (y + 0.5) * bs, lo.z + (z + 0.5) * bs};
            if (tri_box(ctr, h, v[0], v[1], v[2])) lists[((size_t)x * nb + y) * nb + z].push_back((int32_t)f);
          }
    }
    for (int64_t b = 0; b < per_tile; ++b) {
      const int64_t k = (int64_t)lists[(size_t)b].size();
      SNISR_REQUIRE(count + k < (int64_t(1) << 31), SNISR_ERR_CAPACITY, "tile_face_lists: more than 2^31 entries");
      face_num[2 * (t * per_tile + b)] = (int32_t)count;
      face_num[2 * (t * per_tile + b) + 1] = (int32_t)k;
      if (face_idx) {
        SNISR_REQUIRE(count + k <= capacity, SNISR_ERR_CAPACITY, "tile_face_lists: face_idx holds %lld entries",
                      (long long)capacity);
        std::copy(lists[(size_t)b].begin(), lists[(size_t)b].end(), face_idx + count);
      }
      count += k;
    }
  }
  *total = count;
  return SNISR_OK;
}

}  // extern "C"
