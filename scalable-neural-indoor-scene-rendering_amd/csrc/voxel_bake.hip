// The diffuse branch of the reference's tile training around the surface MLP, forward and
// hand-written adjoint:
//   - the occupied cells of a tile's dilated node grid in row-major order (the torch.where of
//     Tile.compute_voxels, training/scripts/tile/tile.py:283-295, over nodes_flag != -1);
//   - the MLP inputs: the cell centres (regular_sample = (index + 0.5) voxel_size + min_corner,
//     tile.py:91-95) or ray samples, normalised as (x - group_center) / group_size
//     (norm_diffuse_point, training/scripts/tile/tileGroup.py:74);
//   - the diffuse head of Tile.inference_diffuse (tile.py:175-194): sigmoid of the first three MLP
//     outputs, ReLU of the fourth, written per row or scattered into the [D][D][D][4] voxel grid
//     (the bake's voxels[non_empty] = ...);
//   - Tile.query_voxel (tile.py:225-236): F.grid_sample (trilinear, align_corners=True, zero padding)
//     of the baked grid at world samples, with the reference's coordinate chain
//     indices = (x - min_corner) / voxel_size - 0.5, grids = indices / (D - 1) * 2 - 1 and the
//     sampler's unnormalisation ((g + 1) / 2) (D - 1); its adjoint scatters into the grid gradient.
// The grid is channel-last ([z][y][x][rgb, sigma], the memory order of the reference's
// voxels[None].permute(0, 4, 1, 2, 3) view), so a corner is one 16-byte load or four atomics.
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {
namespace {

constexpr int kFlagBlock = 1024;  // cells per workgroup of the ordered compaction (256 threads x 4)

__global__ __launch_bounds__(256) void k_flag_count(const int16_t* __restrict__ flags, int64_t n,
                                                    int32_t* __restrict__ counts) {
  __shared__ int wsum[4];
  int c = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t k = (int64_t)blockIdx.x * kFlagBlock + j * 256 + threadIdx.x;
    c += (k < n && flags[k] != -1) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// exclusive scan of the per-block counts in place (one workgroup), total into *count
__global__ __launch_bounds__(1024) void k_flag_scan(int32_t* __restrict__ counts, int nb, int32_t* __restrict__ count) {
  __shared__ int part[1024];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int b = b0 + threadIdx.x;
    const int v = b < nb ? counts[b] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
      const int add = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (b < nb) counts[b] = carry + part[threadIdx.x] - v;
    carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

__global__ __launch_bounds__(256) void k_flag_scatter(const int16_t* __restrict__ flags, int64_t n,
                                                      const int32_t* __restrict__ offsets, int32_t* __restrict__ cells) {
  __shared__ int wcnt[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int base = offsets[blockIdx.x];
  for (int j = 0; j < 4; ++j) {
    const int64_t k = (int64_t)blockIdx.x * kFlagBlock + j * 256 + threadIdx.x;
    const bool on = k < n && flags[k] != -1;
    const uint64_t m = __ballot(on);
    if (lane == 0) wcnt[wv] = __popcll(m);
    __syncthreads();
    int wbase = base;
    for (int w = 0; w < wv; ++w) wbase += wcnt[w];
    if (on) cells[wbase + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)k;
    base += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    __syncthreads();
  }
}

// cell centre of row-major cell index c in a D^3 grid, then (p - center) / scale, in the reference's
// FP32 operation order
__global__ __launch_bounds__(256) void k_bake_points(const int32_t* __restrict__ cells, int64_t n, int D, float3 mc,
                                                     float vs, float3 gc, float gs, float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int c = cells[k];
  const int x = c % D, y = (c / D) % D, z = c / (D * D);
  out[3 * k] = (((float)x + 0.5f) * vs + mc.x - gc.x) / gs;
  out[3 * k + 1] = (((float)y + 0.5f) * vs + mc.y - gc.y) / gs;
  out[3 * k + 2] = (((float)z + 0.5f) * vs + mc.z - gc.z) / gs;
}

__global__ __launch_bounds__(256) void k_normalize_points(const float* __restrict__ x, int64_t n, float3 gc, float gs,
                                                          float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  out[3 * k] = (x[3 * k] - gc.x) / gs;
  out[3 * k + 1] = (x[3 * k + 1] - gc.y) / gs;
  out[3 * k + 2] = (x[3 * k + 2] - gc.z) / gs;
}

__device__ __forceinline__ float sigmoid_f(float v) { return 1.0f / (1.0f + exp_cw(-v)); }

// row k of raw [n][4] -> (sigmoid r0..r2, relu r3) at row cells[k] (or k); a cell index outside
// 0..n_rows-1 writes nothing
__global__ __launch_bounds__(256) void k_head_fwd(const float* __restrict__ raw, int64_t n,
                                                  const int32_t* __restrict__ cells, int64_t n_rows,
                                                  float* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 r = reinterpret_cast<const float4*>(raw)[k];
  const int64_t row = cells ? (int64_t)cells[k] : k;
  if (row < 0 || row >= n_rows) return;
  reinterpret_cast<float4*>(out)[row] = make_float4(sigmoid_f(r.x), sigmoid_f(r.y), sigmoid_f(r.z), r.w > 0.0f ? r.w : 0.0f);
}

// dL/draw of row k from dL/dout at row cells[k] (or k); a cell index outside 0..n_rows-1 reads a
// zero gradient
__global__ __launch_bounds__(256) void k_head_bwd(const float* __restrict__ raw, int64_t n,
                                                  const int32_t* __restrict__ cells, int64_t n_rows,
                                                  const float* __restrict__ g_out, float* __restrict__ g_raw) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 r = reinterpret_cast<const float4*>(raw)[k];
  const int64_t row = cells ? (int64_t)cells[k] : k;
  const float4 g = (row >= 0 && row < n_rows) ? reinterpret_cast<const float4*>(g_out)[row]
                                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float s0 = sigmoid_f(r.x), s1 = sigmoid_f(r.y), s2 = sigmoid_f(r.z);
  reinterpret_cast<float4*>(g_raw)[k] =
      make_float4(g.x * (s0 * (1.0f - s0)), g.y * (s1 * (1.0f - s1)), g.z * (s2 * (1.0f - s2)), r.w > 0.0f ? g.w : 0.0f);
}

// The reference's sample coordinate along one axis: grid coordinate then grid_sample's
// align_corners unnormalisation.
__device__ __forceinline__ float query_coord(float x, float mc, float vs, float dm1) {
  const float idx = (x - mc) / vs - 0.5f;
  const float g = idx / dm1 * 2.0f - 1.0f;
  return (g + 1.0f) / 2.0f * dm1;
}

struct QueryCorners {
  int64_t cell[8];
  float w[8];
};

// The 8 corners (z-major, then y, then x, as grid_sample sums them) with trilinear weights; a
// corner outside the grid has weight 0 and cell -1.
__device__ __forceinline__ QueryCorners query_corners(const float* p, float3 mc, float vs, int D) {
  const float dm1 = (float)(D - 1);
  const float ix = query_coord(p[0], mc.x, vs, dm1), iy = query_coord(p[1], mc.y, vs, dm1),
              iz = query_coord(p[2], mc.z, vs, dm1);
  const float fx = floorf(ix), fy = floorf(iy), fz = floorf(iz);
  // far-away (or NaN) samples: every corner is outside, and the clamp keeps the conversion in range
  const float lo = -2.0f, hi = (float)D + 1.0f;
  const int x0 = (int)fminf(fmaxf(fx, lo), hi), y0 = (int)fminf(fmaxf(fy, lo), hi), z0 = (int)fminf(fmaxf(fz, lo), hi);
  const float wx[2] = {fx + 1.0f - ix, ix - fx}, wy[2] = {fy + 1.0f - iy, iy - fy}, wz[2] = {fz + 1.0f - iz, iz - fz};
  QueryCorners q;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int dz = j >> 2, dy = (j >> 1) & 1, dx = j & 1;
    const int x = x0 + dx, y = y0 + dy, z = z0 + dz;
    const bool in = x >= 0 && x < D && y >= 0 && y < D && z >= 0 && z < D;
    q.cell[j] = in ? ((int64_t)z * D + y) * D + x : -1;
    q.w[j] = in ? wx[dx] * wy[dy] * wz[dz] : 0.0f;
  }
  return q;
}

__global__ __launch_bounds__(256) void k_query_fwd(const float* __restrict__ vox, int D, const float* __restrict__ pts,
                                                   int64_t n, float3 mc, float vs, float* __restrict__ rgb,
                                                   float* __restrict__ sigma) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float p[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
  const QueryCorners q = query_corners(p, mc, vs, D);
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (q.cell[j] >= 0) {
      const float4 v = reinterpret_cast<const float4*>(vox)[q.cell[j]];
      acc.x = acc.x + v.x * q.w[j];
      acc.y = acc.y + v.y * q.w[j];
      acc.z = acc.z + v.z * q.w[j];
      acc.w = acc.w + v.w * q.w[j];
    }
  rgb[3 * k] = acc.x;
  rgb[3 * k + 1] = acc.y;
  rgb[3 * k + 2] = acc.z;
  sigma[k] = acc.w;
}

// The adjoint adds w_j * g into the 8 corners of every sample.  Consecutive samples (lanes) of a ray
// often share a corner cell: for each corner slot the wavefront sums the contributions over runs of
// lanes with the same cell (a segmented scan over the lanes) and only a run's last lane issues the
// four float atomics, which removes most of the contention on the cells of densely sampled voxels.
__global__ __launch_bounds__(256) void k_query_bwd(int D, const float* __restrict__ pts, int64_t n, float3 mc, float vs,
                                                   const float* __restrict__ g_rgb, const float* __restrict__ g_sigma,
                                                   float* __restrict__ g_vox) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  if (k - lane >= n) return;  // the whole wavefront is past the end
  const bool valid = k < n;
  float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  QueryCorners q;
  if (valid) {
    if (g_rgb) {
      g[0] = g_rgb[3 * k];
      g[1] = g_rgb[3 * k + 1];
      g[2] = g_rgb[3 * k + 2];
    }
    if (g_sigma) g[3] = g_sigma[k];
    const float p[3] = {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
    q = query_corners(p, mc, vs, D);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      q.cell[j] = -1;
      q.w[j] = 0.0f;
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int cell = (int)q.cell[j];  // < D^3 < 2^31 (checked on the host)
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = cell >= 0 ? g[c] * q.w[j] : 0.0f;
    const int up = __shfl_up(cell, 1), dn = __shfl_down(cell, 1);
    int f = (lane == 0 || up != cell) ? 1 : 0;  // run head
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {  // inclusive sums within runs
      float t[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) t[c] = __shfl_up(v[c], o);
      const int tf = __shfl_up(f, o);
      if (lane >= o) {
        if (!f) {
#pragma unroll
          for (int c = 0; c < 4; ++c) v[c] = v[c] + t[c];
        }
        f |= tf;
      }
    }
    const bool tail = lane == 63 || dn != cell;
    if (tail && cell >= 0) {
      float* dst = g_vox + 4 * (int64_t)cell;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (v[c] != 0.0f) atomicAdd(dst + c, v[c]);
    }
  }
}

__global__ void k_vox_zero(float4* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_flag_cells_workspace_size(int64_t n_cells) {
  if (n_cells < 0) return -1;
  return 4 * ((n_cells + kFlagBlock - 1) / kFlagBlock) + 256;
}

int snisr_flag_cells(const int16_t* flags, int64_t n_cells, int32_t* cells, int32_t* count, void* workspace,
                     int64_t workspace_bytes, void* stream) {
  SNISR_REQUIRE(n_cells >= 0 && n_cells < (int64_t)1 << 31, SNISR_ERR_SHAPE, "flag_cells: cell count out of range");
  SNISR_REQUIRE(count && (n_cells == 0 || (flags && cells && workspace)), SNISR_ERR_ARG, "flag_cells: NULL pointer");
  SNISR_REQUIRE(n_cells == 0 || workspace_bytes >= snisr_flag_cells_workspace_size(n_cells), SNISR_ERR_CAPACITY,
                "flag_cells: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)((n_cells + kFlagBlock - 1) / kFlagBlock);
  int32_t* counts = (int32_t*)workspace;
  if (nb > 0) {
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(256), 0, st, flags, n_cells, counts);
    SNISR_LAUNCHED();
  }
  hipLaunchKernelGGL(k_flag_scan, dim3(1), dim3(1024), 0, st, counts, nb, count);
  SNISR_LAUNCHED();
  if (nb > 0) {
    hipLaunchKernelGGL(k_flag_scatter, dim3(nb), dim3(256), 0, st, flags, n_cells, (const int32_t*)counts, cells);
    SNISR_LAUNCHED();
  }
  return SNISR_OK;
}

int snisr_bake_points(const int32_t* cells, int64_t n, int32_t D, float mc_x, float mc_y, float mc_z, float voxel_size,
                      float center_x, float center_y, float center_z, float scale, float* points, void* stream) {
  SNISR_REQUIRE(n >= 0 && D >= 2, SNISR_ERR_SHAPE, "bake_points: bad sizes");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(cells && points, SNISR_ERR_ARG, "bake_points: NULL pointer");
  hipLaunchKernelGGL(k_bake_points, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, cells, n, (int)D,
                     make_float3(mc_x, mc_y, mc_z), voxel_size, make_float3(center_x, center_y, center_z), scale, points);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_normalize_points(const float* x, int64_t n, float center_x, float center_y, float center_z, float scale,
                           float* out, void* stream) {
  SNISR_REQUIRE(n >= 0, SNISR_ERR_SHAPE, "normalize_points: negative size");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(x && out, SNISR_ERR_ARG, "normalize_points: NULL pointer");
  hipLaunchKernelGGL(k_normalize_points, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, n,
                     make_float3(center_x, center_y, center_z), scale, out);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_diffuse_head_forward(const float* raw, int64_t n, const int32_t* cells, int64_t n_rows, float* out,
                               void* stream) {
  SNISR_REQUIRE(n >= 0 && n_rows >= 0 && (cells || n_rows == n), SNISR_ERR_SHAPE,
                "diffuse_head_forward: without cells the output has one row per input row");
  SNISR_REQUIRE((n == 0 || raw) && (n_rows == 0 || out), SNISR_ERR_ARG, "diffuse_head_forward: NULL pointer");
  SNISR_REQUIRE((!raw || aligned(raw, 16)) && (!out || aligned(out, 16)), SNISR_ERR_ALIGN,
                "diffuse_head_forward: rows must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  if (cells && n_rows > 0) {  // the grid's empty cells are zero
    hipLaunchKernelGGL(k_vox_zero, dim3(blocks_for(n_rows)), dim3(256), 0, st, (float4*)out, n_rows);
    SNISR_LAUNCHED();
  }
  if (n == 0) return SNISR_OK;
  hipLaunchKernelGGL(k_head_fwd, dim3(blocks_for(n)), dim3(256), 0, st, raw, n, cells, n_rows, out);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_diffuse_head_backward(const float* raw, int64_t n, const int32_t* cells, int64_t n_rows, const float* g_out,
                                float* g_raw, void* stream) {
  SNISR_REQUIRE(n >= 0 && n_rows >= 0 && (cells || n_rows == n), SNISR_ERR_SHAPE,
                "diffuse_head_backward: without cells dL/dout has one row per input row");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(raw && g_out && g_raw, SNISR_ERR_ARG, "diffuse_head_backward: NULL pointer");
  SNISR_REQUIRE(aligned(raw, 16) && aligned(g_out, 16) && aligned(g_raw, 16), SNISR_ERR_ALIGN,
                "diffuse_head_backward: rows must be 16-byte aligned");
  hipLaunchKernelGGL(k_head_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, raw, n, cells, n_rows, g_out,
                     g_raw);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_voxel_query_forward(const float* voxels, int32_t D, const float* points, int64_t n, float mc_x, float mc_y,
                              float mc_z, float voxel_size, float* rgb, float* sigma, void* stream) {
  SNISR_REQUIRE(n >= 0 && D >= 2, SNISR_ERR_SHAPE, "voxel_query_forward: bad sizes");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(voxels && points && rgb && sigma, SNISR_ERR_ARG, "voxel_query_forward: NULL pointer");
  SNISR_REQUIRE(aligned(voxels, 16), SNISR_ERR_ALIGN, "voxel_query_forward: grid must be 16-byte aligned");
  hipLaunchKernelGGL(k_query_fwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, voxels, (int)D, points, n,
                     make_float3(mc_x, mc_y, mc_z), voxel_size, rgb, sigma);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_voxel_query_backward(int32_t D, const float* points, int64_t n, float mc_x, float mc_y, float mc_z,
                               float voxel_size, const float* g_rgb, const float* g_sigma, float* g_voxels,
                               void* stream) {
  SNISR_REQUIRE(n >= 0 && D >= 2 && (int64_t)D * D * D < ((int64_t)1 << 31), SNISR_ERR_SHAPE,
                "voxel_query_backward: bad sizes (D^3 must stay below 2^31)");
  SNISR_REQUIRE(g_voxels, SNISR_ERR_ARG, "voxel_query_backward: NULL gradient grid");
  hipStream_t st = (hipStream_t)stream;
  const int64_t cells = (int64_t)D * D * D;
  hipLaunchKernelGGL(k_vox_zero, dim3(blocks_for(cells)), dim3(256), 0, st, (float4*)g_voxels, cells);
  SNISR_LAUNCHED();
  if (n == 0 || (!g_rgb && !g_sigma)) return SNISR_OK;
  SNISR_REQUIRE(points, SNISR_ERR_ARG, "voxel_query_backward: NULL pointer");
  hipLaunchKernelGGL(k_query_bwd, dim3(blocks_for(n)), dim3(256), 0, st, (int)D, points, n,
                     make_float3(mc_x, mc_y, mc_z), voxel_size, g_rgb, g_sigma, g_voxels);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
