// Reference-format operators of the render path's second half: inverse-depth sampling (A6),
// per-pixel compositing of 16-sample chunks (A5), assign_pixels (A8) and the group boundary.  The
// pixel sort and the device-side regrouping are in regroup.hip.
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

// A6: z_j = 1/((offset+j)*ib + in), step_j = max(z_{j+1} - z_j, 0), point = origin + z d
// (rendering/src/sampling.cu:5-46; __fdiv_rn there is IEEE round-to-nearest division, as here).
__global__ __launch_bounds__(256) void k_inverse_z(const float* __restrict__ origin, const float* __restrict__ rays_d,
                                                   int64_t d_row_stride, int64_t d_col_stride, int W,
                                                   const float* __restrict__ inv_near, const float* __restrict__ inv_bound,
                                                   int64_t npix, int offset, int spr, float* __restrict__ points,
                                                   float* __restrict__ steps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix * spr) return;
  const int64_t p = t / spr;
  const int j = (int)(t - p * spr);
  const float ib = inv_bound[p], in = inv_near[p];
  const float z = 1.0f / ((float)(offset + j) * ib + in);
  const float nz = 1.0f / ((float)(offset + j + 1) * ib + in);
  steps[t] = fmaxx(nz - z, 0.0f);
  const int64_t py = p / W, px = p - py * W;
  const float* d = rays_d + py * d_row_stride + px * d_col_stride;
#pragma unroll
  for (int c = 0; c < 3; ++c) points[t * 3 + c] = origin[c] + z * d[c];
}

// A5: 16 samples per pixel; skip pixels with T < early_stop; alpha = 1 - exp(-sigma*step),
// C += alpha*T*rgb, T *= 1 - alpha; T and colours updated in place
// (rendering/src/volume_rendering.cu:5-58).  One wave per 64 pixels: the 64 x 16 sample records
// and steps are read with contiguous loads and transposed through LDS to sample-major rows
// (padded to 65 entries, so both the stores and the per-pixel reads are bank-conflict free); each
// lane then composites its pixel in sample order.
constexpr int kIntegPix = 64, kIntegPad = kIntegPix + 1;
__global__ __launch_bounds__(kIntegPix) void k_integrate_points(const float4* __restrict__ res,
                                                                const float* __restrict__ steps, int64_t npix,
                                                                float early, float* __restrict__ T,
                                                                float* __restrict__ color) {
  __shared__ float4 rs[16 * kIntegPad];
  __shared__ float ss[16 * kIntegPad];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kIntegPix;
  const int64_t rem = npix - base;
  const int nv = 16 * (int)(rem < kIntegPix ? rem : kIntegPix);  // records of this block
  const float4* rb = res + base * 16;
  const float* sb = steps + base * 16;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int g = k * kIntegPix + lane;  // record g: pixel g / 16, sample g % 16
    if (g < nv) {
      rs[(g & 15) * kIntegPad + (g >> 4)] = rb[g];
      ss[(g & 15) * kIntegPad + (g >> 4)] = sb[g];
    }
  }
  __syncthreads();
  const int64_t p = base + lane;
  if (p >= npix) return;
  float t = T[p];
  if (t < early) return;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const float4 v = rs[i * kIntegPad + lane];
    const float alpha = 1.0f - exp_cw(-1.0f * v.w * ss[i * kIntegPad + lane]);
    const float at = alpha * t;
    c0 = c0 + at * v.x;
    c1 = c1 + at * v.y;
    c2 = c2 + at * v.z;
    t = t * (1.0f - alpha);
  }
  T[p] = t;
  color[p * 3] += c0;
  color[p * 3 + 1] += c1;
  color[p * 3 + 2] += c2;
}

// A8: assign_pixels_kernel (rendering/src/reorder.cu:25-67): group i's pixel range is split into
// block_starts[i+1]-block_starts[i] chunks whose sizes differ by at most one (larger first).  One
// thread per output chunk (the reference loops over a group's chunks in one thread): chunk b finds
// its group by binary search over block_starts, then its start in closed form.
__global__ __launch_bounds__(256) void k_assign_pixels(const int16_t* __restrict__ net, const int32_t* __restrict__ pixel_starts,
                                                       const int32_t* __restrict__ block_starts, int64_t nnet,
                                                       int64_t nblocks, int16_t* __restrict__ new_net,
                                                       int32_t* __restrict__ new_starts) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  // group i with block_starts[i] <= b < block_starts[i+1] (groups with no chunk are skipped)
  int64_t lo = 0, hi = nnet;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (block_starts[mid] <= b) lo = mid; else hi = mid;
  }
  const int64_t i = lo;
  const int nb = block_starts[i + 1] - block_starts[i];
  if (nb <= 0 || b < block_starts[i] || b >= block_starts[i + 1]) return;
  const int ps = pixel_starts[i], npx = pixel_starts[i + 1] - ps;
  const int base = npx / nb, add = npx - base * nb;
  const int r = (int)(b - block_starts[i]);
  new_net[b] = net[i];
  new_starts[b] = ps + r * base + (r < add ? r : add);
}

// compute_group_boundary_kernel (rendering/src/compute_group_boundary.cu:4-50): a pixel is on a
// group boundary when any pixel of its ksize x ksize window (clipped to the image) has another group
// id; the flag is only ever set (an already-set flag stays set).  A 16 x 16 block of pixels stages
// its window rows of group ids in LDS once.
constexpr int kGbTile = 16, kGbMaxRadius = 15;
__global__ __launch_bounds__(kGbTile * kGbTile) void k_group_boundary(const int16_t* __restrict__ net, int H, int W,
                                                                      int radius, uint8_t* __restrict__ boundary) {
  __shared__ int16_t tile[kGbTile + 2 * kGbMaxRadius][kGbTile + 2 * kGbMaxRadius];
  const int x0 = blockIdx.x * kGbTile - radius, y0 = blockIdx.y * kGbTile - radius;
  const int span = kGbTile + 2 * radius;
  for (int q = threadIdx.x; q < span * span; q += blockDim.x) {
    const int ty = q / span, tx = q - ty * span;
    const int gy = min(max(y0 + ty, 0), H - 1), gx = min(max(x0 + tx, 0), W - 1);
    tile[ty][tx] = net[(int64_t)gy * W + gx];
  }
  __syncthreads();
  const int lx = threadIdx.x % kGbTile, ly = threadIdx.x / kGbTile;
  const int px = blockIdx.x * kGbTile + lx, py = blockIdx.y * kGbTile + ly;
  if (px >= W || py >= H) return;
  const int64_t i = (int64_t)py * W + px;
  if (boundary[i]) return;
  const int16_t c = tile[ly + radius][lx + radius];
  // clipped window: rows / columns outside the image are not part of it (clamped copies of the
  // edge would compare equal to the edge pixel, which is itself in the window)
  bool diff = false;
  for (int dy = -radius; dy <= radius && !diff; ++dy)
    for (int dx = -radius; dx <= radius; ++dx)
      if (tile[ly + radius + dy][lx + radius + dx] != c) { diff = true; break; }
  if (diff) boundary[i] = 1;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_inverse_z_sampling(const float* origin, const float* rays_d, int64_t d_row_stride, int64_t d_col_stride,
                             int32_t H, int32_t W, const float* inverse_near, const float* inverse_bound, int32_t offset,
                             int32_t sample_per_ray, float* sample_points, float* integrate_steps, void* stream) {
  SNISR_REQUIRE(H >= 0 && W >= 0 && sample_per_ray > 0, SNISR_ERR_ARG, "inverse_z_sampling: bad size");
  const int64_t total = (int64_t)H * W * sample_per_ray;
  if (total == 0) return SNISR_OK;
  SNISR_REQUIRE(origin && rays_d && inverse_near && inverse_bound && sample_points && integrate_steps, SNISR_ERR_ARG,
                "inverse_z_sampling: NULL pointer");
  hipLaunchKernelGGL(k_inverse_z, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, origin,
                     rays_d, d_row_stride, d_col_stride, W, inverse_near, inverse_bound, (int64_t)H * W, offset,
                     sample_per_ray, sample_points, integrate_steps);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_integrate_points(const float* sample_results, const float* integrate_steps, int64_t npix, float early_stop,
                           float* transparency, float* colors, void* stream) {
  SNISR_REQUIRE(npix >= 0, SNISR_ERR_ARG, "integrate_points: bad size");
  if (npix == 0) return SNISR_OK;
  SNISR_REQUIRE(sample_results && integrate_steps && transparency && colors, SNISR_ERR_ARG,
                "integrate_points: NULL pointer");
  SNISR_REQUIRE(aligned(sample_results, 16), SNISR_ERR_ALIGN, "integrate_points: sample_results must be 16-byte aligned");
  hipLaunchKernelGGL(k_integrate_points, dim3((unsigned)((npix + kIntegPix - 1) / kIntegPix)), dim3(kIntegPix), 0,
                     (hipStream_t)stream,
                     (const float4*)sample_results, integrate_steps, npix, early_stop, transparency, colors);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_compute_group_boundary(const int16_t* net_idx, int32_t H, int32_t W, int32_t ksize, uint8_t* boundary,
                                 void* stream) {
  SNISR_REQUIRE(H >= 0 && W >= 0 && ksize >= 0, SNISR_ERR_ARG, "compute_group_boundary: negative size");
  if ((int64_t)H * W == 0) return SNISR_OK;
  SNISR_REQUIRE(net_idx && boundary, SNISR_ERR_ARG, "compute_group_boundary: NULL pointer");
  SNISR_REQUIRE(ksize / 2 <= kGbMaxRadius, SNISR_ERR_SHAPE, "compute_group_boundary: ksize above %d",
                2 * kGbMaxRadius + 1);
  const dim3 grid((unsigned)((W + kGbTile - 1) / kGbTile), (unsigned)((H + kGbTile - 1) / kGbTile));
  hipLaunchKernelGGL(k_group_boundary, grid, dim3(kGbTile * kGbTile), 0, (hipStream_t)stream, net_idx, H, W,
                     ksize / 2, boundary);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

int snisr_assign_pixels(const int16_t* net, const int32_t* pixel_starts, const int32_t* block_starts, int64_t n_net,
                        int64_t n_blocks, int16_t* new_net, int32_t* new_pixel_starts, void* stream) {
  if (n_net <= 0 || n_blocks <= 0) return SNISR_OK;
  SNISR_REQUIRE(net && pixel_starts && block_starts && new_net && new_pixel_starts, SNISR_ERR_ARG,
                "assign_pixels: NULL pointer");
  hipLaunchKernelGGL(k_assign_pixels, dim3((unsigned)((n_blocks + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     net, pixel_starts, block_starts, n_net, n_blocks, new_net, new_pixel_starts);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
