// Reference-format operators of the render path's second half: inverse-depth sampling (A6),
// per-pixel compositing of 16-sample chunks (A5), and the pixel regrouping scheduler (A8).
#include <hipcub/hipcub.hpp>

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

// A8: start of every run of equal keys in a sorted array: flag[i] = i == 0 || keys[i] != keys[i-1]
__global__ __launch_bounds__(256) void k_run_flags(const int16_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// compacts the run heads: keys_out[r] = keys[i], starts_out[r] = starts_in[i] for run head i with
// exclusive rank r; the last thread also writes the run count
__global__ __launch_bounds__(256) void k_run_compact(const int16_t* __restrict__ keys, const int32_t* __restrict__ starts_in,
                                                     const int32_t* __restrict__ flag, const int32_t* __restrict__ rank,
                                                     int64_t n, int16_t* __restrict__ keys_out,
                                                     int32_t* __restrict__ starts_out, int32_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) {
    keys_out[rank[i]] = keys[i];
    starts_out[rank[i]] = starts_in[i];
  }
  if (i == n - 1) *count = rank[i] + flag[i];
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

// ---- host-sync-free regrouping (snisr_regroup_pixels)
__global__ __launch_bounds__(256) void k_iota(int32_t* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}

// run heads of the sorted keys: head position of run r at heads[r], run count at *m
__global__ __launch_bounds__(256) void k_run_heads(const int32_t* __restrict__ flag, const int32_t* __restrict__ rank,
                                                   int64_t n, int32_t* __restrict__ heads, int32_t* __restrict__ m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) heads[rank[i]] = (int32_t)i;
  if (i == n - 1) *m = rank[i] + flag[i];
}

// blocks of run r (entries r >= m are 0, so the scan over all n entries ends at the block total)
__global__ __launch_bounds__(256) void k_run_blocks(const int32_t* __restrict__ heads, const int32_t* __restrict__ m,
                                                    int64_t n, int32_t per_block, int32_t* __restrict__ nblk) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t M = *m;
  int32_t b = 0;
  if (r < M) {
    const int64_t end = r + 1 < M ? heads[r + 1] : n;
    b = (int32_t)((end - heads[r] + per_block - 1) / per_block);
  }
  nblk[r] = b;
}

// the blocks of run r at off[r]: its group and the first sorted position of each block
__global__ __launch_bounds__(256) void k_emit_blocks(const int16_t* __restrict__ keys, const int32_t* __restrict__ heads,
                                                     const int32_t* __restrict__ m, const int32_t* __restrict__ nblk,
                                                     const int32_t* __restrict__ off, int64_t n, int32_t per_block,
                                                     int64_t max_blocks, int16_t* __restrict__ block_net,
                                                     int32_t* __restrict__ block_starts, int32_t* __restrict__ total) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t M = *m;
  if (r == n - 1) *total = off[r] + nblk[r];
  if (r >= M) return;
  // the run's pixels split into nb blocks whose sizes differ by at most one, as assign_pixels does
  const int16_t key = keys[heads[r]];
  const int32_t nb = nblk[r], cnt = (int32_t)((r + 1 < M ? heads[r + 1] : n) - heads[r]);
  const int32_t base = cnt / nb, add = cnt - base * nb;
  for (int32_t j = 0; j < nb; ++j) {
    const int64_t b = (int64_t)off[r] + j;
    if (b < max_blocks) {
      block_net[b] = key;
      block_starts[b] = heads[r] + j * base + (j < add ? j : add);
    }
  }
}

__global__ void k_zero_i32(int32_t* __restrict__ v) { *v = 0; }

// entries past the last block: group -1, start n (block_starts[total] = n closes the last block)
__global__ __launch_bounds__(256) void k_block_tail(const int32_t* __restrict__ total, int64_t n, int64_t max_blocks,
                                                    int16_t* __restrict__ block_net, int32_t* __restrict__ block_starts,
                                                    int32_t* __restrict__ n_blocks) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t T = *total < max_blocks ? *total : max_blocks;
  if (b == 0) *n_blocks = (int32_t)T;
  if (b < T || b > max_blocks) return;
  if (b < max_blocks) block_net[b] = (int16_t)-1;
  block_starts[b] = (int32_t)n;
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

// Workspace bytes of snisr_sort_by_key for n keys.
int64_t snisr_sort_by_key_workspace_size(int64_t n) {
  size_t tmp = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const int16_t*)nullptr, (int16_t*)nullptr,
                                         (const int32_t*)nullptr, (int32_t*)nullptr, (int)n) != hipSuccess)
    return -1;
  size_t tmp2 = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (const int32_t*)nullptr, (int32_t*)nullptr, (int)n) != hipSuccess)
    return -1;
  const int64_t a = 256;
  auto up = [&](int64_t v) { return (v + a - 1) / a * a; };
  return up((int64_t)std::max(tmp, tmp2)) + up(2 * n) + up(4 * n) + 3 * up(4 * n) + up(4);
}

// A8 sort_by_key (rendering/src/reorder.cu:11-23): stable ascending sort of keys i16[n] carrying
// values i32[n], then the run heads: keys[0..m) = distinct keys, starts[0..m) = starts[] at the head
// of each run (the reference passes starts = arange(n), giving each run's first position).  Returns
// m in *n_unique (device-to-host copy, synchronises the stream as the reference's return value does).
int snisr_sort_by_key(int16_t* keys, int32_t* values, int32_t* starts, int64_t n, void* workspace,
                      int64_t workspace_bytes, int32_t* n_unique, void* stream) {
  SNISR_REQUIRE(n_unique, SNISR_ERR_ARG, "sort_by_key: n_unique is NULL");
  SNISR_REQUIRE(n >= 0 && n < (int64_t(1) << 31), SNISR_ERR_ARG, "sort_by_key: bad size");
  if (n == 0) {
    *n_unique = 0;
    return SNISR_OK;
  }
  SNISR_REQUIRE(keys && values && starts && workspace, SNISR_ERR_ARG, "sort_by_key: NULL pointer");
  const int64_t need = snisr_sort_by_key_workspace_size(n);
  SNISR_REQUIRE(need > 0 && workspace_bytes >= need, SNISR_ERR_CAPACITY, "sort_by_key: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  size_t tmp = 0, tmp2 = 0;
  SNISR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const int16_t*)nullptr, (int16_t*)nullptr,
                                               (const int32_t*)nullptr, (int32_t*)nullptr, (int)n));
  SNISR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (const int32_t*)nullptr, (int32_t*)nullptr, (int)n));
  const int64_t a = 256;
  auto up = [&](int64_t v) { return (v + a - 1) / a * a; };
  uint8_t* base = (uint8_t*)workspace;
  void* t = base;
  int16_t* keys2 = (int16_t*)(base + up((int64_t)std::max(tmp, tmp2)));
  int32_t* vals2 = (int32_t*)((uint8_t*)keys2 + up(2 * n));
  int32_t* flag = (int32_t*)((uint8_t*)vals2 + up(4 * n));
  int32_t* rank = (int32_t*)((uint8_t*)flag + up(4 * n));
  int32_t* starts2 = (int32_t*)((uint8_t*)rank + up(4 * n));
  int32_t* cnt = (int32_t*)((uint8_t*)starts2 + up(4 * n));
  size_t ts = std::max(tmp, tmp2);
  SNISR_HIP(hipcub::DeviceRadixSort::SortPairs(t, ts, keys, keys2, values, vals2, (int)n, 0, 16, st));
  SNISR_HIP(hipMemcpyAsync(values, vals2, 4 * n, hipMemcpyDeviceToDevice, st));
  const dim3 g((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_run_flags, g, dim3(256), 0, st, keys2, n, flag);
  SNISR_LAUNCHED();
  ts = std::max(tmp, tmp2);
  SNISR_HIP(hipcub::DeviceScan::ExclusiveSum(t, ts, flag, rank, (int)n, st));
  SNISR_HIP(hipMemcpyAsync(starts2, starts, 4 * n, hipMemcpyDeviceToDevice, st));
  // keys keep the sorted order beyond the first m entries, as thrust::unique_by_key leaves them
  SNISR_HIP(hipMemcpyAsync(keys, keys2, 2 * n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_run_compact, g, dim3(256), 0, st, keys2, starts2, flag, rank, n, keys, starts, cnt);
  SNISR_LAUNCHED();
  SNISR_HIP(hipMemcpyAsync(n_unique, cnt, 4, hipMemcpyDeviceToHost, st));
  SNISR_HIP(hipStreamSynchronize(st));
  return SNISR_OK;
}

int64_t snisr_regroup_max_blocks(int64_t n, int32_t pixels_per_block) {
  if (n < 0 || pixels_per_block <= 0) return -1;
  // sum over groups of ceil(count / P) <= n / P + (distinct groups), and at most 65,536 i16 groups
  return (n + pixels_per_block - 1) / pixels_per_block + (n < 65536 ? n : 65536);
}

int64_t snisr_regroup_workspace_size(int64_t n) {
  if (n < 0 || n >= (int64_t(1) << 31)) return -1;
  size_t tmp = 0, tmp2 = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const int16_t*)nullptr, (int16_t*)nullptr,
                                         (const int32_t*)nullptr, (int32_t*)nullptr, (int)std::max<int64_t>(n, 1)) !=
          hipSuccess ||
      hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (const int32_t*)nullptr, (int32_t*)nullptr,
                                       (int)std::max<int64_t>(n, 1)) != hipSuccess)
    return -1;
  auto up = [](int64_t v) { return (v + 255) / 256 * 256; };
  return up((int64_t)std::max(tmp, tmp2)) + up(2 * n) + up(4 * n) + 5 * up(4 * n) + 3 * up(4);
}

int snisr_regroup_pixels(const int16_t* net_idx, int64_t n, int32_t pixels_per_block, int64_t max_blocks,
                         int32_t* query, int16_t* block_net, int32_t* block_starts, int32_t* n_blocks,
                         void* workspace, int64_t workspace_bytes, void* stream) {
  SNISR_REQUIRE(n >= 0 && n < (int64_t(1) << 31), SNISR_ERR_ARG, "regroup_pixels: bad size");
  SNISR_REQUIRE(pixels_per_block > 0, SNISR_ERR_ARG, "regroup_pixels: pixels_per_block must be positive");
  SNISR_REQUIRE(max_blocks >= 0 && max_blocks < (int64_t(1) << 31), SNISR_ERR_ARG, "regroup_pixels: bad max_blocks");
  SNISR_REQUIRE(block_net && block_starts && n_blocks && workspace, SNISR_ERR_ARG, "regroup_pixels: NULL pointer");
  SNISR_REQUIRE(n == 0 || (net_idx && query), SNISR_ERR_ARG, "regroup_pixels: NULL pointer");
  const int64_t need = snisr_regroup_workspace_size(n);
  SNISR_REQUIRE(need > 0 && workspace_bytes >= need, SNISR_ERR_CAPACITY, "regroup_pixels: workspace too small");
  const hipStream_t st = (hipStream_t)stream;
  auto up = [](int64_t v) { return (v + 255) / 256 * 256; };
  size_t tmp = 0, tmp2 = 0;
  const int nn = (int)std::max<int64_t>(n, 1);
  SNISR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const int16_t*)nullptr, (int16_t*)nullptr,
                                               (const int32_t*)nullptr, (int32_t*)nullptr, nn));
  SNISR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (const int32_t*)nullptr, (int32_t*)nullptr, nn));
  uint8_t* base = (uint8_t*)workspace;
  void* t = base;
  int16_t* keys = (int16_t*)(base + up((int64_t)std::max(tmp, tmp2)));
  int32_t* iota = (int32_t*)((uint8_t*)keys + up(2 * n));
  int32_t* flag = (int32_t*)((uint8_t*)iota + up(4 * n));
  int32_t* rank = (int32_t*)((uint8_t*)flag + up(4 * n));
  int32_t* heads = (int32_t*)((uint8_t*)rank + up(4 * n));
  int32_t* nblk = (int32_t*)((uint8_t*)heads + up(4 * n));
  int32_t* off = (int32_t*)((uint8_t*)nblk + up(4 * n));
  int32_t* m = (int32_t*)((uint8_t*)off + up(4 * n));
  int32_t* total = (int32_t*)((uint8_t*)m + up(4));
  const dim3 g((unsigned)((n + 255) / 256)), b256(256);
  if (n > 0) {
    hipLaunchKernelGGL(k_iota, g, b256, 0, st, iota, n);
    SNISR_LAUNCHED();
    size_t ts = std::max(tmp, tmp2);
    SNISR_HIP(hipcub::DeviceRadixSort::SortPairs(t, ts, net_idx, keys, iota, query, (int)n, 0, 16, st));
    hipLaunchKernelGGL(k_run_flags, g, b256, 0, st, keys, n, flag);
    SNISR_LAUNCHED();
    ts = std::max(tmp, tmp2);
    SNISR_HIP(hipcub::DeviceScan::ExclusiveSum(t, ts, flag, rank, (int)n, st));
    hipLaunchKernelGGL(k_run_heads, g, b256, 0, st, flag, rank, n, heads, m);
    SNISR_LAUNCHED();
    hipLaunchKernelGGL(k_run_blocks, g, b256, 0, st, heads, m, n, pixels_per_block, nblk);
    SNISR_LAUNCHED();
    ts = std::max(tmp, tmp2);
    SNISR_HIP(hipcub::DeviceScan::ExclusiveSum(t, ts, nblk, off, (int)n, st));
    hipLaunchKernelGGL(k_emit_blocks, g, b256, 0, st, keys, heads, m, nblk, off, n, pixels_per_block, max_blocks,
                       block_net, block_starts, total);
    SNISR_LAUNCHED();
  } else {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, total);
    SNISR_LAUNCHED();
  }
  hipLaunchKernelGGL(k_block_tail, dim3((unsigned)((max_blocks + 1 + 255) / 256)), b256, 0, st, total, n, max_blocks,
                     block_net, block_starts, n_blocks);
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
