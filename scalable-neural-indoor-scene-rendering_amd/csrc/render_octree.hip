// A3 — the reference's octree-format diffuse renderer (rendering_diffuse_octree_fp16).
//
// Scene format (rendering/src/__init__.py:111-121): every tile is cut into 4^3 blocks; an occupied
// block is a node whose dilated num_voxel^3 grid (16^3 voxels plus a one-voxel border, num_voxel =
// 18) maps each cell to a voxel of the node (nodes_IndexMap, -1 = empty) and flags the cells that are
// sampled (nodes_sampleFlag); voxels are half4 (r, g, b, sigma) at data_voxels[voxels_start[node] +
// index].  Per ray (one thread per ray, rays taken in 8x8 pixel packets when the image width allows):
// for each visited tile until -1 or T < early_stop, the tile's ray-AABB entry, the local origin
// clamped into the tile, a DDA over its 4^3 blocks and, inside an occupied block, samples at
// zval = near, near + step, ... while zval < far; a sample whose cell is not flagged is skipped;
// the first sample taken with T < trans_th records inverse_near = 1 / (zval + tile entry) and the
// tile's group (rendering/src/raydiffuse_render_octree_fp16.cu:20-235, half_ope.h:246-284).
//
// Arithmetic: positions in FP32 with the reference's operation order (a float3 divided by a scalar is
// a multiply by the FP32 reciprocal; make_int3 truncates), so the sampled cells, sample counts, the
// first-hit distance and the group are bit-identical to the FP32 oracle.  Interpolation, alpha,
// weights, colour and transmittance, which the reference evaluates in FP16, are FP32 here.
// Three guards the reference lacks: an index of -1 reads an all-zero voxel, cell / corner
// coordinates are clamped into the dilated grid (malformed scenes cannot read out of bounds), and
// the march inside one block takes at most kOctMaxBlockSteps steps (every wave terminates).
#include "snisr_common.hpp"
#include "snisr_device.hpp"

namespace snisr {

constexpr int kOctBlocks = 4;  // blocks per tile edge (num_block, raydiffuse_render_octree_fp16.cu:6)
// march steps per block, taken or skipped; a block of a valid scene needs ~sqrt(3) * 16 / 0.2 < 140.
// The cap bounds the loop for degenerate rays (zero direction: far = 1e8) and for steps below
// zval's FP32 resolution, where zval += step would never reach far.
constexpr int kOctMaxBlockSteps = 1 << 16;
constexpr int kOctBatch = 2;  // samples whose loads are issued together (2: 417 us, 4: 456 us, 8: 576 us per 720p frame)

__device__ __forceinline__ float lerp_ab(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__global__ __launch_bounds__(64) void k_render_octree(snisr_octree O, const float* __restrict__ origin,
                                                      const float* __restrict__ rays_d, const int32_t* __restrict__ visited,
                                                      int64_t n, int K, int packet_w, float trans_th, float early_stop,
                                                      int soft_hit, float* __restrict__ frame,
                                                      float* __restrict__ inverse_near, int16_t* __restrict__ net_idx,
                                                      int32_t* __restrict__ n_samples) {
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (tid >= n) return;
  const int64_t r = ray_of_thread(tid, n, packet_w);
  const int M = O.num_voxel, MM = M * M, total = MM * M;
  const float ts = O.tile_size, vs = O.voxel_size, step = O.sample_step;
  const float bs = ts / (float)kOctBlocks, inv_bs = 1.0f / bs, inv_vs = 1.0f / vs, hs = ts / 2.0f;
  const V3 o = v3(origin[0], origin[1], origin[2]);
  const float d[3] = {rays_d[3 * r], rays_d[3 * r + 1], rays_d[3 * r + 2]};
  const V3 dv = v3(d[0], d[1], d[2]);
  float col[3] = {0.0f, 0.0f, 0.0f};
  float T = 1.0f;
  bool first_hit = false;
  int group = -1, ns = 0;
  float bl = -1.0f, bh = -1.0f;
  const int32_t* vis = visited + r * K;
  for (int i = 0; i < K; ++i) {
    const int tile = vis[i];
    if (tile == -1) break;
    if (T < early_stop) break;
    group = O.group_map[tile];
    const float c[3] = {O.centers[3 * tile], O.centers[3 * tile + 1], O.centers[3 * tile + 2]};
    ray_aabb(o, dv, v3(c[0], c[1], c[2]), v3(hs, hs, hs), bl, bh);
    if (bl == -1.0f) continue;
    float ol[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) ol[a] = clampx(get(o, a) + bl * d[a] - (c[a] - hs), 0.0f, ts - 0.00001f);
    // DDA over the tile's 4^3 blocks
    int cur[3], stp[3];
    float tmax[3], tdel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cur[a] = (int)(ol[a] * inv_bs);
      stp[a] = d[a] >= 0.0f ? 1 : -1;
      float nb = (float)(cur[a] + stp[a]) * bs;
      if (stp[a] < 0) nb += bs;
      tmax[a] = safe_div(nb - ol[a], d[a]);
      tdel[a] = safe_div(bs * (float)stp[a], d[a]);
    }
    float near = 0.0f, far = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
    const int32_t* bmap = O.block_index_map + (int64_t)tile * (kOctBlocks * kOctBlocks * kOctBlocks);
    while (true) {
      const int nidx = bmap[cur[0] * kOctBlocks * kOctBlocks + cur[1] * kOctBlocks + cur[2]];
      if (nidx != -1) {
        const int16_t* nodes = O.nodes_index_map + (int64_t)nidx * total;
        const uint8_t* flags = O.nodes_sample_flag + (int64_t)nidx * total;
        const int64_t start = O.voxels_start[nidx];
        float corner[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) corner[a] = (float)cur[a] * bs - vs;
        // samples are taken kOctBatch at a time: positions (zval advanced by repeated += step, as
        // in the sequential march), then all their flag loads, index loads and voxel loads are
        // issued together, then the samples are composited in order
        float zval = near;
        int guard = 0;
        while (zval < far && guard < kOctMaxBlockSteps) {
          float zs[kOctBatch], rel[kOctBatch][3];
          int cell[kOctBatch];
          bool live[kOctBatch];
          float z = zval;
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b) {
            zs[b] = z;
            live[b] = z < far && guard + b < kOctMaxBlockSteps;
            int vl[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float p = ol[a] + z * d[a];
              rel[b][a] = (p - corner[a]) * inv_vs;
              vl[a] = min(max((int)rel[b][a], 0), M - 1);
            }
            cell[b] = vl[0] * MM + vl[1] * M + vl[2];
            z += step;
          }
          // the flag and voxel-id loads depend only on the sample positions (cell and base are
          // clamped into the block), so they are issued together, unconditionally, and the flag is
          // applied afterwards: two dependent load levels per batch instead of three
          uint8_t flg[kOctBatch];
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b) flg[b] = flags[cell[b]];
          // trilinear corners over the dilated grid (x, then y, then z)
          int base[kOctBatch];
          float w[kOctBatch][3];
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b) {
            int i0[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float f = rel[b][a] - 0.5f;
              i0[a] = min(max((int)f, 0), M - 2);
              const float t = f - (float)(int)f;
              w[b][a] = t < 0.0f ? 0.0f : t;
            }
            base[b] = i0[0] * MM + i0[1] * M + i0[2];
          }
          int16_t vid[kOctBatch][8];
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b)
#pragma unroll
            for (int q = 0; q < 8; ++q) vid[b][q] = nodes[base[b] + (q & 1) * MM + ((q >> 1) & 1) * M + (q >> 2)];
          bool take[kOctBatch];
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b) {
            take[b] = live[b] && flg[b] != 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) vid[b][q] = take[b] ? vid[b][q] : (int16_t)-1;
          }
          uint2 raw[kOctBatch][8];
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b)
#pragma unroll
            for (int q = 0; q < 8; ++q)
              raw[b][q] = vid[b][q] >= 0 ? reinterpret_cast<const uint2*>(O.data_voxels)[start + vid[b][q]]
                                         : make_uint2(0u, 0u);
#pragma unroll
          for (int b = 0; b < kOctBatch; ++b) {
            if (!take[b]) continue;
            if (!first_hit && T < trans_th) {
              first_hit = true;
              inverse_near[r] = 1.0f / (zs[b] + bl);
              net_idx[r] = (int16_t)group;
            }
            float cv[8][4];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const _Float16* hv = reinterpret_cast<const _Float16*>(&raw[b][q]);
#pragma unroll
              for (int c = 0; c < 4; ++c) cv[q][c] = (float)hv[c];
            }
            float v[4];
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
              const float s0 = lerp_ab(lerp_ab(cv[0][ch], cv[1][ch], w[b][0]), lerp_ab(cv[2][ch], cv[3][ch], w[b][0]), w[b][1]);
              const float s1 = lerp_ab(lerp_ab(cv[4][ch], cv[5][ch], w[b][0]), lerp_ab(cv[6][ch], cv[7][ch], w[b][0]), w[b][1]);
              v[ch] = lerp_ab(s0, s1, w[b][2]);
            }
            const float interval = fminf(step, far - zs[b]);
            const float alpha = 1.0f - exp_cw(-1.0f * v[3] * interval);
            const float wgt = alpha * T;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) col[ch] = col[ch] + wgt * v[ch];
            T = T * (1.0f - alpha);
            ++ns;
          }
          zval = z;
          guard += kOctBatch;
        }
      }
      if (tmax[0] < tmax[1]) {
        if (tmax[0] < tmax[2]) { cur[0] += stp[0]; tmax[0] += tdel[0]; }
        else { cur[2] += stp[2]; tmax[2] += tdel[2]; }
      } else {
        if (tmax[1] < tmax[2]) { cur[1] += stp[1]; tmax[1] += tdel[1]; }
        else { cur[2] += stp[2]; tmax[2] += tdel[2]; }
      }
      near = far;
      far = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
      if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= kOctBlocks || cur[1] >= kOctBlocks ||
          cur[2] >= kOctBlocks)
        break;
    }
  }
  if (soft_hit && !first_hit && group != -1) {
    inverse_near[r] = 1.0f / bh;
    net_idx[r] = (int16_t)group;
  }
  frame[3 * r] = col[0];
  frame[3 * r + 1] = col[1];
  frame[3 * r + 2] = col[2];
  if (n_samples) n_samples[r] = ns;
}

int packet_width(const snisr_rays& R);  // render_fp32.hip

}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_render_octree(const snisr_octree* oct, const float* origin, const float* rays_d, const int32_t* visited,
                        int64_t n, int32_t K, int32_t image_width, float trans_th, float early_stop, int32_t soft_hit,
                        float* frame_diffuse, float* inverse_near, int16_t* net_idx, int32_t* n_samples,
                        void* stream) {
  SNISR_REQUIRE(oct, SNISR_ERR_ARG, "render_octree: octree is NULL");
  SNISR_REQUIRE(n >= 0 && K >= 0, SNISR_ERR_ARG, "render_octree: negative size");
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(origin && rays_d && (K == 0 || visited) && frame_diffuse && inverse_near && net_idx, SNISR_ERR_ARG,
                "render_octree: NULL pointer");
  SNISR_REQUIRE(oct->group_map && oct->data_voxels && oct->block_index_map && oct->nodes_index_map &&
                    oct->nodes_sample_flag && oct->voxels_start && oct->centers,
                SNISR_ERR_ARG, "render_octree: octree table is NULL");
  SNISR_REQUIRE(oct->tile_size > 0.0f && oct->voxel_size > 0.0f && oct->sample_step > 0.0f, SNISR_ERR_ARG,
                "render_octree: tile_size, voxel_size and sample_step must be positive");
  SNISR_REQUIRE(oct->num_voxel >= 2 && oct->num_voxel <= 32, SNISR_ERR_ARG,
                "render_octree: num_voxel must lie in [2, 32] (reference: 18)");
  SNISR_REQUIRE(aligned(oct->data_voxels, 8), SNISR_ERR_ALIGN, "render_octree: data_voxels must be 8-byte aligned");
  snisr_rays R{};
  R.n = n;
  R.image_width = image_width;
  hipLaunchKernelGGL(k_render_octree, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, (hipStream_t)stream, *oct, origin,
                     rays_d, visited, n, K, packet_width(R), trans_th, early_stop, soft_hit, frame_diffuse,
                     inverse_near, net_idx, n_samples);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
