/* snisr.h — C ABI of the MI355X-native tiled sparse-voxel ray-marching renderer.
 *
 * Plain C: every argument is a pointer, a size or a POD descriptor.  Device pointers are
 * HIP device memory owned by the caller; `stream` is a hipStream_t passed as void*
 * (NULL = the null stream).  Every entry point returns 0 on success or a negative
 * snisr_status; snisr_last_error() returns a thread-local message for the last failure.
 * Nothing is kept between calls (no __constant__ or global state), so calls on different
 * streams are independent.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference checkout).  The reference is a pybind11/ATen extension whose ops write into
 * caller-allocated tensors and return void (rendering/src/binding.cpp:13-27); the Python
 * layer of this package (snisr.fastrendering) re-exposes the same names on torch tensors.
 */
#ifndef SNISR_H
#define SNISR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum snisr_status {
  SNISR_OK = 0,
  SNISR_ERR_ARG = -1,        /* null pointer, bad size or inconsistent descriptor */
  SNISR_ERR_SHAPE = -2,      /* a shape the kernels do not support */
  SNISR_ERR_ALIGN = -3,      /* a pointer without the alignment the kernel needs */
  SNISR_ERR_DEVICE = -4,     /* a HIP runtime error (launch or query) */
  SNISR_ERR_CAPACITY = -5,   /* the workspace was too small for this call */
  SNISR_ERR_UNSUPPORTED = -6
} snisr_status;

enum { SNISR_F32 = 0, SNISR_F16 = 1, SNISR_BF16 = 2 };

/* Scene in the tiled sparse-brick layout (device pointers).
 *   tile_index_map  i32[Tx*Ty*Tz]  dense tile cell (x*Ty*Tz + y*Tz + z) -> tile id or -1
 *                   (the reference's tile_IndexMap, rendering/src/__init__.py:117)
 *   tile_centers    f32[T,3]       (centers, __init__.py:115)
 *   tile_res        i32[T]         voxels per tile side, a multiple of 8
 *   tile_brick_base i32[T]         offset of the tile's (res/8)^3 brick table in brick_map
 *   brick_map       i32[..]        brick cell ((bx*nb + by)*nb + bz) -> brick id or -1
 *                                  (role of block_IndexMap, __init__.py:118)
 *   sigma           f32[NB,512]    density per voxel of each brick, voxel (x*8+y)*8+z
 *   feats           [NB,512,F]     features, F contiguous per voxel (f32 or f16)
 *   mlp_params      f32[T,P]       per-tile MLP, per layer W[out][in] then b[out];
 *                                  layers in->W, (L-2)x W->W, W->3; input = F features
 *                                  + 15-d view encoding
 *   mlp_f32t        f32 copy of mlp_params from snisr_pack_mlp(SNISR_F32): per layer Wt[in][out4]
 *                   then b[out4], out4 = out rounded up to 4 (required by the FP32 path)
 *   mlp_packed      device buffer from snisr_pack_mlp(SNISR_F16 / SNISR_BF16) (MFMA path)
 */
typedef struct snisr_scene {
  int32_t tile_shape[3];
  float scene_min[3];
  float tile_size;
  int32_t n_tiles;
  int32_t n_bricks;
  int32_t feat_dim;
  int32_t feat_dtype;
  int32_t mlp_in;
  int32_t mlp_width;
  int32_t mlp_layers;
  int32_t mlp_packed_dtype;   /* SNISR_F16 or SNISR_BF16 for mlp_packed */
  int64_t mlp_stride;
  int64_t brick_map_len;
  const int32_t* tile_index_map;
  const float* tile_centers;
  const int32_t* tile_res;
  const int32_t* tile_brick_base;
  const int32_t* brick_map;
  const float* sigma;
  const void* feats;
  const float* mlp_params;
  const float* mlp_f32t;
  const void* mlp_packed;
} snisr_scene;

/* March parameters (rendering/src/__init__.py:37-41; rendering/config/base.yaml:8). */
typedef struct snisr_march_params {
  int32_t max_tiles;     /* K: non-empty tiles per ray */
  float sample_step;     /* world units along the unit direction */
  float early_stop;      /* a ray stops once T < early_stop */
  float trans_th;        /* depth records the first sample reached with T < trans_th */
} snisr_march_params;

/* Rays: origins f32[3] (origin_stride 0) or f32[N,3] (origin_stride 3); dirs f32[N,3]
 * (any length; marching normalises them); t_hit f32[N] or NULL (march bound, e.g. from
 * snisr_mesh_first_hit); visited i32[N,max_tiles] or NULL (NULL: the tile DDA runs inline). */
typedef struct snisr_rays {
  const float* origins;
  int32_t origin_stride;
  int32_t image_width;   /* >0: rays are an image of this width (enables 8x8 packets) */
  int64_t n;
  const float* dirs;
  const float* t_hit;
  const int32_t* visited;
} snisr_rays;

/* Outputs (NULL = not written): rgb f32[N,3]; T f32[N] final transmittance; n_samples i32[N];
 * depth f32[N] (t of the first sample reached with T < trans_th, -1 if none);
 * bricks i32[N,brick_cap] occupied bricks marched in order, -1 padded (debug).
 * stage_events: optional array of n_stage_events hipEvent_t recorded on `stream` at the stage
 * boundaries (render_mfma: start, after the ray schedule, after the march, after the tile histogram
 * and scan, after the tile scatter, after the MLP, after the reduction; render_fp32: before and after
 * its kernel). */
typedef struct snisr_render_out {
  float* rgb;
  float* T;
  int32_t* n_samples;
  float* depth;
  int32_t* bricks;
  int32_t brick_cap;
  int32_t n_stage_events;
  void* const* stage_events;
} snisr_render_out;

const char* snisr_last_error(void);
int snisr_version(void);
/* sizeof of the descriptors (0 scene, 1 march params, 2 rays, 3 render out, 4 mlp query, 5 octree, 6 tuning,
 * 7 dense grid, 8 tile mesh, 9 graph desc) so that foreign
 * bindings can check their struct layouts; no reference counterpart. */
int64_t snisr_abi_sizeof(int32_t which);

/* A1 — compute_rays (rendering/src/compute_rays.cu:46-68, include/compute_rays.h:7-11).
 * intrinsic f32[3] = (cx, cy, 1/f), c2w f32[9] row-major: device pointers.
 * rays_d f32[H,W,3] = R [(px-cx)/f, (py-cy)/f, 1] (not normalised). */
int snisr_compute_rays(const float* intrinsic, const float* c2w, int32_t H, int32_t W, float* rays_d,
                       void* stream);

/* A2 — tracing_tiles (rendering/src/raytile_tracing.cu:63-95, include/raytile_tracing.h:8-13).
 * origin f32[3] device; rays_d f32[n,3]; visited i32[n,K], -1 padded.  Of the scene only the tile
 * grid is read (tile_shape, scene_min, tile_size, tile_index_map); scene_size = tile_shape * tile_size. */
int snisr_trace_tiles(const snisr_scene* scene, const float* origin, const float* rays_d, int64_t n, int32_t K,
                      int32_t* visited, void* stream);

/* A3+A4+A5+A7 — one-thread-per-ray FP32 render: tile DDA, brick DDA, uniform march,
 * trilinear sigma + features, per-tile FP32 MLP, early-terminating compositing.  Replaces
 * rendering_diffuse_octree_fp16 (rendering/src/raydiffuse_render_octree_fp16.cu:238-287,
 * include/raydiffuse_render_octree.h:7-21) and the FP32 marcher render_diffuse_kernel
 * (training/src/cuda_func.cu:1687-1802) for the synthetic feature scenes. */
int snisr_render_fp32(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                      const snisr_render_out* out, void* stream);

/* Launch schedule and record capacity of snisr_render_mfma_tuned (no reference counterpart).
 * ray_order 0: rays in index order (8x8 pixel packets when image_width allows), 1: rays
 * counting-sorted by (first occupied tile, direction octant) before the march; mlp_threads: 256 or
 * 512 threads per MLP workgroup (one workgroup per CU); sort_grid: workgroups of the record / ray
 * counting sorts (chosen per workload by an autotuner; results do not depend on them).
 * records_per_ray: sample-record slots per ray (0 = 64); every other workspace byte goes to the
 * tile-sorted sample arrays.  chunks (0 or 1 = one pass; at most 64; ray_order 0 only): the rays are
 * split into that many contiguous ranges of whole 8-row packet bands (or 64-ray wavefronts); the
 * march of range c runs on the caller's stream while a second stream of the library's sorts and
 * shades range c - 1, and the call joins that stream back before it returns (capturable).  Each
 * range has its own tile-sorted arrays of the full sample capacity.  ray_order 2: rays
 * counting-sorted by (first tile, octant, origin cell of a c^3 grid over the scene box, c as large
 * as 8192 buckets allow); 3: by (first tile, origin cell). */
typedef struct snisr_tuning {
  int32_t ray_order;
  int32_t mlp_threads;
  int32_t sort_grid;
  int32_t records_per_ray;
  int32_t chunks;
} snisr_tuning;

/* Workspace bytes for snisr_render_mfma on n rays with `records_per_ray` record slots per ray
 * (0 = 64; the march writes each ray's records at its own slots, in one pass) and room for
 * `sample_capacity` samples in the tile-sorted arrays (0 = 16 per ray).  A frame with more samples
 * for one ray or in total is reported by snisr_render_mfma_stats (overflow; its rgb is NaN). */
int64_t snisr_render_mfma_workspace_size(const snisr_scene* scene, int64_t n, int64_t records_per_ray,
                                         int64_t sample_capacity);

/* Workspace bytes for snisr_render_mfma_tuned with this schedule (records_per_ray and chunks are
 * read; every chunk gets room for `sample_capacity` samples, 0 = 16 per ray). */
int64_t snisr_render_mfma_workspace_size_tuned(const snisr_scene* scene, int64_t n, const snisr_tuning* tuning,
                                               int64_t sample_capacity);

/* Same render with the per-tile MLP on MFMA (FP16 or BF16 operands, FP32 accumulation,
 * FP32 transmittance).  Stages: one sigma-only march writing every ray's sample records (voxel
 * coordinate, compositing weight, tile) at the ray's own slots, a counting sort of the records by
 * tile, the per-tile MFMA MLP (weights staged in LDS once per tile and workgroup), and an ordered
 * per-ray reduction.  Sample counts, T, depth and the brick lists are bit-identical to
 * snisr_render_fp32; the pixels are deterministic.  Stage events (snisr_render_out): 0 start,
 * 1 after the optional ray schedule, 2 after the march, 3 after the tile histogram, 4 after the
 * record sort, 5 after the MLP (with its trilinear feature gather), 6 after the reduction. */
int snisr_render_mfma(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                      const snisr_render_out* out, void* workspace, int64_t workspace_bytes, void* stream);

/* snisr_render_mfma with an explicit schedule and record capacity (NULL = {0, 512, 512, 0}). */
int snisr_render_mfma_tuned(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                            const snisr_render_out* out, const snisr_tuning* tuning, void* workspace,
                            int64_t workspace_bytes, void* stream);

/* Samples of the last snisr_render_mfma on `workspace`, whether the frame overflowed (a ray had more
 * samples than its record slots, or the frame more than the sample capacity; the affected rgb is
 * NaN) and the largest per-ray sample count (max_count may be NULL; size the next workspace with
 * records_per_ray >= max_count and sample_capacity >= total).  Synchronises `stream`.  No
 * reference counterpart. */
int snisr_render_mfma_stats(const void* workspace, int64_t* total, int32_t* overflow, int32_t* max_count,
                            void* stream);

/* A9 — proxy-mesh nearest hit.  Replaces ray_firstHit_inTile_kernel (training/src/cuda_func.cu:130-282,
 * per-tile per-block face lists BConFaceIdx/BConFaceNum, training/scripts/tools.py:364-379) and its
 * RayTriangleIntersection (rendering/src/include/cuda_utils.h:403-436) with a BVH over the whole mesh.
 *
 * snisr_bvh_build_bounds: host bytes to allocate for snisr_bvh_build's two outputs.
 * snisr_bvh_build (host, synchronous): vertices f32[V,3], faces i32[F,3] -> nodes (64 B each) and
 * leaf-ordered triangles (64 B each); n_nodes and the tree depth are returned.  Upload both arrays
 * to device memory unchanged for snisr_mesh_first_hit, which must be given the same n_nodes, n_tris
 * and depth.
 * snisr_mesh_first_hit: per ray (directions normalised as in the marcher) the smallest t > 0 over
 * all triangles, ties to the smaller triangle id; t = +inf and id = -1 without a hit. */
int snisr_bvh_build_bounds(int64_t n_faces, int64_t* node_bytes, int64_t* tri_bytes);
int snisr_bvh_build(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces, void* nodes_out,
                    int64_t node_bytes, void* tris_out, int64_t tri_bytes, int32_t* n_nodes, int32_t* depth);
int snisr_mesh_first_hit(const void* nodes, int32_t n_nodes, const void* tris, int64_t n_tris, int32_t depth,
                         const snisr_rays* rays, float* t_out, int32_t* id_out, void* stream);

/* A9, per-tile form — ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel
 * (training/src/cuda_func.cu:130-444) with the per-tile per-block face lists BConFaceIdx / BConFaceNum
 * (training/scripts/tools.py:364-379).  Each tile's num_block^3 blocks (block (x*nb + y)*nb + z) list
 * the faces that may be hit inside them: face_num int2[n_tiles * nb^3] = (start, count) into
 * face_idx i32[n_face_idx].  One result per (ray, visited tile): visited i32[n_rays, K] (-1 = none);
 * rays_o / rays_d f32[n_rays, 3] are used as given.  The ray walks the tile's blocks from its AABB
 * entry; hits need t > 0 and a hit point inside the block within 1e-5.
 *   mode 0 (first): hits f32[n_rays, K] = nearest t of the first block with a hit (left unwritten
 *          without a hit, as the reference leaves it), hit_face i32[n_rays, K] its face or -1;
 *   mode 1 (double): hits f32[n_rays, K, 2] = the two nearest t (-1 where absent); hit_face unused.
 * snisr_tile_face_lists (host, synchronous) builds the lists for a mesh: a face is listed in every
 * block whose box, widened by max(2e-5, 1e-4 block edges), it intersects (separating-axis test), so
 * every hit the kernel can accept is listed.  Call it with face_idx NULL to get *total, then again
 * with room for *total entries. */
typedef struct snisr_tile_mesh {
  const float* vertices;     /* [V, 3] */
  const int32_t* faces;      /* [n_faces, 3] */
  int64_t n_faces;
  const float* centers;      /* [n_tiles, 3] tile centres */
  const int32_t* face_idx;   /* [n_face_idx] */
  int64_t n_face_idx;
  const int32_t* face_num;   /* [n_tiles * num_block^3, 2], 8-byte aligned */
  float tile_size;
  int32_t num_block;
} snisr_tile_mesh;

int snisr_mesh_hit_tiles(const snisr_tile_mesh* mesh, const float* rays_o, const float* rays_d, int64_t n_rays,
                         int32_t K, const int32_t* visited, int32_t mode, float* hits, int32_t* hit_face,
                         void* stream);
int snisr_tile_face_lists(const float* vertices, int64_t n_vertices, const int32_t* faces, int64_t n_faces,
                          const float* centers, int32_t n_tiles, float tile_size, int32_t num_block, int32_t* face_num,
                          int32_t* face_idx, int64_t capacity, int64_t* total);

/* A3 — rendering_diffuse_octree_fp16 (rendering/src/raydiffuse_render_octree_fp16.cu:238-287,
 * include/raydiffuse_render_octree.h:7-21) on the reference's octree scene format.  Tiles hold 4^3
 * blocks; an occupied block is a node of num_voxel^3 dilated cells (reference: 18). */
typedef struct snisr_octree {
  const int32_t* group_map;          /* tile2group [n_tiles] */
  const void* data_voxels;           /* half4 (r, g, b, sigma), 8-byte aligned */
  const int32_t* block_index_map;    /* [n_tiles * 64] node of each block, -1 = empty */
  const int16_t* nodes_index_map;    /* [n_nodes * num_voxel^3] voxel of each cell within its node, -1 = empty */
  const uint8_t* nodes_sample_flag;  /* [n_nodes * num_voxel^3] cells that are sampled (bool) */
  const int64_t* voxels_start;       /* [n_nodes] first voxel of each node in data_voxels */
  const float* centers;              /* [n_tiles * 3] tile centres */
  float tile_size, voxel_size, sample_step;
  int32_t num_voxel;
} snisr_octree;

/* Rays: origin f32[3], rays_d f32[n,3] (not necessarily unit), visited i32[n,K] from
 * snisr_trace_tiles.  Outputs per ray: frame_diffuse f32[n,3] (always written), inverse_near f32[n]
 * and net_idx i16[n] (written at the first sample with T < trans_th, or by soft_hit; otherwise left
 * as the caller initialised them: the reference uses 1e-8 and -1); n_samples i32[n] optional (NULL).
 * image_width > 0 lets the kernel walk 8x8 pixel packets (n a multiple of 8 rows of that width). */
int snisr_render_octree(const snisr_octree* octree, const float* origin, const float* rays_d, const int32_t* visited,
                        int64_t n, int32_t K, int32_t image_width, float trans_th, float early_stop, int32_t soft_hit,
                        float* frame_diffuse, float* inverse_near, int16_t* net_idx, int32_t* n_samples,
                        void* stream);

/* A10 — render_diffuse_kernel (training/src/cuda_func.cu:1687-1802): the reference's plain FP32
 * marcher over dense per-tile grids (the training tiles' Voxel-T / Node-T grids, training/scripts/tile/
 * tile.py:415-432).  A grid slot holds num_voxel^3 dilated cells (NUM_VOXEL + 2 per side), cell
 * (x*num_voxel + y)*num_voxel + z, of float4 (r, g, b, sigma) and an int16 flag (only 1 is sampled);
 * sparse_to_group maps a tile id to its slot (-1: the tile is outside the group, which ends the ray
 * as invalid).  Per ray (origin and direction f32[n,3], used as given) over visited i32[n,K]: the
 * tile's AABB chord is marched from its entry in sample_step; flagged cells are interpolated with
 * the OVERLAP rule and composited in FP32 until T < 1e-5.  Outputs: diffuse f32[n,3]; depth f32[n] =
 * the first sample reached with T <= 0.5 (-1 when invalid); valid i32[n]. */
typedef struct snisr_dense_grid {
  const float* voxels;             /* [slots, num_voxel^3, 4], 16-byte aligned */
  const int16_t* nodes;            /* [slots, num_voxel^3] */
  const int32_t* sparse_to_group;  /* [n_tiles] tile -> slot or -1 */
  const float* centers;            /* [n_tiles, 3] tile centres */
  float tile_size, voxel_size, sample_step;
  int32_t num_voxel;
} snisr_dense_grid;

int snisr_render_dense(const snisr_dense_grid* grid, const float* rays_o, const float* rays_d, const int32_t* visited,
                       int64_t n, int32_t K, float* diffuse, float* depth, int32_t* valid, void* stream);

/* A6 — inverse_z_sampling (rendering/src/sampling.cu:48-81, include/sampling.h:7-15).  For pixel p
 * (row-major H x W) and point j < sample_per_ray: z = 1/((offset+j) ib[p] + in[p]), step =
 * max(z_{j+1} - z_j, 0), point = origin + z d[p].  rays_d may be a strided view: element (y, x, c) at
 * rays_d[y*d_row_stride + x*d_col_stride + c] (floats), as the reference's rays_d[::2, ::2]. */
int snisr_inverse_z_sampling(const float* origin, const float* rays_d, int64_t d_row_stride, int64_t d_col_stride,
                             int32_t H, int32_t W, const float* inverse_near, const float* inverse_bound, int32_t offset,
                             int32_t sample_per_ray, float* sample_points, float* integrate_steps, void* stream);

/* A5 — integrate_points (rendering/src/volume_rendering.cu:60-87, include/volume_rendering.h:7-12):
 * 16 samples per pixel of sample_results f32[npix,16,4] (rgb, sigma) and steps f32[npix,16];
 * transparency f32[npix] and colors f32[npix,3] are updated in place; pixels with T < early_stop
 * are left untouched. */
int snisr_integrate_points(const float* sample_results, const float* integrate_steps, int64_t npix, float early_stop,
                           float* transparency, float* colors, void* stream);

/* A8 — sort_by_key (rendering/src/reorder.cu:11-23, include/reorder.h:7): stable ascending sort of
 * keys i16[n] carrying values i32[n], then the run heads (unique_by_key on keys with starts): the
 * first m keys are the distinct keys and starts[0..m) the starts[] entries at the head of each run.
 * m is returned through *n_unique (host pointer; the call synchronises `stream`, like the
 * reference's int return).  Needs snisr_sort_by_key_workspace_size(n) bytes of device workspace. */
int64_t snisr_sort_by_key_workspace_size(int64_t n);
int snisr_sort_by_key(int16_t* keys, int32_t* values, int32_t* starts, int64_t n, void* workspace,
                      int64_t workspace_bytes, int32_t* n_unique, void* stream);

/* compute_group_boundary (rendering/src/compute_group_boundary.cu:53-75, include/compute_group_boundary.h):
 * boundary u8[H*W] (bool) is set to 1 where a pixel of the ksize x ksize window around the pixel
 * (clipped to the image) has another group id than the pixel; set flags stay set.  net_idx i16[H*W]. */
int snisr_compute_group_boundary(const int16_t* net_idx, int32_t H, int32_t W, int32_t ksize, uint8_t* boundary,
                                 void* stream);

/* A8 — assign_pixels (rendering/src/reorder.cu:69-94, include/reorder.h:10-15).  n_blocks =
 * block_starts[n_net], the number of output chunks (the reference's caller sizes the outputs with it). */
int snisr_assign_pixels(const int16_t* net, const int32_t* pixel_starts, const int32_t* block_starts, int64_t n_net,
                        int64_t n_blocks, int16_t* new_net, int32_t* new_pixel_starts, void* stream);

/* A7 — inference_mlp (rendering/src/rayreflection_render.cu:111-335, include/rayreflection_render.h:7-19).
 * The reflection MLP of every sample of the listed pixel blocks:  block b covers sorted pixels
 * [pixel_starts[b], pixel_starts[b+1]) of query_pixel_indices and uses group net_idx[b] (-1: skip);
 * a pixel whose transparency is below trans_stop is skipped; sample j of original pixel o reads
 * samples[o*spr + j] and writes sample_results[o*spr + j] = (rgb, sigma).  Parameters come packed by
 * snisr_pack_inference_mlp from the reference's 22,556 halves per group.  FP16 MFMA operands, FP32
 * accumulation. */
typedef struct snisr_mlp_query {
  int32_t samples_per_ray;            /* max_samples_per_ray: samples per pixel in this call */
  int32_t n_groups;                   /* groups in packed_params / group_centers */
  int64_t n_blocks;                   /* num_block: entries of net_idx; pixel_starts holds n_blocks + 1 */
  float inverse_far, trans_stop;
  const int16_t* net_idx;             /* [n_blocks] */
  const int32_t* pixel_starts;        /* [n_blocks + 1], positions in query_pixel_indices */
  const int32_t* query_pixel_indices; /* [n_query] sorted position -> original pixel (y*W + x) */
  int64_t n_query;
  const float* transparency;          /* [H*W] */
  const float* group_centers;         /* [n_groups, 3] */
  const float* samples;               /* [H*W, samples_per_ray, 3] */
  const float* rays_d;                /* element (y, x, c) at rays_d[y*d_row_stride + x*d_col_stride + c] */
  int64_t d_row_stride, d_col_stride;
  int32_t H, W;
  const void* packed_params;          /* snisr_pack_inference_mlp output, 16-byte aligned */
  float* sample_results;              /* [H*W, samples_per_ray, 4], 16-byte aligned; skipped samples untouched */
  const int32_t* n_blocks_device;     /* NULL, or a device count of the blocks to process (at most n_blocks),
                                         written by snisr_regroup_pixels: no host round trip for it */
} snisr_mlp_query;

int64_t snisr_inference_mlp_packed_size(int64_t n_groups);
int snisr_pack_inference_mlp(const void* params_f16, int64_t n_groups, void* packed, void* stream);
int snisr_inference_mlp(const snisr_mlp_query* query, void* stream);

/* The per-frame pixel regrouping of FRender.render (sort_by_key, the host-side block_starts
 * and assign_pixels: rendering/src/__init__.py:272-291, reorder.cu:11-94) as one device-side pass
 * with no host synchronisation (the reference makes two per frame):  query i32[n] = pixel indices
 * stably sorted by net_idx i16[n]; the blocks of at most pixels_per_block consecutive sorted pixels
 * of one group, in group order: block_net i16[max_blocks] (-1 past the last block),
 * block_starts i32[max_blocks + 1] (positions in query, n past the last block), and the block count
 * in n_blocks (device).  Blocks of group -1 are listed like the reference's (inference_mlp skips
 * them).  max_blocks >= snisr_regroup_max_blocks(n, pixels_per_block) always suffices. */
int64_t snisr_regroup_max_blocks(int64_t n, int32_t pixels_per_block);
int64_t snisr_regroup_workspace_size(int64_t n);
int snisr_regroup_pixels(const int16_t* net_idx, int64_t n, int32_t pixels_per_block, int64_t max_blocks,
                         int32_t* query, int16_t* block_net, int32_t* block_starts, int32_t* n_blocks,
                         void* workspace, int64_t workspace_bytes, void* stream);

/* Allocation-free form of snisr_regroup_pixels.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The same regrouping and outputs, with the workspace in a fixed device array of the
 * library (about 10 MB per GPU, held until the process ends) instead of a caller workspace; n is at
 * most snisr_regroup_pixels_static_max() (1,048,576 pixels), larger n returns SNISR_ERR_CAPACITY.
 * Calls on one GPU share that array: calls that overlap (other streams or threads) race. */
int64_t snisr_regroup_pixels_static_max(void);
int snisr_regroup_pixels_static(const int16_t* net_idx, int64_t n, int32_t pixels_per_block, int64_t max_blocks,
                                int32_t* query, int16_t* block_net, int32_t* block_starts, int32_t* n_blocks,
                                void* stream);

/* A11 — training (FP32 features and MLP).  snisr_train_forward renders like snisr_render_fp32
 * (identical traversal and samples) and keeps per-sample records in the workspace;
 * snisr_train_backward then computes the gradients of L = sum_i <grad_rgb_i, rgb_i> with respect to
 * sigma f32[NB,512], the features f32[NB,512,F] and the MLP parameters f32[T,P] (mlp_params layout),
 * overwriting the three gradient buffers.  The MLP forward is recomputed in the backward rather than
 * stored.  Replaces the forward/autograd backward of Tile.train_batch
 * (training/scripts/tile/tile.py:631-820): compositing of utils.volume_rendering
 * (training/scripts/utils.py:823-837), grid_sample trilinear gathers (tile.py:225-236) and the MLP
 * (training/scripts/networks.py:17-41).  The workspace must not be reused between a forward and its
 * backward; the forward packs the MLP weights of scene->mlp_params into it and the backward uses
 * that copy, so the parameters may change only after the backward.  train_forward's stage_events
 * (4): start, after march+sort, after the MLP, after the reduction. */
int64_t snisr_train_workspace_size(const snisr_scene* scene, int64_t n, int64_t sample_capacity);
int snisr_train_forward(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                        const snisr_render_out* out, void* workspace, int64_t workspace_bytes, void* stream);
int snisr_train_backward(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                         const float* grad_rgb, float* dsigma, float* dfeats, float* dmlp, void* workspace,
                         int64_t workspace_bytes, void* const* stage_events, int32_t n_stage_events, void* stream);

/* A5 training form — utils.volume_rendering (training/scripts/utils.py:823-837), forward and
 * hand-written adjoint.  Per ray r over its n samples (arrays [n_rays, n], rgb [n_rays, n, 3]):
 * alpha_i = 1 - exp(-sigma_i dists_i), trans_i = init_T_r prod_{j<i} (1 - alpha_j + 1e-10) (exclusive
 * cumprod), out_r = sum_i alpha_i trans_i rgb_i.  init_T NULL = 1 (the in-tile pass); the background
 * pass chained behind the in-tile pass takes the in-tile trans[:, n-1] (tile.py:703-704).  Outputs
 * out [n_rays, 3], trans and alpha [n_rays, n] (all three, as the reference returns them).  The
 * backward takes the forward's trans and alpha and dL/dout [n_rays, 3] plus optional dL/dtrans and
 * dL/dalpha (NULL = 0) and writes dL/drgb, dL/dsigma and, when g_init is not NULL, dL/dinit_T. */
int snisr_volume_render_forward(const float* rgb, const float* sigma, const float* dists, const float* init_T,
                                int64_t n_rays, int32_t n, float* out, float* trans, float* alpha, void* stream);
int snisr_volume_render_backward(const float* rgb, const float* dists, const float* trans, const float* alpha,
                                 int64_t n_rays, int32_t n, const float* g_out, const float* g_trans,
                                 const float* g_alpha, float* g_rgb, float* g_sigma, float* g_init, void* stream);

/* A7 training form — the specular head of Tile.inference_specular (training/scripts/tile/tile.py:
 * 198-223) with eval_sh (training/scripts/sh.py:55-111): spec[s][c] = sigmoid(sum_b Y_b(d)
 * coeffs[s][3 b + c]) for the n samples s (rows of coeff_stride floats, the MLP output whose first
 * 3 (deg+1)^2 entries are [basis][rgb]), d the normalised direction dirs[s / samples_per_ray], real
 * SH of degree deg <= 2.  The backward takes the forward's spec and dL/dspec and writes dL/dcoeffs
 * into entries 0..26 of each row of g_coeffs (zero above the degree); other entries are untouched. */
int snisr_sh_sigmoid_forward(const float* coeffs, int32_t coeff_stride, const float* dirs, int64_t n,
                             int32_t samples_per_ray, int32_t deg, float* spec, void* stream);
int snisr_sh_sigmoid_backward(const float* spec, const float* dirs, int64_t n, int32_t samples_per_ray, int32_t deg,
                              const float* g_spec, float* g_coeffs, int32_t coeff_stride, void* stream);

/* A7/A11 training form — networks.MLP (training/scripts/networks.py:4-41), FP32 forward and
 * hand-written adjoint: Positional_Encoding with L = 10 frequencies of a 3-d point ([x, sin(2^0 x),
 * cos(2^0 x), ..., cos(2^9 x)], 63 values), then `depth` Linear layers of width 64 with LeakyReLU(slope)
 * between them (slope 0 = ReLU) and n_out outputs: the surface MLP (depth 8, n_out 4, ReLU) and the
 * reflection MLP (depth 6, n_out 28, LeakyReLU 0.01) of the reference's configuration
 * (training/config/cfg.yaml:38-46).  params: snisr_pemlp_param_count floats in nn.Linear order (per
 * layer W[out][in], then b[out]); slope >= 0.  points f32[n, 3], out f32[n, n_out] (16-byte aligned).
 * Both calls pack the weights into `workspace` first: the forward needs
 * snisr_pemlp_workspace_size(0, ...) bytes, the backward snisr_pemlp_workspace_size(n, ...) (the
 * weights plus each layer's input and delta rows).  The backward recomputes the forward from the same
 * points, takes dL/dout and writes dL/dparams (overwritten). */
int64_t snisr_pemlp_param_count(int32_t L, int32_t depth, int32_t n_out);
int64_t snisr_pemlp_workspace_size(int64_t n, int32_t L, int32_t depth, int32_t n_out);
int snisr_pemlp_forward(const float* params, const float* points, int64_t n, int32_t L, int32_t depth, int32_t n_out,
                        float slope, float* out, void* workspace, int64_t workspace_bytes, void* stream);
int snisr_pemlp_backward(const float* params, const float* points, int64_t n, int32_t L, int32_t depth, int32_t n_out,
                         float slope, const float* g_out, float* g_params, void* workspace, int64_t workspace_bytes,
                         void* stream);

/* A11 training form — the diffuse branch around the surface MLP (Tile.compute_voxels,
 * Tile.inference_diffuse and Tile.query_voxel, training/scripts/tile/tile.py:175-236, 283-295).
 * The voxel grid of a tile is f32[D][D][D][4] channel-last ([z][y][x][r, g, b, sigma]; D = the
 * dilated node count NUM_VOXEL + 2), the memory order of the reference's permuted voxels view.
 *  - flag_cells: the row-major indices of the cells whose node flag (i16[D^3]) is not -1, in order,
 *    into cells i32[D^3 capacity] and their number into *count (device memory); workspace of
 *    snisr_flag_cells_workspace_size bytes;
 *  - bake_points: MLP inputs of those cells, ((index + 0.5) voxel_size + min_corner - center) / scale
 *    per axis (regular_sample, tile.py:91-95, then norm_diffuse_point, tileGroup.py:74), f32[n, 3];
 *  - normalize_points: (x - center) / scale of world points (ray samples), f32[n, 3];
 *  - diffuse_head_forward: raw f32[n, 4] MLP outputs -> (sigmoid(r0..r2), relu(r3)); with cells, row k
 *    goes to row cells[k] of out f32[n_rows, 4] after the n_rows rows are zeroed (the bake), without
 *    cells out has n_rows = n rows; a cell index outside 0..n_rows-1 is skipped.  The backward reads
 *    dL/dout at the same rows (zero for a skipped index) and writes dL/draw;
 *  - voxel_query_forward: F.grid_sample (trilinear, align_corners, zero padding) of the grid at world
 *    points f32[n, 3] through the reference's coordinate chain ((x - min_corner) / voxel_size - 0.5,
 *    then / (D - 1) * 2 - 1, then the sampler's ((g + 1) / 2) (D - 1)); rgb f32[n, 3], sigma f32[n].
 *    The backward zeroes g_voxels f32[D^3, 4] and accumulates dL/dgrid (either gradient may be NULL). */
int64_t snisr_flag_cells_workspace_size(int64_t n_cells);
int snisr_flag_cells(const int16_t* flags, int64_t n_cells, int32_t* cells, int32_t* count, void* workspace,
                     int64_t workspace_bytes, void* stream);
int snisr_bake_points(const int32_t* cells, int64_t n, int32_t D, float mc_x, float mc_y, float mc_z, float voxel_size,
                      float center_x, float center_y, float center_z, float scale, float* points, void* stream);
int snisr_normalize_points(const float* x, int64_t n, float center_x, float center_y, float center_z, float scale,
                           float* out, void* stream);
int snisr_diffuse_head_forward(const float* raw, int64_t n, const int32_t* cells, int64_t n_rows, float* out,
                               void* stream);
int snisr_diffuse_head_backward(const float* raw, int64_t n, const int32_t* cells, int64_t n_rows, const float* g_out,
                                float* g_raw, void* stream);
int snisr_voxel_query_forward(const float* voxels, int32_t D, const float* points, int64_t n, float mc_x, float mc_y,
                              float mc_z, float voxel_size, float* rgb, float* sigma, void* stream);
int snisr_voxel_query_backward(int32_t D, const float* points, int64_t n, float mc_x, float mc_y, float mc_z,
                               float voxel_size, const float* g_rgb, const float* g_sigma, float* g_voxels,
                               void* stream);

/* Training samplers of the reference's CUDA_EXT (training/src/Sample.cpp:97-158, kernels in
 * training/src/Sample_kernel.cu): per ray of one tile (centre cx, cy, cz; edge tile_size), the sample
 * depths of its training batch, written in place into z_vals f32[n, n_samples] (and dists for
 * sparse); rays the reference skips keep their values.  node_flags i16[num_voxel^3] (1 = sampled
 * voxel, index (x*nv + y)*nv + z); voxels f32[num_voxel^3][4] (r, g, b, sigma).
 *  - sample_sparse      Sample_sparse (Sample_kernel.cu:181-434): length-proportional samples over
 *                       runs of flagged voxels, 4 before the first run and 4 after the last;
 *  - sample_reflection  Sample_reflection (16-179): inverse-depth samples from the exit of the voxel
 *                       where transmittance reaches T_th to `far`;
 *  - sample_uniform     Sample_uniform (436-479, 544-578): uniform over the tile's chord;
 *  - sample_bg          Sample_bg (481-542, 580-616): uniform over sample_range around bg_zval[i],
 *                       behind the tile's exit; tile_center is a device f32[3]. */
int snisr_sample_sparse(const float* rays_o, const float* rays_d, int64_t n, const int16_t* node_flags, float cx,
                        float cy, float cz, float tile_size, int32_t num_voxel, int32_t n_samples, float* z_vals,
                        float* dists, void* stream);
int snisr_sample_reflection(const float* rays_o, const float* rays_d, int64_t n, const int16_t* node_flags,
                            const float* voxels, float cx, float cy, float cz, float tile_size, float far,
                            int32_t num_voxel, float T_th, int32_t n_samples, float* z_vals, void* stream);
int snisr_sample_uniform(const float* rays_o, const float* rays_d, int64_t n, float cx, float cy, float cz,
                         float tile_size, int32_t n_samples, float* z_vals, void* stream);
int snisr_sample_bg(const float* rays_o, const float* rays_d, const float* bg_zval, int64_t n, const float* tile_center,
                    float tile_size, float sample_range, int32_t n_samples, float* z_vals, void* stream);

/* Bytes of a packed weight buffer of the scene's MLP shape for `dtype`, and the packing itself
 * (device to device, from scene->mlp_params into `packed`): SNISR_F32 gives the mlp_f32t layout,
 * SNISR_F16 / SNISR_BF16 the MFMA fragment layout of mlp_packed. */
int64_t snisr_pack_mlp_size(const snisr_scene* scene, int32_t dtype);
int snisr_pack_mlp(const snisr_scene* scene, int32_t dtype, void* packed, void* stream);

/* One Adam step on a flat FP32 tensor (the training chain's optimizer: torch.optim.Adam(params,
 * weight_decay=1e-5), training/scripts/tile/tileGroup.py:342, stepped at tile.py:780-785), in one
 * pass over param, grad and the two moment buffers (all f32[n], updated in place), with torch's
 * formulas and operation order; step counts from 1, the scalars are formed in double precision as
 * torch forms them. */
int snisr_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, double lr,
                    double beta1, double beta2, double eps, double weight_decay, int64_t step, void* stream);

/* Screen-space refinement U-Net (rendering/scripts/networks.py:5-73: CNN / Unet with 5 input
 * channels; applied in FRender.render at rendering/src/__init__.py:336-339, where the reference runs
 * it through torch2trt in FP16, __init__.py:85-88).  Replaces that TensorRT engine with FP16 MFMA
 * convolutions, FP32 accumulation.
 * snisr_unet_pack: weights[l] / biases[l] are device pointers to conv{l+1}.weight f32[cout, cin, 3, 3]
 *   and conv{l+1}.bias f32[cout] of the module (cin/cout of conv1..conv10: 5/64, 64/32, 32/64, 64/64,
 *   64/128, 128/128, 192/64, 64/64, 96/32, 32/3); packed (device, 256-byte aligned) receives
 *   snisr_unet_packed_size() bytes.
 * snisr_unet_refine: out f32[H, W, 3] = min(ReLU(Unet(x)), 1) for x = (frame rgb f32[H, W, 3],
 *   inverse_near f32[H, W], boundary u8[H, W] as 0/1); the reference clamps the network output to
 *   [0, 1].  H and W are multiples of 4; the workspace (256-byte aligned) holds
 *   snisr_unet_workspace_size(H, W) bytes (-1 with an error for bad sizes). */
int64_t snisr_unet_packed_size(void);
int snisr_unet_pack(const float* const* weights, const float* const* biases, void* packed, void* stream);
int64_t snisr_unet_workspace_size(int32_t H, int32_t W);
int snisr_unet_refine(const void* packed, const float* frame, const float* inverse_near, const uint8_t* boundary,
                      int32_t H, int32_t W, void* workspace, int64_t workspace_bytes, float* out, void* stream);

/* Scene handle and captured frames: the native host side of the renderer.
 *
 * snisr_scene_create uploads a scene whose pointers are HOST memory (same descriptor, same layout and
 * sizes as the device form: tile tables, brick_map_len brick entries, n_bricks bricks of sigma and
 * feats, n_tiles x mlp_stride MLP parameters) into device buffers owned by the handle, packs the FP32
 * MLP and, with mfma_dtype SNISR_F16 / SNISR_BF16, the MFMA fragments (the role of FRender.
 * render_prepare, rendering/src/__init__.py:98-135).  snisr_scene_get_desc returns the device
 * descriptor for the stateless operators above.
 *
 * snisr_graph_create builds a frame driver for H x W camera frames of that scene (the role of
 * FRender.render, rendering/src/__init__.py:247-341): device buffers for the camera, rays, mesh
 * bound and outputs; for the MFMA path a workspace sized from an eager render of the given camera;
 * then the step (snisr_compute_rays, snisr_mesh_first_hit when a BVH is given, snisr_render_fp32 or
 * snisr_render_mfma_tuned) captured once into a HIP graph on the driver's own stream.
 * snisr_graph_render copies a new camera (host intrinsic[3], c2w[9], origin[3]; all NULL keeps the
 * last one) and replays the graph, ordered after and before `stream` (NULL: the driver's stream).
 * snisr_graph_outputs returns the device outputs (rgb f32[H*W,3], T, n_samples, depth, t_hit or NULL),
 * overwritten by every replay.  snisr_graph_stats reports the last replay's samples and overflow
 * (a view with more samples than the sized workspace renders NaN; create a new graph for it) and
 * synchronises.  No reference counterpart for the handle types themselves. */
typedef struct snisr_scene_handle snisr_scene_handle;
typedef struct snisr_graph snisr_graph;

typedef struct snisr_graph_desc {
  int32_t H, W;
  int32_t mlp_dtype;          /* SNISR_F32: the plain FP32 path; the scene's MFMA dtype: the MFMA render */
  int32_t bvh_n_nodes;
  const void* bvh_nodes;      /* device BVH from snisr_bvh_build, or NULL (no mesh bound) */
  const void* bvh_tris;
  int64_t bvh_n_tris;
  int32_t bvh_depth;
  int32_t records_per_ray;    /* 0: sized by the first render */
  int64_t sample_capacity;    /* 0: sized by the first render */
  float sample_headroom;      /* sorted-sample room over the first render's samples (e.g. 0.05) */
  snisr_march_params params;
} snisr_graph_desc;

int snisr_scene_create(const snisr_scene* host, int32_t mfma_dtype, snisr_scene_handle** out);
int snisr_scene_destroy(snisr_scene_handle* scene);
const snisr_scene* snisr_scene_get_desc(const snisr_scene_handle* scene);
int snisr_graph_create(const snisr_scene_handle* scene, const snisr_graph_desc* desc, const float* intrinsic,
                       const float* c2w, const float* origin, snisr_graph** out);
int snisr_graph_render(snisr_graph* graph, const float* intrinsic, const float* c2w, const float* origin,
                       void* stream);
int snisr_graph_outputs(const snisr_graph* graph, float** rgb, float** T, int32_t** n_samples, float** depth,
                        float** t_hit);
int snisr_graph_stats(snisr_graph* graph, int64_t* total, int32_t* overflow, int32_t* records_per_ray,
                      int64_t* sample_capacity);
int snisr_graph_synchronize(snisr_graph* graph);
int snisr_graph_destroy(snisr_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* SNISR_H */
