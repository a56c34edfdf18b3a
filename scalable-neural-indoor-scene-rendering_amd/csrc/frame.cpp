// Native host side of the renderer: a scene handle that owns the device copies of a scene (the role
// of FRender.render_prepare, rendering/src/__init__.py:98-135) and a frame driver that captures one
// camera step (camera rays, proxy-mesh nearest hit, march + MLP + compositing) into a HIP graph once
// and replays it per frame with the camera as the only per-frame input (the role of FRender.render /
// render_one_frame, rendering/src/__init__.py:236-341, for the feature scenes).  Everything here is
// host C++ over the library's own C ABI; the graph holds no memset and no host synchronisation.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "snisr_common.hpp"

struct snisr_scene_handle {
  snisr_scene desc;                 // device descriptor handed to the stateless operators
  std::vector<void*> allocations;   // device buffers owned by the handle
};

struct snisr_graph {
  const snisr_scene_handle* scene = nullptr;
  snisr_graph_desc cfg{};
  hipStream_t stream = nullptr;     // the library's own stream (capture needs a non-null stream)
  float* camera = nullptr;          // device: intrinsic[3], c2w[9], origin[3]
  float* rays = nullptr;
  float* t_hit = nullptr;
  int32_t* tri = nullptr;
  float *rgb = nullptr, *T = nullptr, *depth = nullptr;
  int32_t* n_samples = nullptr;
  void* workspace = nullptr;
  int64_t workspace_bytes = 0;
  int32_t records_per_ray = 0;
  int64_t sample_capacity = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace snisr {
namespace {

int dev_alloc(std::vector<void*>& owned, void** p, int64_t bytes) {
  *p = nullptr;
  if (bytes <= 0) bytes = 16;
  SNISR_HIP(hipMalloc(p, (size_t)bytes));
  owned.push_back(*p);
  return SNISR_OK;
}

int upload(std::vector<void*>& owned, const void* host, int64_t bytes, const void** dev) {
  void* p = nullptr;
  if (int e = dev_alloc(owned, &p, bytes)) return e;
  if (bytes > 0) SNISR_HIP(hipMemcpy(p, host, (size_t)bytes, hipMemcpyHostToDevice));
  *dev = p;
  return SNISR_OK;
}

void release(std::vector<void*>& owned) {
  for (void* p : owned) (void)hipFree(p);
  owned.clear();
}

// the render step of a graph (also run eagerly to size the workspace)
int graph_step(snisr_graph* g) {
  const snisr_graph_desc& c = g->cfg;
  const hipStream_t st = g->stream;
  const int64_t n = (int64_t)c.H * c.W;
  if (int e = snisr_compute_rays(g->camera, g->camera + 3, c.H, c.W, g->rays, st)) return e;
  snisr_rays r{};
  r.origins = g->camera + 12;
  r.origin_stride = 0;
  r.image_width = c.W;
  r.n = n;
  r.dirs = g->rays;
  if (c.bvh_nodes) {
    if (int e = snisr_mesh_first_hit(c.bvh_nodes, c.bvh_n_nodes, c.bvh_tris, c.bvh_n_tris, c.bvh_depth, &r, g->t_hit,
                                     g->tri, st))
      return e;
    r.t_hit = g->t_hit;
  }
  snisr_render_out o{};
  o.rgb = g->rgb;
  o.T = g->T;
  o.n_samples = g->n_samples;
  o.depth = g->depth;
  if (c.mlp_dtype == SNISR_F32) return snisr_render_fp32(&g->scene->desc, &c.params, &r, &o, st);
  snisr_tuning tn{0, 512, 512, g->records_per_ray, 1};
  return snisr_render_mfma_tuned(&g->scene->desc, &c.params, &r, &o, &tn, g->workspace, g->workspace_bytes, st);
}

int set_camera(snisr_graph* g, const float* intrinsic, const float* c2w, const float* origin) {
  float cam[15];
  std::memcpy(cam, intrinsic, 3 * sizeof(float));
  std::memcpy(cam + 3, c2w, 9 * sizeof(float));
  std::memcpy(cam + 12, origin, 3 * sizeof(float));
  SNISR_HIP(hipMemcpyAsync(g->camera, cam, sizeof cam, hipMemcpyHostToDevice, g->stream));
  SNISR_HIP(hipStreamSynchronize(g->stream));   // cam lives on this stack frame
  return SNISR_OK;
}

void destroy_graph(snisr_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  void* bufs[] = {g->camera, g->rays, g->t_hit, g->tri, g->rgb, g->T, g->depth, g->n_samples, g->workspace};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // namespace
}  // namespace snisr

using namespace snisr;

extern "C" {

int snisr_scene_create(const snisr_scene* host, int32_t mfma_dtype, snisr_scene_handle** out) {
  SNISR_REQUIRE(host && out, SNISR_ERR_ARG, "scene_create: NULL pointer");
  *out = nullptr;
  SNISR_REQUIRE(mfma_dtype == SNISR_F32 || mfma_dtype == SNISR_F16 || mfma_dtype == SNISR_BF16, SNISR_ERR_ARG,
                "scene_create: mfma_dtype must be SNISR_F32 (no MFMA packing), SNISR_F16 or SNISR_BF16");
  if (int e = check_scene(host, true)) return e;
  SNISR_REQUIRE(host->brick_map_len > 0 && host->mlp_stride > 0, SNISR_ERR_ARG,
                "scene_create: brick_map_len and mlp_stride must be positive");
  const int64_t T = host->n_tiles, nb = host->n_bricks;
  const int64_t shape = (int64_t)host->tile_shape[0] * host->tile_shape[1] * host->tile_shape[2];
  const int64_t fe = host->feat_dtype == SNISR_F16 ? 2 : 4;
  snisr_scene_handle* h = new (std::nothrow) snisr_scene_handle();
  SNISR_REQUIRE(h, SNISR_ERR_DEVICE, "scene_create: out of host memory");
  h->desc = *host;
  std::vector<void*>& own = h->allocations;
  const void* p = nullptr;
  int e = SNISR_OK;
  auto up = [&](const void* src, int64_t bytes) {
    if (!e) e = upload(own, src, bytes, &p);
    return e ? nullptr : p;
  };
  h->desc.tile_index_map = (const int32_t*)up(host->tile_index_map, 4 * shape);
  h->desc.tile_centers = (const float*)up(host->tile_centers, 12 * T);
  h->desc.tile_res = (const int32_t*)up(host->tile_res, 4 * T);
  h->desc.tile_brick_base = (const int32_t*)up(host->tile_brick_base, 4 * T);
  h->desc.brick_map = (const int32_t*)up(host->brick_map, 4 * host->brick_map_len);
  // at least one brick of sigma and features: the MFMA gather reads brick 0 for empty corners
  const int64_t nb1 = nb > 0 ? nb : 1;
  auto up_bricks = [&](const void* src, int64_t per_brick) -> const void* {
    void* q = nullptr;
    if (!e) e = dev_alloc(own, &q, per_brick * nb1);
    if (!e && nb == 0 && hipMemset(q, 0, (size_t)per_brick) != hipSuccess) e = SNISR_ERR_DEVICE;
    if (!e && nb > 0 && hipMemcpy(q, src, (size_t)(per_brick * nb), hipMemcpyHostToDevice) != hipSuccess)
      e = SNISR_ERR_DEVICE;
    if (e == SNISR_ERR_DEVICE) set_error("scene_create: brick upload failed");
    return e ? nullptr : q;
  };
  h->desc.sigma = (const float*)up_bricks(host->sigma, 2048);
  h->desc.feats = up_bricks(host->feats, 512 * host->feat_dim * fe);
  h->desc.mlp_params = (const float*)up(host->mlp_params, 4 * T * host->mlp_stride);
  h->desc.mlp_f32t = nullptr;
  h->desc.mlp_packed = nullptr;
  if (!e) {   // FP32 packing (FP32 path) and, if asked, the MFMA fragments
    void* q = nullptr;
    const int64_t b32 = snisr_pack_mlp_size(&h->desc, SNISR_F32);
    e = b32 > 0 ? dev_alloc(own, &q, b32) : SNISR_ERR_ARG;
    if (!e) e = snisr_pack_mlp(&h->desc, SNISR_F32, q, nullptr);
    if (!e) h->desc.mlp_f32t = (const float*)q;
    if (!e && mfma_dtype != SNISR_F32) {
      const int64_t bm = snisr_pack_mlp_size(&h->desc, mfma_dtype);
      e = bm > 0 ? dev_alloc(own, &q, bm) : SNISR_ERR_SHAPE;
      if (!e) e = snisr_pack_mlp(&h->desc, mfma_dtype, q, nullptr);
      if (!e) {
        h->desc.mlp_packed = q;
        h->desc.mlp_packed_dtype = mfma_dtype;
      }
    }
    if (!e && hipDeviceSynchronize() != hipSuccess) {
      set_error("scene_create: packing failed");
      e = SNISR_ERR_DEVICE;
    }
  }
  if (e) {
    release(own);
    delete h;
    return e;
  }
  *out = h;
  return SNISR_OK;
}

int snisr_scene_destroy(snisr_scene_handle* scene) {
  if (!scene) return SNISR_OK;
  (void)hipDeviceSynchronize();
  release(scene->allocations);
  delete scene;
  return SNISR_OK;
}

const snisr_scene* snisr_scene_get_desc(const snisr_scene_handle* scene) { return scene ? &scene->desc : nullptr; }

int snisr_graph_create(const snisr_scene_handle* scene, const snisr_graph_desc* desc, const float* intrinsic,
                       const float* c2w, const float* origin, snisr_graph** out) {
  SNISR_REQUIRE(scene && desc && intrinsic && c2w && origin && out, SNISR_ERR_ARG, "graph_create: NULL pointer");
  *out = nullptr;
  SNISR_REQUIRE(desc->H > 0 && desc->W > 0 && (int64_t)desc->H * desc->W < (int64_t(1) << 31), SNISR_ERR_ARG,
                "graph_create: bad image size");
  SNISR_REQUIRE(desc->mlp_dtype == SNISR_F32 || desc->mlp_dtype == scene->desc.mlp_packed_dtype, SNISR_ERR_ARG,
                "graph_create: mlp_dtype must be SNISR_F32 or the scene's MFMA packing");
  SNISR_REQUIRE(desc->mlp_dtype == SNISR_F32 || scene->desc.mlp_packed, SNISR_ERR_ARG,
                "graph_create: the scene has no MFMA packing");
  if (int e = check_params(&desc->params)) return e;
  snisr_graph* g = new (std::nothrow) snisr_graph();
  SNISR_REQUIRE(g, SNISR_ERR_DEVICE, "graph_create: out of host memory");
  g->scene = scene;
  g->cfg = *desc;
  const int64_t n = (int64_t)desc->H * desc->W;
  int e = SNISR_OK;
  auto fail = [&](int code) {
    destroy_graph(g);
    return code;
  };
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("graph_create: hipStreamCreate failed");
    return fail(SNISR_ERR_DEVICE);
  }
  void* p[8] = {};
  const int64_t sizes[8] = {15 * 4, 12 * n, 4 * n, 4 * n, 12 * n, 4 * n, 4 * n, 4 * n};
  for (int k = 0; k < 8; ++k)
    if (hipMalloc(&p[k], (size_t)sizes[k]) != hipSuccess) {
      for (int j = 0; j < k; ++j) (void)hipFree(p[j]);
      set_error("graph_create: hipMalloc of %lld bytes failed", (long long)sizes[k]);
      return fail(SNISR_ERR_DEVICE);
    }
  g->camera = (float*)p[0];
  g->rays = (float*)p[1];
  g->t_hit = (float*)p[2];
  g->tri = (int32_t*)p[3];
  g->rgb = (float*)p[4];
  g->T = (float*)p[5];
  g->depth = (float*)p[6];
  g->n_samples = (int32_t*)p[7];
  if ((e = set_camera(g, intrinsic, c2w, origin))) return fail(e);
  if (desc->mlp_dtype != SNISR_F32) {
    // size the workspace from an eager render of this view: record slots for the largest ray (+25%)
    // and sorted-sample room for the frame (+ sample_headroom), unless the caller fixed them
    g->records_per_ray = desc->records_per_ray;
    g->sample_capacity = desc->sample_capacity;
    if (!g->records_per_ray || !g->sample_capacity) {
      const int64_t probe = snisr_render_mfma_workspace_size(&scene->desc, n, 0, 0);
      if (probe < 0 || hipMalloc(&g->workspace, (size_t)probe) != hipSuccess) {
        set_error("graph_create: workspace allocation failed");
        return fail(SNISR_ERR_DEVICE);
      }
      g->workspace_bytes = probe;
      g->records_per_ray = 64;
      if ((e = graph_step(g))) return fail(e);
      int64_t total = 0;
      int32_t over = 0, mx = 0;
      if ((e = snisr_render_mfma_stats(g->workspace, &total, &over, &mx, g->stream))) return fail(e);
      (void)hipFree(g->workspace);
      g->workspace = nullptr;
      if (!desc->records_per_ray) g->records_per_ray = std::max(8, ((int)(mx * 1.25f) + 8) / 8 * 8);
      if (!desc->sample_capacity) g->sample_capacity = (int64_t)(total * (1.0 + desc->sample_headroom)) + 64;
    }
    g->workspace_bytes = snisr_render_mfma_workspace_size(&scene->desc, n, g->records_per_ray, g->sample_capacity);
    if (g->workspace_bytes < 0 || hipMalloc(&g->workspace, (size_t)g->workspace_bytes) != hipSuccess) {
      set_error("graph_create: workspace allocation of %lld bytes failed", (long long)g->workspace_bytes);
      return fail(SNISR_ERR_DEVICE);
    }
  }
  // warm-up on the stream, then capture the step
  if ((e = graph_step(g))) return fail(e);
  if (hipStreamSynchronize(g->stream) != hipSuccess) {
    set_error("graph_create: warm-up render failed");
    return fail(SNISR_ERR_DEVICE);
  }
  if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    set_error("graph_create: hipStreamBeginCapture failed");
    return fail(SNISR_ERR_DEVICE);
  }
  e = graph_step(g);
  const hipError_t ec = hipStreamEndCapture(g->stream, &g->graph);
  if (e) return fail(e);
  if (ec != hipSuccess || hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
    set_error("graph_create: capture / instantiation failed");
    return fail(SNISR_ERR_DEVICE);
  }
  *out = g;
  return SNISR_OK;
}

int snisr_graph_render(snisr_graph* g, const float* intrinsic, const float* c2w, const float* origin, void* stream) {
  SNISR_REQUIRE(g, SNISR_ERR_ARG, "graph_render: NULL graph");
  if (intrinsic || c2w || origin) {
    SNISR_REQUIRE(intrinsic && c2w && origin, SNISR_ERR_ARG, "graph_render: give all of the camera or none");
    if (int e = set_camera(g, intrinsic, c2w, origin)) return e;
  }
  const hipStream_t st = stream ? (hipStream_t)stream : g->stream;
  if (st != g->stream) {   // order the replay after the caller's stream and the caller's work after it
    hipEvent_t ev;
    SNISR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    SNISR_HIP(hipEventRecord(ev, st));
    SNISR_HIP(hipStreamWaitEvent(g->stream, ev, 0));
    SNISR_HIP(hipGraphLaunch(g->exec, g->stream));
    SNISR_HIP(hipEventRecord(ev, g->stream));
    SNISR_HIP(hipStreamWaitEvent(st, ev, 0));
    SNISR_HIP(hipEventDestroy(ev));
    return SNISR_OK;
  }
  SNISR_HIP(hipGraphLaunch(g->exec, g->stream));
  return SNISR_OK;
}

int snisr_graph_outputs(const snisr_graph* g, float** rgb, float** T, int32_t** n_samples, float** depth,
                        float** t_hit) {
  SNISR_REQUIRE(g, SNISR_ERR_ARG, "graph_outputs: NULL graph");
  if (rgb) *rgb = g->rgb;
  if (T) *T = g->T;
  if (n_samples) *n_samples = g->n_samples;
  if (depth) *depth = g->depth;
  if (t_hit) *t_hit = g->cfg.bvh_nodes ? g->t_hit : nullptr;
  return SNISR_OK;
}

int snisr_graph_stats(snisr_graph* g, int64_t* total, int32_t* overflow, int32_t* records_per_ray,
                      int64_t* sample_capacity) {
  SNISR_REQUIRE(g && total && overflow, SNISR_ERR_ARG, "graph_stats: NULL pointer");
  if (records_per_ray) *records_per_ray = g->records_per_ray;
  if (sample_capacity) *sample_capacity = g->sample_capacity;
  if (g->cfg.mlp_dtype == SNISR_F32) {
    SNISR_HIP(hipStreamSynchronize(g->stream));
    *total = -1;
    *overflow = 0;
    return SNISR_OK;
  }
  return snisr_render_mfma_stats(g->workspace, total, overflow, nullptr, g->stream);
}

int snisr_graph_synchronize(snisr_graph* g) {
  SNISR_REQUIRE(g, SNISR_ERR_ARG, "graph_synchronize: NULL graph");
  SNISR_HIP(hipStreamSynchronize(g->stream));
  return SNISR_OK;
}

int snisr_graph_destroy(snisr_graph* g) {
  if (!g) return SNISR_OK;
  (void)hipStreamSynchronize(g->stream);
  destroy_graph(g);
  return SNISR_OK;
}

}  // extern "C"
