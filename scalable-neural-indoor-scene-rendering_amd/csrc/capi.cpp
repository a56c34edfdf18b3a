// C-ABI plumbing: thread-local error message, descriptor validation, version.
#include <cstring>

#include "snisr_common.hpp"

namespace snisr {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int check_scene(const snisr_scene* s, bool need_feats) {
  SNISR_REQUIRE(s, SNISR_ERR_ARG, "scene descriptor is NULL");
  SNISR_REQUIRE(s->tile_shape[0] > 0 && s->tile_shape[1] > 0 && s->tile_shape[2] > 0, SNISR_ERR_ARG,
                "tile_shape must be positive");
  SNISR_REQUIRE(s->tile_size > 0.0f, SNISR_ERR_ARG, "tile_size must be positive");
  SNISR_REQUIRE(s->n_tiles > 0, SNISR_ERR_ARG, "n_tiles must be positive");
  SNISR_REQUIRE(s->tile_index_map && s->tile_centers && s->tile_res && s->tile_brick_base && s->brick_map,
                SNISR_ERR_ARG, "scene tile tables must be non-NULL");
  SNISR_REQUIRE(s->n_bricks >= 0, SNISR_ERR_ARG, "n_bricks must be >= 0");
  SNISR_REQUIRE(s->n_bricks == 0 || s->sigma, SNISR_ERR_ARG, "sigma is NULL");
  SNISR_REQUIRE(aligned(s->sigma, 16), SNISR_ERR_ALIGN, "sigma must be 16-byte aligned");
  if (need_feats) {
    SNISR_REQUIRE(s->feat_dtype == SNISR_F32 || s->feat_dtype == SNISR_F16, SNISR_ERR_ARG,
                  "feat_dtype must be SNISR_F32 or SNISR_F16");
    SNISR_REQUIRE(s->n_bricks == 0 || s->feats, SNISR_ERR_ARG, "feats is NULL");
    SNISR_REQUIRE(aligned(s->feats, 16), SNISR_ERR_ALIGN, "feats must be 16-byte aligned");
    SNISR_REQUIRE(s->mlp_in == s->feat_dim + 15, SNISR_ERR_ARG, "mlp_in must be feat_dim + 15 (got %d, F=%d)",
                  s->mlp_in, s->feat_dim);
    SNISR_REQUIRE(s->mlp_layers >= 2 && s->mlp_width > 0, SNISR_ERR_ARG, "mlp needs >= 2 layers");
    SNISR_REQUIRE(s->mlp_params, SNISR_ERR_ARG, "mlp_params is NULL");
  }
  return SNISR_OK;
}

int check_params(const snisr_march_params* p) {
  SNISR_REQUIRE(p, SNISR_ERR_ARG, "march params are NULL");
  SNISR_REQUIRE(p->max_tiles >= 0 && p->max_tiles <= 4096, SNISR_ERR_ARG, "max_tiles out of range (%d)",
                p->max_tiles);
  SNISR_REQUIRE(p->sample_step > 0.0f, SNISR_ERR_ARG, "sample_step must be positive");
  return SNISR_OK;
}

int check_rays(const snisr_rays* r, const snisr_march_params* p) {
  SNISR_REQUIRE(r, SNISR_ERR_ARG, "rays descriptor is NULL");
  SNISR_REQUIRE(r->n >= 0, SNISR_ERR_ARG, "ray count must be >= 0");
  if (r->n == 0) return SNISR_OK;
  SNISR_REQUIRE(r->origins && r->dirs, SNISR_ERR_ARG, "origins and dirs must be non-NULL");
  SNISR_REQUIRE(r->origin_stride == 0 || r->origin_stride == 3, SNISR_ERR_ARG, "origin_stride must be 0 or 3");
  SNISR_REQUIRE(r->n < (int64_t(1) << 31), SNISR_ERR_SHAPE, "at most 2^31-1 rays per call");
  (void)p;
  return SNISR_OK;
}

int64_t mlp_f32t_stride(int in, int width, int layers) {
  int64_t s = 0;
  int i = in;
  for (int l = 0; l < layers; ++l) {
    const int o = (l == layers - 1) ? 3 : width;
    const int o4 = (o + 3) & ~3;
    s += (int64_t)i * o4 + o4;
    i = o;
  }
  return s;
}

}  // namespace snisr

extern "C" {

const char* snisr_last_error(void) { return snisr::g_last_error.c_str(); }
int snisr_version(void) { return 1; }

int64_t snisr_abi_sizeof(int32_t which) {
  switch (which) {
    case 0: return (int64_t)sizeof(snisr_scene);
    case 1: return (int64_t)sizeof(snisr_march_params);
    case 2: return (int64_t)sizeof(snisr_rays);
    case 3: return (int64_t)sizeof(snisr_render_out);
    case 4: return (int64_t)sizeof(snisr_mlp_query);
    case 5: return (int64_t)sizeof(snisr_octree);
    case 6: return (int64_t)sizeof(snisr_tuning);
    case 7: return (int64_t)sizeof(snisr_dense_grid);
    case 8: return (int64_t)sizeof(snisr_tile_mesh);
    case 9: return (int64_t)sizeof(snisr_graph_desc);
    default: return -1;
  }
}

}  // extern "C"
