// MFMA render path (in progress).
#include "snisr_common.hpp"
namespace snisr {
int64_t mfma_pack_bytes(const snisr_scene*, int) { set_error("MFMA packing not built"); return -1; }
int mfma_pack(const snisr_scene*, int, void*, void*) { set_error("MFMA packing not built"); return SNISR_ERR_UNSUPPORTED; }
}
extern "C" {
int64_t snisr_render_mfma_workspace_size(const snisr_scene*, int64_t, int64_t) { snisr::set_error("MFMA path not built"); return -1; }
int snisr_render_mfma(const snisr_scene*, const snisr_march_params*, const snisr_rays*, const snisr_render_out*, void*, int64_t, void*) {
  snisr::set_error("MFMA path not built"); return SNISR_ERR_UNSUPPORTED; }
}
