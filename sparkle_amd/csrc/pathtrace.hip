// Wavefront path tracer (BasePass on the device). Filled in below the closest-hit path.
#include <hip/hip_runtime.h>

#include "../../include/spk_pt.h"
#include "spk_context.h"

namespace spk {
struct PathState {
    int unused = 0;
};
}  // namespace spk

void spk_context::on_scene_changed() {}

void spk_context::release_all() {
    delete paths;
    paths = nullptr;
    scene.nodes.release();
    scene.tris.release();
    scene.shade.release();
    scene.materials.release();
    scene.emitters.release();
    scene.valid = false;
}

extern "C" {
int spk_set_camera(spk_context *ctx, const spk_camera_desc *camera) {
    if (!ctx || !camera) return SPK_ERR_INVALID_ARGUMENT;
    ctx->camera = *camera;
    ctx->has_camera = true;
    return SPK_OK;
}
int spk_render_frame(spk_context *ctx, const spk_frame_desc *, float *, float *, spk_frame_stats *) {
    return spk::set_error(ctx, SPK_ERR_UNSUPPORTED, "not built yet");
}
int spk_accumulate(spk_context *ctx, const float *, float *, uint64_t, uint32_t, uint32_t) {
    return spk::set_error(ctx, SPK_ERR_UNSUPPORTED, "not built yet");
}
int spk_tonemap(spk_context *ctx, const float *, float *, uint32_t, uint32_t, float) {
    return spk::set_error(ctx, SPK_ERR_UNSUPPORTED, "not built yet");
}
}
