// Host-side renderer over the C ABI (see gpu_renderer.h).
#include "gpu_renderer.h"

#include <cmath>
#include <cstring>

namespace sparkle_amd {

GpuPathRenderer::GpuPathRenderer(int device) : device_(device) {}

GpuPathRenderer::~GpuPathRenderer() {
    Release();
    if (ctx_) spk_destroy(ctx_);
}

void GpuPathRenderer::Release() {
    if (!ctx_) return;
    spk_device_free(ctx_, gbuffer_);
    spk_device_free(ctx_, frame_buffer_);
    spk_device_free(ctx_, image_);
    spk_device_free(ctx_, normals_);
    spk_device_free(ctx_, ping_pong_);
    gbuffer_ = frame_buffer_ = image_ = normals_ = ping_pong_ = nullptr;
}

std::string GpuPathRenderer::LastError() const { return ctx_ ? spk_last_error(ctx_) : "no context"; }

int GpuPathRenderer::InitRenderResources(const SceneDesc &scene, const RenderConfig &config,
                                         const spk_camera_desc &camera) {
    int rc;
    if (!ctx_ && (rc = spk_create(device_, &ctx_)) != SPK_OK) return rc;
    Release();
    config_ = config;
    // CPURenderer.cpp:102-104: the frame is weighted as round(sqrt(spp))^2 samples
    const unsigned sub = (unsigned)std::lround(std::sqrt((float)config.sample_per_pixel));
    actual_sample_per_pixel_ = std::max(1u, sub * sub);
    if ((rc = spk_scene_upload(ctx_, scene.meshes.data(), (uint32_t)scene.meshes.size(), scene.instances.data(),
                               (uint32_t)scene.instances.size(), scene.materials.data(),
                               (uint32_t)scene.materials.size(), nullptr)) != SPK_OK)
        return rc;
    const size_t px = (size_t)config.width * config.height;
    if ((rc = spk_device_alloc(ctx_, px * 16, (void **)&gbuffer_)) != SPK_OK ||
        (rc = spk_device_alloc(ctx_, px * 16, (void **)&frame_buffer_)) != SPK_OK ||
        (rc = spk_device_alloc(ctx_, px * 16, (void **)&image_)) != SPK_OK ||
        (config.spatial_denoise && ((rc = spk_device_alloc(ctx_, px * 12, (void **)&normals_)) != SPK_OK ||
                                    (rc = spk_device_alloc(ctx_, px * 16, (void **)&ping_pong_)) != SPK_OK)))
        return rc;
    dispatched_sample_count_ = 0;
    cumulated_sample_count_ = 0;
    pending_sample_count_ = 0;
    return SetCamera(camera);
}

int GpuPathRenderer::SetCamera(const spk_camera_desc &camera) {
    if (!ctx_) return SPK_ERR_INVALID_ARGUMENT;
    pixels_dirty_ = true;
    return spk_set_camera(ctx_, &camera);
}

int GpuPathRenderer::Render() {
    if (!ctx_ || !gbuffer_) return SPK_ERR_INVALID_ARGUMENT;
    // CameraRenderProxy::Update (CameraRenderProxy.cpp:66-75): reset on dirty, then fold in the
    // samples of the previous frame, capped at max_spp
    if (pixels_dirty_) {
        cumulated_sample_count_ = 0;
        pending_sample_count_ = 0;
    }
    cumulated_sample_count_ = std::min(config_.max_sample_per_pixel, cumulated_sample_count_ + pending_sample_count_);
    pending_sample_count_ = 0;
    // CPURenderer::Render: clear the frame buffer when the camera asks for it
    if (pixels_dirty_) {
        std::vector<float> zeros((size_t)config_.width * config_.height * 4, 0.f);
        int rc = spk_copy_to_device(ctx_, frame_buffer_, zeros.data(), zeros.size() * 4);
        if (rc != SPK_OK) return rc;
        dispatched_sample_count_ = 0;
        pixels_dirty_ = false;
    }
    int rc;
    if ((rc = BasePass()) != SPK_OK || (rc = DenoisePass()) != SPK_OK || (rc = ToneMappingPass()) != SPK_OK) return rc;
    dispatched_sample_count_ += actual_sample_per_pixel_;
    pending_sample_count_ += actual_sample_per_pixel_;
    return SPK_OK;
}

int GpuPathRenderer::BasePass() {
    spk_frame_desc f;
    std::memset(&f, 0, sizeof(f));
    f.width = config_.width;
    f.height = config_.height;
    f.frame_seed = dispatched_sample_count_;  // CPURenderer.cpp:432
    f.max_bounce = config_.max_bounce;
    f.debug_mode = config_.debug_mode;
    f.sky_enabled = config_.sky ? 1u : 0u;
    for (int c = 0; c < 3; ++c) f.sky_color[c] = config_.sky_color[c];
    f.flags = config_.frame_flags;
    spk_frame_stats st;
    int rc = spk_render_frame(ctx_, &f, gbuffer_, normals_, &st);
    if (rc == SPK_OK) {
        last_frame_ms_ = st.milliseconds;
        last_frame_rays_ = st.rays_traced + st.shadow_rays;
    }
    return rc;
}

int GpuPathRenderer::DenoisePass() {
    const float *pass_input = gbuffer_;
    if (config_.spatial_denoise) {
        int rc = spk_spatial_fill(ctx_, gbuffer_, normals_, ping_pong_, config_.width, config_.height);
        if (rc != SPK_OK) return rc;
        pass_input = ping_pong_;
    }
    return spk_accumulate(ctx_, pass_input, frame_buffer_, (uint64_t)config_.width * config_.height,
                          cumulated_sample_count_, actual_sample_per_pixel_);
}

int GpuPathRenderer::ToneMappingPass() {
    return spk_tonemap(ctx_, frame_buffer_, image_, config_.width, config_.height, config_.exposure);
}

int GpuPathRenderer::ReadbackFrameBuffer(std::vector<float> &rgba) const {
    rgba.resize((size_t)config_.width * config_.height * 4);
    return spk_copy_to_host(ctx_, rgba.data(), frame_buffer_, rgba.size() * 4);
}

int GpuPathRenderer::ReadbackImage(std::vector<float> &rgba) const {
    rgba.resize((size_t)config_.width * config_.height * 4);
    return spk_copy_to_host(ctx_, rgba.data(), image_, rgba.size() * 4);
}

}  // namespace sparkle_amd
