// Host-side renderer that plays CPURenderer's role (libraries/include/renderer/renderer/
// CPURenderer.h:13-72) on top of the C ABI in include/spk_pt.h: the scene, camera and render
// configuration keep Sparkle's meaning, and only the per-pixel sample loop runs on the MI355X.
//
//   InitRenderResources  CPURenderer::InitRenderResources (CPURenderer.cpp:38-105): sub-pixel
//                        count round(sqrt(spp))^2, G-buffer and frame buffer
//   Render               CPURenderer::Render (CPURenderer.cpp:119-208): clear on dirty, BasePass,
//                        DenoisePass (temporal), ToneMappingPass, sample bookkeeping
//   BasePass             CPURenderer::BasePass (CPURenderer.cpp:419-450) -> spk_render_frame
//   DenoisePass          CPURenderer::DenoisePass (CPURenderer.cpp:506-552): optional spatial fill of
//                        invalid pixels, then the temporal running mean
//   ToneMappingPass      CPURenderer::ToneMappingPass (CPURenderer.cpp:554-576)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spk_pt.h"

namespace sparkle_amd {

struct RenderConfig {
    uint32_t width = 256;
    uint32_t height = 256;
    uint32_t max_bounce = 4;        // packaged config.json value
    uint32_t sample_per_pixel = 1;  // per frame
    uint32_t max_sample_per_pixel = 16;
    uint32_t debug_mode = SPK_DEBUG_COLOR;
    bool sky = false;
    float sky_color[3] = {0.5f, 0.7f, 1.0f};
    float exposure = 1.f;
    uint32_t frame_flags = 0;  // SPK_FRAME_* (NEE, cache) on top of the reference estimator
    bool spatial_denoise = false;  // RenderConfig::spatial_denoise (packaged config: off)
};

struct SceneDesc {
    std::vector<spk_mesh_desc> meshes;
    std::vector<spk_instance_desc> instances;
    std::vector<spk_material_desc> materials;
};

class GpuPathRenderer {
public:
    explicit GpuPathRenderer(int device = 0);
    ~GpuPathRenderer();
    GpuPathRenderer(const GpuPathRenderer &) = delete;
    GpuPathRenderer &operator=(const GpuPathRenderer &) = delete;

    // Returns SPK_OK or an SPK_ERR_* code; LastError() has the message.
    int InitRenderResources(const SceneDesc &scene, const RenderConfig &config, const spk_camera_desc &camera);
    int SetCamera(const spk_camera_desc &camera);  // marks the accumulation dirty, as a camera move does
    int Render();
    int ReadbackFrameBuffer(std::vector<float> &rgba) const;  // accumulated, before tone mapping
    int ReadbackImage(std::vector<float> &rgba) const;        // tone mapped, vertically flipped
    uint32_t CumulatedSampleCount() const { return cumulated_sample_count_; }
    uint32_t DispatchedSampleCount() const { return dispatched_sample_count_; }
    double LastFrameMilliseconds() const { return last_frame_ms_; }
    uint64_t LastFrameRays() const { return last_frame_rays_; }
    std::string LastError() const;

private:
    int BasePass();
    int DenoisePass();
    int ToneMappingPass();
    void Release();

    spk_context *ctx_ = nullptr;
    int device_ = 0;
    RenderConfig config_;
    float *gbuffer_ = nullptr;       // one frame's samples, RGBA (valid flag in alpha)
    float *normals_ = nullptr;       // first-hit world normals (spatial fill only)
    float *ping_pong_ = nullptr;     // spatially filled frame (spatial fill only)
    float *frame_buffer_ = nullptr;  // accumulation across frames
    float *image_ = nullptr;         // tone-mapped output
    uint32_t actual_sample_per_pixel_ = 1;
    uint32_t dispatched_sample_count_ = 0;
    uint32_t cumulated_sample_count_ = 0;
    uint32_t pending_sample_count_ = 0;
    bool pixels_dirty_ = true;
    double last_frame_ms_ = 0.0;
    uint64_t last_frame_rays_ = 0;
};

}  // namespace sparkle_amd
