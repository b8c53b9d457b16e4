// Host-side state of the learned radiance cache (kernels in cache.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/spk_pt.h"
#include "spk_context.h"

namespace spk {

// Optimiser state kept on the device, so that a training step can follow a frame without a host
// round trip: the frame's sample count stays in device memory and decides there whether the step runs.
struct AdamStep {
    uint32_t step;    // Adam steps taken
    uint32_t active;  // 1 when the current launch's step runs
    float c1, c2;     // bias corrections 1 - beta^step of the current step
    float loss;       // loss of the last gradient pass
};
constexpr uint32_t kMinTrainSamples = 64;  // a frame with fewer training samples takes no step

struct CacheState {
    spk_cache_config cfg{};
    float cfg_beta1 = 0.9f, cfg_beta2 = 0.99f, cfg_eps = 1e-8f;
    uint32_t res[16] = {0};
    uint32_t dense[16] = {0};
    float lo[3] = {0, 0, 0}, inv_ext[3] = {1, 1, 1};
    DeviceBuffer table_h, table_f, mt, vt, gtable;  // hash table (FP16 copy, FP32 master, Adam, BF16 gradient)
    DeviceBuffer w, mw, vw, gw, wf, wb;             // MLP weights (FP32 master, Adam, gradient, BF16 copies)
    DeviceBuffer act_t, grad_t, state;              // training scratch, AdamStep
    // per-frame records of the path tracer
    DeviceBuffer cq_points, cq_attrs, cq_meta, cq_out;  // cache queries: point, attributes, (throughput, path),
                                                        // radiance / albedo
    DeviceBuffer tr_points, tr_attrs, tr_meta0, tr_meta1, tr_targets;  // training samples
    // the last frame's training samples, handed to the training step that runs inside the next
    // frame's graph beside its first kernels (points, attributes, targets, count)
    DeviceBuffer st_points, st_attrs, st_targets, st_count;
    uint32_t capacity = 0;
    // the table gradient is all zero (k_adam_table clears what it consumes); false between a gradient
    // pass and the step that consumes it
    bool gtable_clean = false;
    uint64_t table_params() const { return (uint64_t)cfg.levels * (1ull << cfg.log2_table) * cfg.features; }
};

size_t cache_num_weights();
hipError_t cache_pack(CacheState &cs, hipStream_t s);
// attrs (device, may be null): SPK_CACHE_ATTRS floats per query (normal, albedo, roughness, 1); null: zeros
hipError_t cache_infer(CacheState &cs, const float *points, const float *attrs, uint32_t n, int world, float *out,
                       hipStream_t s);
// n_dev (device, may be null): the sample count, capped at n; with it, fewer than kMinTrainSamples
// samples make the gradient pass and the following cache_adam (given the same n_dev) no-ops.
hipError_t cache_gradients(CacheState &cs, const float *points, const float *attrs, const float *targets, uint32_t n,
                           int world, float *out, hipStream_t s, const uint32_t *n_dev = nullptr);
hipError_t cache_adam(CacheState &cs, hipStream_t s, const uint32_t *n_dev = nullptr, uint32_t cap = 0);
// Sizes the training scratch for n samples (call before capturing a graph that trains).
hipError_t cache_prepare(CacheState &cs, uint32_t n);
inline AdamStep *cache_state(CacheState &cs) { return static_cast<AdamStep *>(cs.state.ptr); }
// Inference over a queue whose length lives in device memory (capturable in a graph): `grid`
// persistent single-wave blocks walk the 64-query tiles.
hipError_t cache_infer_queue(CacheState &cs, const float *points, const float *attrs, const uint32_t *count_dev,
                             uint32_t capacity, float *out, int grid, hipStream_t s);

}  // namespace spk
