// Host-side launchers of the HIP kernels (each defined next to its kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spk {

struct TraceVariant {
    int width;  // 4 or 8
    int stack;  // LDS stack entries per lane
    int block;  // threads per block
    int refill; // 0: one ray per thread; k > 0: persistent lanes, a wave refills when k lanes are idle
};

// Closest hit of n rays (32-byte records) into n 8-byte hits. invalid_count (device) is
// incremented once per NaN/inf/zero-direction ray. cnt_nodes/cnt_tris (device, may be null)
// receive per-ray traversal counts when non-null. work (device, 8 bytes) is the ray counter of
// the persistent variants; it is cleared by a kernel launched just before them.
// spheres (device, may be null): 2 float4 per sphere, tested after the BVH with ids id_base + k.
// insts (device, may be null): instance records of a two-level scene (root = TLAS root); only
// one-ray-per-thread variants (refill 0) exist for two-level scenes.
hipError_t launch_trace(const TraceVariant &v, const void *rays, uint64_t n, void *hits, const void *nodes,
                        const void *tris, uint32_t root, uint32_t *invalid_count, uint32_t *cnt_nodes,
                        uint32_t *cnt_tris, unsigned long long *work, const void *spheres, uint32_t num_spheres,
                        uint32_t sphere_id_base, const void *insts, hipStream_t stream);

// Largest stack depth any compiled variant supports.
int max_trace_stack();
// Smallest compiled stack size >= depth (two-level variants: 24, 32, 64), or -1.
int pick_trace_stack(int depth, bool two = false);
// Compiled (stack, block, refill) configurations of one width, for the autotuner.
int trace_variants(TraceVariant *out, int max_out, int width, bool two = false);

}  // namespace spk
