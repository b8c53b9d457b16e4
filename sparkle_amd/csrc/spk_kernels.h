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
    int sort = 0;  // 1: rays traced in coherence order (counting sort by direction bin and origin cell)
};

// Device scratch of the coherence order (a two-pass counting sort): the key of each ray, the keys
// and input indices in direction-bin order (pass 1), the final permutation (pass 2), per chunk the
// offset of its run in each bin (pass 1: rows of kRayDirBins per kRaySortChunk1 rays; pass 2: rows
// of kRaySortBins per kRaySortChunk rays), the bin totals then bases (kRayDirBins for pass 1, then
// kRaySortBins); lo / inv: the scene box (origin cells are eighths of it).
constexpr uint32_t kRayDirBins = 96u;    // cube face of the major axis x 4 x 4
constexpr uint32_t kRayCellBins = 512u;  // 8^3 origin cells
constexpr uint32_t kRaySortBins = kRayDirBins * kRayCellBins;
struct RaySortScratch {
    uint32_t *keys, *keys1, *perm1, *perm, *offs1, *offs, *total;
    float lo[3], inv[3];
};
constexpr uint32_t kRaySortChunk = 32768u;  // rays per block of pass 2 (16-bit LDS counters)
// rays per block of pass 1 (96 bins: more blocks per CU than pass 2; 4096 was slower, its count pass
// doubling the device atomics on the 96 bin totals: 91 against 64 us per 8M rays)
constexpr uint32_t kRaySortChunk1 = 8192u;

// Closest hit of n rays (32-byte records) into n 8-byte hits. invalid_count (device) is
// incremented once per NaN/inf/zero-direction ray. cnt_nodes/cnt_tris (device, may be null)
// receive per-ray traversal counts when non-null. work (device, 8 bytes) is the ray counter of
// the persistent variants; it is cleared by a kernel launched just before them.
// spheres (device, may be null): 2 float4 per sphere, tested after the BVH with ids id_base + k.
// insts (device, may be null): instance records of a two-level scene (root = TLAS root); only
// one-ray-per-thread variants (refill 0) exist for two-level scenes.
// sort (may be null for variants without ray_sort): scratch of the coherence order.
// timed (may be null): two events recorded just before and after the traversal kernel.
hipError_t launch_trace(const TraceVariant &v, const void *rays, uint64_t n, void *hits, const void *nodes,
                        const void *tris, uint32_t root, uint32_t *invalid_count, uint32_t *cnt_nodes,
                        uint32_t *cnt_tris, unsigned long long *work, const void *spheres, uint32_t num_spheres,
                        uint32_t sphere_id_base, const void *insts, const RaySortScratch *sort,
                        hipStream_t stream, const hipEvent_t *timed = nullptr);

// Largest stack depth any compiled variant supports.
int max_trace_stack();
// Smallest compiled stack size >= depth (two-level variants: 24, 32, 64), or -1.
int pick_trace_stack(int depth, bool two = false);
// Compiled (stack, block, refill) configurations of one width, for the autotuner.
int trace_variants(TraceVariant *out, int max_out, int width, bool two = false);

}  // namespace spk
