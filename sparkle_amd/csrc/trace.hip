// Closest-hit kernels over the quantised wide BVH (one ray per lane, LDS traversal stack).
#include <algorithm>

#include "spk_kernels.h"
#include "traverse.h"

namespace spk {

// Analytic spheres of the scene, tested after the BVH (few, debug primitives in the reference).
struct SphereSet {
    const float4 *spheres;
    uint32_t count, id_base;
};

// TWO: a two-level scene (TLAS over the instance records `insts`, one BLAS per mesh).
template <int WIDTH, int STACK, int BLOCK, bool COUNT, bool TWO>
__global__ __launch_bounds__(BLOCK) void k_trace(const float4 *__restrict__ rays, uint64_t n,
                                                 uint2 *__restrict__ hits, const uint4 *__restrict__ nodes,
                                                 const uint4 *__restrict__ tris, uint32_t root,
                                                 uint32_t *__restrict__ invalid_count,
                                                 uint32_t *__restrict__ cnt_nodes, uint32_t *__restrict__ cnt_tris,
                                                 SphereSet sph, const uint4 *__restrict__ insts) {
    __shared__ uint2 stack[STACK * BLOCK];
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    RayIn r{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint2 out = make_uint2(0xFFFFFFFFu, 0x7f800000u);
    TraverseCounters c{0u, 0u};
    if (!ray_valid(r)) {
        atomicAdd(invalid_count, 1u);
    } else {
        BestHit best;
        best.t = r.tmax;
        best.prim = 0xFFFFFFFFu;
        best.tri = 0u;
        best.U = best.V = best.W = 0.f;
        best.det = 1.f;
        best.inst = 0u;
        traverse<WIDTH, STACK, COUNT, false, TWO>(nodes, tris, root, r, stack + threadIdx.x, BLOCK, best, c, insts);
        if (sph.count) spheres_closest(sph.spheres, sph.count, sph.id_base, r, best);
        if (best.prim != 0xFFFFFFFFu)
            out = make_uint2(best.prim, __float_as_uint(final_t(tris, r, best, TWO ? insts : nullptr)));
    }
    hits[i] = out;
    if (COUNT) {
        if (cnt_nodes) cnt_nodes[i] = c.nodes;
        if (cnt_tris) cnt_tris[i] = c.tris;
    }
}

// Persistent variant: a grid sized to the resident capacity of the device; each lane traces one
// ray at a time and, whenever at least `refill_min` lanes of its wave are idle, the wave takes the
// next rays from a global counter (one atomic per refill) so that lanes stay busy instead of
// waiting for the wave's slowest ray. Results are identical to k_trace.
__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int src) {
    const uint32_t lo = __shfl((uint32_t)v, src), hi = __shfl((uint32_t)(v >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}

template <int WIDTH, int STACK, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace_persistent(
    const float4 *__restrict__ rays, uint64_t n, uint2 *__restrict__ hits, const uint4 *__restrict__ nodes,
    const uint4 *__restrict__ tris, uint32_t root, uint32_t *__restrict__ invalid_count,
    uint32_t *__restrict__ cnt_nodes, uint32_t *__restrict__ cnt_tris, unsigned long long *__restrict__ next_ray,
    uint32_t refill_min, SphereSet sph) {
    __shared__ uint2 stack[STACK * BLOCK];
    uint2 *st = stack + threadIdx.x;
    const uint32_t lane = __lane_id();
    const unsigned long long below = lane ? (~0ull >> (64u - lane)) : 0ull;
    Traversal<WIDTH, STACK, COUNT> tr;
    uint64_t id = 0;
    bool active = false;
    bool drained = false;  // wave-uniform: the counter has passed n
    for (;;) {
        if (!drained) {
            const unsigned long long idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= refill_min) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(next_ray, (unsigned long long)nidle);
                base = shfl_u64(base, 0);
                drained = base + nidle >= n;
                if (!active) {
                    const uint64_t my = base + (uint64_t)__popcll(idle & below);
                    if (my < n) {
                        const float4 a = rays[2 * my], b = rays[2 * my + 1];
                        const RayIn r{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                        if (ray_valid(r)) {
                            tr.init(r, root);
                            id = my;
                            active = true;
                        } else {
                            atomicAdd(invalid_count, 1u);
                            hits[my] = make_uint2(0xFFFFFFFFu, 0x7f800000u);
                            if (COUNT) {
                                if (cnt_nodes) cnt_nodes[my] = 0u;
                                if (cnt_tris) cnt_tris[my] = 0u;
                            }
                        }
                    }
                }
            }
        }
        if (!__any(active)) {
            if (drained) break;
            continue;
        }
        if (active && !tr.step(nodes, tris, st, BLOCK)) {
            uint2 out = make_uint2(0xFFFFFFFFu, 0x7f800000u);
            const float4 a = rays[2 * id], b = rays[2 * id + 1];
            const RayIn r{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            if (sph.count) spheres_closest(sph.spheres, sph.count, sph.id_base, r, tr.best);
            if (tr.best.prim != 0xFFFFFFFFu) out = make_uint2(tr.best.prim, __float_as_uint(final_t(tris, r, tr.best)));
            hits[id] = out;
            if (COUNT) {
                if (cnt_nodes) cnt_nodes[id] = tr.cnt.nodes;
                if (cnt_tris) cnt_tris[id] = tr.cnt.tris;
            }
            active = false;
        }
    }
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

namespace {
using TraceFn = void (*)(const float4 *, uint64_t, uint2 *, const uint4 *, const uint4 *, uint32_t, uint32_t *,
                         uint32_t *, uint32_t *, SphereSet, const uint4 *);
using PersistentFn = void (*)(const float4 *, uint64_t, uint2 *, const uint4 *, const uint4 *, uint32_t, uint32_t *,
                              uint32_t *, uint32_t *, unsigned long long *, uint32_t, SphereSet);
struct Entry {
    int width, stack, block;
    bool count;
    TraceFn fn;
    PersistentFn pfn;  // null: no persistent variant (two-level scenes)
    bool two;
};

#define SPK_TV(W, S, B)                                                                                   \
    {W, S, B, false, k_trace<W, S, B, false, false>, k_trace_persistent<W, S, B, false>, false},          \
    {                                                                                                     \
        W, S, B, true, k_trace<W, S, B, true, false>, k_trace_persistent<W, S, B, true>, false            \
    }
#define SPK_TV2(W, S, B)                                                                                  \
    {W, S, B, false, k_trace<W, S, B, false, true>, nullptr, true},                                       \
    {                                                                                                     \
        W, S, B, true, k_trace<W, S, B, true, true>, nullptr, true                                        \
    }

// LDS per block = stack * block * 8 bytes, kept at or below 64 KiB
const Entry kEntries[] = {
    SPK_TV(8, 12, 64),  SPK_TV(8, 12, 128), SPK_TV(4, 12, 64),  SPK_TV(4, 12, 128), SPK_TV(8, 16, 64),  SPK_TV(8, 16, 128), SPK_TV(8, 16, 256), SPK_TV(8, 24, 64),  SPK_TV(8, 24, 128),
    SPK_TV(8, 24, 256), SPK_TV(8, 32, 64),  SPK_TV(8, 32, 128), SPK_TV(8, 32, 256), SPK_TV(8, 64, 64),
    SPK_TV(8, 64, 128), SPK_TV(4, 16, 64),  SPK_TV(4, 16, 128), SPK_TV(4, 16, 256), SPK_TV(4, 24, 64),
    SPK_TV(4, 24, 128), SPK_TV(4, 24, 256), SPK_TV(4, 32, 64),  SPK_TV(4, 32, 128), SPK_TV(4, 32, 256),
    SPK_TV(4, 64, 64),  SPK_TV(4, 64, 128),
    // two-level scenes: the stack also holds a TLAS level's group and pending instances
    SPK_TV2(8, 24, 64), SPK_TV2(8, 24, 128), SPK_TV2(8, 32, 64), SPK_TV2(8, 32, 128), SPK_TV2(8, 64, 64),
    SPK_TV2(4, 24, 64), SPK_TV2(4, 24, 128), SPK_TV2(4, 32, 64), SPK_TV2(4, 32, 128), SPK_TV2(4, 64, 64),
};
#undef SPK_TV
#undef SPK_TV2
constexpr int kNumEntries = (int)(sizeof(kEntries) / sizeof(kEntries[0]));
int g_resident_blocks[kNumEntries];  // resident blocks per CU of each persistent kernel, filled on first use

// refill thresholds tried by the autotuner for the persistent kernels
const int kRefill[] = {1, 16, 32};
}  // namespace

hipError_t launch_trace(const TraceVariant &v, const void *rays, uint64_t n, void *hits, const void *nodes,
                        const void *tris, uint32_t root, uint32_t *invalid_count, uint32_t *cnt_nodes,
                        uint32_t *cnt_tris, unsigned long long *work, const void *spheres, uint32_t num_spheres,
                        uint32_t sphere_id_base, const void *insts, hipStream_t stream) {
    const SphereSet sph{(const float4 *)spheres, spheres ? num_spheres : 0u, sphere_id_base};
    const bool count = cnt_nodes != nullptr || cnt_tris != nullptr;
    const bool two = insts != nullptr;
    for (int k = 0; k < kNumEntries; ++k) {
        const Entry &e = kEntries[k];
        if (e.width != v.width || e.stack != v.stack || e.block != v.block || e.count != count || e.two != two)
            continue;
        if (n == 0) return hipSuccess;
        const uint64_t blocks = (n + (uint64_t)v.block - 1) / (uint64_t)v.block;
        if (v.refill > 0) {
            if (!work || !e.pfn) return hipErrorInvalidValue;
            if (g_resident_blocks[k] == 0) {
                int nb = 0;
                hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, e.pfn, v.block, 0);
                if (oe != hipSuccess) return oe;
                g_resident_blocks[k] = nb > 0 ? nb : 1;
            }
            int dev = 0, cus = 0;
            hipError_t de = hipGetDevice(&dev);
            if (de == hipSuccess) de = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (de != hipSuccess) return de;
            const uint64_t grid = std::min<uint64_t>(blocks, (uint64_t)g_resident_blocks[k] * (uint64_t)cus);
            const uint32_t refill = (uint32_t)std::min(std::max(v.refill, 1), 64);
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, work);
            hipLaunchKernelGGL(e.pfn, dim3((uint32_t)grid), dim3(v.block), 0, stream, (const float4 *)rays, n,
                               (uint2 *)hits, (const uint4 *)nodes, (const uint4 *)tris, root, invalid_count,
                               cnt_nodes, cnt_tris, work, refill, sph);
            return hipGetLastError();
        }
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(e.fn, dim3((uint32_t)blocks), dim3(v.block), 0, stream, (const float4 *)rays, n,
                           (uint2 *)hits, (const uint4 *)nodes, (const uint4 *)tris, root, invalid_count, cnt_nodes,
                           cnt_tris, sph, (const uint4 *)insts);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

int max_trace_stack() { return 64; }

int pick_trace_stack(int depth, bool two) {
    const int sizes[] = {12, 16, 24, 32, 64};
    for (int s : sizes)
        if (s >= depth && (!two || s >= 24)) return s;
    return -1;
}

int trace_variants(TraceVariant *out, int max_out, int width, bool two) {
    int k = 0;
    for (const Entry &e : kEntries) {
        if (e.count || e.width != width || e.two != two) continue;
        if (k < max_out) out[k] = TraceVariant{e.width, e.stack, e.block, 0};
        ++k;
        if (!e.pfn) continue;
        for (int r : kRefill) {
            if (k < max_out) out[k] = TraceVariant{e.width, e.stack, e.block, r};
            ++k;
        }
    }
    return k;
}

}  // namespace spk
