// Closest-hit kernels over the quantised wide BVH (one ray per lane, LDS traversal stack).
#include "spk_kernels.h"
#include "traverse.h"

namespace spk {

template <int WIDTH, int STACK, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace(const float4 *__restrict__ rays, uint64_t n,
                                                 uint2 *__restrict__ hits, const uint4 *__restrict__ nodes,
                                                 const uint4 *__restrict__ tris, uint32_t root,
                                                 uint32_t *__restrict__ invalid_count,
                                                 uint32_t *__restrict__ cnt_nodes, uint32_t *__restrict__ cnt_tris) {
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
        traverse<WIDTH, STACK, COUNT>(nodes, tris, root, r, stack + threadIdx.x, BLOCK, best, c);
        if (best.prim != 0xFFFFFFFFu) out = make_uint2(best.prim, __float_as_uint(final_t(tris, r, best)));
    }
    hits[i] = out;
    if (COUNT) {
        if (cnt_nodes) cnt_nodes[i] = c.nodes;
        if (cnt_tris) cnt_tris[i] = c.tris;
    }
}

namespace {
using TraceFn = void (*)(const float4 *, uint64_t, uint2 *, const uint4 *, const uint4 *, uint32_t, uint32_t *,
                         uint32_t *, uint32_t *);
struct Entry {
    int width, stack, block;
    bool count;
    TraceFn fn;
};

#define SPK_TV(W, S, B)                                        \
    {W, S, B, false, k_trace<W, S, B, false>}, {W, S, B, true, k_trace<W, S, B, true>}

// LDS per block = stack * block * 8 bytes, kept at or below 64 KiB
const Entry kEntries[] = {
    SPK_TV(8, 16, 64),  SPK_TV(8, 16, 128), SPK_TV(8, 16, 256), SPK_TV(8, 24, 64),  SPK_TV(8, 24, 128),
    SPK_TV(8, 24, 256), SPK_TV(8, 32, 64),  SPK_TV(8, 32, 128), SPK_TV(8, 32, 256), SPK_TV(8, 64, 64),
    SPK_TV(8, 64, 128), SPK_TV(4, 16, 64),  SPK_TV(4, 16, 128), SPK_TV(4, 16, 256), SPK_TV(4, 24, 64),
    SPK_TV(4, 24, 128), SPK_TV(4, 24, 256), SPK_TV(4, 32, 64),  SPK_TV(4, 32, 128), SPK_TV(4, 32, 256),
    SPK_TV(4, 64, 64),  SPK_TV(4, 64, 128),
};
#undef SPK_TV
}  // namespace

hipError_t launch_trace(const TraceVariant &v, const void *rays, uint64_t n, void *hits, const void *nodes,
                        const void *tris, uint32_t root, uint32_t *invalid_count, uint32_t *cnt_nodes,
                        uint32_t *cnt_tris, hipStream_t stream) {
    const bool count = cnt_nodes != nullptr || cnt_tris != nullptr;
    for (const Entry &e : kEntries) {
        if (e.width != v.width || e.stack != v.stack || e.block != v.block || e.count != count) continue;
        if (n == 0) return hipSuccess;
        const uint64_t blocks = (n + (uint64_t)v.block - 1) / (uint64_t)v.block;
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(e.fn, dim3((uint32_t)blocks), dim3(v.block), 0, stream, (const float4 *)rays, n,
                           (uint2 *)hits, (const uint4 *)nodes, (const uint4 *)tris, root, invalid_count, cnt_nodes,
                           cnt_tris);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

int max_trace_stack() { return 64; }

int pick_trace_stack(int depth) {
    const int sizes[] = {16, 24, 32, 64};
    for (int s : sizes)
        if (s >= depth) return s;
    return -1;
}

int trace_variants(TraceVariant *out, int max_out, int width) {
    int k = 0;
    for (const Entry &e : kEntries) {
        if (e.count || e.width != width) continue;
        if (k < max_out) out[k] = TraceVariant{e.width, e.stack, e.block};
        ++k;
    }
    return k;
}

}  // namespace spk
