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
// perm (may be null): thread k traces ray perm[k] (a coherence order) and writes its slot.
template <int WIDTH, int STACK, int BLOCK, bool COUNT, bool TWO>
__global__ __launch_bounds__(BLOCK) void k_trace(const float4 *__restrict__ rays, uint64_t n,
                                                 uint2 *__restrict__ hits, const uint4 *__restrict__ nodes,
                                                 const uint4 *__restrict__ tris, uint32_t root,
                                                 uint32_t *__restrict__ invalid_count,
                                                 uint32_t *__restrict__ cnt_nodes, uint32_t *__restrict__ cnt_tris,
                                                 SphereSet sph, const uint4 *__restrict__ insts,
                                                 const uint32_t *__restrict__ perm) {
    __shared__ uint2 stack[STACK * BLOCK];
    const uint64_t k = (uint64_t)xcd_group_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (k >= n) return;
    const uint64_t i = perm ? (uint64_t)perm[k] : k;
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

// ---- coherence order: counting sort of ray indices by (direction bin, origin cell) ----
// Direction bin: the cube face of the major axis (6) x a 4 x 4 grid of the other two components
// divided by the major one; origin cell: 8^3 cells of the scene box (clamped). Invalid rays go to
// bin 0 (the trace kernel reports them as before).
__device__ __forceinline__ uint32_t ray_bin(const RayIn &r, const RaySortScratch &s) {
    if (!ray_valid(r)) return 0u;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    uint32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = fminf(fmaxf((o[a] - s.lo[a]) * s.inv[a] * 8.f, 0.f), 7.f);
        c[a] = (uint32_t)u;
    }
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    const int m = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    const float dm = m == 0 ? d[0] : (m == 1 ? d[1] : d[2]);
    const float am = fabsf(dm);
    const float s1 = (m == 0 ? d[1] : (m == 1 ? d[2] : d[0])) / am, s2 = (m == 0 ? d[2] : (m == 1 ? d[0] : d[1])) / am;
    const uint32_t i1 = (uint32_t)fminf(fmaxf((s1 + 1.f) * 2.f, 0.f), 3.f);
    const uint32_t i2 = (uint32_t)fminf(fmaxf((s2 + 1.f) * 2.f, 0.f), 3.f);
    const uint32_t face = (uint32_t)m * 2u + (dm < 0.f ? 1u : 0u);
    return ((face * 4u + i1) * 4u + i2) * kRayCellBins + (c[0] + 8u * (c[1] + 8u * c[2]));
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) p[i] = v;
}

// The sort runs in two counting passes over chunks of rays (kRaySortChunk1 in pass 1,
// kRaySortChunk in pass 2), one 1024-thread block per chunk. Pass 1 orders the rays by direction bin alone (96 bins), pass 2 orders that sequence by the
// full key (direction bin x origin cell, 49,152 bins). A chunk of pass 1 holds ~340 rays per bin and
// one of pass 2 only a few direction bins, so the scatters write runs of neighbouring positions;
// one pass over the full key wrote single 4-byte entries all over the permutation. Per pass: count
// the chunk's bins with LDS atomics on 16-bit halves (two bins per word; a chunk holds fewer than
// 2^16 rays), reserve the chunk's run in every non-empty bin with one device atomic on the bin's
// total (the chunk order inside a bin is whatever order the atomics ran in: results do not depend on
// it), scan the totals into bin bases, and scatter each ray to base + run + its rank in the chunk
// (LDS atomics again). Four rays per thread are in flight: a 96 KiB block per CU leaves 16 waves.
constexpr uint32_t kSortBatch = 4;

template <int PASS>
__global__ __launch_bounds__(1024) void k_ray_count(const float4 *__restrict__ rays, uint64_t n, RaySortScratch s) {
    constexpr uint32_t kBins = PASS == 1 ? kRayDirBins : kRaySortBins;
    constexpr uint32_t kChunk = PASS == 1 ? kRaySortChunk1 : kRaySortChunk;
    __shared__ uint32_t cnt[kBins / 2];
    for (uint32_t b = threadIdx.x; b < kBins / 2; b += 1024u) cnt[b] = 0u;
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * kChunk;
    const uint64_t last = min(first + kChunk, n);
    for (uint64_t base = first + threadIdx.x; base < last; base += 1024u * kSortBatch) {
        uint32_t bin[kSortBatch];
        if (PASS == 1) {
            float4 a[kSortBatch], b[kSortBatch];
#pragma unroll
            for (uint32_t q = 0; q < kSortBatch; ++q) {
                const uint64_t i = base + 1024u * q;
                if (i < last) a[q] = rays[2 * i], b[q] = rays[2 * i + 1];
            }
#pragma unroll
            for (uint32_t q = 0; q < kSortBatch; ++q) {
                const uint64_t i = base + 1024u * q;
                bin[q] = 0u;
                if (i < last) {
                    const uint32_t key =
                        ray_bin(RayIn{a[q].x, a[q].y, a[q].z, a[q].w, b[q].x, b[q].y, b[q].z, b[q].w}, s);
                    s.keys[i] = key;
                    bin[q] = key / kRayCellBins;
                }
            }
        } else {
#pragma unroll
            for (uint32_t q = 0; q < kSortBatch; ++q) {
                const uint64_t i = base + 1024u * q;
                bin[q] = i < last ? s.keys1[i] : 0u;
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; ++q)
            if (base + 1024u * q < last) atomicAdd(&cnt[bin[q] >> 1], (bin[q] & 1u) ? 0x10000u : 1u);
    }
    __syncthreads();
    uint32_t *row = PASS == 1 ? s.offs1 + (size_t)blockIdx.x * kRayDirBins : s.offs + (size_t)blockIdx.x * kRaySortBins;
    uint32_t *total = s.total + (PASS == 1 ? 0u : kRayDirBins);
    for (uint32_t b = threadIdx.x; b < kBins / 2; b += 1024u) {
        const uint32_t w = cnt[b];
        if (w & 0xFFFFu) row[2 * b] = atomicAdd(&total[2 * b], w & 0xFFFFu);
        if (w >> 16) row[2 * b + 1] = atomicAdd(&total[2 * b + 1], w >> 16);
    }
}

// exclusive scan of kBins bin totals in place (one block; each thread owns a contiguous run)
template <uint32_t kBins>
__global__ __launch_bounds__(1024) void k_ray_bin_scan(uint32_t *hist) {
    __shared__ uint32_t part[1024];
    constexpr uint32_t per = (kBins + 1023u) / 1024u;
    const uint32_t t = threadIdx.x;
    const uint32_t b0 = min(t * per, kBins), b1 = min(b0 + per, kBins);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) sum += hist[b];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = hist[b];
        hist[b] = run;
        run += c;
    }
}

// pass 1: keys1[pos] = key, perm1[pos] = i; pass 2: perm[pos] = perm1[j]
template <int PASS>
__global__ __launch_bounds__(1024) void k_ray_scatter(uint64_t n, RaySortScratch s) {
    constexpr uint32_t kBins = PASS == 1 ? kRayDirBins : kRaySortBins;
    constexpr uint32_t kChunk = PASS == 1 ? kRaySortChunk1 : kRaySortChunk;
    __shared__ uint32_t cnt[kBins / 2];
    for (uint32_t b = threadIdx.x; b < kBins / 2; b += 1024u) cnt[b] = 0u;
    __syncthreads();
    const uint64_t first = (uint64_t)blockIdx.x * kChunk;
    const uint64_t last = min(first + kChunk, n);
    const uint32_t *row =
        PASS == 1 ? s.offs1 + (size_t)blockIdx.x * kRayDirBins : s.offs + (size_t)blockIdx.x * kRaySortBins;
    const uint32_t *bases = s.total + (PASS == 1 ? 0u : kRayDirBins);
    for (uint64_t base = first + threadIdx.x; base < last; base += 1024u * kSortBatch) {
        uint32_t key[kSortBatch], src[kSortBatch], pos[kSortBatch];
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; ++q) {
            const uint64_t i = base + 1024u * q;
            key[q] = 0u;
            src[q] = 0u;
            if (i < last) {
                key[q] = PASS == 1 ? s.keys[i] : s.keys1[i];
                src[q] = PASS == 1 ? (uint32_t)i : s.perm1[i];
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; ++q) {
            pos[q] = 0u;
            if (base + 1024u * q < last) {
                const uint32_t b = PASS == 1 ? key[q] / kRayCellBins : key[q];
                const uint32_t old = atomicAdd(&cnt[b >> 1], (b & 1u) ? 0x10000u : 1u);
                pos[q] = bases[b] + row[b] + ((b & 1u) ? old >> 16 : old & 0xFFFFu);
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kSortBatch; ++q) {
            if (base + 1024u * q >= last) continue;
            if (PASS == 1) {
                s.keys1[pos[q]] = key[q];
                s.perm1[pos[q]] = src[q];
            } else {
                s.perm[pos[q]] = src[q];
            }
        }
    }
}

namespace {
using TraceFn = void (*)(const float4 *, uint64_t, uint2 *, const uint4 *, const uint4 *, uint32_t, uint32_t *,
                         uint32_t *, uint32_t *, SphereSet, const uint4 *, const uint32_t *);
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
                        uint32_t sphere_id_base, const void *insts, const RaySortScratch *sort,
                        hipStream_t stream, const hipEvent_t *timed) {
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
            if (!work || !e.pfn || v.sort) return hipErrorInvalidValue;
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
            if (timed && (de = hipEventRecord(timed[0], stream)) != hipSuccess) return de;
            hipLaunchKernelGGL(e.pfn, dim3((uint32_t)grid), dim3(v.block), 0, stream, (const float4 *)rays, n,
                               (uint2 *)hits, (const uint4 *)nodes, (const uint4 *)tris, root, invalid_count,
                               cnt_nodes, cnt_tris, work, refill, sph);
            if ((de = hipGetLastError()) != hipSuccess) return de;
            return timed ? hipEventRecord(timed[1], stream) : hipSuccess;
        }
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        const uint32_t *perm = nullptr;
        if (v.sort && n <= 0xFFFFFFFFull) {  // 32-bit ray indices; larger calls trace in input order
            if (!sort) return hipErrorInvalidValue;
            const uint32_t chunks = (uint32_t)((n + kRaySortChunk - 1) / kRaySortChunk);
            const uint32_t chunks1 = (uint32_t)((n + kRaySortChunk1 - 1) / kRaySortChunk1);
            const float4 *r4 = (const float4 *)rays;
            hipLaunchKernelGGL(k_fill_u32, dim3((kRayDirBins + kRaySortBins + 255u) / 256u), dim3(256), 0, stream,
                               sort->total, kRayDirBins + kRaySortBins, 0u);
            hipLaunchKernelGGL(k_ray_count<1>, dim3(chunks1), dim3(1024), 0, stream, r4, n, *sort);
            hipLaunchKernelGGL(k_ray_bin_scan<kRayDirBins>, dim3(1), dim3(1024), 0, stream, sort->total);
            hipLaunchKernelGGL(k_ray_scatter<1>, dim3(chunks1), dim3(1024), 0, stream, n, *sort);
            hipLaunchKernelGGL(k_ray_count<2>, dim3(chunks), dim3(1024), 0, stream, r4, n, *sort);
            hipLaunchKernelGGL(k_ray_bin_scan<kRaySortBins>, dim3(1), dim3(1024), 0, stream, sort->total + kRayDirBins);
            hipLaunchKernelGGL(k_ray_scatter<2>, dim3(chunks), dim3(1024), 0, stream, n, *sort);
            perm = sort->perm;
        }
        hipError_t te;
        if (timed && (te = hipEventRecord(timed[0], stream)) != hipSuccess) return te;
        hipLaunchKernelGGL(e.fn, dim3((uint32_t)blocks), dim3(v.block), 0, stream, (const float4 *)rays, n,
                           (uint2 *)hits, (const uint4 *)nodes, (const uint4 *)tris, root, invalid_count, cnt_nodes,
                           cnt_tris, sph, (const uint4 *)insts, perm);
        if ((te = hipGetLastError()) != hipSuccess) return te;
        return timed ? hipEventRecord(timed[1], stream) : hipSuccess;
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
        if (k < max_out) out[k] = TraceVariant{e.width, e.stack, e.block, 0, 0};
        ++k;
        if (k < max_out) out[k] = TraceVariant{e.width, e.stack, e.block, 0, 1};
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
