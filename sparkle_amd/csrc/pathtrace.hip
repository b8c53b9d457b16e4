// Wavefront path tracer: CPURenderer::BasePass (CPURenderer.cpp:419-450) on the device.
//
// One path per pixel per frame, as in the reference. Per bounce the live paths pass through
//   extend  - closest hit of every queued ray (shared traversal code), plus a material histogram
//   sort    - counting sort of the queue by material id (scan + scatter)
//   shade   - SamplePixel's loop body (CPURenderer.cpp:259-368) for each hit, in material order;
//             survivors are appended to the next bounce's queue
//   shadow  - occlusion rays of next-event estimation (opt-in, not in the reference)
// With the learned radiance cache enabled, non-training paths stop at the first non-metal vertex
// from the cache bounce on and take the cache's prediction of the reflected radiance; a fraction
// of paths (training paths) run on and record, at each such vertex, the radiance gathered after
// that vertex as a training target, until they too end in the cache a few hit indices later
// (self-training; one in 16 is traced to its end).
// The queue kernels are persistent (the grid is sized to the device and waves fetch work from
// counters in device memory); camera rays, and every bounce under a schedule without lane refill,
// run one ray per thread over a grid that covers every pixel. Work sizes live in device memory, so a
// frame is captured once as a hipGraph and replayed; per-frame values (the frame seed) travel through
// a small device buffer written before each replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/spk_pt.h"
#include "shade.h"
#include "spk_cache.h"
#include "spk_context.h"
#include "traverse.h"

namespace spk {

constexpr int kMaxBounce = 64;
constexpr int kCtrStride = 8;  // counters per bounce
constexpr int kCtrGlobal = kMaxBounce * kCtrStride;
constexpr int kCtrTotal = kCtrGlobal + 16;
constexpr int kMaxSortKeys = 1024;
constexpr int kPtBlock = 128;

enum : uint32_t {
    CTR_QUEUE = 0,    // paths entering this bounce
    CTR_EXTEND = 1,   // extend work fetch
    CTR_SHADE = 2,    // shade work fetch
    CTR_SHADOWS = 3,  // shadow rays emitted at this bounce
    CTR_SHADOWW = 4,  // shadow work fetch
    CTR_SCATTER = 5,  // scatter work fetch
};
enum : uint32_t {
    GCTR_CACHE_Q = 0,   // cache queries of the frame
    GCTR_TRAIN = 1,     // training samples recorded (may exceed the capacity)
    GCTR_APPLY = 2,     // cache apply work fetch
    GCTR_TARGETS = 3,   // target work fetch
    GCTR_TAIL_RAYS = 4,     // segments traced by the tail kernel
    GCTR_TAIL_SHADOWS = 5,  // shadow rays traced by the tail kernel
};
// feature sets a kernel variant is compiled for
enum PtFeat : uint32_t { kFeatNee = 1u, kFeatCache = 2u, kFeatAll = 3u };

enum : uint32_t {
    PF_NEE_PREV = 1u,  // the previous vertex sampled the emitters directly (MIS-weight emission hits)
    PF_TRAIN = 2u,          // a training path of the radiance cache
    PF_TRAIN_FULL = 4u,     // a training path traced to its end (its targets never see the cache)
};

// Per-scene record of a frame: BVH root, image size, first pixel of the scene's image, camera.
// spk_render_frame uses one record, rewritten with the frame seed before each replay (a camera move
// needs no new graph); spk_batch_render one per scene, uploaded with the batch.
struct SceneRec {
    uint32_t root, width, height, pix_base;
    spk_camera_desc cam;
    uint32_t pad;
};
static_assert(sizeof(SceneRec) == 112, "SceneRec layout");

struct FrameDev {
    uint32_t frame_seed;
    uint32_t pad[3];
    SceneRec scene0;
};

struct PtArgs {
    const uint4 *nodes;
    const uint4 *tris;   // traversal triangles (a two-level scene's are in object space)
    const uint4 *wtris;  // world triangles indexed like `shade` (the same array as tris when flattened)
    const uint4 *insts;  // two-level scenes: instance records of the TLAS leaves
    const TriShade *shade;
    const MaterialGpu *mats;
    const uint32_t *emitters;
    const float *emitter_cdf;
    const TexRecord *tex_table;  // texture table and float4 texels
    const float4 *texels;
    const float4 *spheres;  // analytic spheres, tested after the BVH
    uint32_t num_spheres, sphere_base;
    uint32_t num_emitters;
    float emitter_total;
    float4 *rays;      // 2 per path: (org, tmin), (dir, tmax)
    float4 *thr;       // throughput rgb, w = path flags (PF_* bits)
    float *prev_pdf;   // solid-angle pdf of the direction sampled at the previous (diffuse) vertex
    float4 *col;       // colour rgb, w = valid flag
    uint4 *rng;        // xoshiro128+ state per path
    uint4 *nee_rng;    // second stream for light sampling
    uint32_t *queue0;  // ping-pong path queues
    uint32_t *queue1;
    uint4 *hits;       // per queue slot: leaf triangle (or miss), t, u, v
    uint32_t *sorted;  // queue slots in material order
    float4 *shadow;    // 3 per shadow ray: (org, tmax), (dir, path), (contribution)
    uint32_t *counters;
    uint32_t *hist;    // per bounce: counts[nkeys], offsets[nkeys]
    uint32_t nkeys;
    float *rgba_out;
    float *normal_out;
    const FrameDev *frame;
    const SceneRec *scenes;  // num_scenes records, sorted by pix_base
    uint32_t *path_scene;    // scene of each path (written only when num_scenes > 1)
    uint32_t num_scenes, grid;
    uint32_t refill;  // extend: a wave takes new paths once this many of its lanes are idle (1..64)
    uint32_t direct;  // 1: wavefront extends and shadow rays run one per thread over a full grid
    uint32_t tail_from;  // segments from this index on run in k_tail (max_bounce: none)
    uint32_t npix, max_bounce, debug_mode, sky_enabled, flags, sort;
    int32_t sky_faces[6];  // cube-map faces +X, -X, +Y, -Y, +Z, -Z in the texture table, or -1
    float sky[3];
    float output_limit;
    // learned radiance cache
    uint32_t cache_on, cache_bounce, train_cap;
    uint32_t train_suffix;  // hit indices a training path runs past cache_bounce before it ends in the cache
    float train_prob;
    float *cq_points;  // 3 per query
    float4 *cq_attrs;  // 2 per query: (normal, albedo.r), (albedo.gb, roughness, 1)
    float4 *cq_meta;   // (throughput, path)
    float *cq_out;     // 3 per query
    float *tr_points;  // 3 per sample
    float4 *tr_attrs;  // 2 per sample, as cq_attrs
    float4 *tr_meta0;  // (colour before the vertex, path)
    float4 *tr_meta1;  // (throughput at the vertex, 0)
    float *tr_targets; // 3 per sample
    float *st_points;  // the frame's training samples handed to the next frame's training step
    float4 *st_attrs;
    float *st_targets;
    uint32_t *st_count;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Device-scope atomics on one counter serialise (a few ns each), so work distribution avoids one
// atomic per 64 items: kernels with even per-item cost stride statically over 64-item chunks
// (wave w takes chunks w, w + waves, ...); traversal kernels take kFetchChunk slots per atomic.
// 16: a wave that ends up holding unprocessed slots when the queue runs dry delays the kernel's end
// (1080p frames of the many-instance scene: 128 slots 2.62 ms, 64 2.50, 16 2.35; 8, 4 and 1 no
// faster than 16; with cache + NEE 3.34, 3.20, 3.06-3.08 ms; 64-scene batch 1414 -> 1482-1488 M
// samples/s)
constexpr uint32_t kFetchChunk = 16;
// shadow rays are taken in whole waves of 64 (k_shadow): a multiple of 64 per atomic
constexpr uint32_t kShadowChunk = 128;

__device__ __forceinline__ uint32_t grid_wave() { return (blockIdx.x * kPtBlock + threadIdx.x) >> 6; }
__device__ __forceinline__ uint32_t grid_waves() { return gridDim.x * (uint32_t)(kPtBlock >> 6); }

// one atomic per wave for n items: the wave's lanes take consecutive items from the returned base
__device__ __forceinline__ uint32_t wave_fetch(uint32_t *ctr, uint32_t n) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, n);
    return __shfl(base, 0);
}

// A wave's private run of queue slots [next, end), refilled kFetchChunk at a time from the shared
// counter. take() hands one slot to each idle lane (lanes in `idle`, rank = idle lanes below it);
// a slot >= count means no work. Wave-uniform state; every lane of the wave must call take().
struct SlotRun {
    uint32_t next = 0u, end = 0u;
    bool global_done = false;
    __device__ __forceinline__ uint32_t take(uint32_t *ctr, uint32_t count, uint32_t nidle, uint32_t rank) {
        const uint32_t avail = end > next ? end - next : 0u;
        uint32_t fresh = 0u;
        const bool fetch = nidle > avail && !global_done;
        // at least the idle lanes the run cannot serve, so every rank gets a slot of this wave's own
        const uint32_t n = fetch ? max(kFetchChunk, nidle - avail) : kFetchChunk;
        if (fetch) {
            fresh = wave_fetch(ctr, n);
            global_done = fresh + n >= count;
        }
        uint32_t my = 0xFFFFFFFFu;
        if (rank < avail) my = next + rank;
        else if (fetch) my = fresh + (rank - avail);
        if (nidle <= avail) {
            next += nidle;
        } else if (fetch) {
            next = fresh + (nidle - avail);
            end = min(fresh + n, count);
        } else {
            next = end;
        }
        return my;
    }
    __device__ __forceinline__ bool drained() const { return global_done && next >= end; }
};

// wave-aggregated append over the lanes that execute it: returns the slot for lanes with `want`
// (the lowest active lane takes the slots, so it also works inside divergent code)
__device__ __forceinline__ uint32_t wave_append(uint32_t *ctr, bool want) {
    const uint64_t mask = __ballot(want);
    const uint32_t cnt = __popcll(mask);
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)__ballot(true)) - 1u;
    uint32_t base = 0;
    if (lane_id() == leader && cnt) base = atomicAdd(ctr, cnt);
    base = __shfl(base, (int)leader);
    const uint64_t below = lane_id() == 0 ? 0ull : (mask & ((~0ull) >> (64 - lane_id())));
    return base + (uint32_t)__popcll(below);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// material of a hit record's primitive (leaf triangle, or SPK_SPHERE_BIT | sphere)
__device__ __forceinline__ uint32_t hit_material(const PtArgs &a, uint32_t prim) {
    if (prim & SPK_SPHERE_BIT) return __float_as_uint(a.spheres[2 * (prim & ~SPK_SPHERE_BIT) + 1].x);
    return a.shade[prim].material;
}

// last scene whose first pixel is <= p
__device__ __forceinline__ uint32_t scene_of_pixel(const PtArgs &a, uint32_t p) {
    uint32_t lo = 0u, hi = a.num_scenes - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (a.scenes[mid].pix_base <= p) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

__device__ __forceinline__ const SceneRec &scene_of_path(const PtArgs &a, uint32_t path) {
    return a.scenes[a.num_scenes > 1u ? a.path_scene[path] : 0u];
}

// ---- generate: reseed per pixel, jitter, lens, camera ray (CPURenderer.cpp:220-238, 398-445) ----
// Pixel p of the frame is pixel p - pix_base of its scene's image; seeds, jitter and camera are the
// scene's own, so a scene renders the same alone or in a batch.
// Position of pixel (i, j) of a W x H image in 8x8-tile order (tiles row by row, pixels row by row
// inside a tile; tiles on the right and bottom edges are narrower or shorter, so the order is a
// bijection onto [0, W*H)). Bounce 0 takes its paths in this order, so a wave's 64 camera rays
// leave one tile of the image instead of a 64x1 strip.
__device__ __forceinline__ uint32_t tile_order(uint32_t i, uint32_t j, uint32_t W, uint32_t H) {
    const uint32_t tx = i >> 3, ty = j >> 3;
    const uint32_t w = min(8u, W - (tx << 3)), h = min(8u, H - (ty << 3));
    return (ty << 3) * W + (tx << 3) * h + (j & 7u) * w + (i & 7u);
}

__global__ __launch_bounds__(kPtBlock) void k_generate(PtArgs a) {
    const uint32_t tid = blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t stride = gridDim.x * kPtBlock;
    // clear this frame's counters and histograms (no memset nodes in the graph)
    for (uint32_t i = tid; i < (uint32_t)kCtrTotal; i += stride) a.counters[i] = 0u;
    for (uint32_t i = tid; i < a.max_bounce * 2u * a.nkeys; i += stride) a.hist[i] = 0u;
    const uint32_t frame_seed = a.frame->frame_seed;
    for (uint32_t p = tid; p < a.npix; p += stride) {
        uint32_t sid = 0u;
        if (a.num_scenes > 1u) {
            sid = scene_of_pixel(a, p);
            a.path_scene[p] = sid;
        }
        const SceneRec &S = a.scenes[sid];
        const uint32_t W = S.width, H = S.height, q = p - S.pix_base;
        const float pw = 1.f / (float)(W - 1);
        const float ph = 1.f / (float)(H - 1);
        const uint32_t j = q / W, i = q - j * W;
        uint64_t seed;
        if (a.flags & SPK_FRAME_SEED64)
            seed = (uint64_t)j * W + i + (uint64_t)frame_seed * W * H;
        else
            seed = (uint32_t)(j * W + i + frame_seed * W * H);
        Rng rng = rng_seed(seed);
        const float u = ((float)i + rng_unit(rng)) * pw;
        const float v = ((float)j + rng_unit(rng)) * ph;
        float nx, ny;
        rng_disk(rng, nx, ny);
        const spk_camera_desc &cam = S.cam;
        nx = nx * cam.lens_radius;
        ny = ny * cam.lens_radius;
        const F3 lens = ld3(cam.right) * nx + ld3(cam.up) * ny;
        const F3 o = ld3(cam.position) + lens;
        const F3 focus = ld3(cam.lower_left) + ld3(cam.max_u) * u + ld3(cam.max_v) * v;
        const F3 d = normalized(focus - o);
        uint32_t pf = 0u;
        if (a.cache_on) {
            // training paths: an independent hash of pixel and frame, so the path itself is unchanged
            const float h = (float)(hash32(p * 0x9E3779B9u ^ hash32(frame_seed + 0x632BE5ABu)) >> 8) * (1.f / 16777216.f);
            if (h < a.train_prob) pf |= PF_TRAIN;
            if (h < a.train_prob * (1.f / 16.f)) pf |= PF_TRAIN_FULL;
        }
        a.rays[2 * p] = make_float4(o.x, o.y, o.z, 0.f);
        a.rays[2 * p + 1] = make_float4(d.x, d.y, d.z, INFINITY);
        a.thr[p] = make_float4(1.f, 1.f, 1.f, __uint_as_float(pf));
        a.col[p] = make_float4(0.f, 0.f, 0.f, 1.f);
        a.rng[p] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
        if (a.flags & SPK_FRAME_NEE) {
            const Rng r2 = rng_seed(seed ^ 0x5bd1e9955bd1e995ull);
            a.nee_rng[p] = make_uint4(r2.s0, r2.s1, r2.s2, r2.s3);
        }
        a.queue0[S.pix_base + tile_order(i, j, W, H)] = p;
        if (a.normal_out) {
            a.normal_out[3 * (size_t)p] = 0.f;
            a.normal_out[3 * (size_t)p + 1] = 0.f;
            a.normal_out[3 * (size_t)p + 2] = 0.f;
        }
    }
    if (tid == 0) a.counters[CTR_QUEUE] = a.npix;
}

// hit record of a finished closest-hit query: the triangle's shading index (leaf order in a
// flattened scene, global id in a two-level one) or SPK_SPHERE_BIT | sphere, t and the
// barycentrics, or a miss; spheres are tested after the BVH
template <bool TWO>
__device__ __forceinline__ uint4 hit_record(const PtArgs &a, const RayIn &r, BestHit &best) {
    if (a.num_spheres) spheres_closest(a.spheres, a.num_spheres, a.sphere_base, r, best);
    if (best.prim == 0xFFFFFFFFu) return make_uint4(0xFFFFFFFFu, __float_as_uint(INFINITY), 0u, 0u);
    const float t = final_t(a.tris, r, best, TWO ? a.insts : nullptr);
    if (best.tri & SPK_SPHERE_BIT) return make_uint4(best.tri, __float_as_uint(t), 0u, 0u);
    return make_uint4(TWO ? best.prim : best.tri, __float_as_uint(t), __float_as_uint(best.V / best.det),
                      __float_as_uint(best.W / best.det));
}

// ---- extend: closest hit for every queued path + material histogram ----
// Lanes trace one path each and a wave takes new queue slots (one atomic) once `refill` of its
// lanes are idle, so a wave does not wait for its slowest ray before fetching more work.
template <int WIDTH, int STACK, bool TWO>
__global__ __launch_bounds__(kPtBlock) void k_extend(PtArgs a, uint32_t bounce) {
    __shared__ uint2 stack[STACK * kPtBlock];
    __shared__ uint32_t lhist[kMaxSortKeys];
    uint32_t *ctr = a.counters + bounce * kCtrStride;
    const uint32_t count = ctr[CTR_QUEUE];
    const uint32_t *queue = (bounce & 1u) ? a.queue1 : a.queue0;
    if (a.sort)
        for (uint32_t k = threadIdx.x; k < a.nkeys; k += kPtBlock) lhist[k] = 0u;
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    // camera rays of neighbouring pixels are coherent: their waves refill whole, so lanes stay in step
    const uint32_t refill = bounce == 0u ? 64u : a.refill;
    Traversal<WIDTH, STACK, false, false, TWO> tr;
    uint32_t slot = 0u, path = 0u;
    bool active = false;
    SlotRun run;
    for (;;) {
        if (!run.drained()) {
            const uint64_t idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= refill) {
                const uint32_t my = run.take(ctr + CTR_EXTEND, count, nidle, (uint32_t)__popcll(idle & below));
                if (!active) {
                    if (my < count) {
                        path = queue[my];
                        const float4 ro = a.rays[2 * path], rd = a.rays[2 * path + 1];
                        const RayIn r{ro.x, ro.y, ro.z, ro.w, rd.x, rd.y, rd.z, rd.w};
                        const uint32_t root = a.num_scenes > 1u ? a.scenes[a.path_scene[path]].root : 0u;
                        tr.init(r, root, a.insts);
                        slot = my;
                        active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (run.drained()) break;
            continue;
        }
        if (active && !tr.step(a.nodes, a.tris, stack + threadIdx.x, kPtBlock)) {
            uint32_t key = a.nkeys - 1u;  // misses sort last
            const float4 ro = a.rays[2 * path], rd = a.rays[2 * path + 1];
            const RayIn r{ro.x, ro.y, ro.z, ro.w, rd.x, rd.y, rd.z, rd.w};
            const uint4 h = hit_record<TWO>(a, r, tr.best);
            if (h.x != 0xFFFFFFFFu) {
                key = hit_material(a, h.x);
                if (key >= a.nkeys - 1u) key = a.nkeys - 2u;
            }
            a.hits[slot] = h;
            if (a.sort) atomicAdd(&lhist[key], 1u);
            active = false;
        }
    }
    if (a.sort) {
        __syncthreads();
        uint32_t *gh = a.hist + (size_t)bounce * 2u * a.nkeys;
        for (uint32_t k = threadIdx.x; k < a.nkeys; k += kPtBlock)
            if (lhist[k]) atomicAdd(&gh[k], lhist[k]);
    }
}

// Extend without the persistent grid: one queue slot per thread over a grid that covers every
// pixel (a queue never holds more paths). Always used for bounce 0: camera rays of a tile are
// coherent, so lane refill buys nothing there, and a full grid keeps more waves resident than the
// persistent one. Later bounces use it when the schedule's refill is 0. Same hit records and
// histogram as k_extend.
// SORT: the frame sorts by material (the histogram's 4 KiB of LDS and the key lookup only then;
// without them a 128-thread block needs 12 KiB and a 1080p camera-ray pass takes ~50 us less).
template <int WIDTH, int STACK, bool TWO, bool SORT>
__global__ __launch_bounds__(kPtBlock) void k_extend_first(PtArgs a, uint32_t bounce) {
    __shared__ uint2 stack[STACK * kPtBlock];
    __shared__ uint32_t lhist[SORT ? kMaxSortKeys : 1];
    const uint32_t *ctr = a.counters + bounce * kCtrStride;
    const uint32_t count = ctr[CTR_QUEUE];
    const uint32_t *queue = (bounce & 1u) ? a.queue1 : a.queue0;
    if (SORT) {
        for (uint32_t k = threadIdx.x; k < a.nkeys; k += kPtBlock) lhist[k] = 0u;
        __syncthreads();
    }
    const uint32_t slot = blockIdx.x * kPtBlock + threadIdx.x;
    if (slot < count) {
        const uint32_t path = queue[slot];
        const float4 ro = a.rays[2 * path], rd = a.rays[2 * path + 1];
        const RayIn r{ro.x, ro.y, ro.z, ro.w, rd.x, rd.y, rd.z, rd.w};
        const uint32_t root = a.num_scenes > 1u ? a.scenes[a.path_scene[path]].root : 0u;
        Traversal<WIDTH, STACK, false, false, TWO> tr;
        tr.init(r, root, a.insts);
        while (tr.step(a.nodes, a.tris, stack + threadIdx.x, kPtBlock)) {
        }
        const uint4 h = hit_record<TWO>(a, r, tr.best);
        a.hits[slot] = h;
        if (SORT) {
            uint32_t key = a.nkeys - 1u;  // misses sort last
            if (h.x != 0xFFFFFFFFu) {
                key = hit_material(a, h.x);
                if (key >= a.nkeys - 1u) key = a.nkeys - 2u;
            }
            atomicAdd(&lhist[key], 1u);
        }
    }
    if (SORT) {
        __syncthreads();
        uint32_t *gh = a.hist + (size_t)bounce * 2u * a.nkeys;
        for (uint32_t k = threadIdx.x; k < a.nkeys; k += kPtBlock)
            if (lhist[k]) atomicAdd(&gh[k], lhist[k]);
    }
}

// ---- sort: exclusive scan of the histogram (one block) and scatter of queue slots ----
__global__ __launch_bounds__(1024) void k_sort_scan(PtArgs a, uint32_t bounce) {
    __shared__ uint32_t part[1024];
    uint32_t *gh = a.hist + (size_t)bounce * 2u * a.nkeys;
    const uint32_t k = threadIdx.x;
    const uint32_t v = k < a.nkeys ? gh[k] : 0u;
    part[k] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t add = k >= off ? part[k - off] : 0u;
        __syncthreads();
        part[k] += add;
        __syncthreads();
    }
    if (k < a.nkeys) gh[a.nkeys + k] = part[k] - v;  // exclusive offsets
}

// One atomic per distinct key of a wave (not per slot): most slots of a wave share a few materials,
// and per-slot atomics on one hot key serialise in L2.
__global__ __launch_bounds__(kPtBlock) void k_sort_scatter(PtArgs a, uint32_t bounce) {
    uint32_t *ctr = a.counters + bounce * kCtrStride;
    const uint32_t count = ctr[CTR_QUEUE];
    uint32_t *offs = a.hist + (size_t)bounce * 2u * a.nkeys + a.nkeys;
    const uint32_t lane = lane_id();
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    for (uint32_t base = grid_wave() * 64u; base < count; base += grid_waves() * 64u) {
        const uint32_t slot = base + lane;
        const bool valid = slot < count;
        uint32_t key = a.nkeys - 1u;
        if (valid) {
            const uint4 h = a.hits[slot];
            if (h.x != 0xFFFFFFFFu) {
                key = hit_material(a, h.x);
                if (key >= a.nkeys - 1u) key = a.nkeys - 2u;
            }
        }
        uint64_t pending = __ballot(valid);
        uint32_t pos = 0u;
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t k = __shfl(key, leader);
            const uint64_t peers = __ballot(valid && key == k);
            uint32_t first = 0u;
            if ((int)lane == leader) first = atomicAdd(&offs[k], (uint32_t)__popcll(peers));
            first = __shfl(first, leader);
            if (valid && key == k) pos = first + (uint32_t)__popcll(peers & below);
            pending &= ~peers;
        }
        if (valid) a.sorted[pos] = slot;
    }
}

__device__ __forceinline__ float4 lerp_attr(const float *p0, const float *p1, const float *p2, float u, float v) {
    // v0 + (v1 - v0) * u + (v2 - v0) * v (Utilities.h:114-117)
    return make_float4(p0[0] + (p1[0] - p0[0]) * u + (p2[0] - p0[0]) * v,
                       p0[1] + (p1[1] - p0[1]) * u + (p2[1] - p0[1]) * v,
                       p0[2] + (p1[2] - p0[2]) * u + (p2[2] - p0[2]) * v,
                       p0[3] + (p1[3] - p0[3]) * u + (p2[3] - p0[3]) * v);
}

// Image2DCube::Sample (io/Image.cpp:298-357): face by the largest axis, uv in [0, 1]^2, then the
// face image's bilinear sample
__device__ __forceinline__ F3 sample_sky_cube(const PtArgs &a, F3 dir) {
    const float x = dir.x, y = dir.y, z = dir.z;
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    int face;
    float u, v;
    if (ax >= ay && ax >= az) {
        if (x > 0.f) { face = 0; u = -z / x; v = -y / x; } else { face = 1; u = z / -x; v = -y / -x; }
    } else if (ay >= ax && ay >= az) {
        if (y > 0.f) { face = 2; u = x / y; v = z / y; } else { face = 3; u = x / -y; v = -z / -y; }
    } else {
        if (z > 0.f) { face = 4; u = x / z; v = -y / z; } else { face = 5; u = -x / -z; v = -y / -z; }
    }
    u = (u + 1.f) * 0.5f;
    v = (v + 1.f) * 0.5f;
    return tex_sample(a.tex_table, a.texels, a.sky_faces[face], u, v);
}

// GetPossibleMajorAxis (Utilities.h:222-234)
__device__ __forceinline__ F3 major_axis(F3 n) {
    const float inv_sqrt3 = 0.577350269189625764509f;
    if (fabsf(n.x) < inv_sqrt3) return f3(1.f, 0.f, 0.f);
    if (fabsf(n.y) < inv_sqrt3) return f3(0.f, 1.f, 0.f);
    return f3(0.f, 0.f, 1.f);
}

// Area pdf with which next-event estimation picks a point of an emitter of this material: triangles
// are drawn by area x mean emitted radiance (factor), points uniformly, so the pdf is lum / total.
__device__ __forceinline__ float light_pdf_area(const PtArgs &a, const MaterialGpu &m) {
    return (m.emissive[0] + m.emissive[1] + m.emissive[2]) / 3.f / a.emitter_total;
}

// Material getters at a uv (MaterialRenderProxy.h:117-162)
struct MatEval {
    F3 base, emissive;
    float metallic, roughness;
    bool emits;
};

__device__ __forceinline__ MatEval eval_material(const PtArgs &a, const MaterialGpu &m, float tu, float tv) {
    MatEval e;
    e.base = f3(m.base[0], m.base[1], m.base[2]);
    e.emissive = f3(m.emissive[0], m.emissive[1], m.emissive[2]);
    e.metallic = m.metallic;
    e.roughness = m.roughness;
    e.emits = m.emissive_flag != 0u;
    if (m.tex_base >= 0) e.base = cmul(tex_sample(a.tex_table, a.texels, m.tex_base, tu, tv), e.base);
    if (m.tex_mr >= 0) {
        const F3 t = tex_sample(a.tex_table, a.texels, m.tex_mr, tu, tv);
        e.metallic = t.z * m.metallic;
        e.roughness = t.y * m.roughness;
    }
    if (m.tex_emissive >= 0) {
        e.emissive = cmul(tex_sample(a.tex_table, a.texels, m.tex_emissive, tu, tv), e.emissive);
        e.emits = sqrtf(sqnorm(e.emissive)) > SPK_EPS;
    }
    return e;
}

// ---- debug modes (RenderConfig::DebugMode, RenderConfig.h:33-47) ----
// Modes that return at the first non-emissive hit (CPURenderer.cpp:296-327): Debug, Normal,
// RayDirection, Metallic, Roughness, Albedo, Emissive, Depth.
__device__ __forceinline__ bool first_hit_debug(uint32_t m) { return m == 1u || m == 3u || m == 4u || (m >= 7u && m <= 11u); }
// Modes that rewrite the colour once the loop ends (CPURenderer.cpp:370-391): Debug (paths that end
// before a non-emissive first hit), RayDepth, IndirectLighting, DirectionalLighting.
__device__ __forceinline__ bool post_loop_debug(uint32_t m) { return m == 1u || m == 2u || m == 5u || m == 6u; }
__device__ __forceinline__ F3 debug_after_loop(uint32_t m, uint32_t end_bounce, F3 C) {
    if (m == 1u) return f3(0.f, 0.f, 0.f);
    if (m == 5u) return end_bounce <= 1u ? f3(0.f, 0.f, 0.f) : C;
    if (m == 6u) return end_bounce > 1u ? f3(0.f, 0.f, 0.f) : C;
    // RayDepth: VisualizeInteger (Utilities.h:93-112), a 7-colour palette / 255, black from 7 on
    const float r[7] = {137.f, 242.f, 255.f, 0.f, 135.f, 225.f, 255.f};
    const float g[7] = {49.f, 202.f, 0.f, 87.f, 233.f, 24.f, 255.f};
    const float b[7] = {239.f, 25.f, 189.f, 233.f, 17.f, 69.f, 255.f};
    if (end_bounce >= 7u) return f3(0.f, 0.f, 0.f);
    return f3(r[end_bounce] / 255.0f, g[end_bounce] / 255.0f, b[end_bounce] / 255.0f);
}

// ---- shade: SamplePixel's loop body (CPURenderer.cpp:261-368) ----
// What shading one path segment produced: whether the path continues, its shadow ray (NEE), and
// the cache query or training vertex it recorded.
struct ShadeOut {
    bool cont, shadow, cache_q, train;
    float4 sh0, sh1, sh2;
};

// Cache query and training vertex records, written where shading decides on them (wave-aggregated over
// the lanes taking that branch), so their values need no registers past that point.
__device__ __forceinline__ void record_cache_query(const PtArgs &a, uint32_t path, const F3 loc, const F3 T,
                                                   const float4 at0, const float4 at1) {
    const uint32_t cs = wave_append(a.counters + kCtrGlobal + GCTR_CACHE_Q, true);
    a.cq_points[3 * (size_t)cs] = loc.x;
    a.cq_points[3 * (size_t)cs + 1] = loc.y;
    a.cq_points[3 * (size_t)cs + 2] = loc.z;
    a.cq_meta[cs] = make_float4(T.x, T.y, T.z, __uint_as_float(path));
    a.cq_attrs[2 * (size_t)cs] = at0;
    a.cq_attrs[2 * (size_t)cs + 1] = at1;
}

// C: the path's colour before this vertex adds anything (the target is what the path gathers after it)
__device__ __forceinline__ void record_training(const PtArgs &a, uint32_t path, const F3 loc, const F3 C, const F3 T,
                                                const float4 at0, const float4 at1) {
    const uint32_t ts = wave_append(a.counters + kCtrGlobal + GCTR_TRAIN, true);
    if (ts >= a.train_cap) return;
    a.tr_points[3 * (size_t)ts] = loc.x;
    a.tr_points[3 * (size_t)ts + 1] = loc.y;
    a.tr_points[3 * (size_t)ts + 2] = loc.z;
    a.tr_meta0[ts] = make_float4(C.x, C.y, C.z, __uint_as_float(path));
    a.tr_meta1[ts] = make_float4(T.x, T.y, T.z, 0.f);
    a.tr_attrs[2 * (size_t)ts] = at0;
    a.tr_attrs[2 * (size_t)ts + 1] = at1;
}

// Shades path `path` at hit record h of segment `bounce`: adds emission/sky to its colour, samples the
// next direction (rays/thr/rng/prev_pdf written when it continues) and prepares the NEE shadow ray.
// The colour is written back before the shadow ray's contribution is added, so the sum order per
// path is the same whichever kernel runs the segment. FEAT (PtFeat bits) compiles in next-event
// estimation and the cache; a frame that uses neither runs the same arithmetic with fewer registers.
template <uint32_t FEAT = kFeatAll>
__device__ __forceinline__ void shade_path(const PtArgs &a, uint32_t bounce, bool last, bool nee_on, uint32_t path,
                                           const uint4 h, ShadeOut &o) {
    const bool nee = (FEAT & kFeatNee) && nee_on;
    const bool cache_on = (FEAT & kFeatCache) && a.cache_on;
    bool &cont = o.cont, &shadow = o.shadow, &cache_q = o.cache_q, &train = o.train;
    float4 &sh0 = o.sh0, &sh1 = o.sh1, &sh2 = o.sh2;
    F3 loc = f3(0.f, 0.f, 0.f);
    cont = shadow = cache_q = train = false;
    sh0 = sh1 = sh2 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 rd4 = a.rays[2 * path + 1], ro4 = a.rays[2 * path];
    const F3 d = f3(rd4.x, rd4.y, rd4.z);
    const float4 T4 = a.thr[path];
    F3 T = f3(T4.x, T4.y, T4.z);
    const uint32_t pf = __float_as_uint(T4.w);
    uint32_t end_b = bounce;  // the segment the path ends in (max_bounce when the loop runs out)
    bool first_done = false;  // a first-hit debug mode wrote the colour
    float4 C4 = a.col[path];
    F3 C = f3(C4.x, C4.y, C4.z);
    if (h.x == 0xFFFFFFFFu) {
        if (a.sky_enabled) {
            F3 sky;
            if (a.sky_faces[0] >= 0) {
                sky = sample_sky_cube(a, d);
            } else {
                // top: colour, bottom: white (SkyRenderProxy.cpp:58-61)
                const float t = 0.5f * d.z + 0.5f;
                sky = f3(1.f, 1.f, 1.f) + (f3(a.sky[0], a.sky[1], a.sky[2]) - f3(1.f, 1.f, 1.f)) * t;
            }
            C = C + cmul(sky, T);
        }
    } else {
        const float u = __uint_as_float(h.z), v = __uint_as_float(h.w), t = __uint_as_float(h.y);
        loc = f3(ro4.x, ro4.y, ro4.z) + d * t;
        const bool is_sphere = (h.x & SPK_SPHERE_BIT) != 0u;
        F3 normal, tangent;
        float tu = 0.f, tv = 0.f;
        MaterialGpu mat;
        if (is_sphere) {
            // SphereRenderProxy::GetIntersection (SphereRenderProxy.h:22-31): uv 0, no normal map
            const uint32_t k = h.x & ~SPK_SPHERE_BIT;
            const float4 sc = a.spheres[2 * k];
            mat = a.mats[__float_as_uint(a.spheres[2 * k + 1].x)];
            normal = normalized(loc - f3(sc.x, sc.y, sc.z));
            tangent = major_axis(normal);
        } else {
            const TriShade &s = a.shade[h.x];
            const float4 n4 = lerp_attr(s.n0, s.n1, s.n2, u, v);
            const float4 t4 = lerp_attr(s.t0, s.t1, s.t2, u, v);
            // texture coordinate (Mesh.h:119-132)
            tu = s.uv0[0] + (s.uv1[0] - s.uv0[0]) * u + (s.uv2[0] - s.uv0[0]) * v;
            tv = s.uv0[1] + (s.uv1[1] - s.uv0[1]) * u + (s.uv2[1] - s.uv0[1]) * v;
            mat = a.mats[s.material];
            F3 nrm = f3(n4.x, n4.y, n4.z);
            tangent = f3(t4.x, t4.y, t4.z);
            if (mat.tex_normal >= 0) {
                // normal map in the tangent frame (MeshRenderProxy.cpp:350-356, Utilities.h:398-404)
                const F3 tn =
                    tex_sample(a.tex_table, a.texels, mat.tex_normal, tu, tv) * 2.f - f3(1.f, 1.f, 1.f);
                const F3 bit = cross(nrm, tangent) * t4.w;
                nrm = normalized(tangent * tn.x + bit * tn.y + nrm * tn.z);
            }
            normal = normalized(nrm);
        }
        const MatEval me = eval_material(a, mat, tu, tv);
        if (bounce == 0u && a.normal_out) {
            a.normal_out[3 * (size_t)path] = normal.x;
            a.normal_out[3 * (size_t)path + 1] = normal.y;
            a.normal_out[3 * (size_t)path + 2] = normal.z;
        }
        // the cache is a function of position only, so it stands in for the path tail only at
        // vertices whose reflected radiance barely depends on direction (non-metal PBR); paths
        // continue through metal and glass vertices to the next such vertex
        const bool cache_vertex = cache_on && bounce >= a.cache_bounce && mat.type == 0u &&
                                  me.metallic < 0.5f;
        if (me.emits) {
            if ((FEAT & kFeatNee) && (pf & PF_NEE_PREV) && !is_sphere) {  // spheres are not sampled emitters
                // the previous vertex also sampled this emitter: balance-heuristic weight of
                // the BSDF-sampled direction against light sampling of this point
                const uint4 *tp = a.wtris + (size_t)h.x * 3u;
                const uint4 va = tp[0], vb = tp[1], vc = tp[2];
                const F3 p0 = f3(__uint_as_float(va.x), __uint_as_float(va.y), __uint_as_float(va.z));
                const F3 ng = cross(f3(__uint_as_float(vb.x), __uint_as_float(vb.y), __uint_as_float(vb.z)) - p0,
                                    f3(__uint_as_float(vc.x), __uint_as_float(vc.y), __uint_as_float(vc.z)) - p0);
                const float cos_y = fabsf(dot(ng, d)) / sqrtf(sqnorm(ng));
                const float p_light = light_pdf_area(a, mat) * t * t / cos_y;
                const float p_bsdf = a.prev_pdf[path];
                if (cos_y > 0.f) C = C + cmul(me.emissive, T) * (p_bsdf / (p_bsdf + p_light));
            } else {
                C = C + cmul(me.emissive, T);
            }
        } else if (cache_vertex && (!(pf & PF_TRAIN) ||
                                    (!(pf & PF_TRAIN_FULL) && bounce >= a.cache_bounce + a.train_suffix))) {
            // terminate into the cache: radiance leaving this vertex comes from the network (a
            // training path past its suffix too: its targets then include the network's prediction
            // here, which carries light over more bounces than the suffix traces)
            cache_q = true;
            record_cache_query(a, path, loc, T, make_float4(normal.x, normal.y, normal.z, me.base.x),
                               make_float4(me.base.y, me.base.z, me.roughness, 1.f));
        } else {
            if (cache_vertex) {  // a training path: record the vertex, its target follows
                train = true;
                record_training(a, path, loc, C, T, make_float4(normal.x, normal.y, normal.z, me.base.x),
                                make_float4(me.base.y, me.base.z, me.roughness, 1.f));
            }
            uint4 rs = a.rng[path];
            Rng rng{rs.x, rs.y, rs.z, rs.w};
            F3 wi;
            bool diffuse = false;
            const F3 f = sample_surface(rng, mat, me.base, me.metallic, me.roughness, d, normal, tangent, wi,
                                        diffuse);
            bool done = false;
            if (first_hit_debug(a.debug_mode) && bounce == 0u) {
                switch (a.debug_mode) {
                    case 1: C = f3(0.f, 0.f, 0.f); break;  // Debug: the reference's debug colour (zero)
                    case 3: C = (normal + f3(1.f, 1.f, 1.f)) * 0.5f; break;
                    case 4: C = (wi + f3(1.f, 1.f, 1.f)) * 0.5f; break;
                    case 7: C = f3(1.f, 1.f, 1.f) * me.metallic; break;
                    case 8: C = f3(1.f, 1.f, 1.f) * me.roughness; break;
                    case 9: C = me.base; break;
                    case 10: C = me.emissive; break;
                    case 11: {
                        const spk_camera_desc &cam = scene_of_path(a, path).cam;
                        C = f3(1.f, 1.f, 1.f) * (dot(loc - ld3(cam.position), ld3(cam.front)) / cam.far_plane);
                        break;
                    }
                    default: break;
                }
                done = true;
            }
            if (!done && nee && diffuse && !last) {
                // next-event estimation for the Lambertian lobe: one area-sampled point on an
                // emissive triangle, Lambert BRDF base/pi, two-sided emitters. Not at the last
                // vertex: the BSDF-only estimator never reaches an emitter from there.
                uint4 ns = a.nee_rng[path];
                Rng nr{ns.x, ns.y, ns.z, ns.w};
                const float r0 = rng_unit(nr) * a.emitter_total;
                uint32_t lo = 0, hi = a.num_emitters - 1u;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (a.emitter_cdf[mid] < r0) lo = mid + 1u; else hi = mid;
                }
                const uint32_t et = a.emitters[lo];
                const float w_prev = lo ? a.emitter_cdf[lo - 1u] : 0.f;
                const float w_tri = a.emitter_cdf[lo] - w_prev;
                const float r1 = rng_unit(nr), r2 = rng_unit(nr);
                a.nee_rng[path] = make_uint4(nr.s0, nr.s1, nr.s2, nr.s3);
                const uint4 *tp = a.wtris + (size_t)et * 3u;
                const uint4 va = tp[0], vb = tp[1], vc = tp[2];
                const F3 p0 = f3(__uint_as_float(va.x), __uint_as_float(va.y), __uint_as_float(va.z));
                const F3 p1 = f3(__uint_as_float(vb.x), __uint_as_float(vb.y), __uint_as_float(vb.z));
                const F3 p2 = f3(__uint_as_float(vc.x), __uint_as_float(vc.y), __uint_as_float(vc.z));
                const float su = sqrtf(r1);
                const float b0 = 1.f - su, b1 = r2 * su;
                const F3 y = p0 * b0 + p1 * b1 + p2 * (1.f - b0 - b1);
                const F3 ng = cross(p1 - p0, p2 - p0);
                const float area2 = sqrtf(sqnorm(ng));
                const F3 dl = y - loc;
                const float dist2 = sqnorm(dl);
                const float dist = sqrtf(dist2);
                const F3 wl = dl / dist;
                const float cos_x = dot(normal, wl);
                const float cos_y = fabsf(dot(ng, wl)) / area2;
                const MaterialGpu &em = a.mats[vb.w];  // v1.w carries the material id
                // emitted radiance at the sampled point (uv from its barycentrics)
                const TriShade &es = a.shade[et];
                const float bu = b1, bv = 1.f - b0 - b1;
                const MatEval lme = eval_material(a, em, es.uv0[0] + (es.uv1[0] - es.uv0[0]) * bu +
                                                             (es.uv2[0] - es.uv0[0]) * bv,
                                                  es.uv0[1] + (es.uv1[1] - es.uv0[1]) * bu +
                                                      (es.uv2[1] - es.uv0[1]) * bv);
                if (lme.emits && cos_x > 0.f && cos_y > 0.f && dist > 2e-4f && w_tri > 0.f) {
                    // pdf over area: (w_tri / total) / area; contribution f*cos_x*Le*cos_y/(d^2 pdf),
                    // weighted by the balance heuristic against the cosine-sampled direction
                    const float area = 0.5f * area2;
                    const float p_light = light_pdf_area(a, em) * dist2 / cos_y;
                    const float p_bsdf = cos_x / SPK_PI;
                    const float w_mis = p_light / (p_light + p_bsdf);
                    const float g = cos_x * cos_y / dist2 * area * a.emitter_total / w_tri / SPK_PI * w_mis;
                    const F3 base3 = me.base;
                    const F3 Le = lme.emissive;
                    const F3 contrib = cmul(cmul(T, base3), Le) * g;
                    shadow = true;
                    sh0 = make_float4(loc.x + wl.x * SPK_TOLERANCE, loc.y + wl.y * SPK_TOLERANCE,
                                      loc.z + wl.z * SPK_TOLERANCE, dist - 2.f * SPK_TOLERANCE);
                    sh1 = make_float4(wl.x, wl.y, wl.z, __uint_as_float(path));
                    sh2 = make_float4(contrib.x, contrib.y, contrib.z, 0.f);
                }
            }
            if (!done) {
                T = cmul(T, f);
                if (sqnorm(T) < SPK_EPS) {
                    C4.w = -1.f;
                } else {
                    bool alive = true;
                    if (bounce >= 3u) {
                        float p = fmaxf(fmaxf(T.x, T.y), T.z);
                        p = fminf(fmaxf(p, 0.05f), 1.0f);
                        if (rng_unit(rng) > p) alive = false; else T = T / p;
                    }
                    if (alive && last) end_b = bounce + 1u;  // the loop ran out
                    if (alive && !last) {
                        cont = true;
                        const F3 o2 = loc + wi * SPK_TOLERANCE;
                        const uint32_t npf = (pf & (PF_TRAIN | PF_TRAIN_FULL)) | ((nee && diffuse) ? PF_NEE_PREV : 0u);
                        if (nee && diffuse) a.prev_pdf[path] = fmaxf(dot(wi, normal), 0.f) / SPK_PI;
                        a.rays[2 * path] = make_float4(o2.x, o2.y, o2.z, SPK_EPS);
                        a.rays[2 * path + 1] = make_float4(wi.x, wi.y, wi.z, INFINITY);
                        a.thr[path] = make_float4(T.x, T.y, T.z, __uint_as_float(npf));
                        a.rng[path] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
                    }
                }
            }
            first_done = done;
        }
    }
    if (post_loop_debug(a.debug_mode) && !cont && !cache_q && !first_done) C = debug_after_loop(a.debug_mode, end_b, C);
    a.col[path] = make_float4(C.x, C.y, C.z, C4.w);
}

// cache query / training vertex records of a shaded segment (wave-aggregated over the active lanes)
__global__ __launch_bounds__(kPtBlock) void k_shade(PtArgs a, uint32_t bounce) {
    uint32_t *ctr = a.counters + bounce * kCtrStride;
    const uint32_t count = ctr[CTR_QUEUE];
    const uint32_t *queue = (bounce & 1u) ? a.queue1 : a.queue0;
    uint32_t *next_queue = (bounce & 1u) ? a.queue0 : a.queue1;
    const bool last = bounce + 1u >= a.max_bounce;
    const bool nee = (a.flags & SPK_FRAME_NEE) && a.num_emitters > 0;
    for (uint32_t base = grid_wave() * 64u; base < count; base += grid_waves() * 64u) {
        const uint32_t idx = base + lane_id();
        uint32_t path = 0;
        ShadeOut o;
        o.cont = o.shadow = o.cache_q = o.train = false;
        if (idx < count) {
            const uint32_t slot = a.sort ? a.sorted[idx] : idx;
            path = queue[slot];
            shade_path(a, bounce, last, nee, path, a.hits[slot], o);
        }
        const uint32_t qs = wave_append(a.counters + (bounce + 1u) * kCtrStride + CTR_QUEUE, o.cont);
        if (o.cont) next_queue[qs] = path;
        const uint32_t ss = wave_append(ctr + CTR_SHADOWS, o.shadow);
        if (o.shadow) {
            a.shadow[3 * (size_t)ss] = o.sh0;
            a.shadow[3 * (size_t)ss + 1] = o.sh1;
            a.shadow[3 * (size_t)ss + 2] = o.sh2;
        }
    }
}

// occlusion query of a NEE segment (s0 = origin, tmax; s1 = direction): true when nothing blocks it
template <int WIDTH, int STACK, bool TWO>
__device__ __forceinline__ bool shadow_unoccluded(const PtArgs &a, const float4 s0, const float4 s1, uint2 *stack) {
    RayIn r{s0.x, s0.y, s0.z, SPK_EPS, s1.x, s1.y, s1.z, s0.w};
    BestHit best;
    best.t = r.tmax;
    best.prim = 0xFFFFFFFFu;
    best.tri = 0u;
    best.U = best.V = best.W = 0.f;
    best.det = 1.f;
    best.inst = 0u;
    TraverseCounters c{0u, 0u};
    traverse<WIDTH, STACK, false, true, TWO>(a.nodes, a.tris, 0u, r, stack, kPtBlock, best, c, a.insts);
    if (best.prim == 0xFFFFFFFFu && a.num_spheres) spheres_closest<true>(a.spheres, a.num_spheres, a.sphere_base, r, best);
    return best.prim == 0xFFFFFFFFu;
}

__device__ __forceinline__ void add_color(const PtArgs &a, uint32_t path, const float4 c) {
    float4 C = a.col[path];
    C.x += c.x;
    C.y += c.y;
    C.z += c.z;
    a.col[path] = C;
}

// ---- shadow: occlusion of the NEE segments; unoccluded contributions added to the path colour ----
template <int WIDTH, int STACK, bool TWO>
__global__ __launch_bounds__(kPtBlock) void k_shadow(PtArgs a, uint32_t bounce) {
    __shared__ uint2 stack[STACK * kPtBlock];
    uint32_t *ctr = a.counters + bounce * kCtrStride;
    const uint32_t count = ctr[CTR_SHADOWS];
    for (;;) {
        const uint32_t chunk = wave_fetch(ctr + CTR_SHADOWW, kShadowChunk);
        if (chunk >= count) break;
        for (uint32_t base = chunk; base < min(chunk + kShadowChunk, count); base += 64u) {
            const uint32_t i = base + lane_id();
            if (i >= count) continue;
            const float4 s0 = a.shadow[3 * (size_t)i], s1 = a.shadow[3 * (size_t)i + 1];
            if (shadow_unoccluded<WIDTH, STACK, TWO>(a, s0, s1, stack + threadIdx.x))
                add_color(a, __float_as_uint(s1.w), a.shadow[3 * (size_t)i + 2]);
        }
    }
}

// The same over a full grid, one shadow ray per thread (a bounce holds at most one per path, so at
// most npix): used with the direct extends (schedule refill 0), where the persistent grid of a small
// blocks_per_cu would leave few waves per SIMD for a latency-bound any-hit traversal.
template <int WIDTH, int STACK, bool TWO>
__global__ __launch_bounds__(kPtBlock) void k_shadow_direct(PtArgs a, uint32_t bounce) {
    __shared__ uint2 stack[STACK * kPtBlock];
    const uint32_t count = a.counters[bounce * kCtrStride + CTR_SHADOWS];
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i >= count) return;  // no barrier below
    const float4 s0 = a.shadow[3 * (size_t)i], s1 = a.shadow[3 * (size_t)i + 1];
    if (shadow_unoccluded<WIDTH, STACK, TWO>(a, s0, s1, stack + threadIdx.x))
        add_color(a, __float_as_uint(s1.w), a.shadow[3 * (size_t)i + 2]);
}

// ---- tail: the segments from `from` on, one path per lane until the path ends ----
// Late segments hold few paths with long incoherent rays, and a bounce-synchronous frame waits for
// the slowest ray of every bounce; here a lane carries its path through extend, shade and shadow and
// takes the next queued path when it ends, so the frame waits for the slowest path instead. Every
// segment runs the wavefront kernels' per-path code in the same order (closest hit, shading with
// the colour written back, then the shadow contribution), so images do not depend on `from`.
// One variant per feature set (FEAT, PtFeat bits) a frame uses, with the other branches compiled out:
// the per-path state of NEE and of the cache's records costs registers, and the tail is bound by the
// occupancy they leave (k_tail<8, 12>: 109 VGPRs and 4 waves/SIMD for the reference's estimator,
// 138 and 3 with NEE and the cache).
template <int WIDTH, int STACK, bool TWO, uint32_t FEAT>
__global__ __launch_bounds__(kPtBlock) void k_tail(PtArgs a, uint32_t from) {
    __shared__ uint2 stack[STACK * kPtBlock];
    uint32_t *ctr = a.counters + from * kCtrStride;
    const uint32_t count = ctr[CTR_QUEUE];
    const uint32_t *queue = (from & 1u) ? a.queue1 : a.queue0;
    const bool nee = (FEAT & kFeatNee) && (a.flags & SPK_FRAME_NEE) && a.num_emitters > 0;
    const uint32_t lane = lane_id();
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
    uint32_t path = 0u, bounce = from, rays = 0u, shadows = 0u;
    bool active = false;
    SlotRun run;
    for (;;) {
        if (!run.drained()) {
            const uint64_t idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= a.refill) {
                const uint32_t my = run.take(ctr + CTR_EXTEND, count, nidle, (uint32_t)__popcll(idle & below));
                if (!active) {
                    if (my < count) {
                        path = queue[my];
                        bounce = from;
                        active = true;
                    }
                }
            }
        }
        if (!__any(active)) {
            if (run.drained()) break;
            continue;
        }
        if (active) {
            const float4 ro = a.rays[2 * path], rd = a.rays[2 * path + 1];
            const RayIn r{ro.x, ro.y, ro.z, ro.w, rd.x, rd.y, rd.z, rd.w};
            const uint32_t root = a.num_scenes > 1u ? a.scenes[a.path_scene[path]].root : 0u;
            Traversal<WIDTH, STACK, false, false, TWO> tr;
            tr.init(r, root, a.insts);
            while (tr.step(a.nodes, a.tris, stack + threadIdx.x, kPtBlock)) {
            }
            const uint4 h = hit_record<TWO>(a, r, tr.best);
            ShadeOut o;
            shade_path<FEAT>(a, bounce, bounce + 1u >= a.max_bounce, nee, path, h, o);
            ++rays;
            if (o.shadow) {
                ++shadows;
                if (shadow_unoccluded<WIDTH, STACK, TWO>(a, o.sh0, o.sh1, stack + threadIdx.x)) add_color(a, path, o.sh2);
            }
            if (o.cont) ++bounce; else active = false;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        rays += __shfl_xor(rays, off);
        shadows += __shfl_xor(shadows, off);
    }
    if (lane == 0u) {
        if (rays) atomicAdd(a.counters + kCtrGlobal + GCTR_TAIL_RAYS, rays);
        if (shadows) atomicAdd(a.counters + kCtrGlobal + GCTR_TAIL_SHADOWS, shadows);
    }
}

// ---- cache: add throughput x predicted radiance; build training targets ----
// The network predicts outgoing radiance divided by the vertex albedo (floored), which varies less
// across materials and textures than the radiance itself; the albedo is multiplied back here.
__device__ __forceinline__ F3 demod_albedo(const float4 at0, const float4 at1) {
    return f3(fmaxf(at0.w, SPK_CACHE_ALBEDO_FLOOR), fmaxf(at1.x, SPK_CACHE_ALBEDO_FLOOR),
              fmaxf(at1.y, SPK_CACHE_ALBEDO_FLOOR));
}

__global__ __launch_bounds__(kPtBlock) void k_cache_apply(PtArgs a) {
    const uint32_t count = min(a.counters[kCtrGlobal + GCTR_CACHE_Q], a.npix);
    for (uint32_t base = grid_wave() * 64u; base < count; base += grid_waves() * 64u) {
        const uint32_t i = base + lane_id();
        if (i >= count) continue;
        const float4 m = a.cq_meta[i];
        const uint32_t path = __float_as_uint(m.w);
        const F3 alb = demod_albedo(a.cq_attrs[2 * (size_t)i], a.cq_attrs[2 * (size_t)i + 1]);
        float4 C = a.col[path];
        C.x += m.x * (fmaxf(a.cq_out[3 * (size_t)i], 0.f) * alb.x);
        C.y += m.y * (fmaxf(a.cq_out[3 * (size_t)i + 1], 0.f) * alb.y);
        C.z += m.z * (fmaxf(a.cq_out[3 * (size_t)i + 2], 0.f) * alb.z);
        a.col[path] = C;
    }
}

// Largest training target (radiance after the vertex). Roulette survivors carry weights up to 20, so
// a single sample past a vertex can reach 20x the brightest emitter; clamping them low biases the
// cache dark (at 64, -0.6% on the many-instance scene's mean).
constexpr float kTargetClamp = 1024.f;

// target = radiance gathered after the vertex / throughput at the vertex (clamped to [0, kTargetClamp]), divided
// by the vertex albedo (floored); written with the sample's point and attributes to the staging set
// the next frame's training step reads
__global__ __launch_bounds__(kPtBlock) void k_cache_targets(PtArgs a) {
    const uint32_t count = min(a.counters[kCtrGlobal + GCTR_TRAIN], a.train_cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.st_count = count;
    for (uint32_t base = grid_wave() * 64u; base < count; base += grid_waves() * 64u) {
        const uint32_t i = base + lane_id();
        if (i >= count) continue;
        a.st_points[3 * (size_t)i] = a.tr_points[3 * (size_t)i];
        a.st_points[3 * (size_t)i + 1] = a.tr_points[3 * (size_t)i + 1];
        a.st_points[3 * (size_t)i + 2] = a.tr_points[3 * (size_t)i + 2];
        a.st_attrs[2 * (size_t)i] = a.tr_attrs[2 * (size_t)i];
        a.st_attrs[2 * (size_t)i + 1] = a.tr_attrs[2 * (size_t)i + 1];
        const float4 m0 = a.tr_meta0[i], m1 = a.tr_meta1[i];
        const float4 C = a.col[__float_as_uint(m0.w)];
        const float g[3] = {C.x - m0.x, C.y - m0.y, C.z - m0.z};
        const float tt[3] = {m1.x, m1.y, m1.z};
        const F3 alb = demod_albedo(a.tr_attrs[2 * (size_t)i], a.tr_attrs[2 * (size_t)i + 1]);
        const float al[3] = {alb.x, alb.y, alb.z};
#pragma unroll
        for (int k = 0; k < 3; ++k)
            a.st_targets[3 * (size_t)i + k] = fminf(fmaxf(g[k] / fmaxf(tt[k], 1e-4f), 0.f), kTargetClamp) / al[k];
    }
}

// ---- finalize: per-sample clamp and the valid flag into the G-buffer (CPURenderer.cpp:408-411) ----
__global__ __launch_bounds__(kPtBlock) void k_finalize(PtArgs a) {
    const uint32_t stride = gridDim.x * kPtBlock;
    for (uint32_t p = blockIdx.x * kPtBlock + threadIdx.x; p < a.npix; p += stride) {
        float4 C = a.col[p];
        if (a.max_bounce == 0u && post_loop_debug(a.debug_mode)) {  // the loop never ran: bounce 0
            const F3 c = debug_after_loop(a.debug_mode, 0u, f3(C.x, C.y, C.z));
            C.x = c.x, C.y = c.y, C.z = c.z;
        }
        float4 o;
        o.x = fminf(C.x, a.output_limit);
        o.y = fminf(C.y, a.output_limit);
        o.z = fminf(C.z, a.output_limit);
        o.w = C.w;
        reinterpret_cast<float4 *>(a.rgba_out)[p] = o;
    }
}

// ---- accumulation and tone mapping (CPURenderer.cpp:527-576) ----
__global__ void k_accumulate(const float4 *__restrict__ rgba, float4 *__restrict__ acc, uint64_t n, float ma) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = rgba[i];
    float4 y = acc[i];
    y.x = x.x + (y.x - x.x) * ma;
    y.y = x.y + (y.y - x.y) * ma;
    y.z = x.z + (y.z - x.z) * ma;
    y.w = x.w + (y.w - x.w) * ma;
    acc[i] = y;
}

__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    const float v = (x * (x * a + b)) / (x * (x * c + d) + e);
    return fminf(fmaxf(v, 0.f), 1.f);
}

__global__ void k_tonemap(const float4 *__restrict__ acc, float4 *__restrict__ out, uint32_t w, uint32_t h,
                          float exposure) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t j = i / w, x = i - j * w;
    const float4 c = acc[i];
    out[(size_t)(h - 1u - j) * w + x] = make_float4(aces(c.x * exposure), aces(c.y * exposure), aces(c.z * exposure),
                                                    1.f);
}

// ---- spatial fill of invalid pixels (CPURenderer.cpp:452-504) as a gather ----
// The reference scatters: each valid, non-sky source copies its colour into the first of its 8
// neighbours (direction order below) that is invalid with a normal within dot >= 0.99. Here each
// invalid pixel looks at its 8 possible sources and takes the last one in row-major order whose
// first such neighbour it is, which is the single-threaded order of the reference's loop.
__constant__ int kFillDir[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
// source offsets (source = pixel - direction) in decreasing row-major order of the source
__constant__ int kFillSrc[8][2] = {{1, 1}, {0, 1}, {-1, 1}, {1, 0}, {-1, 0}, {1, -1}, {0, -1}, {-1, -1}};

__device__ __forceinline__ float fill_dot(const float *a, const float *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__global__ void k_spatial_fill(const float4 *__restrict__ rgba, const float *__restrict__ normal,
                               float4 *__restrict__ out, uint32_t w, uint32_t h) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint64_t)w * h) return;
    const float4 c = rgba[p];
    if (c.w > 0.f) {
        out[p] = c;
        return;
    }
    const int j = (int)(p / w), i = (int)(p - (uint64_t)j * w);
    for (int k = 0; k < 8; ++k) {
        const int si = i + kFillSrc[k][0], sj = j + kFillSrc[k][1];
        if (si < 0 || si >= (int)w || sj < 0 || sj >= (int)h) continue;
        const size_t s = (size_t)sj * w + si;
        const float4 sc = rgba[s];
        const float *n = normal + 3 * s;
        if (!(sc.w > 0.f) || (n[0] == 0.f && n[1] == 0.f && n[2] == 0.f)) continue;
        // the first qualifying neighbour of this source
        for (int d = 0; d < 8; ++d) {
            const int ti = si + kFillDir[d][0], tj = sj + kFillDir[d][1];
            if (ti < 0 || ti >= (int)w || tj < 0 || tj >= (int)h) continue;
            const size_t t = (size_t)tj * w + ti;
            if (rgba[t].w > 0.f) continue;
            if (fill_dot(normal + 3 * t, n) < 0.99f) continue;
            if (t == p) {
                out[p] = make_float4(sc.x, sc.y, sc.z, 1.f);
                return;
            }
            break;
        }
    }
    out[p] = c;
}

// ---------------------------------------------------------------------------------------------
using ExtendFn = void (*)(PtArgs, uint32_t);
struct PtEntry {
    int width, stack;
    bool two;
    ExtendFn extend, shadow, first_sorted, shadow_direct;
    ExtendFn first;  // without the material histogram
    ExtendFn tail[4];  // by feature set (PtFeat bits)
};
#define SPK_PT(W, S, T)                                                                                  \
    {W, S, T, k_extend<W, S, T>, k_shadow<W, S, T>, k_extend_first<W, S, T, true>, k_shadow_direct<W, S, T>, \
     k_extend_first<W, S, T, false>,                                                                    \
     {k_tail<W, S, T, 0u>, k_tail<W, S, T, 1u>, k_tail<W, S, T, 2u>, k_tail<W, S, T, 3u>}}
const PtEntry kPtEntries[] = {
    SPK_PT(8, 12, false), SPK_PT(4, 12, false), SPK_PT(8, 16, false), SPK_PT(8, 24, false),
    SPK_PT(8, 32, false), SPK_PT(8, 64, false), SPK_PT(4, 16, false), SPK_PT(4, 24, false),
    SPK_PT(4, 32, false), SPK_PT(4, 64, false),
    // two-level scenes
    SPK_PT(8, 24, true),  SPK_PT(8, 32, true),  SPK_PT(8, 64, true),  SPK_PT(4, 24, true),
    SPK_PT(4, 32, true),  SPK_PT(4, 64, true),
};
#undef SPK_PT

struct PathState {
    DeviceBuffer rays, thr, col, rng, nee_rng, queue0, queue1, hits, sorted, shadow, counters, hist, path_scene;
    DeviceBuffer prev_pdf;
    DeviceBuffer emitters, emitter_cdf, frame;
    FrameDev *frame_host = nullptr;  // pinned
    uint32_t capacity = 0;
    uint32_t num_emitters = 0;
    float emitter_total = 0.f;
    bool emitters_ready = false;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    PtArgs graph_args{};
    hipStream_t graph_stream = nullptr;
    // the branch of a cache frame's graph that trains on the previous frame's samples beside the
    // frame's first kernels (default priority, like every stream here)
    hipStream_t train_stream = nullptr;
    hipEvent_t train_fork = nullptr, train_join = nullptr;
    ~PathState() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_def) (void)hipGraphDestroy(graph_def);
        if (frame_host) (void)hipHostFree(frame_host);
        if (train_fork) (void)hipEventDestroy(train_fork);
        if (train_join) (void)hipEventDestroy(train_join);
        if (train_stream) (void)hipStreamDestroy(train_stream);
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_def) (void)hipGraphDestroy(graph_def);
        graph = nullptr;
        graph_def = nullptr;
    }
};

// A batch of independent scenes: every scene's BVH rebased into one node array and one triangle
// array (child and triangle offsets shifted, material ids offset by the scene's first material),
// plus one SceneRec per scene.
struct BatchState {
    DeviceScene scene;
    DeviceBuffer table;
    std::vector<SceneRec> host_table;
    std::vector<uint64_t> offsets;  // num_scenes + 1 pixel offsets
    uint32_t num_scenes = 0;
    uint32_t npix = 0;
    spk_schedule schedule{8u, 1u, 16u, 0u};
    uint64_t num_tris = 0;
};

static int device_cus(spk_context *ctx) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
        cus = 256;
    return cus;
}

static int ensure_paths(spk_context *ctx, uint32_t npix) {
    if (!ctx->paths) ctx->paths = new PathState();
    PathState &ps = *ctx->paths;
    hipError_t e;
    if (!ps.frame_host) {
        if ((e = hipHostMalloc((void **)&ps.frame_host, sizeof(FrameDev), hipHostMallocDefault)) != hipSuccess)
            return hip_fail(ctx, e, "pinned frame buffer");
        if ((e = ps.frame.ensure(sizeof(FrameDev))) != hipSuccess) return hip_fail(ctx, e, "frame buffer");
        if ((e = ps.counters.ensure(kCtrTotal * sizeof(uint32_t))) != hipSuccess) return hip_fail(ctx, e, "counters");
        if ((e = ps.hist.ensure((size_t)kMaxBounce * 2 * kMaxSortKeys * sizeof(uint32_t))) != hipSuccess)
            return hip_fail(ctx, e, "histograms");
    }
    if (npix > ps.capacity) {
        ps.drop_graph();
        const size_t n = npix;
        if ((e = ps.rays.ensure(n * 32)) != hipSuccess || (e = ps.thr.ensure(n * 16)) != hipSuccess ||
            (e = ps.col.ensure(n * 16)) != hipSuccess || (e = ps.rng.ensure(n * 16)) != hipSuccess ||
            (e = ps.nee_rng.ensure(n * 16)) != hipSuccess || (e = ps.queue0.ensure(n * 4)) != hipSuccess ||
            (e = ps.queue1.ensure(n * 4)) != hipSuccess || (e = ps.hits.ensure(n * 16)) != hipSuccess ||
            (e = ps.sorted.ensure(n * 4)) != hipSuccess || (e = ps.shadow.ensure(n * 48)) != hipSuccess ||
            (e = ps.path_scene.ensure(n * 4)) != hipSuccess || (e = ps.prev_pdf.ensure(n * 4)) != hipSuccess)
            return hip_fail(ctx, e, "path buffers");
        ps.capacity = npix;
    }
    if (!ps.emitters_ready) {
        // emitter list weighted by area x mean emitted radiance
        const std::vector<uint32_t> &em = ctx->scene.host_emitters;
        std::vector<float> cdf(em.size());
        double acc = 0.0;
        const HostScene &hs = ctx->host_scene;
        for (size_t k = 0; k < em.size(); ++k) {
            // world vertices and material: leaf-order records, or global-id order in a two-level scene
            float v0[3], v1[3], v2[3];
            uint32_t mat;
            if (hs.two_level) {
                const float *v = &hs.verts[9 * (size_t)em[k]];
                for (int c = 0; c < 3; ++c) v0[c] = v[c], v1[c] = v[3 + c], v2[c] = v[6 + c];
                mat = hs.shade_in[em[k]].material;
            } else {
                const TriRecord &r = hs.bvh.tris[em[k]];
                for (int c = 0; c < 3; ++c) v0[c] = r.v0[c], v1[c] = r.v1[c], v2[c] = r.v2[c];
                std::memcpy(&mat, &r.v1[3], 4);
            }
            const double e1[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
            const double e2[3] = {(double)v2[0] - v0[0], (double)v2[1] - v0[1], (double)v2[2] - v0[2]};
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                         cz = e1[0] * e2[1] - e1[1] * e2[0];
            const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            const MaterialGpu &m = ctx->host_mats[mat];
            const double lum = (m.emissive[0] + m.emissive[1] + m.emissive[2]) / 3.0;
            acc += area * lum;
            cdf[k] = (float)acc;
        }
        ps.num_emitters = (uint32_t)em.size();
        ps.emitter_total = em.empty() ? 0.f : cdf.back();
        if (!em.empty()) {
            if ((e = ps.emitters.ensure(em.size() * 4)) != hipSuccess ||
                (e = ps.emitter_cdf.ensure(em.size() * 4)) != hipSuccess)
                return hip_fail(ctx, e, "emitter buffers");
            if ((e = hipMemcpy(ps.emitters.ptr, em.data(), em.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(ps.emitter_cdf.ptr, cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
                return hip_fail(ctx, e, "emitter upload");
        }
        ps.emitters_ready = true;
    }
    if (ctx->cache) {
        CacheState &cs = *ctx->cache;
        if (cs.capacity < npix) {
            const size_t n = npix, tcap = cs.cfg.batch;
            if ((e = cs.cq_points.ensure(n * 12)) != hipSuccess || (e = cs.cq_meta.ensure(n * 16)) != hipSuccess ||
                (e = cs.cq_attrs.ensure(n * 32)) != hipSuccess || (e = cs.tr_attrs.ensure(tcap * 32)) != hipSuccess ||
                (e = cs.st_points.ensure(tcap * 12)) != hipSuccess || (e = cs.st_attrs.ensure(tcap * 32)) != hipSuccess ||
                (e = cs.st_targets.ensure(tcap * 12)) != hipSuccess || (e = cache_prepare(cs, (uint32_t)tcap)) != hipSuccess ||
                (e = cs.cq_out.ensure(n * 12)) != hipSuccess || (e = cs.tr_points.ensure(tcap * 12)) != hipSuccess ||
                (e = cs.tr_meta0.ensure(tcap * 16)) != hipSuccess || (e = cs.tr_meta1.ensure(tcap * 16)) != hipSuccess ||
                (e = cs.tr_targets.ensure(tcap * 12)) != hipSuccess)
                return hip_fail(ctx, e, "cache frame buffers");
            cs.capacity = npix;
            ps.drop_graph();
        }
        if (!ps.train_stream) {
            if ((e = hipStreamCreateWithFlags(&ps.train_stream, hipStreamNonBlocking)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&ps.train_fork, hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&ps.train_join, hipEventDisableTiming)) != hipSuccess)
                return hip_fail(ctx, e, "cache training stream");
        }
    }
    return SPK_OK;
}

// One frame as a sequence of launches on s (captured into the frame graph). A cache frame forks a
// branch onto the training stream after k_generate that takes one training step on the previous
// frame's samples (staged by k_cache_targets; frames with fewer than kMinTrainSamples take none)
// beside the bounce-0 traversal, and joins it before the frame's cache queries, which then see the
// updated network, as they would with the step run between the frames.
static hipError_t enqueue_frame(const PtArgs &a, const PtEntry &pe, hipStream_t s, int grid, CacheState *cache,
                                const PathState &ps) {
    hipError_t e;
    auto fork_training = [&]() -> hipError_t {
        if (!a.cache_on) return hipSuccess;
        CacheState &cs = *cache;
        hipError_t f;
        if ((f = hipEventRecord(ps.train_fork, s)) != hipSuccess ||
            (f = hipStreamWaitEvent(ps.train_stream, ps.train_fork, 0)) != hipSuccess ||
            (f = cache_gradients(cs, a.st_points, (const float *)a.st_attrs, a.st_targets, cs.cfg.batch, 1, nullptr,
                                 ps.train_stream, a.st_count)) != hipSuccess ||
            (f = cache_adam(cs, ps.train_stream, a.st_count, cs.cfg.batch)) != hipSuccess ||
            (f = hipEventRecord(ps.train_join, ps.train_stream)) != hipSuccess)
            return f;
        return hipSuccess;
    };
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kPtBlock), 0, s, a);
    // forked after the camera rays: forked before them, the training kernels took the CUs first and
    // k_generate ran 5x longer (1080p cache frame: 2.86-2.89 ms; forked here 2.80 ms; forked before
    // the per-path tail, where the training fills the CUs the tail's finished waves leave, 2.81 ms)
    if ((e = fork_training()) != hipSuccess) return e;
    for (uint32_t b = 0; b < a.tail_from; ++b) {
        if (b == 0u || a.direct)  // a full grid, one queue slot per thread
            hipLaunchKernelGGL(a.sort ? pe.first_sorted : pe.first, dim3((a.npix + kPtBlock - 1) / kPtBlock),
                               dim3(kPtBlock), 0, s, a, b);
        else
            hipLaunchKernelGGL(pe.extend, dim3(grid), dim3(kPtBlock), 0, s, a, b);
        if (a.sort) {
            hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, a, b);
            hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(kPtBlock), 0, s, a, b);
        }
        hipLaunchKernelGGL(k_shade, dim3(grid), dim3(kPtBlock), 0, s, a, b);
        if ((a.flags & SPK_FRAME_NEE) && a.num_emitters) {
            if (a.direct)
                hipLaunchKernelGGL(pe.shadow_direct, dim3((a.npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, s,
                                   a, b);
            else
                hipLaunchKernelGGL(pe.shadow, dim3(grid), dim3(kPtBlock), 0, s, a, b);
        }
    }
    if (a.tail_from < a.max_bounce) {
        const uint32_t feat = (((a.flags & SPK_FRAME_NEE) && a.num_emitters > 0) ? kFeatNee : 0u) |
                              (a.cache_on ? kFeatCache : 0u);
        hipLaunchKernelGGL(pe.tail[feat], dim3(grid), dim3(kPtBlock), 0, s, a, a.tail_from);
    }
    if (a.cache_on) {
        if ((e = hipStreamWaitEvent(s, ps.train_join, 0)) != hipSuccess) return e;
        // one 64-query tile per block over the largest possible count (one query per path), blocks past
        // the frame's count exit at once: the gather-bound kernel gets every wave the chip holds (a
        // persistent grid of 8 one-wave blocks per CU, 2 waves/SIMD, made cache frames 0.6-2.4% longer)
        e = cache_infer_queue(*cache, a.cq_points, (const float *)a.cq_attrs,
                                         a.counters + kCtrGlobal + GCTR_CACHE_Q, a.npix, a.cq_out,
                                         (int)((a.npix + 63u) / 64u), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_cache_apply, dim3(grid), dim3(kPtBlock), 0, s, a);
        hipLaunchKernelGGL(k_cache_targets, dim3(grid), dim3(kPtBlock), 0, s, a);
    }
    hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(kPtBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace spk

using namespace spk;

void spk_context::on_scene_changed() {
    if (paths) {
        paths->drop_graph();
        paths->emitters_ready = false;
    }
}

void spk_context::drop_batch() {
    if (paths) paths->drop_graph();
    delete batch;
    batch = nullptr;
}

void spk_context::release_all() {
    delete paths;
    paths = nullptr;
    delete cache;
    cache = nullptr;
    delete batch;
    batch = nullptr;
    spk::destroy_host_staging(staging);
    staging = nullptr;
    scene.nodes.release();
    scene.tris.release();
    scene.shade.release();
    scene.materials.release();
    scene.emitters.release();
    scene.valid = false;
    texels.release();
    tex_table.release();
    num_textures = 0;
    for (hipEvent_t ev : kernel_events) (void)hipEventDestroy(ev);
    kernel_events.clear();
    kernel_timed = 0;
}

extern "C" {

int spk_set_sky_cubemap(spk_context *ctx, const uint32_t faces[6]) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!faces) {
        for (int f = 0; f < 6; ++f) ctx->sky_faces[f] = -1;
        return SPK_OK;
    }
    const bool none = faces[0] == 0;
    for (int f = 0; f < 6; ++f) {
        if ((faces[f] == 0) != none)
            return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "either all six faces or none");
        if (faces[f] > ctx->num_textures)
            return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "face %d names texture %u of %u", f, faces[f] - 1,
                             ctx->num_textures);
    }
    for (int f = 0; f < 6; ++f) ctx->sky_faces[f] = (int32_t)faces[f] - 1;
    return SPK_OK;
}

int spk_set_camera(spk_context *ctx, const spk_camera_desc *camera) {
    if (!ctx || !camera) return SPK_ERR_INVALID_ARGUMENT;
    for (float v : {camera->position[0], camera->position[1], camera->position[2], camera->lens_radius})
        if (!std::isfinite(v)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "camera has a non-finite value");
    ctx->camera = *camera;
    ctx->has_camera = true;
    return SPK_OK;
}

}  // extern "C"

namespace spk {

// What one frame traces: a device scene and its scene records.
struct FrameSource {
    const DeviceScene *ds;
    const SceneRec *scenes_dev;  // null: the single record in the frame buffer (filled from single)
    SceneRec single;
    uint32_t num_scenes;
    uint32_t npix;
    uint32_t blocks_per_cu;
    uint32_t sort;
    uint32_t refill;
    uint32_t tail;  // first segment run by k_tail (0: none)
};

static int run_frame(spk_context *ctx, const spk_frame_desc *frame, const FrameSource &src, float *rgba_dev,
                     float *normal_dev, spk_frame_stats *stats) {
    if (frame->max_bounce > (uint32_t)kMaxBounce)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "max_bounce above %d", kMaxBounce);
    if (frame->debug_mode > 11u)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "unknown debug mode %u", frame->debug_mode);
    // the modes that rewrite the colour after the loop see the reference's estimator only
    if ((frame->debug_mode == 2u || frame->debug_mode == 5u || frame->debug_mode == 6u) &&
        (frame->flags & (SPK_FRAME_NEE | SPK_FRAME_CACHE)))
        return set_error(ctx, SPK_ERR_UNSUPPORTED, "debug mode %u takes neither NEE nor the cache", frame->debug_mode);
    const bool cache_on = (frame->flags & SPK_FRAME_CACHE) != 0;
    if (cache_on && !ctx->cache)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "SPK_FRAME_CACHE needs spk_cache_create first");
    const uint32_t npix = src.npix;
    int rc = ensure_paths(ctx, npix);
    if (rc != SPK_OK) return rc;
    PathState &ps = *ctx->paths;
    const DeviceScene &ds = *src.ds;
    const PtEntry *pe = nullptr;
    for (const PtEntry &e : kPtEntries)
        if (e.width == ds.width && e.two == ds.two_level && e.stack == pick_trace_stack(ds.depth, ds.two_level))
            pe = &e;
    if (!pe) return set_error(ctx, SPK_ERR_STACK_DEPTH, "no path-tracer kernel for depth %d", ds.depth);

    PtArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nodes = (const uint4 *)ds.nodes.ptr;
    a.tris = (const uint4 *)ds.tris.ptr;
    a.wtris = ds.two_level ? (const uint4 *)ds.wtris.ptr : a.tris;
    a.insts = ds.two_level ? (const uint4 *)ds.insts.ptr : nullptr;
    a.shade = (const TriShade *)ds.shade.ptr;
    a.mats = (const MaterialGpu *)ds.materials.ptr;
    a.tex_table = (const TexRecord *)ctx->tex_table.ptr;
    a.spheres = (const float4 *)ds.spheres.ptr;
    a.num_spheres = ds.spheres.ptr ? ds.num_spheres : 0u;
    a.sphere_base = ds.num_tris;
    a.texels = (const float4 *)ctx->texels.ptr;
    a.emitters = (const uint32_t *)ps.emitters.ptr;
    a.emitter_cdf = (const float *)ps.emitter_cdf.ptr;
    a.num_emitters = ps.num_emitters;
    a.emitter_total = ps.emitter_total;
    a.rays = (float4 *)ps.rays.ptr;
    a.thr = (float4 *)ps.thr.ptr;
    a.prev_pdf = (float *)ps.prev_pdf.ptr;
    a.col = (float4 *)ps.col.ptr;
    a.rng = (uint4 *)ps.rng.ptr;
    a.nee_rng = (uint4 *)ps.nee_rng.ptr;
    a.queue0 = (uint32_t *)ps.queue0.ptr;
    a.queue1 = (uint32_t *)ps.queue1.ptr;
    a.hits = (uint4 *)ps.hits.ptr;
    a.sorted = (uint32_t *)ps.sorted.ptr;
    a.shadow = (float4 *)ps.shadow.ptr;
    a.counters = (uint32_t *)ps.counters.ptr;
    a.hist = (uint32_t *)ps.hist.ptr;
    a.nkeys = ds.num_materials + 1u;
    a.sort = src.sort && a.nkeys <= (uint32_t)kMaxSortKeys && !(frame->flags & SPK_FRAME_NO_SORT) ? 1u : 0u;
    if (!a.sort) a.nkeys = 2u;
    a.rgba_out = rgba_dev;
    a.normal_out = normal_dev;
    a.frame = (const FrameDev *)ps.frame.ptr;
    a.scenes = src.scenes_dev ? src.scenes_dev
                              : (const SceneRec *)((const char *)ps.frame.ptr + offsetof(FrameDev, scene0));
    a.path_scene = (uint32_t *)ps.path_scene.ptr;
    a.num_scenes = src.num_scenes;
    const int cus = device_cus(ctx);
    const int grid = cus * (int)src.blocks_per_cu;
    a.grid = (uint32_t)grid;
    // refill 0: direct extends; the per-path tail then refills at 8 idle lanes (1080p frames of the
    // many-instance scene against 16: plain 2.59-2.65 ms against 2.63-2.69, cache + NEE 3.32-3.33
    // against 3.35-3.36, the 64-scene batch 1419 against 1414 M samples/s; cache from bounce 1
    // 2.45-2.47 against 2.44-2.45; 4, 6 and 12 within noise of 8, 32 and 48 slower; measured with
    // 128 queue slots per atomic, with 16 the thresholds 4, 8 and 16 are within noise of each other)
    a.refill = src.refill ? std::min(src.refill, 64u) : 8u;
    a.direct = src.refill == 0u ? 1u : 0u;
    a.npix = npix;
    a.max_bounce = frame->max_bounce;
    a.tail_from = src.tail && src.tail < frame->max_bounce ? src.tail : frame->max_bounce;
    a.debug_mode = frame->debug_mode;
    a.sky_enabled = frame->sky_enabled;
    for (int f = 0; f < 6; ++f) a.sky_faces[f] = ctx->sky_faces[f];
    a.flags = frame->flags & ~SPK_FRAME_NO_SORT;
    for (int c = 0; c < 3; ++c) a.sky[c] = frame->sky_color[c];
    a.output_limit = frame->output_limit > 0.f ? frame->output_limit : SPK_OUTPUT_LIMIT;
    if (cache_on) {
        CacheState &cs = *ctx->cache;
        a.cache_on = 1u;
        a.cache_bounce = cs.cfg.terminate_bounce;
        a.train_cap = cs.cfg.batch;
        a.train_suffix = cs.cfg.train_suffix ? cs.cfg.train_suffix : 0xFFFFu;  // 0: traced to the end
        a.train_prob = cs.cfg.train_fraction > 0.f ? cs.cfg.train_fraction
                                                   : std::min(1.f, (float)cs.cfg.batch / (float)npix / 2.f);
        a.cq_points = (float *)cs.cq_points.ptr;
        a.cq_meta = (float4 *)cs.cq_meta.ptr;
        a.cq_attrs = (float4 *)cs.cq_attrs.ptr;
        a.tr_attrs = (float4 *)cs.tr_attrs.ptr;
        a.cq_out = (float *)cs.cq_out.ptr;
        a.tr_points = (float *)cs.tr_points.ptr;
        a.tr_meta0 = (float4 *)cs.tr_meta0.ptr;
        a.tr_meta1 = (float4 *)cs.tr_meta1.ptr;
        a.tr_targets = (float *)cs.tr_targets.ptr;
        a.st_points = (float *)cs.st_points.ptr;
        a.st_attrs = (float4 *)cs.st_attrs.ptr;
        a.st_targets = (float *)cs.st_targets.ptr;
        a.st_count = (uint32_t *)cs.st_count.ptr;
    }

    hipError_t e;
    ps.frame_host->frame_seed = frame->frame_seed;
    ps.frame_host->scene0 = src.single;
    if ((e = hipMemcpyAsync(ps.frame.ptr, ps.frame_host, sizeof(FrameDev), hipMemcpyHostToDevice, ctx->stream)) !=
        hipSuccess)
        return hip_fail(ctx, e, "frame parameters");
    const bool reuse = ps.graph && ps.graph_stream == ctx->stream && std::memcmp(&ps.graph_args, &a, sizeof(a)) == 0;
    if (!reuse && frame->max_bounce > 0) {
        ps.drop_graph();
        if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(ctx, e, "begin capture");
        hipError_t le = enqueue_frame(a, *pe, ctx->stream, grid, ctx->cache, ps);
        e = hipStreamEndCapture(ctx->stream, &ps.graph_def);
        if (le != hipSuccess) return hip_fail(ctx, le, "frame launch during capture");
        if (e != hipSuccess) return hip_fail(ctx, e, "end capture");
        if ((e = hipGraphInstantiate(&ps.graph, ps.graph_def, nullptr, nullptr, 0)) != hipSuccess)
            return hip_fail(ctx, e, "graph instantiate");
        ps.graph_args = a;
        ps.graph_stream = ctx->stream;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (stats) {
        (void)hipEventCreate(&t0);
        (void)hipEventCreate(&t1);
        (void)hipEventRecord(t0, ctx->stream);
    }
    if (frame->max_bounce == 0) {
        // no segment is traced: the G-buffer is the untouched path state
        hipLaunchKernelGGL(k_generate, dim3(grid), dim3(kPtBlock), 0, ctx->stream, a);
        hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(kPtBlock), 0, ctx->stream, a);
        e = hipGetLastError();
    } else {
        e = hipGraphLaunch(ps.graph, ctx->stream);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "frame launch");
    // the online training of the radiance cache runs inside the graph (on the previous frame's
    // samples); its loss is read back below
    const bool train = cache_on && frame->max_bounce > 0;
    if (stats) (void)hipEventRecord(t1, ctx->stream);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "frame");
    if (train) {
        AdamStep st{};
        if ((e = hipMemcpy(&st, ctx->cache->state.ptr, sizeof(st), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(ctx, e, "cache state");
        if (st.active) ctx->cache_last_loss = st.loss;
    }
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        std::vector<uint32_t> c(kCtrTotal);
        if ((e = hipMemcpy(c.data(), ps.counters.ptr, c.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(ctx, e, "frame counters");
        uint64_t rays = c[kCtrGlobal + GCTR_TAIL_RAYS], shadows = c[kCtrGlobal + GCTR_TAIL_SHADOWS];
        for (uint32_t b = 0; b < a.tail_from; ++b) {
            rays += c[b * kCtrStride + CTR_QUEUE];
            shadows += c[b * kCtrStride + CTR_SHADOWS];
        }
        stats->rays_traced = rays;
        stats->shadow_rays = shadows;
        stats->milliseconds = ms;
    }
    return SPK_OK;
}

// Rebases one scene's BVH into the batch arrays: child indices by node_off, triangle indices by
// tri_off, material ids by mat_off.
template <class Node>
static void append_nodes(const std::vector<uint8_t> &src, uint32_t count, uint32_t node_off, uint32_t tri_off,
                         std::vector<uint8_t> &dst) {
    const size_t at = dst.size();
    dst.resize(at + (size_t)count * sizeof(Node));
    for (uint32_t k = 0; k < count; ++k) {
        Node nd;
        std::memcpy(&nd, src.data() + (size_t)k * sizeof(Node), sizeof(Node));
        nd.child_base += node_off;
        nd.tri_base += tri_off;
        std::memcpy(dst.data() + at + (size_t)k * sizeof(Node), &nd, sizeof(Node));
    }
}

}  // namespace spk

extern "C" {

int spk_render_frame(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, float *normal_dev,
                     spk_frame_stats *stats) {
    if (!ctx || !frame) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (!ctx->has_camera) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "no camera set");
    if (!rgba_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null output");
    if (frame->width < 2 || frame->height < 2)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "frame must be at least 2x2 (pixel size 1/(W-1))");
    if ((uint64_t)frame->width * frame->height >= (1ull << 31))
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "frame too large");
    FrameSource src;
    std::memset(&src, 0, sizeof(src));
    src.ds = &ctx->scene;
    src.scenes_dev = nullptr;
    src.single.root = 0u;
    src.single.width = frame->width;
    src.single.height = frame->height;
    src.single.pix_base = 0u;
    src.single.cam = ctx->camera;
    src.num_scenes = 1u;
    src.npix = frame->width * frame->height;
    src.blocks_per_cu = ctx->schedule.blocks_per_cu;
    src.sort = ctx->schedule.material_sort;
    src.refill = ctx->schedule.refill;
    src.tail = ctx->schedule.tail_bounce;
    return run_frame(ctx, frame, src, rgba_dev, normal_dev, stats);
}

int spk_batch_upload(spk_context *ctx, const spk_batch_scene *scenes, uint32_t num_scenes,
                     const spk_build_options *options) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!scenes || num_scenes == 0) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "batch needs at least one scene");
    if (options && options->two_level)
        return set_error(ctx, SPK_ERR_UNSUPPORTED, "batches are flattened (two_level is for spk_scene_upload)");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx->batch;
    ctx->batch = nullptr;
    if (ctx->paths) ctx->paths->drop_graph();
    auto bs = std::make_unique<BatchState>();
    std::vector<uint8_t> nodes;
    std::vector<TriRecord> tris;
    std::vector<TriShade> shade;
    std::vector<MaterialGpu> mats;
    uint64_t pix = 0;
    int width = 0, depth = 0;
    uint32_t num_nodes = 0;
    for (uint32_t s = 0; s < num_scenes; ++s) {
        const spk_batch_scene &sc = scenes[s];
        if (sc.width < 2 || sc.height < 2)
            return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "scene %u: image must be at least 2x2", s);
        for (float v : {sc.camera.position[0], sc.camera.position[1], sc.camera.position[2], sc.camera.lens_radius})
            if (!std::isfinite(v)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "scene %u: non-finite camera", s);
        HostScene hs;
        std::vector<MaterialGpu> sm;
        std::string err;
        int rc = prepare_scene(sc.meshes, sc.num_meshes, sc.instances, sc.num_instances, sc.materials,
                               sc.num_materials, ctx->num_textures, options, hs, sm, err);
        if (rc != SPK_OK) return set_error(ctx, rc, "scene %u: %s", s, err.c_str());
        if (s == 0) width = (int)hs.input.width;
        depth = std::max(depth, (int)hs.bvh.depth);
        const uint64_t node_off = num_nodes, tri_off = tris.size(), mat_off = mats.size();
        if (node_off + hs.bvh.num_nodes >= (1ull << 32) || tri_off + hs.bvh.tris.size() >= (1ull << 32) ||
            mat_off + sm.size() >= (1ull << 31))
            return set_error(ctx, SPK_ERR_INVALID_SCENE, "batch exceeds 2^32 nodes or triangles");
        if (width == 8)
            append_nodes<NodeW8>(hs.bvh.nodes, hs.bvh.num_nodes, (uint32_t)node_off, (uint32_t)tri_off, nodes);
        else
            append_nodes<NodeW4>(hs.bvh.nodes, hs.bvh.num_nodes, (uint32_t)node_off, (uint32_t)tri_off, nodes);
        num_nodes += hs.bvh.num_nodes;
        for (const TriRecord &t : hs.bvh.tris) {
            TriRecord r = t;
            uint32_t m;
            std::memcpy(&m, &r.v1[3], 4);
            m += (uint32_t)mat_off;
            std::memcpy(&r.v1[3], &m, 4);
            tris.push_back(r);
        }
        for (TriShade sh : leaf_order_shade(hs)) {
            sh.material += (uint32_t)mat_off;
            shade.push_back(sh);
        }
        mats.insert(mats.end(), sm.begin(), sm.end());
        SceneRec rec;
        std::memset(&rec, 0, sizeof(rec));
        rec.root = (uint32_t)node_off;
        rec.width = sc.width;
        rec.height = sc.height;
        rec.pix_base = (uint32_t)pix;
        rec.cam = sc.camera;
        bs->host_table.push_back(rec);
        bs->offsets.push_back(pix);
        pix += (uint64_t)sc.width * sc.height;
        if (pix >= (1ull << 31)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "batch has 2^31 or more pixels");
    }
    bs->offsets.push_back(pix);
    if (pick_trace_stack(depth) < 0)
        return set_error(ctx, SPK_ERR_STACK_DEPTH, "BVH depth %d exceeds the largest traversal stack", depth);
    DeviceScene &ds = bs->scene;
    hipError_t e;
    if ((e = ds.nodes.ensure(nodes.size())) != hipSuccess ||
        (e = ds.tris.ensure(tris.size() * sizeof(TriRecord))) != hipSuccess ||
        (e = ds.shade.ensure(shade.size() * sizeof(TriShade))) != hipSuccess ||
        (e = ds.materials.ensure(mats.size() * sizeof(MaterialGpu))) != hipSuccess ||
        (e = bs->table.ensure(bs->host_table.size() * sizeof(SceneRec))) != hipSuccess)
        return hip_fail(ctx, e, "batch allocation");
    if ((e = hipMemcpy(ds.nodes.ptr, nodes.data(), nodes.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds.tris.ptr, tris.data(), tris.size() * sizeof(TriRecord), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.shade.ptr, shade.data(), shade.size() * sizeof(TriShade), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.materials.ptr, mats.data(), mats.size() * sizeof(MaterialGpu), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(bs->table.ptr, bs->host_table.data(), bs->host_table.size() * sizeof(SceneRec),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(ctx, e, "batch upload");
    ds.num_tris = (uint32_t)tris.size();
    ds.num_nodes = num_nodes;
    ds.num_materials = (uint32_t)mats.size();
    ds.width = width;
    ds.depth = depth;
    ds.variant = TraceVariant{width, pick_trace_stack(depth), 128};
    ds.valid = true;
    bs->num_scenes = num_scenes;
    bs->npix = (uint32_t)pix;
    bs->num_tris = tris.size();
    ctx->batch = bs.release();
    return SPK_OK;
}

int spk_batch_offsets(spk_context *ctx, uint64_t *offsets, uint32_t max_count) {
    if (!ctx || !offsets) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->batch) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    if (max_count < ctx->batch->num_scenes + 1u)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "offsets need %u entries", ctx->batch->num_scenes + 1u);
    std::copy(ctx->batch->offsets.begin(), ctx->batch->offsets.end(), offsets);
    return SPK_OK;
}

int spk_batch_render(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_frame_stats *stats) {
    if (!ctx || !frame) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->batch) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    if (!rgba_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null output");
    if (frame->flags & (SPK_FRAME_NEE | SPK_FRAME_CACHE))
        return set_error(ctx, SPK_ERR_UNSUPPORTED, "NEE and the radiance cache are single-scene features");
    const BatchState &bs = *ctx->batch;
    FrameSource src;
    std::memset(&src, 0, sizeof(src));
    src.ds = &bs.scene;
    src.scenes_dev = (const SceneRec *)bs.table.ptr;
    src.num_scenes = bs.num_scenes;
    src.npix = bs.npix;
    src.blocks_per_cu = bs.schedule.blocks_per_cu;
    src.sort = bs.schedule.material_sort;
    src.refill = bs.schedule.refill;
    src.tail = bs.schedule.tail_bounce;
    return run_frame(ctx, frame, src, rgba_dev, nullptr, stats);
}

static int check_schedule(spk_context *ctx, const spk_schedule &sc) {
    if (sc.blocks_per_cu < 1 || sc.blocks_per_cu > 32 || sc.material_sort > 1 || sc.refill > 64 ||
        sc.tail_bounce > (uint32_t)kMaxBounce)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT,
                         "schedule out of range (blocks 1..32, sort 0/1, refill 0..64, tail 0..max bounces)");
    return SPK_OK;
}

static spk_schedule *schedule_of(spk_context *ctx, int batch) {
    if (!batch) return &ctx->schedule;
    return ctx->batch ? &ctx->batch->schedule : nullptr;
}

int spk_get_schedule(spk_context *ctx, int batch, spk_schedule *out) {
    if (!ctx || !out) return SPK_ERR_INVALID_ARGUMENT;
    spk_schedule *sc = schedule_of(ctx, batch);
    if (!sc) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    *out = *sc;
    return SPK_OK;
}

int spk_set_schedule(spk_context *ctx, int batch, const spk_schedule *schedule) {
    if (!ctx || !schedule) return SPK_ERR_INVALID_ARGUMENT;
    spk_schedule *sc = schedule_of(ctx, batch);
    if (!sc) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    int rc = check_schedule(ctx, *schedule);
    if (rc != SPK_OK) return rc;
    *sc = *schedule;
    return SPK_OK;
}

// Times every candidate schedule on one frame (a capture frame first, then a timed one).
// Frame time of the current schedule: one frame to capture its graph, then the median of five
// (single frames vary by a few percent, as much as neighbouring schedules differ; the fastest of a
// few frames rewarded lucky frames and picked schedules slower than the default on the median).
static int time_schedule(spk_context *ctx, int batch, const spk_frame_desc *frame, float *rgba_dev, double &ms) {
    double t[5];
    for (int k = 0; k < 6; ++k) {
        spk_frame_stats st{};
        const int rc = batch ? spk_batch_render(ctx, frame, rgba_dev, k ? &st : nullptr)
                             : spk_render_frame(ctx, frame, rgba_dev, nullptr, k ? &st : nullptr);
        if (rc != SPK_OK) return rc;
        if (k) t[k - 1] = st.milliseconds;
    }
    std::sort(t, t + 5);
    ms = t[2];
    return SPK_OK;
}

static int autotune_schedule(spk_context *ctx, int batch, const spk_frame_desc *frame, float *rgba_dev,
                             spk_schedule *chosen, double *best_ms) {
    spk_schedule *target = schedule_of(ctx, batch);
    if (!target) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    static const uint32_t kBlocks[] = {2u, 4u, 8u, 16u};
    static const uint32_t kRefill[] = {0u, 16u, 64u};
    double best = 1e30;
    spk_schedule best_s = *target;
    std::vector<std::pair<double, spk_schedule>> base;
    for (uint32_t srt = 0; srt < 2; ++srt) {
        for (uint32_t b : kBlocks) {
            for (uint32_t r : kRefill) {
                *target = spk_schedule{b, srt, r, 0u};
                double ms = 0.0;
                const int rc = time_schedule(ctx, batch, frame, rgba_dev, ms);
                if (rc != SPK_OK) return rc;
                base.emplace_back(ms, *target);
                if (ms < best) {
                    best = ms;
                    best_s = *target;
                }
            }
        }
    }
    // then the first segment of the per-path tail kernel, on the three fastest grid / sort / refill
    // choices: the grid that is best without the tail is not always best with it
    std::sort(base.begin(), base.end(),
              [](const std::pair<double, spk_schedule> &x, const std::pair<double, spk_schedule> &y) {
                  return x.first < y.first;
              });
    static const uint32_t kTail[] = {1u, 2u, 3u, 4u, 5u};
    for (size_t c = 0; c < std::min<size_t>(3, base.size()); ++c) {
        const spk_schedule bs = base[c].second;
        for (uint32_t t : kTail) {
            if (t >= frame->max_bounce) continue;
            *target = spk_schedule{bs.blocks_per_cu, bs.material_sort, bs.refill, t};
            double ms = 0.0;
            const int rc = time_schedule(ctx, batch, frame, rgba_dev, ms);
            if (rc != SPK_OK) return rc;
            if (ms < best) {
                best = ms;
                best_s = *target;
            }
        }
    }
    *target = best_s;
    if (chosen) *chosen = best_s;
    if (best_ms) *best_ms = best;
    return SPK_OK;
}

int spk_render_autotune(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_schedule *chosen,
                        double *best_ms) {
    if (!ctx || !frame) return SPK_ERR_INVALID_ARGUMENT;
    return autotune_schedule(ctx, 0, frame, rgba_dev, chosen, best_ms);
}

int spk_batch_autotune(spk_context *ctx, const spk_frame_desc *frame, float *rgba_dev, spk_schedule *chosen,
                       double *best_ms) {
    if (!ctx || !frame) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->batch) return set_error(ctx, SPK_ERR_NO_SCENE, "no batch uploaded");
    return autotune_schedule(ctx, 1, frame, rgba_dev, chosen, best_ms);
}

int spk_accumulate(spk_context *ctx, const float *rgba_dev, float *accum_dev, uint64_t num_pixels,
                   uint32_t cumulated, uint32_t spp) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (num_pixels && (!rgba_dev || !accum_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null buffer");
    if (cumulated == 0 && spp == 0) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (num_pixels == 0) return SPK_OK;
    const float ma = (float)cumulated / (float)(cumulated + spp);
    const uint64_t blocks = (num_pixels + 255) / 256;
    hipLaunchKernelGGL(k_accumulate, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, (const float4 *)rgba_dev,
                       (float4 *)accum_dev, num_pixels, ma);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "accumulate");
    return SPK_OK;
}

int spk_spatial_fill(spk_context *ctx, const float *rgba_dev, const float *normal_dev, float *out_dev, uint32_t width,
                     uint32_t height) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!rgba_dev || !normal_dev || !out_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null buffer");
    if (out_dev == rgba_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "the fill cannot run in place");
    const uint64_t n = (uint64_t)width * height;
    if (n == 0) return SPK_OK;
    if (n >= (1ull << 40)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "image too large");
    hipLaunchKernelGGL(k_spatial_fill, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float4 *)rgba_dev, normal_dev, (float4 *)out_dev, width, height);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "spatial fill");
    return SPK_OK;
}

int spk_tonemap(spk_context *ctx, const float *accum_dev, float *out_dev, uint32_t width, uint32_t height,
                float exposure) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!accum_dev || !out_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null buffer");
    const uint64_t n = (uint64_t)width * height;
    if (n == 0) return SPK_OK;
    hipLaunchKernelGGL(k_tonemap, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float4 *)accum_dev, (float4 *)out_dev, width, height, exposure);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "tonemap");
    return SPK_OK;
}

}  // extern "C"
