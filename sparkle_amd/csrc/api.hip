// C ABI of the path-tracer hot path (include/spk_pt.h): context, scene upload, closest-hit queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spk_pt.h"
#include "bvh_build.h"
#include "spk_context.h"
#include "spk_kernels.h"

using namespace spk;

namespace spk {

constexpr int kMaxVariants = 256;

// ---- scratch of spk_trace_closest_noalloc ------------------------------------------------------
// The coherence sort of up to SPK_NOALLOC_MAX_RAYS rays in one file-scope device array, so that
// call allocates nothing: keys, keys in direction order, indices in direction order, the final
// permutation (4 B per ray each), the pass-2 run offsets (one row of kRaySortBins per 32K-ray
// chunk), the pass-1 run offsets (kRayDirBins per 8K-ray chunk) and the bin totals. 44.3 MiB, held
// by every GPU the library runs on until the process ends. The host never uses the array's name as
// an address: each device's address comes from hipGetSymbolAddress at spk_create.
constexpr size_t align256(size_t x) { return (x + 255u) & ~size_t(255u); }
constexpr uint64_t kNoAllocRays = SPK_NOALLOC_MAX_RAYS;
constexpr size_t kScrKeys = 0;
constexpr size_t kScrKeys1 = kScrKeys + align256(kNoAllocRays * 4u);
constexpr size_t kScrPerm1 = kScrKeys1 + align256(kNoAllocRays * 4u);
constexpr size_t kScrPerm = kScrPerm1 + align256(kNoAllocRays * 4u);
constexpr size_t kScrOffs = kScrPerm + align256(kNoAllocRays * 4u);
constexpr size_t kScrOffs1 =
    kScrOffs + align256((kNoAllocRays + kRaySortChunk - 1u) / kRaySortChunk * kRaySortBins * 4u);
constexpr size_t kScrTotal =
    kScrOffs1 + align256((kNoAllocRays + kRaySortChunk1 - 1u) / kRaySortChunk1 * kRayDirBins * 4u);
constexpr size_t kScrBytes = kScrTotal + align256((size_t)(kRayDirBins + kRaySortBins) * 4u);

static __device__ __align__(256) uint8_t scratch[kScrBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// This device's address of `scratch`, looked up once per device with that device current.
hipError_t noalloc_scratch_address(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

int set_error(spk_context *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->error = buf;
    }
    return code;
}

int hip_fail(spk_context *ctx, hipError_t e, const char *what) {
    return set_error(ctx, e == hipErrorOutOfMemory ? SPK_ERR_OUT_OF_MEMORY : SPK_ERR_HIP, "%s: %s", what,
                     hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer() { release(); }
void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
hipError_t DeviceBuffer::ensure(size_t want) {
    if (want <= bytes && ptr) return hipSuccess;
    release();
    if (want == 0) return hipSuccess;
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return e;
    }
    bytes = want;
    return hipSuccess;
}

// World position of an object-space point: M (row-major 3x4) x p, copied exactly for the identity.
static inline void transform_point(const float *M, const float *p, float *dst) {
    static const float I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    if (std::memcmp(M, I, sizeof(I)) == 0) {
        dst[0] = p[0], dst[1] = p[1], dst[2] = p[2];
        return;
    }
    dst[0] = M[0] * p[0] + M[1] * p[1] + M[2] * p[2] + M[3];
    dst[1] = M[4] * p[0] + M[5] * p[1] + M[6] * p[2] + M[7];
    dst[2] = M[8] * p[0] + M[9] * p[1] + M[10] * p[2] + M[11];
}

// Flattens instances into world-space triangles and (unless `build` is false) builds the wide BVH
// over them on the host.
static int flatten_and_build(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                             uint32_t num_instances, const spk_build_options *options, HostScene &hs,
                             std::string &err, bool build = true) {
    if (!meshes || num_meshes == 0 || !instances || num_instances == 0) {
        err = "scene needs at least one mesh and one instance";
        return SPK_ERR_INVALID_ARGUMENT;
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        if (instances[i].mesh >= num_meshes) {
            err = "instance " + std::to_string(i) + " names mesh " + std::to_string(instances[i].mesh) +
                  " which does not exist";
            return SPK_ERR_INVALID_SCENE;
        }
        total += meshes[instances[i].mesh].num_triangles;
    }
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const spk_mesh_desc &md = meshes[m];
        if (md.num_triangles && (!md.positions || !md.indices)) {
            err = "mesh " + std::to_string(m) + " has no positions or indices";
            return SPK_ERR_INVALID_ARGUMENT;
        }
        for (uint64_t k = 0; k < 3ull * md.num_triangles; ++k)
            if (md.indices[k] >= md.num_vertices) {
                err = "mesh " + std::to_string(m) + " index " + std::to_string(k) + " is out of range";
                return SPK_ERR_INVALID_SCENE;
            }
    }
    if (total == 0) {
        err = "scene has no triangles";
        return SPK_ERR_INVALID_SCENE;
    }
    if (total >= 0xFFFFFFFFull) {
        err = "scene has 2^32 or more triangles";
        return SPK_ERR_INVALID_SCENE;
    }
    const uint32_t n = (uint32_t)total;
    hs.verts.resize(9 * (size_t)n);
    hs.global_ids.resize(n);
    hs.materials.resize(n);
    hs.local_ids.resize(n);
    hs.shade_src.resize(n);
    uint32_t g = 0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        const spk_instance_desc &ins = instances[i];
        const spk_mesh_desc &md = meshes[ins.mesh];
        const float *M = ins.world_from_local;
        for (uint32_t f = 0; f < md.num_triangles; ++f, ++g) {
            for (int k = 0; k < 3; ++k)
                transform_point(M, md.positions + 3 * (size_t)md.indices[3 * (size_t)f + k],
                                &hs.verts[9 * (size_t)g + 3 * k]);
            hs.global_ids[g] = g;
            hs.materials[g] = ins.material;
            hs.local_ids[g] = f;
            hs.shade_src[g] = {i, f};
        }
    }
    BuildInput in;
    in.tri_verts = hs.verts.data();
    in.global_ids = hs.global_ids.data();
    in.materials = hs.materials.data();
    in.local_ids = hs.local_ids.data();
    in.num_tris = n;
    in.width = options && options->bvh_width ? options->bvh_width : 8u;
    in.max_leaf_tris = options && options->max_leaf_tris ? options->max_leaf_tris : 3u;
    in.num_threads = options ? options->num_threads : 0u;
    hs.input = in;
    if (!build) return SPK_OK;
    err = build_wide_bvh(in, hs.bvh);
    if (!err.empty()) return SPK_ERR_INVALID_SCENE;
    return SPK_OK;
}

// Shading attributes of face f of mesh md placed by M (row-major 3x4): normals and tangents through
// the 3x3 part, the geometric normal of the world triangle v9 when the mesh has no normals, the
// default tangent (1, 0, 0) through the 3x3 part when it has no tangents.
static void shade_triangle(const spk_mesh_desc &md, const float *M, uint32_t f, const float *v, uint32_t material,
                           TriShade &s) {
    std::memset(&s, 0, sizeof(s));
    float *ns[3] = {s.n0, s.n1, s.n2}, *ts[3] = {s.t0, s.t1, s.t2}, *uvs[3] = {s.uv0, s.uv1, s.uv2};
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]}, e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    float gn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float gl = std::sqrt(gn[0] * gn[0] + gn[1] * gn[1] + gn[2] * gn[2]);
    if (gl > 0.f) gn[0] /= gl, gn[1] /= gl, gn[2] /= gl;
    for (int c = 0; c < 3; ++c) {
        const uint32_t vi = md.indices[3 * (size_t)f + c];
        if (md.normals) {
            const float *nn = md.normals + 3 * (size_t)vi;
            for (int r = 0; r < 3; ++r) ns[c][r] = M[4 * r] * nn[0] + M[4 * r + 1] * nn[1] + M[4 * r + 2] * nn[2];
        } else {
            for (int r = 0; r < 3; ++r) ns[c][r] = gn[r];
        }
        if (md.tangents) {
            const float *tt = md.tangents + 4 * (size_t)vi;
            for (int r = 0; r < 3; ++r) ts[c][r] = M[4 * r] * tt[0] + M[4 * r + 1] * tt[1] + M[4 * r + 2] * tt[2];
            ts[c][3] = tt[3];
        } else {
            ts[c][0] = M[0], ts[c][1] = M[4], ts[c][2] = M[8];
            ts[c][3] = 1.f;
        }
        if (md.uvs) {
            uvs[c][0] = md.uvs[2 * (size_t)vi];
            uvs[c][1] = md.uvs[2 * (size_t)vi + 1];
        }
    }
    s.material = material;
}

static spk_mesh_desc mesh_view(const MeshCopy &mc) {
    spk_mesh_desc d;
    std::memset(&d, 0, sizeof(d));
    d.positions = mc.positions.data();
    d.normals = mc.normals.empty() ? nullptr : mc.normals.data();
    d.tangents = mc.tangents.empty() ? nullptr : mc.tangents.data();
    d.uvs = mc.uvs.empty() ? nullptr : mc.uvs.data();
    d.indices = mc.indices.data();
    d.num_vertices = mc.num_vertices;
    d.num_triangles = mc.num_triangles;
    return d;
}

// Scene box of the world vertices (the ray sort's origin cells), kept with the device scene.
static void set_scene_box(const HostScene &hs, DeviceScene &ds) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < hs.verts.size(); i += 3)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], hs.verts[i + a]);
            hi[a] = std::max(hi[a], hs.verts[i + a]);
        }
    for (int a = 0; a < 3; ++a) {
        ds.box_lo[a] = std::isfinite(lo[a]) ? lo[a] : 0.f;
        const float ext = std::isfinite(hi[a] - lo[a]) ? hi[a] - lo[a] : 0.f;
        ds.box_inv[a] = ext > 0.f ? 1.f / ext : 1.f;
    }
}

// ---- two-level scenes ----
// BLAS per mesh over its object-space triangles (v0.w = face index), built once.
static int build_blases(HostScene &hs, uint32_t width, uint32_t max_leaf, uint32_t threads, std::string &err) {
    const uint32_t nm = (uint32_t)hs.mesh_copies.size();
    std::vector<char> used(nm, 0);
    for (uint32_t m : hs.inst_mesh) used[m] = 1;
    hs.blas.assign(nm, BuildOutput());
    for (uint32_t m = 0; m < nm; ++m) {
        const MeshCopy &mc = hs.mesh_copies[m];
        if (!used[m] || mc.num_triangles == 0) continue;
        std::vector<float> v(9 * (size_t)mc.num_triangles);
        std::vector<uint32_t> ids(mc.num_triangles), zero(mc.num_triangles, 0u);
        for (uint32_t f = 0; f < mc.num_triangles; ++f) {
            ids[f] = f;
            for (int k = 0; k < 3; ++k)
                std::memcpy(&v[9 * (size_t)f + 3 * k], &mc.positions[3 * (size_t)mc.indices[3 * (size_t)f + k]], 12);
        }
        BuildInput in;
        in.tri_verts = v.data();
        in.global_ids = ids.data();
        in.materials = zero.data();
        in.local_ids = ids.data();
        in.num_tris = mc.num_triangles;
        in.width = width;
        in.max_leaf_tris = max_leaf;
        in.num_threads = threads;
        err = build_wide_bvh(in, hs.blas[m]);
        if (!err.empty()) {
            err = "mesh " + std::to_string(m) + ": " + err;
            return SPK_ERR_INVALID_SCENE;
        }
    }
    return SPK_OK;
}

// TLAS over the instances' world boxes (each box as the pseudo-triangle lo, hi, hi, one per leaf
// slot), then the device arrays: TLAS nodes first, every BLAS after them with its child and
// triangle bases rebased, and one instance record per TLAS leaf slot.
static int assemble_two_level(HostScene &hs, uint32_t width, uint32_t threads, std::string &err) {
    const uint32_t ni = (uint32_t)hs.inst_mesh.size();
    std::vector<uint32_t> live;  // instances with triangles
    std::vector<float> pv;
    for (uint32_t i = 0; i < ni; ++i) {
        const uint32_t nt = hs.mesh_copies[hs.inst_mesh[i]].num_triangles;
        if (nt == 0) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        const float *v = &hs.verts[9 * (size_t)hs.inst_id_base[i]];
        for (size_t k = 0; k < 3 * (size_t)nt; ++k)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], v[3 * k + a]);
                hi[a] = std::max(hi[a], v[3 * k + a]);
            }
        // the BLAS is hit with a rounded object-space ray: pad the world box by a few ulps of its
        // coordinates so that the TLAS never culls a hit the object-space test reports
        for (int a = 0; a < 3; ++a) {
            const float pad = 1e-5f * std::max({std::fabs(lo[a]), std::fabs(hi[a]), hi[a] - lo[a], 1e-6f});
            lo[a] -= pad;
            hi[a] += pad;
        }
        live.push_back(i);
        pv.insert(pv.end(), {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], hi[0], hi[1], hi[2]});
    }
    std::vector<uint32_t> ids(live.size()), zero(live.size(), 0u);
    for (uint32_t k = 0; k < live.size(); ++k) ids[k] = k;
    BuildInput tin;
    tin.tri_verts = pv.data();
    tin.global_ids = ids.data();
    tin.materials = zero.data();
    tin.local_ids = ids.data();
    tin.num_tris = (uint32_t)live.size();
    tin.width = width;
    tin.max_leaf_tris = 1;
    tin.num_threads = threads;
    err = build_wide_bvh(tin, hs.tlas);
    if (!err.empty()) {
        err = "TLAS: " + err;
        return SPK_ERR_INVALID_SCENE;
    }
    const uint32_t nb = hs.tlas.node_bytes;
    const uint32_t slots = width;
    // the traversal keeps a TLAS node's leaf slots as one bit mask below bit 24
    for (uint32_t k = 0; k < hs.tlas.num_nodes; ++k) {
        const uint8_t *meta = &hs.tlas.nodes[(size_t)k * nb + 24];
        for (uint32_t sl = 0; sl < slots; ++sl)
            if (meta[sl] && (meta[sl] & 31u) + (meta[sl] >> 5) > 24u) {
                err = "TLAS leaf layout exceeds the traversal's pending-instance mask";
                return SPK_ERR_INVALID_SCENE;
            }
    }
    uint32_t total_nodes = hs.tlas.num_nodes, total_tris = 0, max_depth = 0;
    std::vector<uint32_t> root(hs.blas.size(), 0u), tri_off(hs.blas.size(), 0u);
    for (size_t m = 0; m < hs.blas.size(); ++m) {
        if (hs.blas[m].num_nodes == 0) continue;
        root[m] = total_nodes;
        tri_off[m] = total_tris;
        total_nodes += hs.blas[m].num_nodes;
        total_tris += (uint32_t)hs.blas[m].tris.size();
        max_depth = std::max(max_depth, hs.blas[m].depth);
    }
    hs.nodes2.assign((size_t)total_nodes * nb, 0);
    std::memcpy(hs.nodes2.data(), hs.tlas.nodes.data(), (size_t)hs.tlas.num_nodes * nb);
    hs.tris2.resize(total_tris);
    for (size_t m = 0; m < hs.blas.size(); ++m) {
        const BuildOutput &b = hs.blas[m];
        if (b.num_nodes == 0) continue;
        uint8_t *dst = &hs.nodes2[(size_t)root[m] * nb];
        std::memcpy(dst, b.nodes.data(), (size_t)b.num_nodes * nb);
        for (uint32_t k = 0; k < b.num_nodes; ++k) {
            uint32_t cb, tb;
            std::memcpy(&cb, dst + (size_t)k * nb + 16, 4);
            std::memcpy(&tb, dst + (size_t)k * nb + 20, 4);
            cb += root[m];
            tb += tri_off[m];
            std::memcpy(dst + (size_t)k * nb + 16, &cb, 4);
            std::memcpy(dst + (size_t)k * nb + 20, &tb, 4);
        }
        std::memcpy(&hs.tris2[tri_off[m]], b.tris.data(), b.tris.size() * sizeof(TriRecord));
    }
    hs.inst_rec.assign(16 * hs.tlas.tris.size(), 0.f);
    for (size_t j = 0; j < hs.tlas.tris.size(); ++j) {
        const uint32_t i = live[hs.tlas.tri_input[j]];
        const float *M = &hs.inst_xf[12 * (size_t)i];
        // object_from_world = inverse of [A | t] in double, rounded once
        const double a[9] = {M[0], M[1], M[2], M[4], M[5], M[6], M[8], M[9], M[10]};
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                           a[2] * (a[3] * a[7] - a[4] * a[6]);
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
            err = "instance " + std::to_string(i) + " has a singular transform";
            return SPK_ERR_INVALID_SCENE;
        }
        const double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                               (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                               (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                               (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                               (a[0] * a[4] - a[1] * a[3]) / det};
        const double t[3] = {M[3], M[7], M[11]};
        float *r = &hs.inst_rec[16 * j];
        for (int row = 0; row < 3; ++row) {
            r[4 * row] = (float)inv[3 * row];
            r[4 * row + 1] = (float)inv[3 * row + 1];
            r[4 * row + 2] = (float)inv[3 * row + 2];
            r[4 * row + 3] = (float)-(inv[3 * row] * t[0] + inv[3 * row + 1] * t[1] + inv[3 * row + 2] * t[2]);
        }
        const uint32_t tail[4] = {root[hs.inst_mesh[i]], hs.inst_id_base[i], i, 0u};
        std::memcpy(r + 12, tail, 16);
    }
    hs.num_nodes2 = total_nodes;
    hs.depth2 = hs.tlas.depth + max_depth + 2u;
    return SPK_OK;
}

// Two-level form of a validated, flattened scene: object-space mesh copies, BLASes and the TLAS.
static int build_two_level(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                           uint32_t num_instances, const spk_build_options *options, HostScene &hs,
                           std::string &err) {
    hs.two_level = true;
    hs.mesh_copies.assign(num_meshes, MeshCopy());
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const spk_mesh_desc &md = meshes[m];
        MeshCopy &mc = hs.mesh_copies[m];
        mc.num_vertices = md.num_vertices;
        mc.num_triangles = md.num_triangles;
        if (md.num_triangles == 0) continue;
        mc.positions.assign(md.positions, md.positions + 3 * (size_t)md.num_vertices);
        if (md.normals) mc.normals.assign(md.normals, md.normals + 3 * (size_t)md.num_vertices);
        if (md.tangents) mc.tangents.assign(md.tangents, md.tangents + 4 * (size_t)md.num_vertices);
        if (md.uvs) mc.uvs.assign(md.uvs, md.uvs + 2 * (size_t)md.num_vertices);
        mc.indices.assign(md.indices, md.indices + 3 * (size_t)md.num_triangles);
    }
    hs.inst_mesh.resize(num_instances);
    hs.inst_material.resize(num_instances);
    hs.inst_id_base.resize(num_instances);
    hs.inst_xf.resize(12 * (size_t)num_instances);
    uint32_t g = 0;
    for (uint32_t i = 0; i < num_instances; ++i) {
        hs.inst_mesh[i] = instances[i].mesh;
        hs.inst_material[i] = instances[i].material;
        hs.inst_id_base[i] = g;
        std::memcpy(&hs.inst_xf[12 * (size_t)i], instances[i].world_from_local, 48);
        g += meshes[instances[i].mesh].num_triangles;
    }
    const uint32_t width = options && options->bvh_width ? options->bvh_width : 8u;
    const uint32_t leaf = options && options->max_leaf_tris ? options->max_leaf_tris : 3u;
    const uint32_t threads = options ? options->num_threads : 0u;
    hs.input.width = width;
    hs.input.max_leaf_tris = leaf;
    hs.input.num_threads = threads;
    int rc = build_blases(hs, width, leaf, threads, err);
    if (rc != SPK_OK) return rc;
    return assemble_two_level(hs, width, threads, err);
}

static void fill_stats(const HostScene &hs, spk_scene_stats *st, uint64_t device_bytes) {
    std::memset(st, 0, sizeof(*st));
    st->num_triangles = hs.global_ids.size();
    st->bvh_width = hs.input.width;
    st->triangle_bytes = sizeof(TriRecord);
    st->device_bytes = device_bytes;
    if (hs.two_level) {
        // nodes of the TLAS and every BLAS; depth = TLAS depth + deepest BLAS; costs summed
        st->num_nodes = hs.num_nodes2;
        st->depth = hs.tlas.depth + (hs.depth2 - hs.tlas.depth - 2u);
        st->node_bytes = hs.tlas.node_bytes;
        st->sah_cost = hs.tlas.sah_cost;
        st->build_seconds = hs.tlas.seconds;
        for (const BuildOutput &b : hs.blas) {
            st->degenerate_tris += b.degenerate;
            st->build_seconds += b.seconds;
        }
        return;
    }
    st->num_nodes = hs.bvh.num_nodes;
    st->depth = hs.bvh.depth;
    st->node_bytes = hs.bvh.node_bytes;
    st->degenerate_tris = hs.bvh.degenerate;
    st->sah_cost = hs.bvh.sah_cost;
    st->build_seconds = hs.bvh.seconds;
}

// World triangles in global-id order with the flattened records' tags (v0.w global id, v1.w
// material, v2.w face index): what a two-level scene's shading reads instead of leaf-order records.
static std::vector<TriRecord> world_records(const HostScene &hs, uint32_t first, uint32_t count) {
    std::vector<TriRecord> out(count);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t g = first + k;
        TriRecord &r = out[k];
        const float *v = &hs.verts[9 * (size_t)g];
        float *dst[3] = {r.v0, r.v1, r.v2};
        for (int c = 0; c < 3; ++c) dst[c][0] = v[3 * c], dst[c][1] = v[3 * c + 1], dst[c][2] = v[3 * c + 2];
        std::memcpy(&r.v0[3], &g, 4);
        std::memcpy(&r.v1[3], &hs.shade_in[g].material, 4);
        std::memcpy(&r.v2[3], &hs.shade_src[g].face, 4);
    }
    return out;
}

static int upload_two_level(spk_context *ctx) {
    HostScene &hs = ctx->host_scene;
    const std::vector<MaterialGpu> &mats = ctx->host_mats;
    const uint32_t n = (uint32_t)hs.global_ids.size();
    const std::vector<TriRecord> wt = world_records(hs, 0, n);
    hipError_t e;
    DeviceScene &ds = ctx->scene;
    ds.valid = false;
    if ((e = ds.nodes.ensure(hs.nodes2.size())) != hipSuccess ||
        (e = ds.tris.ensure(std::max<size_t>(1, hs.tris2.size()) * sizeof(TriRecord))) != hipSuccess ||
        (e = ds.insts.ensure(hs.inst_rec.size() * 4)) != hipSuccess ||
        (e = ds.wtris.ensure(wt.size() * sizeof(TriRecord))) != hipSuccess ||
        (e = ds.shade.ensure(hs.shade_in.size() * sizeof(TriShade))) != hipSuccess ||
        (e = ds.materials.ensure(mats.size() * sizeof(MaterialGpu))) != hipSuccess)
        return hip_fail(ctx, e, "scene allocation");
    if ((e = hipMemcpy(ds.nodes.ptr, hs.nodes2.data(), hs.nodes2.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds.tris.ptr, hs.tris2.data(), hs.tris2.size() * sizeof(TriRecord), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.insts.ptr, hs.inst_rec.data(), hs.inst_rec.size() * 4, hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.wtris.ptr, wt.data(), wt.size() * sizeof(TriRecord), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.shade.ptr, hs.shade_in.data(), hs.shade_in.size() * sizeof(TriShade),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds.materials.ptr, mats.data(), mats.size() * sizeof(MaterialGpu), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return hip_fail(ctx, e, "scene upload");
    ds.two_level = true;
    set_scene_box(hs, ds);
    ds.num_tris = n;
    ds.num_nodes = hs.num_nodes2;
    ds.num_materials = (uint32_t)mats.size();
    ds.width = (int)hs.input.width;
    ds.depth = (int)hs.depth2;
    ds.variant = TraceVariant{ds.width, pick_trace_stack((int)hs.depth2, true), 128};
    // emissive triangles by global id for next-event estimation
    ds.host_emitters.clear();
    for (uint32_t g = 0; g < n; ++g)
        if (mats[hs.shade_in[g].material].emissive_flag) ds.host_emitters.push_back(g);
    ds.num_spheres = 0;
    ds.valid = true;
    fill_stats(hs, &ctx->stats,
               ds.nodes.bytes + ds.tris.bytes + ds.insts.bytes + ds.wtris.bytes + ds.shade.bytes + ds.materials.bytes);
    return SPK_OK;
}

// Uploads the host scene's current BVH (nodes, triangles, leaf-order shading) and materials.
static int upload_built(spk_context *ctx, int stack) {
    HostScene &hs = ctx->host_scene;
    if (hs.two_level) return upload_two_level(ctx);
    ctx->scene.two_level = false;
    const std::vector<MaterialGpu> &mats = ctx->host_mats;
    const uint32_t n = (uint32_t)hs.bvh.tris.size();
    const std::vector<TriShade> shade = leaf_order_shade(hs);
    hipError_t e;
    DeviceScene &ds = ctx->scene;
    ds.valid = false;
    if ((e = ds.nodes.ensure(hs.bvh.nodes.size())) != hipSuccess ||
        (e = ds.tris.ensure(hs.bvh.tris.size() * sizeof(TriRecord))) != hipSuccess ||
        (e = ds.shade.ensure(shade.size() * sizeof(TriShade))) != hipSuccess ||
        (e = ds.materials.ensure(mats.size() * sizeof(MaterialGpu))) != hipSuccess)
        return hip_fail(ctx, e, "scene allocation");
    if ((e = hipMemcpy(ds.nodes.ptr, hs.bvh.nodes.data(), hs.bvh.nodes.size(), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.tris.ptr, hs.bvh.tris.data(), hs.bvh.tris.size() * sizeof(TriRecord),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds.shade.ptr, shade.data(), shade.size() * sizeof(TriShade), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(ds.materials.ptr, mats.data(), mats.size() * sizeof(MaterialGpu), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return hip_fail(ctx, e, "scene upload");
    set_scene_box(hs, ds);
    ds.num_tris = n;
    ds.num_nodes = hs.bvh.num_nodes;
    ds.num_materials = (uint32_t)mats.size();
    ds.width = (int)hs.input.width;
    ds.depth = (int)hs.bvh.depth;
    ds.variant = TraceVariant{ds.width, stack, 128};
    // emissive triangles (leaf order) for next-event estimation
    ds.host_emitters.clear();
    for (uint32_t k = 0; k < n; ++k)
        if (mats[shade[k].material].emissive_flag) ds.host_emitters.push_back(k);
    ds.num_spheres = 0;  // spheres are set after each upload (spk_scene_set_spheres)
    ds.valid = true;
    fill_stats(hs, &ctx->stats, ds.nodes.bytes + ds.tris.bytes + ds.shade.bytes + ds.materials.bytes);
    return SPK_OK;
}

// Validates and flattens one scene, builds its BVH and converts its shading attributes and
// materials (everything spk_scene_upload and spk_batch_upload do on the host).
int prepare_scene(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                  uint32_t num_instances, const spk_material_desc *materials, uint32_t num_materials,
                  uint32_t num_textures, const spk_build_options *options, HostScene &hs,
                  std::vector<MaterialGpu> &mats, std::string &err) {
    if (!materials || num_materials == 0) {
        err = "scene needs at least one material";
        return SPK_ERR_INVALID_ARGUMENT;
    }
    for (uint32_t i = 0; instances && i < num_instances; ++i)
        if (instances[i].material >= num_materials) {
            err = "instance " + std::to_string(i) + " names material " + std::to_string(instances[i].material) +
                  " of " + std::to_string(num_materials);
            return SPK_ERR_INVALID_SCENE;
        }
    mats.assign(num_materials, MaterialGpu());
    for (uint32_t m = 0; m < num_materials; ++m) {
        const spk_material_desc &d = materials[m];
        MaterialGpu &g = mats[m];
        std::memset(&g, 0, sizeof(g));
        g.type = d.type;
        for (int c = 0; c < 3; ++c) g.base[c] = d.base_color[c], g.emissive[c] = d.emissive[c];
        g.metallic = d.metallic;
        g.roughness = d.roughness;
        g.eta = d.eta;
        const float en = std::sqrt(d.emissive[0] * d.emissive[0] + d.emissive[1] * d.emissive[1] +
                                   d.emissive[2] * d.emissive[2]);
        g.emissive_flag = en > 1e-6f ? 1u : 0u;
        // texture fields: 0 = none, k + 1 = table entry k
        const uint32_t tex[4] = {d.base_color_texture, d.metallic_roughness_texture, d.emissive_texture,
                                 d.normal_texture};
        int32_t idx[4];
        for (int k = 0; k < 4; ++k) {
            if (tex[k] > num_textures) {
                err = "material " + std::to_string(m) + " names texture " + std::to_string(tex[k] - 1) + " of " +
                      std::to_string(num_textures);
                return SPK_ERR_INVALID_SCENE;
            }
            idx[k] = (int32_t)tex[k] - 1;
        }
        g.tex_base = idx[0];
        g.tex_mr = idx[1];
        g.tex_emissive = idx[2];
        g.tex_normal = idx[3];
        if (d.type != SPK_MATERIAL_PBR && d.type != SPK_MATERIAL_DIELECTRIC)
            {
            err = "material " + std::to_string(m) + " has unknown type " + std::to_string(d.type);
            return SPK_ERR_INVALID_SCENE;
        }
    }

    const bool two = options && options->two_level;
    int rc = flatten_and_build(meshes, num_meshes, instances, num_instances, options, hs, err, !two);
    if (rc != SPK_OK) return rc;
    if (two) {
        rc = build_two_level(meshes, num_meshes, instances, num_instances, options, hs, err);
        if (rc != SPK_OK) return rc;
    }
    const uint32_t need = two ? hs.depth2 : hs.bvh.depth;
    if (pick_trace_stack((int)need, two) < 0) {
        err = "BVH depth " + std::to_string(need) + " exceeds the largest traversal stack (" +
              std::to_string(max_trace_stack()) + ")";
        return SPK_ERR_STACK_DEPTH;
    }

    // shading attributes in input (global id) order; a flattened scene permutes them to leaf order
    const uint32_t n = (uint32_t)hs.global_ids.size();
    hs.shade_in.assign(n, TriShade());
    for (uint32_t src = 0; src < n; ++src) {
        const uint32_t inst = hs.shade_src[src].inst, f = hs.shade_src[src].face;
        const spk_instance_desc &ins = instances[inst];
        shade_triangle(meshes[ins.mesh], ins.world_from_local, f, &hs.verts[9 * (size_t)src], ins.material,
                       hs.shade_in[src]);
    }
    return SPK_OK;
}

// Shading attributes permuted to the BVH's leaf order.
std::vector<TriShade> leaf_order_shade(const HostScene &hs) {
    const uint32_t n = (uint32_t)hs.bvh.tris.size();
    std::vector<TriShade> shade(n);
    for (uint32_t k = 0; k < n; ++k) shade[k] = hs.shade_in[hs.bvh.tri_input[k]];
    return shade;
}

}  // namespace spk

extern "C" {

int spk_create(int device, spk_context **out) {
    if (!out) return SPK_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SPK_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return SPK_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return SPK_ERR_HIP;
    auto *ctx = new spk_context();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SPK_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    // the no-allocation entry point's scratch on this device (its calls report an error if absent)
    if (noalloc_scratch_address(device, &ctx->noalloc_scratch) != hipSuccess) ctx->noalloc_scratch = nullptr;
    if (ctx->counters.ensure(64 * sizeof(uint32_t)) != hipSuccess ||
        hipMemsetAsync(ctx->counters.ptr, 0, 64 * sizeof(uint32_t), ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return SPK_ERR_HIP;
    }
    *out = ctx;
    return SPK_OK;
}

int spk_destroy(spk_context *ctx) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_all();
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return SPK_OK;
}

const char *spk_last_error(const spk_context *ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int spk_set_stream(spk_context *ctx, void *hip_stream) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return SPK_OK;
}

void *spk_get_stream(spk_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int spk_synchronize(spk_context *ctx) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return ctx->check_invalid_rays();
}

int spk_device_alloc(spk_context *ctx, size_t bytes, void **out_dev) {
    if (!ctx || !out_dev) return SPK_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(out_dev, bytes ? bytes : 1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    return SPK_OK;
}

int spk_device_free(spk_context *ctx, void *dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!dev) return SPK_OK;
    hipError_t e = hipFree(dev);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    return SPK_OK;
}

int spk_copy_to_device(spk_context *ctx, void *dst_dev, const void *src, size_t bytes) {
    if (!ctx || (!dst_dev && bytes) || (!src && bytes)) return SPK_ERR_INVALID_ARGUMENT;
    hipError_t e = hipMemcpyAsync(dst_dev, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "copy to device");
    return SPK_OK;
}

int spk_copy_to_host(spk_context *ctx, void *dst, const void *src_dev, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src_dev && bytes)) return SPK_ERR_INVALID_ARGUMENT;
    hipError_t e = hipMemcpyAsync(dst, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "copy to host");
    return SPK_OK;
}

int spk_build_check(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                    uint32_t num_instances, const spk_build_options *options, spk_scene_stats *stats,
                    char *errmsg, size_t errmsg_size) {
    HostScene hs;
    std::string err;
    int rc = flatten_and_build(meshes, num_meshes, instances, num_instances, options, hs, err);
    if (rc == SPK_OK) {
        err = validate_wide_bvh(hs.input, hs.bvh);
        if (!err.empty()) rc = SPK_ERR_INVALID_SCENE;
    }
    if (rc == SPK_OK && stats) fill_stats(hs, stats, 0);
    if (errmsg && errmsg_size) {
        std::snprintf(errmsg, errmsg_size, "%s", err.c_str());
    }
    return rc;
}

int spk_scene_upload(spk_context *ctx, const spk_mesh_desc *meshes, uint32_t num_meshes,
                     const spk_instance_desc *instances, uint32_t num_instances,
                     const spk_material_desc *materials, uint32_t num_materials,
                     const spk_build_options *options) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    HostScene hs;
    std::vector<MaterialGpu> mats;
    std::string err;
    int rc = prepare_scene(meshes, num_meshes, instances, num_instances, materials, num_materials,
                           ctx->num_textures, options, hs, mats, err);
    if (rc != SPK_OK) return set_error(ctx, rc, "%s", err.c_str());
    const int stack = pick_trace_stack((int)hs.bvh.depth);
    ctx->host_mats = mats;
    ctx->host_scene = std::move(hs);
    int rc2 = upload_built(ctx, stack);
    if (rc2 != SPK_OK) return rc2;
    ctx->on_scene_changed();
    return SPK_OK;
}

int spk_textures_upload(spk_context *ctx, const spk_texture_desc *textures, uint32_t num_textures) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (num_textures && !textures) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null texture array");
    (void)hipSetDevice(ctx->device);
    std::vector<TexRecord> table(num_textures);
    uint64_t total = 0;
    for (uint32_t k = 0; k < num_textures; ++k) {
        const spk_texture_desc &t = textures[k];
        if (t.width == 0 || t.height == 0 || t.width > 65536 || t.height > 65536 || !t.pixels)
            return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "texture %u: empty or larger than 65536^2", k);
        if (t.format > SPK_TEXTURE_RGBA32F)
            return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "texture %u: unknown format %u", k, t.format);
        std::memset(&table[k], 0, sizeof(TexRecord));
        table[k].offset = total;
        table[k].width = t.width;
        table[k].height = t.height;
        table[k].srgb = t.format == SPK_TEXTURE_RGBA8_SRGB ? 1u : 0u;
        total += (uint64_t)t.width * t.height;
    }
    // texels as float4: 8-bit channels become value / 255 (ColorToVec, Utilities.h:377-380)
    std::vector<float> texels(4 * std::max<uint64_t>(total, 1));
    for (uint32_t k = 0; k < num_textures; ++k) {
        const spk_texture_desc &t = textures[k];
        const uint64_t n = (uint64_t)t.width * t.height;
        float *dst = texels.data() + 4 * table[k].offset;
        if (t.format == SPK_TEXTURE_RGBA32F) {
            std::memcpy(dst, t.pixels, n * 16);
        } else {
            const uint8_t *src = static_cast<const uint8_t *>(t.pixels);
            for (uint64_t i = 0; i < 4 * n; ++i) dst[i] = static_cast<float>(src[i]) / 255.f;
        }
    }
    hipError_t e;
    if ((e = ctx->texels.ensure(texels.size() * 4)) != hipSuccess ||
        (e = ctx->tex_table.ensure(std::max<size_t>(1, table.size()) * sizeof(TexRecord))) != hipSuccess)
        return hip_fail(ctx, e, "texture allocation");
    if ((e = hipMemcpy(ctx->texels.ptr, texels.data(), texels.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (num_textures && (e = hipMemcpy(ctx->tex_table.ptr, table.data(), table.size() * sizeof(TexRecord),
                                        hipMemcpyHostToDevice)) != hipSuccess))
        return hip_fail(ctx, e, "texture upload");
    ctx->num_textures = num_textures;
    for (int f = 0; f < 6; ++f) ctx->sky_faces[f] = -1;  // the sky cube map is set again after a new table
    // materials of an uploaded scene or batch may name textures that no longer exist
    ctx->scene.valid = false;
    ctx->on_scene_changed();
    ctx->drop_batch();
    return SPK_OK;
}

int spk_scene_set_spheres(spk_context *ctx, const spk_sphere_desc *spheres, uint32_t num_spheres) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (num_spheres && !spheres) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null sphere array");
    if ((uint64_t)ctx->scene.num_tris + num_spheres >= 0xFFFFFFFFull || num_spheres >= (1u << 31))
        return set_error(ctx, SPK_ERR_INVALID_SCENE, "too many primitives");
    std::vector<float> rec(8 * (size_t)std::max(num_spheres, 1u), 0.f);
    for (uint32_t k = 0; k < num_spheres; ++k) {
        const spk_sphere_desc &sd = spheres[k];
        if (sd.material >= ctx->scene.num_materials)
            return set_error(ctx, SPK_ERR_INVALID_SCENE, "sphere %u names material %u of %u", k, sd.material,
                             ctx->scene.num_materials);
        if (!(sd.radius > 0.f) || !std::isfinite(sd.radius) || !std::isfinite(sd.center[0]) ||
            !std::isfinite(sd.center[1]) || !std::isfinite(sd.center[2]))
            return set_error(ctx, SPK_ERR_INVALID_SCENE, "sphere %u has a non-finite centre or radius", k);
        rec[8 * k] = sd.center[0];
        rec[8 * k + 1] = sd.center[1];
        rec[8 * k + 2] = sd.center[2];
        rec[8 * k + 3] = sd.radius * sd.radius;  // scaled_radius_2_ (SphereRenderProxy.h:29-32)
        std::memcpy(&rec[8 * k + 4], &sd.material, 4);
    }
    hipError_t e;
    if ((e = ctx->scene.spheres.ensure(rec.size() * 4)) != hipSuccess ||
        (e = hipMemcpy(ctx->scene.spheres.ptr, rec.data(), rec.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(ctx, e, "sphere upload");
    ctx->scene.num_spheres = num_spheres;
    ctx->on_scene_changed();
    return SPK_OK;
}

int spk_scene_stats_get(spk_context *ctx, spk_scene_stats *out) {
    if (!ctx || !out) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    *out = ctx->stats;
    return SPK_OK;
}

int spk_scene_update_instances(spk_context *ctx, const float *world_from_local, uint32_t num_instances) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    HostScene &hs = ctx->host_scene;
    if (!hs.two_level)
        return set_error(ctx, SPK_ERR_UNSUPPORTED, "instances move only in two-level scenes (spk_build_options)");
    if (!world_from_local || num_instances != (uint32_t)hs.inst_mesh.size())
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "expected %zu transforms", hs.inst_mesh.size());
    for (size_t k = 0; k < 12 * (size_t)num_instances; ++k)
        if (!std::isfinite(world_from_local[k]))
            return set_error(ctx, SPK_ERR_INVALID_SCENE, "instance %zu has a non-finite transform", k / 12);
    (void)hipSetDevice(ctx->device);
    // world geometry and shading of the moved instances, as a fresh upload would compute them
    for (uint32_t i = 0; i < num_instances; ++i) {
        const float *M = world_from_local + 12 * (size_t)i;
        const double det = (double)M[0] * ((double)M[5] * M[10] - (double)M[6] * M[9]) -
                           (double)M[1] * ((double)M[4] * M[10] - (double)M[6] * M[8]) +
                           (double)M[2] * ((double)M[4] * M[9] - (double)M[5] * M[8]);
        if (!(std::fabs(det) > 0.0))
            return set_error(ctx, SPK_ERR_INVALID_SCENE, "instance %u has a singular transform", i);
    }
    std::vector<uint32_t> moved;
    const std::vector<float> old_xf = hs.inst_xf;
    for (uint32_t i = 0; i < num_instances; ++i) {
        const float *M = world_from_local + 12 * (size_t)i;
        if (std::memcmp(M, &hs.inst_xf[12 * (size_t)i], 48) == 0) continue;
        moved.push_back(i);
    }
    // world vertices and shading of each moved instance under the transforms in hs.inst_xf
    // (independent ranges, spread over the build's thread count; 0: every hardware thread)
    auto rebake_moved = [&]() {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t m; (m = next.fetch_add(1)) < moved.size();) {
                const uint32_t i = moved[m];
                const float *M = &hs.inst_xf[12 * (size_t)i];
                const MeshCopy &mc = hs.mesh_copies[hs.inst_mesh[i]];
                const spk_mesh_desc md = mesh_view(mc);
                const uint32_t g0 = hs.inst_id_base[i];
                for (uint32_t f = 0; f < mc.num_triangles; ++f) {
                    float *v = &hs.verts[9 * (size_t)(g0 + f)];
                    for (int k = 0; k < 3; ++k)
                        transform_point(M, &mc.positions[3 * (size_t)mc.indices[3 * (size_t)f + k]], v + 3 * k);
                    shade_triangle(md, M, f, v, hs.inst_material[i], hs.shade_in[g0 + f]);
                }
            }
        };
        const unsigned hw =
            hs.input.num_threads ? hs.input.num_threads : std::max(1u, std::thread::hardware_concurrency());
        const unsigned nt = std::min<unsigned>(hw, std::max<unsigned>(1u, (unsigned)moved.size()));
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
        work();
        for (std::thread &t : pool) t.join();
    };
    for (uint32_t i : moved) std::memcpy(&hs.inst_xf[12 * (size_t)i], world_from_local + 12 * (size_t)i, 48);
    rebake_moved();
    std::string err;
    int rc = assemble_two_level(hs, hs.input.width, hs.input.num_threads, err);
    int stack = rc == SPK_OK ? pick_trace_stack((int)hs.depth2, true) : -1;
    if (rc == SPK_OK && stack < 0) {
        rc = SPK_ERR_STACK_DEPTH;
        err = "two-level depth " + std::to_string(hs.depth2) + " exceeds the deepest compiled traversal stack";
    }
    if (rc != SPK_OK) {
        // nothing has reached the device: put the host copy back as it was (the previous transforms,
        // their world geometry and shading, and the TLAS assembled from them, which built before)
        hs.inst_xf = old_xf;
        rebake_moved();
        std::string err2;
        if (assemble_two_level(hs, hs.input.width, hs.input.num_threads, err2) != SPK_OK)
            return set_error(ctx, rc, "%s (and the previous TLAS did not rebuild: %s)", err.c_str(), err2.c_str());
        return set_error(ctx, rc, "%s", err.c_str());
    }
    DeviceScene &ds = ctx->scene;
    hipError_t e;
    if ((e = ds.nodes.ensure(hs.nodes2.size())) != hipSuccess || (e = ds.insts.ensure(hs.inst_rec.size() * 4)) != hipSuccess)
        return hip_fail(ctx, e, "scene allocation");
    if ((e = hipMemcpy(ds.nodes.ptr, hs.nodes2.data(), hs.nodes2.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ds.insts.ptr, hs.inst_rec.data(), hs.inst_rec.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(ctx, e, "TLAS upload");
    for (uint32_t i : moved) {
        const uint32_t g0 = hs.inst_id_base[i], nt = hs.mesh_copies[hs.inst_mesh[i]].num_triangles;
        if (nt == 0) continue;
        const std::vector<TriRecord> wt = world_records(hs, g0, nt);
        if ((e = hipMemcpy((TriRecord *)ds.wtris.ptr + g0, wt.data(), nt * sizeof(TriRecord),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy((TriShade *)ds.shade.ptr + g0, &hs.shade_in[g0], nt * sizeof(TriShade),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(ctx, e, "instance shading upload");
    }
    ds.num_nodes = hs.num_nodes2;
    ds.depth = (int)hs.depth2;
    set_scene_box(hs, ds);  // the ray sort's origin cells follow the moved instances
    if (ds.variant.stack < ds.depth) {
        // a deeper stack; the block size and the coherence order stay where a variant exists
        const int block = stack == 64 ? 64 : (ds.variant.block == 64 || ds.variant.block == 128 ? ds.variant.block : 128);
        ds.variant = TraceVariant{ds.width, stack, block, 0, ds.variant.sort};
    }
    fill_stats(hs, &ctx->stats,
               ds.nodes.bytes + ds.tris.bytes + ds.insts.bytes + ds.wtris.bytes + ds.shade.bytes + ds.materials.bytes);
    ctx->on_scene_changed();  // frame graphs hold the node pointer; emitter areas may have changed
    return SPK_OK;
}

// Scratch of the coherence order for n rays (grown on demand); null when the variant does not sort.
static hipError_t sort_scratch(spk_context *ctx, uint64_t n, RaySortScratch &s, const RaySortScratch *&out) {
    out = nullptr;
    if (!ctx->scene.variant.sort || n > 0xFFFFFFFFull) return hipSuccess;  // such calls trace unsorted
    hipError_t e;
    const uint64_t m = std::max<uint64_t>(n, 1);
    const uint64_t chunks = (m + kRaySortChunk - 1) / kRaySortChunk;
    if ((e = ctx->sort_keys.ensure(m * 4)) != hipSuccess || (e = ctx->sort_keys1.ensure(m * 4)) != hipSuccess ||
        (e = ctx->sort_perm1.ensure(m * 4)) != hipSuccess || (e = ctx->sort_perm.ensure(m * 4)) != hipSuccess ||
        (e = ctx->sort_offs.ensure(chunks * kRaySortBins * 4)) != hipSuccess ||
        (e = ctx->sort_offs1.ensure((m + kRaySortChunk1 - 1) / kRaySortChunk1 * kRayDirBins * 4)) != hipSuccess ||
        (e = ctx->sort_total.ensure((size_t)(kRayDirBins + kRaySortBins) * 4)) != hipSuccess)
        return e;
    s.keys = (uint32_t *)ctx->sort_keys.ptr;
    s.keys1 = (uint32_t *)ctx->sort_keys1.ptr;
    s.perm1 = (uint32_t *)ctx->sort_perm1.ptr;
    s.perm = (uint32_t *)ctx->sort_perm.ptr;
    s.offs = (uint32_t *)ctx->sort_offs.ptr;
    s.offs1 = (uint32_t *)ctx->sort_offs1.ptr;
    s.total = (uint32_t *)ctx->sort_total.ptr;
    for (int a = 0; a < 3; ++a) s.lo[a] = ctx->scene.box_lo[a], s.inv[a] = ctx->scene.box_inv[a];
    out = &s;
    return hipSuccess;
}

int spk_trace_closest_async(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (n && (!rays_dev || !hits_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray or hit array");
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    const DeviceScene &ds = ctx->scene;
    RaySortScratch ss;
    const RaySortScratch *sp = nullptr;
    hipError_t e = sort_scratch(ctx, n, ss, sp);
    if (e != hipSuccess) return hip_fail(ctx, e, "ray sort buffers");
    const hipEvent_t *timed = nullptr;
    if (n && 2 * (size_t)ctx->kernel_timed < ctx->kernel_events.size())
        timed = &ctx->kernel_events[2 * (size_t)ctx->kernel_timed];
    e = launch_trace(ds.variant, rays_dev, n, hits_dev, ds.nodes.ptr, ds.tris.ptr, 0u, ctx->invalid_counter(), nullptr,
                     nullptr, ctx->work_counter(), ds.spheres.ptr, ds.num_spheres, ds.num_tris,
                     ds.two_level ? ds.insts.ptr : nullptr, sp, ctx->stream, timed);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    if (timed) ++ctx->kernel_timed;  // a call counts once both of its events are recorded
    return SPK_OK;
}

int spk_trace_kernel_timing(spk_context *ctx, uint32_t capacity) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (capacity > (1u << 16)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "at most 65536 timed calls");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);  // no recorded event may still be pending
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    for (hipEvent_t ev : ctx->kernel_events) (void)hipEventDestroy(ev);
    ctx->kernel_events.clear();
    ctx->kernel_timed = 0;
    for (uint32_t k = 0; k < 2 * capacity; ++k) {
        hipEvent_t ev;
        if ((e = hipEventCreate(&ev)) != hipSuccess) {
            for (hipEvent_t x : ctx->kernel_events) (void)hipEventDestroy(x);
            ctx->kernel_events.clear();
            return hip_fail(ctx, e, "hipEventCreate");
        }
        ctx->kernel_events.push_back(ev);
    }
    return SPK_OK;
}

int spk_trace_kernel_times(spk_context *ctx, float *ms, uint32_t max, uint32_t *count) {
    if (!ctx || (max && !ms)) return SPK_ERR_INVALID_ARGUMENT;
    const uint32_t n = ctx->kernel_timed;
    if (count) *count = n;
    if (n == 0) return SPK_OK;
    hipError_t e = hipEventSynchronize(ctx->kernel_events[2 * (size_t)n - 1]);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventSynchronize");
    for (uint32_t k = 0; k < std::min(n, max); ++k) {
        if ((e = hipEventElapsedTime(&ms[k], ctx->kernel_events[2 * (size_t)k], ctx->kernel_events[2 * (size_t)k + 1])) !=
            hipSuccess)
            return hip_fail(ctx, e, "hipEventElapsedTime");
    }
    return SPK_OK;
}

int spk_trace_closest(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev) {
    int rc = spk_trace_closest_async(ctx, rays_dev, n, hits_dev);
    if (rc != SPK_OK) return rc;
    return spk_synchronize(ctx);
}

int spk_trace_closest_noalloc(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (n && (!rays_dev || !hits_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray or hit array");
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (n > kNoAllocRays)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT,
                         "spk_trace_closest_noalloc takes at most %llu rays per call (got %llu)",
                         (unsigned long long)kNoAllocRays, (unsigned long long)n);
    if (!ctx->noalloc_scratch)
        return set_error(ctx, SPK_ERR_HIP, "no-allocation scratch is not available on device %d", ctx->device);
    const DeviceScene &ds = ctx->scene;
    RaySortScratch ss;
    const RaySortScratch *sp = nullptr;
    if (ds.variant.sort) {
        // every part of the sort scratch that must start at zero (the bin totals) is cleared by a fill
        // kernel inside launch_trace; the rest is written before it is read
        uint8_t *base = static_cast<uint8_t *>(ctx->noalloc_scratch);
        ss.keys = reinterpret_cast<uint32_t *>(base + kScrKeys);
        ss.keys1 = reinterpret_cast<uint32_t *>(base + kScrKeys1);
        ss.perm1 = reinterpret_cast<uint32_t *>(base + kScrPerm1);
        ss.perm = reinterpret_cast<uint32_t *>(base + kScrPerm);
        ss.offs = reinterpret_cast<uint32_t *>(base + kScrOffs);
        ss.offs1 = reinterpret_cast<uint32_t *>(base + kScrOffs1);
        ss.total = reinterpret_cast<uint32_t *>(base + kScrTotal);
        for (int a = 0; a < 3; ++a) ss.lo[a] = ds.box_lo[a], ss.inv[a] = ds.box_inv[a];
        sp = &ss;
    }
    const hipError_t e = launch_trace(ds.variant, rays_dev, n, hits_dev, ds.nodes.ptr, ds.tris.ptr, 0u,
                                      ctx->invalid_counter(), nullptr, nullptr, ctx->work_counter(), ds.spheres.ptr,
                                      ds.num_spheres, ds.num_tris, ds.two_level ? ds.insts.ptr : nullptr, sp,
                                      ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    return spk_synchronize(ctx);
}

int spk_trace_closest_host(spk_context *ctx, const spk_ray *rays, uint64_t n, spk_hit *hits) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (n == 0) return SPK_OK;
    if (!rays || !hits) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray or hit array");
    return trace_closest_host_staged(ctx, rays, n, hits);
}

int spk_trace_count(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, uint32_t *nodes_dev, uint32_t *tris_dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (n && !rays_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray array");
    if (!nodes_dev && !tris_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "no count array");
    hipError_t e;
    if ((e = ctx->tmp_hits.ensure(n * sizeof(spk_hit))) != hipSuccess) return hip_fail(ctx, e, "hit buffer");
    const DeviceScene &ds = ctx->scene;
    RaySortScratch ss;
    const RaySortScratch *sp = nullptr;
    if ((e = sort_scratch(ctx, n, ss, sp)) != hipSuccess) return hip_fail(ctx, e, "ray sort buffers");
    e = launch_trace(ds.variant, rays_dev, n, ctx->tmp_hits.ptr, ds.nodes.ptr, ds.tris.ptr, 0u, ctx->invalid_counter(),
                     nodes_dev, tris_dev, ctx->work_counter(), ds.spheres.ptr, ds.num_spheres, ds.num_tris,
                     ds.two_level ? ds.insts.ptr : nullptr, sp, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "count launch");
    return spk_synchronize(ctx);
}

}  // extern "C"

int spk_context::check_invalid_rays() {
    uint32_t bad = 0;
    hipError_t e = hipMemcpy(&bad, counters.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return spk::hip_fail(this, e, "read invalid-ray counter");
    if (bad == 0) return SPK_OK;
    e = hipMemset(counters.ptr, 0, sizeof(uint32_t));
    if (e != hipSuccess) return spk::hip_fail(this, e, "reset invalid-ray counter");
    return spk::set_error(this, SPK_ERR_INVALID_RAY, "%u rays had a NaN/inf component or zero direction", bad);
}

extern "C" {

int spk_get_trace_config(spk_context *ctx, spk_trace_config *out) {
    if (!ctx || !out) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    out->bvh_width = (uint32_t)ctx->scene.variant.width;
    out->stack = (uint32_t)ctx->scene.variant.stack;
    out->block = (uint32_t)ctx->scene.variant.block;
    out->refill = (uint32_t)ctx->scene.variant.refill;
    out->ray_sort = (uint32_t)ctx->scene.variant.sort;
    return SPK_OK;
}

int spk_list_trace_configs(spk_context *ctx, spk_trace_config *out, uint32_t max_out, uint32_t *count) {
    if (!ctx || !count) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    TraceVariant vs[kMaxVariants];
    const int nv = std::min(kMaxVariants, trace_variants(vs, kMaxVariants, ctx->scene.width, ctx->scene.two_level));
    uint32_t k = 0;
    for (int i = 0; i < nv; ++i) {
        if (vs[i].stack < ctx->scene.depth) continue;
        if (out && k < max_out)
            out[k] = spk_trace_config{(uint32_t)vs[i].width, (uint32_t)vs[i].stack, (uint32_t)vs[i].block,
                                      (uint32_t)vs[i].refill, (uint32_t)vs[i].sort};
        ++k;
    }
    *count = k;
    return SPK_OK;
}

int spk_set_trace_config(spk_context *ctx, const spk_trace_config *config) {
    if (!ctx || !config) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (config->bvh_width != 4 && config->bvh_width != 8)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "bvh width must be 4 or 8");
    if ((int)config->bvh_width != ctx->scene.width) {
        HostScene &hs = ctx->host_scene;
        hs.input.width = config->bvh_width;
        std::string err;
        int stack;
        if (hs.two_level) {
            int rc = build_blases(hs, hs.input.width, hs.input.max_leaf_tris, hs.input.num_threads, err);
            if (rc == SPK_OK) rc = assemble_two_level(hs, hs.input.width, hs.input.num_threads, err);
            if (rc != SPK_OK) return set_error(ctx, rc, "%s", err.c_str());
            stack = pick_trace_stack((int)hs.depth2, true);
        } else {
            err = build_wide_bvh(hs.input, hs.bvh);
            if (!err.empty()) return set_error(ctx, SPK_ERR_INVALID_SCENE, "%s", err.c_str());
            stack = pick_trace_stack((int)hs.bvh.depth);
        }
        if (stack < 0) return set_error(ctx, SPK_ERR_STACK_DEPTH, "BVH too deep for the compiled stacks");
        int rc = upload_built(ctx, stack);
        if (rc != SPK_OK) return rc;
        ctx->on_scene_changed();
    }
    if ((int)config->stack < ctx->scene.depth)
        return set_error(ctx, SPK_ERR_STACK_DEPTH, "stack %u is smaller than the tree depth %d", config->stack,
                         ctx->scene.depth);
    TraceVariant vs[kMaxVariants];
    const int nv = std::min(kMaxVariants, trace_variants(vs, kMaxVariants, ctx->scene.width, ctx->scene.two_level));
    for (int i = 0; i < nv; ++i) {
        if (vs[i].stack == (int)config->stack && vs[i].block == (int)config->block &&
            vs[i].refill == (int)config->refill && vs[i].sort == (int)(config->ray_sort != 0)) {
            ctx->scene.variant = vs[i];
            ctx->on_scene_changed();
            return SPK_OK;
        }
    }
    return set_error(ctx, SPK_ERR_UNSUPPORTED, "no compiled variant with stack %u, block %u, refill %u, ray sort %u",
                     config->stack, config->block, config->refill, config->ray_sort);
}

static int time_variant(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, double &ms) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return set_error(ctx, SPK_ERR_HIP, "hipEventCreate");
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return set_error(ctx, SPK_ERR_HIP, "hipEventCreate");
    }
    hipError_t e = ctx->tmp_hits.ensure(n * sizeof(spk_hit));
    RaySortScratch ss;
    const RaySortScratch *sp = nullptr;
    if (e == hipSuccess) e = sort_scratch(ctx, n, ss, sp);
    double best = 1e300;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
        e = hipEventRecord(a, ctx->stream);
        if (e == hipSuccess)
            e = launch_trace(ctx->scene.variant, rays_dev, n, ctx->tmp_hits.ptr, ctx->scene.nodes.ptr,
                             ctx->scene.tris.ptr, 0u, ctx->invalid_counter(), nullptr, nullptr, ctx->work_counter(),
                             ctx->scene.spheres.ptr, ctx->scene.num_spheres, ctx->scene.num_tris,
                             ctx->scene.two_level ? ctx->scene.insts.ptr : nullptr, sp, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
        if (rep > 0) best = std::min(best, (double)t);  // first launch warms caches
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(ctx, e, "autotune timing");
    ms = best;
    return SPK_OK;
}

int spk_autotune(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, uint32_t flags, spk_trace_config *chosen,
                 double *best_ms) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (!rays_dev || n == 0) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "autotune needs rays");
    const int widths[2] = {ctx->scene.width, ctx->scene.width == 8 ? 4 : 8};
    const int nw = (flags & SPK_TUNE_WIDTH) ? 2 : 1;
    spk_trace_config best{0, 0, 0, 0, 0};
    double best_t = 1e300;
    for (int wi = 0; wi < nw; ++wi) {
        spk_trace_config cfg{(uint32_t)widths[wi], 64u, 64u, 0u, 0u};
        if ((int)cfg.bvh_width != ctx->scene.width) {
            int rc = spk_set_trace_config(ctx, &cfg);
            if (rc != SPK_OK) return rc;
        }
        spk_trace_config list[kMaxVariants];
        uint32_t cnt = 0;
        int rc = spk_list_trace_configs(ctx, list, kMaxVariants, &cnt);
        if (rc != SPK_OK) return rc;
        for (uint32_t i = 0; i < std::min<uint32_t>(cnt, kMaxVariants); ++i) {
            rc = spk_set_trace_config(ctx, &list[i]);
            if (rc != SPK_OK) return rc;
            double ms = 0;
            rc = time_variant(ctx, rays_dev, n, ms);
            if (rc != SPK_OK) return rc;
            if (ms < best_t) {
                best_t = ms;
                best = list[i];
            }
        }
    }
    int rc = spk_set_trace_config(ctx, &best);
    if (rc != SPK_OK) return rc;
    rc = ctx->check_invalid_rays();
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return rc;
}

}  // extern "C"
