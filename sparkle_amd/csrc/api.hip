// C ABI of the path-tracer hot path (include/spk_pt.h): context, scene upload, closest-hit queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spk_pt.h"
#include "bvh_build.h"
#include "spk_context.h"
#include "spk_kernels.h"

using namespace spk;

namespace spk {

constexpr int kMaxVariants = 256;

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

// Flattens instances into world-space triangles and builds the wide BVH on the host.
static int flatten_and_build(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                             uint32_t num_instances, const spk_build_options *options, HostScene &hs,
                             std::string &err) {
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
        static const float I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        const bool ident = std::memcmp(M, I, sizeof(I)) == 0;
        for (uint32_t f = 0; f < md.num_triangles; ++f, ++g) {
            for (int k = 0; k < 3; ++k) {
                const float *p = md.positions + 3 * (size_t)md.indices[3 * (size_t)f + k];
                float *dst = &hs.verts[9 * (size_t)g + 3 * k];
                if (ident) {
                    dst[0] = p[0], dst[1] = p[1], dst[2] = p[2];
                } else {
                    dst[0] = M[0] * p[0] + M[1] * p[1] + M[2] * p[2] + M[3];
                    dst[1] = M[4] * p[0] + M[5] * p[1] + M[6] * p[2] + M[7];
                    dst[2] = M[8] * p[0] + M[9] * p[1] + M[10] * p[2] + M[11];
                }
            }
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
    err = build_wide_bvh(in, hs.bvh);
    if (!err.empty()) return SPK_ERR_INVALID_SCENE;
    return SPK_OK;
}

static void fill_stats(const HostScene &hs, spk_scene_stats *st, uint64_t device_bytes) {
    std::memset(st, 0, sizeof(*st));
    st->num_triangles = hs.bvh.tris.size();
    st->num_nodes = hs.bvh.num_nodes;
    st->bvh_width = hs.input.width;
    st->depth = hs.bvh.depth;
    st->node_bytes = hs.bvh.node_bytes;
    st->triangle_bytes = sizeof(TriRecord);
    st->device_bytes = device_bytes;
    st->degenerate_tris = hs.bvh.degenerate;
    st->sah_cost = hs.bvh.sah_cost;
    st->build_seconds = hs.bvh.seconds;
}

// Uploads the host scene's current BVH (nodes, triangles, leaf-order shading) and materials.
static int upload_built(spk_context *ctx, int stack) {
    HostScene &hs = ctx->host_scene;
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

    int rc = flatten_and_build(meshes, num_meshes, instances, num_instances, options, hs, err);
    if (rc != SPK_OK) return rc;
    if (pick_trace_stack((int)hs.bvh.depth) < 0) {
        err = "BVH depth " + std::to_string(hs.bvh.depth) + " exceeds the largest traversal stack (" +
              std::to_string(max_trace_stack()) + ")";
        return SPK_ERR_STACK_DEPTH;
    }

    // shading attributes in input order (permuted to leaf order at upload)
    const uint32_t n = (uint32_t)hs.bvh.tris.size();
    hs.shade_in.assign(n, TriShade());
    for (uint32_t src = 0; src < n; ++src) {
        const uint32_t inst = hs.shade_src[src].inst, f = hs.shade_src[src].face;
        const spk_instance_desc &ins = instances[inst];
        const spk_mesh_desc &md = meshes[ins.mesh];
        const float *M = ins.world_from_local;
        TriShade &s = hs.shade_in[src];
        std::memset(&s, 0, sizeof(s));
        float *ns[3] = {s.n0, s.n1, s.n2}, *ts[3] = {s.t0, s.t1, s.t2}, *uvs[3] = {s.uv0, s.uv1, s.uv2};
        const float *v = &hs.verts[9 * (size_t)src];
        // geometric normal when the mesh has none
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
                // default tangent (1, 0, 0) through the 3x3 part
                ts[c][0] = M[0], ts[c][1] = M[4], ts[c][2] = M[8];
                ts[c][3] = 1.f;
            }
            if (md.uvs) {
                uvs[c][0] = md.uvs[2 * (size_t)vi];
                uvs[c][1] = md.uvs[2 * (size_t)vi + 1];
            }
        }
        s.material = ins.material;
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

int spk_trace_closest_async(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (n && (!rays_dev || !hits_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray or hit array");
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    const DeviceScene &ds = ctx->scene;
    hipError_t e = launch_trace(ds.variant, rays_dev, n, hits_dev, ds.nodes.ptr, ds.tris.ptr, 0u, ctx->invalid_counter(),
                                nullptr, nullptr, ctx->work_counter(), ds.spheres.ptr, ds.num_spheres, ds.num_tris,
                                ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    return SPK_OK;
}

int spk_trace_closest(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, spk_hit *hits_dev) {
    int rc = spk_trace_closest_async(ctx, rays_dev, n, hits_dev);
    if (rc != SPK_OK) return rc;
    return spk_synchronize(ctx);
}

int spk_trace_closest_host(spk_context *ctx, const spk_ray *rays, uint64_t n, spk_hit *hits) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (n == 0) return SPK_OK;
    if (!rays || !hits) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray or hit array");
    hipError_t e;
    if ((e = ctx->tmp_rays.ensure(n * sizeof(spk_ray))) != hipSuccess ||
        (e = ctx->tmp_hits.ensure(n * sizeof(spk_hit))) != hipSuccess)
        return hip_fail(ctx, e, "ray buffers");
    if ((e = hipMemcpyAsync(ctx->tmp_rays.ptr, rays, n * sizeof(spk_ray), hipMemcpyHostToDevice, ctx->stream)) !=
        hipSuccess)
        return hip_fail(ctx, e, "ray upload");
    int rc = spk_trace_closest_async(ctx, (const spk_ray *)ctx->tmp_rays.ptr, n, (spk_hit *)ctx->tmp_hits.ptr);
    if (rc != SPK_OK) return rc;
    if ((e = hipMemcpyAsync(hits, ctx->tmp_hits.ptr, n * sizeof(spk_hit), hipMemcpyDeviceToHost, ctx->stream)) !=
        hipSuccess)
        return hip_fail(ctx, e, "hit download");
    return spk_synchronize(ctx);
}

int spk_trace_count(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, uint32_t *nodes_dev, uint32_t *tris_dev) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    if (n && !rays_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null ray array");
    if (!nodes_dev && !tris_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "no count array");
    hipError_t e;
    if ((e = ctx->tmp_hits.ensure(n * sizeof(spk_hit))) != hipSuccess) return hip_fail(ctx, e, "hit buffer");
    const DeviceScene &ds = ctx->scene;
    e = launch_trace(ds.variant, rays_dev, n, ctx->tmp_hits.ptr, ds.nodes.ptr, ds.tris.ptr, 0u, ctx->invalid_counter(),
                     nodes_dev, tris_dev, ctx->work_counter(), ds.spheres.ptr, ds.num_spheres, ds.num_tris,
                     ctx->stream);
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
    return SPK_OK;
}

int spk_list_trace_configs(spk_context *ctx, spk_trace_config *out, uint32_t max_out, uint32_t *count) {
    if (!ctx || !count) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "no scene uploaded");
    TraceVariant vs[kMaxVariants];
    const int nv = std::min(kMaxVariants, trace_variants(vs, kMaxVariants, ctx->scene.width));
    uint32_t k = 0;
    for (int i = 0; i < nv; ++i) {
        if (vs[i].stack < ctx->scene.depth) continue;
        if (out && k < max_out)
            out[k] = spk_trace_config{(uint32_t)vs[i].width, (uint32_t)vs[i].stack, (uint32_t)vs[i].block,
                                      (uint32_t)vs[i].refill};
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
        std::string err = build_wide_bvh(hs.input, hs.bvh);
        if (!err.empty()) return set_error(ctx, SPK_ERR_INVALID_SCENE, "%s", err.c_str());
        const int stack = pick_trace_stack((int)hs.bvh.depth);
        if (stack < 0) return set_error(ctx, SPK_ERR_STACK_DEPTH, "BVH depth %u too deep", hs.bvh.depth);
        int rc = upload_built(ctx, stack);
        if (rc != SPK_OK) return rc;
        ctx->on_scene_changed();
    }
    if ((int)config->stack < ctx->scene.depth)
        return set_error(ctx, SPK_ERR_STACK_DEPTH, "stack %u is smaller than the tree depth %d", config->stack,
                         ctx->scene.depth);
    TraceVariant vs[kMaxVariants];
    const int nv = std::min(kMaxVariants, trace_variants(vs, kMaxVariants, ctx->scene.width));
    for (int i = 0; i < nv; ++i) {
        if (vs[i].stack == (int)config->stack && vs[i].block == (int)config->block &&
            vs[i].refill == (int)config->refill) {
            ctx->scene.variant = vs[i];
            ctx->on_scene_changed();
            return SPK_OK;
        }
    }
    return set_error(ctx, SPK_ERR_UNSUPPORTED, "no compiled variant with stack %u, block %u and refill %u",
                     config->stack, config->block, config->refill);
}

static int time_variant(spk_context *ctx, const spk_ray *rays_dev, uint64_t n, double &ms) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return set_error(ctx, SPK_ERR_HIP, "hipEventCreate");
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return set_error(ctx, SPK_ERR_HIP, "hipEventCreate");
    }
    hipError_t e = ctx->tmp_hits.ensure(n * sizeof(spk_hit));
    double best = 1e300;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
        e = hipEventRecord(a, ctx->stream);
        if (e == hipSuccess)
            e = launch_trace(ctx->scene.variant, rays_dev, n, ctx->tmp_hits.ptr, ctx->scene.nodes.ptr,
                             ctx->scene.tris.ptr, 0u, ctx->invalid_counter(), nullptr, nullptr, ctx->work_counter(),
                             ctx->scene.spheres.ptr, ctx->scene.num_spheres, ctx->scene.num_tris, ctx->stream);
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
    spk_trace_config best{0, 0, 0, 0};
    double best_t = 1e300;
    for (int wi = 0; wi < nw; ++wi) {
        spk_trace_config cfg{(uint32_t)widths[wi], 64u, 64u, 0u};
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
