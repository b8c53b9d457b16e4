// Internal state behind the opaque spk_context handle.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/spk_pt.h"
#include "bvh_build.h"
#include "spk_kernels.h"

namespace spk {

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer();
    void release();
    hipError_t ensure(size_t want);
    template <class T>
    T *as() const {
        return static_cast<T *>(ptr);
    }
};

struct ShadeSource {
    uint32_t inst, face;
};

// Host copy of the flattened scene and its BVH (kept for statistics and re-validation).
// Object-space copy of one mesh (two-level scenes re-flatten their shading data when instances move).
struct MeshCopy {
    std::vector<float> positions, normals, tangents, uvs;
    std::vector<uint32_t> indices;
    uint32_t num_vertices = 0, num_triangles = 0;
};

struct HostScene {
    std::vector<float> verts;
    std::vector<uint32_t> global_ids, materials, local_ids;
    std::vector<ShadeSource> shade_src;
    std::vector<TriShade> shade_in;  // per input triangle
    BuildInput input;
    BuildOutput bvh;
    // two-level scenes (spk_build_options.two_level): one BLAS per mesh in object space under a TLAS
    // whose leaf slots are instances; shading data stays per global triangle id in world space
    bool two_level = false;
    std::vector<MeshCopy> mesh_copies;
    std::vector<BuildOutput> blas;          // per mesh (empty for meshes no instance uses)
    BuildOutput tlas;                       // over the instances' world boxes, one per leaf slot
    std::vector<uint32_t> inst_mesh, inst_material, inst_id_base;
    std::vector<float> inst_xf;             // world_from_local, 12 floats per instance
    std::vector<uint8_t> nodes2;            // TLAS nodes, then every BLAS (child and triangle bases rebased)
    std::vector<TriRecord> tris2;           // object-space BLAS triangles, v0.w = face index in the mesh
    std::vector<float> inst_rec;            // 16 floats per TLAS leaf slot (see traverse.h)
    uint32_t num_nodes2 = 0, depth2 = 0;    // depth2: stack entries a two-level query may need
};

struct DeviceScene {
    DeviceBuffer nodes, tris, shade, materials, emitters;
    DeviceBuffer insts, wtris;  // two-level: instance records; world triangles in global-id order
    bool two_level = false;
    float box_lo[3] = {0, 0, 0}, box_inv[3] = {1, 1, 1};  // scene box (origin cells of the ray sort)
    DeviceBuffer spheres;  // 2 x float4 per sphere: (center, r^2), (material bits, 0, 0, 0)
    uint32_t num_tris = 0, num_nodes = 0, num_materials = 0, num_emitters = 0, num_spheres = 0;
    int width = 8, depth = 0;
    TraceVariant variant{8, 16, 128};
    bool valid = false;
    std::vector<uint32_t> host_emitters;
};

struct PathState;   // path-tracer buffers, defined with the path-tracer kernels
struct CacheState;  // learned radiance cache (spk_cache.h)
struct BatchState;  // batch of independent scenes, defined with the path-tracer kernels
struct HostStaging; // pinned ring buffers and copy streams of host-buffer closest-hit calls

void destroy_host_staging(HostStaging *hs);
// spk_trace_closest_host over the pinned staging pipeline (n > 0, arrays checked by the caller).
int trace_closest_host_staged(spk_context *ctx, const spk_ray *rays, uint64_t n, spk_hit *hits);

// Validates and flattens one scene, builds its BVH, converts shading attributes (input order) and
// materials. Returns an SPK_* code and a message in err.
int prepare_scene(const spk_mesh_desc *meshes, uint32_t num_meshes, const spk_instance_desc *instances,
                  uint32_t num_instances, const spk_material_desc *materials, uint32_t num_materials,
                  uint32_t num_textures, const spk_build_options *options, HostScene &hs,
                  std::vector<MaterialGpu> &mats, std::string &err);
// Shading attributes permuted to the BVH's leaf order.
std::vector<TriShade> leaf_order_shade(const HostScene &hs);

}  // namespace spk

struct spk_context {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string error;
    spk::DeviceBuffer counters;  // [0] invalid rays, [2..3] ray counter of the persistent kernels
    spk::DeviceBuffer tmp_rays, tmp_hits;
    spk::DeviceBuffer sort_keys, sort_keys1, sort_perm1, sort_perm, sort_offs1, sort_offs, sort_total;  // coherence order
    void *noalloc_scratch = nullptr;  // this device's copy of the no-allocation entry point's scratch
    std::vector<hipEvent_t> kernel_events;  // spk_trace_kernel_timing: begin / end per timed call
    uint32_t kernel_timed = 0;              // calls recorded so far
    spk::DeviceScene scene;
    spk::DeviceBuffer texels, tex_table;  // texture table (spk_textures_upload)
    uint32_t num_textures = 0;
    int32_t sky_faces[6] = {-1, -1, -1, -1, -1, -1};  // sky cube map faces in the texture table
    spk::HostScene host_scene;
    std::vector<spk::MaterialGpu> host_mats;
    spk_scene_stats stats{};
    spk_camera_desc camera{};
    bool has_camera = false;
    spk_schedule schedule{8u, 1u, 16u, 0u};  // spk_render_frame's launch schedule
    spk::PathState *paths = nullptr;
    spk::CacheState *cache = nullptr;
    spk::BatchState *batch = nullptr;
    spk::HostStaging *staging = nullptr;
    float cache_last_loss = 0.f;

    uint32_t *invalid_counter() { return static_cast<uint32_t *>(counters.ptr); }
    unsigned long long *work_counter() {
        return reinterpret_cast<unsigned long long *>(static_cast<uint32_t *>(counters.ptr) + 2);
    }
    int check_invalid_rays();
    void on_scene_changed();
    void drop_batch();
    void release_all();
};

namespace spk {
int set_error(spk_context *ctx, int code, const char *fmt, ...);
int hip_fail(spk_context *ctx, hipError_t e, const char *what);
}  // namespace spk
