// Device-resident data layout of the path-tracer hot path, shared by the host builder and the
// HIP kernels. Everything here is plain data: no HIP or C++ library types.
#pragma once

#include <stdint.h>

namespace spk {

// Wide BVH node with 8-bit quantised child boxes, 80 bytes = five 16-byte loads.
// Field set follows the compressed wide BVH of Ylitie, Karras and Laine (HPG 2017): a per-node
// origin p and power-of-two scale 2^e per axis; child c spans [p + qlo[c]*2^e, p + qhi[c]*2^e].
// imask bit s marks slot s as an inner node; inner children are stored contiguously from
// child_base in slot order. For a leaf slot meta[s] = (count << 5) | offset: `count` triangles
// starting at tri_base + offset. meta[s] == 0 marks an empty slot.
// Slots are assigned so that, for a ray whose direction-sign bits are `oct` (bit a set when
// dir[a] < 0), visiting slots in the order s = r ^ oct, r = 0..7, is roughly front to back.
struct alignas(16) NodeW8 {
    float px, py, pz;
    int8_t ex, ey, ez;
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo_x[8], qlo_y[8], qlo_z[8];
    uint8_t qhi_x[8], qhi_y[8], qhi_z[8];
};
static_assert(sizeof(NodeW8) == 80, "NodeW8 must be 80 bytes");

// 4-wide variant of the same scheme, 64 bytes = four 16-byte loads.
struct alignas(16) NodeW4 {
    float px, py, pz;
    int8_t ex, ey, ez;
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[4];
    uint8_t qlo_x[4], qlo_y[4], qlo_z[4];
    uint8_t qhi_x[4], qhi_y[4], qhi_z[4];
    uint32_t pad[3];
};
static_assert(sizeof(NodeW4) == 64, "NodeW4 must be 64 bytes");

// Triangle record in leaf order: three float4 vertices; v0.w carries the global triangle id
// (bit pattern), v1.w the material id, v2.w the scene-local triangle id.
struct alignas(16) TriRecord {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(TriRecord) == 48, "TriRecord must be 48 bytes");

// Shading attributes of one triangle (fetched once per hit, never during traversal).
struct alignas(16) TriShade {
    float n0[4], n1[4], n2[4];  // vertex normals (w unused)
    float t0[4], t1[4], t2[4];  // vertex tangents
    float uv0[2], uv1[2], uv2[2];
    uint32_t material;
    uint32_t pad;
};
static_assert(sizeof(TriShade) == 128, "TriShade must be 128 bytes");

struct MaterialGpu {
    uint32_t type;
    float base[3];
    float emissive[3];
    float metallic;
    float roughness;
    float eta;
    uint32_t emissive_flag;  // |emissive factor| > 1e-6: may emit (with a texture: decided per hit)
    uint32_t pad;
    int32_t tex_base, tex_mr, tex_emissive, tex_normal;  // texture table indices, -1 = none
};
static_assert(sizeof(MaterialGpu) == 64, "MaterialGpu must be 64 bytes");

// One texture of the texture table: texels (float4, converted at upload) start at `offset` of the
// texel array, row-major.
struct TexRecord {
    uint64_t offset;
    uint32_t width, height;
    uint32_t srgb;  // apply the sRGB decode after filtering
    uint32_t pad[3];
};
static_assert(sizeof(TexRecord) == 32, "TexRecord must be 32 bytes");

}  // namespace spk
