// Host BVH builder: binned-SAH binary tree, greedily collapsed to a 4- or 8-wide tree whose child
// boxes are quantised conservatively to 8 bits. Replaces the madmann91/bvh DefaultBuilder calls
// of TLAS::Build / BLAS::Build (SceneRenderProxy.cpp:31-69, MeshRenderProxy.cpp:31-65) for the
// flattened device scene.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "spk_layout.h"

namespace spk {

struct BuildInput {
    const float *tri_verts = nullptr;   // 9 floats per triangle (v0, v1, v2), world space
    const uint32_t *global_ids = nullptr;
    const uint32_t *materials = nullptr;
    const uint32_t *local_ids = nullptr;
    uint32_t num_tris = 0;
    uint32_t width = 8;          // 4 or 8
    uint32_t max_leaf_tris = 3;  // 1..3
    uint32_t num_threads = 0;
};

struct BuildOutput {
    std::vector<uint8_t> nodes;      // num_nodes * node_bytes
    uint32_t node_bytes = 0;
    uint32_t num_nodes = 0;
    std::vector<TriRecord> tris;     // leaf order
    std::vector<uint32_t> tri_input; // leaf slot -> input triangle index
    uint32_t depth = 0;
    double sah_cost = 0.0;
    uint64_t degenerate = 0;
    double seconds = 0.0;
};

// Returns an empty string on success, else an error message.
std::string build_wide_bvh(const BuildInput &in, BuildOutput &out);

// Checks the invariants the traversal relies on: every triangle is referenced exactly once and
// lies inside the dequantised box of every ancestor slot; child indices are in range; the tree
// depth is as reported. Returns an empty string when all hold.
std::string validate_wide_bvh(const BuildInput &in, const BuildOutput &out);

}  // namespace spk
