// Host-side scene construction (see scene_build.cpp).
#pragma once
#include <string>
#include <vector>

#include "omg_layout.h"
#include "omg_numerics.h"

namespace omg {

struct HostScene {
    std::vector<Prim> prims;       // leaf order
    std::vector<Node2> nodes;      // internal nodes, preorder
    std::vector<TriInfo> tris;     // per original triangle
    uint32_t root_ref = 0;
    float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
    uint32_t max_depth = 0;
    uint32_t n_bin_nodes = 0;
    bool empty = false;
    // 4-wide collapse of the same tree (breadth-first); has4 when its
    // worst-case traversal stack fits kStackCap
    std::vector<Node4> nodes4;
    uint32_t root4 = 0, stack4 = 0;
    bool has4 = false;
};

// MeshBuilder::vertex dedup over a triangle soup; returns the vertex count.
uint32_t dedup_soup(const float* tri_verts, uint32_t nt, std::vector<float>& positions,
                    std::vector<uint32_t>& indices);

// Tessellate, build the SAH tree and flatten it.  Returns 0 or -1 (err set).
int build_scene(const float* positions, uint32_t nv, const uint32_t* indices, uint32_t nt,
                const uint16_t* tri_material, const uint16_t* tri_surface,
                const float* tri_max_edge, HostScene& out, std::string& err);

// Tracer constants of one material (see MatParams).
void mat_params(const float* absorption, float scattering, const float* transmission,
                uint32_t nbands, MatParams& out);

}  // namespace omg
