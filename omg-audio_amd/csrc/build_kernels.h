// Device-side scene build: tessellation of oversized triangles, the 16-bin
// SAH tree built level by level, its breadth-first flatten into paired
// nodes and the 4-wide collapse -- the same primitives, tree and node
// numbering as the host build (scene_build.cpp), computed on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "omg_layout.h"

namespace omg {

// deepest tessellation recursion of one triangle (each level halves its
// longest edge)
constexpr uint32_t kTessMaxDepth = 60;
constexpr uint32_t kSegInactive = 0xFFFFFFFFu;
constexpr uint32_t kNoCut = 0xFFFFFFFFu;

// error bits of a device build
constexpr uint32_t kBuildErrIndex = 1u;      // triangle index >= nv
constexpr uint32_t kBuildErrMaterial = 2u;   // tri_material >= nm
constexpr uint32_t kBuildErrDepth = 4u;      // tessellation deeper than kTessMaxDepth
constexpr uint32_t kBuildErrFinite = 8u;     // non-finite vertex coordinate
constexpr uint32_t kBuildErrRefit = 16u;     // refit changed a triangle's patch count

// One primitive during the build: its box and centroid, its index in the
// tessellated order, its segment at the current level and its SAH bin.
struct alignas(16) BRec {
    float lo[3];
    float cx;
    float hi[3];
    float cy;
    float cz;
    uint32_t idx;
    uint32_t seg;
    uint32_t bin;
};
static_assert(sizeof(BRec) == 48, "BRec is 48 B");

// One segment that splits at the current level.
struct ISplit {
    uint32_t start, count, axis, mode;  // mode 0: SAH bins, 1: median
    float cmin_a, cwidth;
    uint32_t cut, mid;                  // cut = last left bin, kNoCut for a median split
};

// Device results of a build, read back once at the end.
struct BuildMeta {
    uint32_t err;
    uint32_t root_ref;
    float root_min[3], root_max[3];
    uint32_t stack4;
    uint32_t pad;
};

// One patch of a triangle that splits, during the breadth-first
// tessellation: its corners, the triangle's slot (nt-1-t, the host's pop
// order) and whether it is final.
struct alignas(16) TessItem {
    float a[3];
    uint32_t slot;
    float b[3];
    uint32_t done;
    float c[3];
    uint32_t pad;
};
static_assert(sizeof(TessItem) == 48, "TessItem is 48 B");

struct TessArgs {
    const float* pos;
    uint32_t nv;
    const uint32_t* idx;
    uint32_t nt;
    const uint16_t* tri_mat;
    const uint16_t* tri_surf;      // may be null (surface 0)
    const float* tri_max_edge;     // may be null (4 m)
    uint32_t nm;
    uint32_t* cnt_rev;             // [nt + 1]: patches of triangle nt-1-i at i
    uint32_t* big;                 // [nt + 1]: 1 when that triangle splits
    const uint32_t* big_off;       // exclusive scan of big: its first work item
    uint32_t* first_of;            // [nt]: its first patch in the final item list
    const uint32_t* off_rev;       // exclusive scan of cnt_rev
    TriInfo* tris;                 // [nt]
    uint32_t* tri_surf32;          // [nt], may be null
    Prim* prims;                   // [n] tessellated order (build) or leaf order (refit)
    const uint32_t* leaf_of;       // refit: leaf position of each tessellated patch
    BRec* recs;                    // build only
    BuildMeta* meta;
};

// One breadth-first round over the work list: every open item either
// stays (no split: final) or becomes its two halves, second half first.
struct TessRound {
    const TessItem* in;
    TessItem* out;
    uint32_t m;                    // items in
    uint32_t* oc;                  // [m + 1] items each one becomes
    const uint32_t* pos;           // exclusive scan of oc
    const float* tri_max_edge;
    uint32_t nt;
};

// Arguments of one level of the tree build.
struct LevelArgs {
    uint32_t n;                   // primitives
    uint32_t level;               // 0 = root
    uint32_t n_seg, n_int;        // segments and splitting segments at this level
    uint32_t base, base_prev;     // first node index of this level / the previous one
    const uint32_t* seg_start;
    const uint32_t* seg_count;
    const uint32_t* seg_irank;    // exclusive scan of (count > 4)
    uint32_t* seg_box;            // [n_seg][12] order-preserving keys
    ISplit* split;                // [n_int]
    uint32_t* bins;               // [n_int][16][8]
    uint32_t* nseg_start;         // [2 n_int] next level
    uint32_t* nseg_count;
    uint32_t* nseg_int;           // [2 n_int + 1]
    BRec* cur;                    // this level's records (k_bin stores each bin)
    BRec* out;                    // the next level's records
    uint32_t* flags;              // [n + 1] left-side flags of SAH-split primitives
    const uint32_t* lscan;        // exclusive scan of flags
    uint32_t* aR;                 // [n] front right-side positions per segment
    uint32_t* bL;                 // [n] back left-side positions per segment
    Node2* nodes;
    BuildMeta* meta;
};

struct Collapse4Args {
    const Node2* nodes;
    Node4* nodes4;
    uint32_t* src4;               // [n4][4]: (node2 << 1 | side) of each child slot (refit)
    uint32_t q, base4, base4_next;
    const uint32_t* queue;        // node2 of each 4-wide node at this level
    uint32_t* queue_next;
    const uint32_t* path;         // stack entries above this level's nodes
    uint32_t* path_next;
    uint32_t* n_int;              // [q + 1] internal children per node
    const uint32_t* int_scan;     // exclusive scan of n_int
    BuildMeta* meta;
};

// One level of the 16-wide derivation (Node16 from the 4-wide nodes).
struct Derive16Args {
    const Node4* nodes4;
    Node16* nodes16;
    uint32_t q, base16, base16_next;
    const uint32_t* queue;        // 4-wide node of each 16-wide node at this level
    uint32_t* queue_next;
    const uint32_t* path;         // stack entries above this level's nodes
    uint32_t* path_next;
    uint32_t* n_int;              // [q + 1] internal children per node
    const uint32_t* int_scan;     // exclusive scan of n_int
    uint32_t* stack16;            // worst-case stack (atomic max)
};

hipError_t launch_derive16_expand(const Derive16Args& A, hipStream_t s);
hipError_t launch_derive16_link(const Derive16Args& A, hipStream_t s);

hipError_t launch_tess_count(const TessArgs& A, hipStream_t s);
hipError_t launch_tess_seed(const TessArgs& A, TessItem* items, hipStream_t s);
hipError_t launch_tess_round_count(const TessRound& R, hipStream_t s);
hipError_t launch_tess_round_write(const TessRound& R, hipStream_t s);
hipError_t launch_tess_leaves(const TessArgs& A, const TessItem* items, uint32_t m, hipStream_t s);
hipError_t launch_tess_check(const uint32_t* cnt, const uint32_t* expect, uint32_t nt,
                             BuildMeta* meta, hipStream_t s);
hipError_t launch_tess_emit(const TessArgs& A, const TessItem* items, uint32_t m, hipStream_t s);
hipError_t launch_box_init(const LevelArgs& A, hipStream_t s);
hipError_t launch_box(const LevelArgs& A, hipStream_t s);
hipError_t launch_small(const LevelArgs& A, hipStream_t s);
hipError_t launch_seg(const LevelArgs& A, hipStream_t s);
hipError_t launch_bin(const LevelArgs& A, hipStream_t s);
hipError_t launch_split(const LevelArgs& A, hipStream_t s);
hipError_t launch_flag(const LevelArgs& A, hipStream_t s);
hipError_t launch_ptable(const LevelArgs& A, hipStream_t s);
hipError_t launch_permute(const LevelArgs& A, hipStream_t s);
hipError_t launch_gather_prims(const Prim* tess, const BRec* recs, uint32_t n, Prim* out,
                               uint32_t* leaf_of, hipStream_t s);
hipError_t launch_collapse_expand(const Collapse4Args& A, hipStream_t s);
hipError_t launch_collapse_link(const Collapse4Args& A, hipStream_t s);

// refit: boxes of one level of paired nodes from their children, deepest first
hipError_t launch_refit_level(Node2* nodes, const Prim* prims, uint32_t first, uint32_t count,
                              hipStream_t s);
hipError_t launch_refit_root(const Node2* nodes, const Prim* prims, uint32_t root_ref,
                             BuildMeta* meta, hipStream_t s);
hipError_t launch_refit4(const Node2* nodes, const uint32_t* src4, Node4* nodes4, uint32_t n4,
                         hipStream_t s);

// exclusive prefix sum of n words (device-wide); tmp grows as needed
hipError_t scan_u32(void*& tmp, size_t& tmp_bytes, const uint32_t* in, uint32_t* out, size_t n,
                    hipStream_t s);

}  // namespace omg
