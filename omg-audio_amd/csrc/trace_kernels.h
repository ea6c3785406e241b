// Device traversal and tracer kernels (wave64, gfx950).
#pragma once
#include "omg_layout.h"
#include "omg_numerics.h"

namespace omg {

// Geometry of one job: a flattened mesh BVH, or the analytic Shoebox.
struct GeomArgs {
    const Node2* nodes;
    const Prim* prims;
    const TriInfo* tris;
    const Node4* nodes4;     // 4-wide collapse of the same tree, or null
    uint32_t root_ref;
    uint32_t root4;          // root of nodes4 (a leaf ref for one-leaf trees)
    uint32_t empty;          // mesh without primitives: every ray escapes
    float box[3];            // Shoebox size (box launches)
    float root_min[3];       // bounds of the whole tree (the root's own box)
    float root_max[3];
};

// Everything one tracer job (one trace_world call, or one query of a batch)
// needs per ray.
struct JobView {
    GeomArgs g;
    const MatParams* mats;   // scene materials, then overlay materials (box: 6 walls first)
    const Overlay* overlays;
    uint32_t n_ovl;
    const float* sources;    // [n_src][3]
    const float* listeners;  // [n_lis][3]
    uint32_t n_src, n_lis;
    uint32_t n_rays, nbands, max_bounces, dir_band;
    uint32_t key0, key1;     // Philox key = job seed
    float bin_dt, max_time, floor_e, radius;
    float energy0[kMaxBands];  // per_ray * source energy (tracer.rs:95); 0 beyond nbands
    uint32_t nbins;
    float* bins;             // [n_src][n_lis][nbins][nbands]
    float* dirs;             // [n_src][n_lis][nbins][3]
    uint32_t* records;       // [n_src * n_rays][4] or null
};

struct TraceArgs : JobView {
    unsigned long long* counters;  // [3] or null
    uint32_t ray_base;       // first global ray of this launch (chunked launches)
    uint32_t ray_count;      // rays in this launch
    uint32_t node_cache;     // BVH nodes (top levels) staged in LDS per block
    uint32_t stack_lds;      // traversal stack entries per ray kept in LDS (16 or 32)
    uint32_t lanes;          // lanes per ray: 1, or 4 (a DPP quad)
    uint32_t width;          // BVH width walked: 2 (paired nodes) or 4 (nodes4)
    uint32_t* hit_ids;       // [rays][max_bounces] hit id per bounce (0xFFFFFFFF none), or null
    float* hit_t;            // [rays][max_bounces] hit distance per bounce, or null
};

// LDS stack entries per ray of the quad 4-wide walk that never spills: used
// when the tree's worst-case stack fits (every scene measured so far needs
// at most 39)
constexpr uint32_t kDeepStack = 40;

// One unit of batched work: up to kBatchItemRays consecutive rays of one
// source of one query.
constexpr uint32_t kBatchItemRays = 64;

struct BatchItem {
    uint32_t query;
    uint32_t src;
    uint32_t ray_start;
    uint32_t count;
};

struct BatchArgs {
    const JobView* jobs;     // [n_queries]
    const BatchItem* items;  // [n_items]
    uint32_t n_items;
    uint32_t* next_item;     // work-stealing cursor, zero before each launch
    unsigned long long* counters;  // [3] or null
};

// Launch shape of the batched kernel (chosen by the autotuner).
struct BatchShape {
    uint32_t lanes;           // lanes per ray: 1 or 4
    uint32_t stack_lds;       // stack entries per ray in LDS (see valid_batch_shape)
    uint32_t waves_per_simd;  // resident waves per SIMD the grid is sized for
    uint32_t regen;           // 1: finished rays are replaced between bounces
    uint32_t width;           // BVH width walked: 2 or 4 (every scene of the batch has nodes4)
};

// The compiled shapes.  Without regen: lanes 1|4 with 16|32 stack entries
// over the binary tree, 16 over the 4-wide tree, or a quad with kDeepStack
// entries over the 4-wide tree.  With regen: one lane per ray with 8|12|16
// entries over either tree, or a quad per ray with 16 (binary tree) or
// kDeepStack (4-wide tree).  A kDeepStack shape is valid only for scenes
// whose 4-wide tree needs at most kDeepStack stack entries.
inline bool valid_batch_shape(const BatchShape& s)
{
    if (s.waves_per_simd < 1 || s.waves_per_simd > 8 || s.regen > 1) return false;
    if (s.width != 2 && s.width != 4) return false;
    if (s.lanes == 4 && s.width == 4 && s.stack_lds == kDeepStack) return true;
    if (s.regen && s.lanes == 4) return s.width == 2 && s.stack_lds == 16;
    if (s.regen) return s.lanes == 1 && (s.stack_lds == 8 || s.stack_lds == 12 || s.stack_lds == 16);
    if (s.width == 4) return (s.lanes == 1 || s.lanes == 4) && s.stack_lds == 16;
    return (s.lanes == 1 || s.lanes == 4) && (s.stack_lds == 16 || s.stack_lds == 32);
}

hipError_t launch_trace(const TraceArgs& A, bool mesh, bool count, hipStream_t s);
// persistent work-stealing launch over n_cu compute units; nbm = the largest
// band count of the batch
hipError_t launch_trace_batch(const BatchArgs& B, uint32_t nbm, const BatchShape& shape,
                              uint32_t n_cu, bool count, hipStream_t s, uint32_t max_blocks = 0);
// A plain job run by the persistent regenerating kernel: d_job (one
// JobView), d_items (ceil(n_rays / 64) x n_src items) and d_cursor (one
// word) are device memory the caller keeps until the launch has run.
hipError_t launch_trace_routed(const TraceArgs& A, uint32_t nbm, uint32_t width, uint32_t n_cu,
                               JobView* d_job, BatchItem* d_items, uint32_t* d_cursor,
                               bool count, hipStream_t s);

}  // namespace omg
