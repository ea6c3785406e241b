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
    const Node16* nodes16;   // 16-wide nodes derived from nodes4, or null
    uint32_t root_ref;
    uint32_t root4;          // root of nodes4 (a leaf ref for one-leaf trees)
    uint32_t root16;         // root of nodes16 (likewise)
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
    uint32_t lanes;          // lanes per ray: 1, 4 (a DPP quad) or 16 (a DPP row, width 16)
    uint32_t width;          // BVH width walked: 2 (paired nodes), 4 (nodes4) or 16 (nodes16)
    uint32_t* hit_ids;       // [rays][max_bounces] hit id per bounce (0xFFFFFFFF none), or null
    float* hit_t;            // [rays][max_bounces] hit distance per bounce, or null
    // Hand-off to the 16-lane walk: with handoff_bounce > 0 a ray still
    // alive after that many bounces saves its state (kHandoffF4 float4s at
    // handoff + kHandoffF4 * k, k from handoff_ctl[0]) instead of going on,
    // and k_trace_resume16 (resume_blocks blocks, node_cache16 16-wide nodes
    // in LDS) continues the saved rays, one 16-lane row each, taking them
    // from the cursor handoff_ctl[1]; handoff_ctl[2 + s] counts the working
    // waves of SIMD s (kResumeAdmitWords of them).  All kHandoffCtlWords
    // words start at zero.
    uint32_t handoff_bounce;
    float4* handoff;
    uint32_t* handoff_ctl;
    uint32_t resume_blocks;
    uint32_t node_cache16;
    // lanes per saved ray in k_trace_resume16: 16 (one slot of the 16-wide
    // node per lane) or 8 (two slots per lane); 0 means 16
    uint32_t resume_lanes;
    // LDS stack entries per row: the scene's worst-case 16-wide stack
    // (<= kStack16); 0 means kStack16
    uint32_t resume_stack;
    // SIMDs of the device (4 per CU): k_trace_resume16 admits at most
    // ceil(saved rays / (SIMDs x rays per wave)) working waves per SIMD, so
    // a few saved rays spread over the chip instead of piling onto the
    // SIMDs whose waves reach the queue first; 0 admits every wave
    uint32_t resume_simds;
};

// Hand-off control words: the saved-ray count, the cursor, then one
// working-wave counter per SIMD, indexed by the SIMD's hardware ids (XCC,
// shader engine, shader array, CU, SIMD: 13 bits).
constexpr uint32_t kResumeAdmitWords = 1u << 13;
constexpr uint32_t kHandoffCtlWords = 2u + kResumeAdmitWords;

// float4s of one saved ray state (position, distance, direction, bounce,
// termination, hash, flags, source, ray, up to kMaxBands energies)
constexpr uint32_t kHandoffF4 = 6;

// LDS stack entries per ray of the quad 4-wide walk that never spills: used
// when the tree's worst-case stack fits (every scene measured so far needs
// at most 39)
constexpr uint32_t kDeepStack = 40;

// LDS stack entries per ray of the 16-lane walk over the 16-wide nodes: a
// scene walks it when its worst-case stack fits
constexpr uint32_t kStack16 = 96;

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
    uint32_t lanes;           // lanes per ray: 1, 4 or 16
    uint32_t stack_lds;       // stack entries per ray in LDS (see valid_batch_shape)
    uint32_t waves_per_simd;  // resident waves per SIMD the grid is sized for
    uint32_t regen;           // 1: finished rays are replaced between bounces; 2: a lane's
                              // bounces follow each other at walk steps (one lane, 4-wide)
    uint32_t width;           // BVH width walked: 2, 4 or 16 (every scene of the batch has it)
};

// The compiled shapes.  Without regen: lanes 1|4 with 16|32 stack entries
// over the binary tree, 16 over the 4-wide tree, or a quad with kDeepStack
// entries over the 4-wide tree.  With regen: one lane per ray with 8|12|16
// entries over either tree, or a quad per ray with 16 (binary tree) or
// kDeepStack (4-wide tree), or a 16-lane row per ray over the 16-wide
// tree with kStack16 entries.  Regen 2 (rays shaded and replaced at walk
// steps): one lane per ray over the 4-wide tree with 12|16 entries.  A
// kDeepStack shape is valid only for scenes whose 4-wide tree needs at
// most kDeepStack stack entries.
inline bool valid_batch_shape(const BatchShape& s)
{
    if (s.waves_per_simd < 1 || s.waves_per_simd > 8 || s.regen > 2) return false;
    if (s.regen == 2) return s.lanes == 1 && s.width == 4 && (s.stack_lds == 12 || s.stack_lds == 16);
    if (s.lanes == 16 || s.width == 16)
        return s.lanes == 16 && s.width == 16 && s.regen == 1 && s.stack_lds == kStack16;
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
