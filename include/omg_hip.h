/*
 * omg_hip.h -- C ABI of the MI355X stochastic energy path tracer.
 *
 * A drop-in for omg-audio's tracer seams: a Rust `omg-hip` crate binds
 * these symbols with `extern "C"` and implements the reference's traits
 * over them (paths below are relative to the reference checkout).  Plain
 * pointers and sizes only; every call returns an int status (OMG_OK = 0,
 * negative = refused or failed: the caller keeps its CPU path, as the
 * reference's backends do by returning `false` / `None`).
 *
 * Conventions kept from the reference seams:
 *   - out-buffers are caller-owned and overwritten completely;
 *   - the static mesh is uploaded once (omg_scene_create), overlays,
 *     sources and listeners travel per call;
 *   - caps are refused, never clamped or silently truncated;
 *   - an omg_ctx is used by one thread at a time (the reference calls its
 *     backends from one sim thread under a Mutex, late.rs:51, :93);
 *     separate contexts may be used from separate threads.
 *
 * Echogram layout, per (source, listener) pair: bins [nbins][nbands] f32
 * then dirs [nbins][3] f32, bin-major, exactly the reference Echogram
 * (crates/omg-core/src/tracer.rs:18-25) with nbins = max_time / bin_dt.
 */
#ifndef OMG_HIP_H
#define OMG_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMG_LAYOUT_VERSION 1u   /* bump on any struct change (cf. layout.rs:12, :91) */
#define OMG_MAX_BANDS 8
#define OMG_MAX_OVERLAYS 64     /* panel overlay cap (layout.rs:94): refused above */
#define OMG_MAX_LISTENERS 64
#define OMG_DOME_BINS 9         /* dome.rs:30 */

#define OMG_OK 0
#define OMG_ERR_NO_DEVICE (-1)  /* no usable GPU (omg-gpu/src/lib.rs:33-47 returns None) */
#define OMG_ERR_INVALID (-2)    /* bad argument */
#define OMG_ERR_CAP (-3)        /* a cap was exceeded: refused (omg-gpu/src/lib.rs:519-528) */
#define OMG_ERR_VERSION (-4)    /* layout version mismatch (web/gpu.js:17-22) */
#define OMG_ERR_DEVICE (-5)     /* HIP runtime failure */
#define OMG_ERR_BUDGET (-6)     /* serving: HBM budget refusal */
#define OMG_ERR_NOT_READY (-7)  /* async: no fresh result yet (late.rs:46-48) */
#define OMG_ERR_OVERFLOW (-8)   /* a bounded device list overflowed; results not valid */

typedef struct omg_ctx omg_ctx;
typedef struct omg_scene omg_scene;

/* material.rs:7-17, generalised to nbands <= OMG_MAX_BANDS */
typedef struct {
    float absorption[OMG_MAX_BANDS];
    float scattering;
    float transmission[OMG_MAX_BANDS];
} omg_material;

/* An axis-aligned overlay box with inline acoustics: door leaves, panes,
 * furniture (scene.rs:84-94, WithPanels; GpuPanel in layout.rs:124-137). */
typedef struct {
    float min[3];
    float max[3];
    omg_material mat;
} omg_overlay;

/* Per-call tracer job (tracer.rs:12-16 constants become parameters). */
typedef struct {
    uint32_t n_rays;          /* rays per source */
    uint32_t nbands;          /* 1..OMG_MAX_BANDS (the reference has 3) */
    uint32_t max_bounces;     /* 64 in tracer.rs:103 */
    uint32_t dir_band;        /* band weighting the arrival direction (1 in tracer.rs:123) */
    uint64_t seed;            /* Philox key of the per-ray counter stream */
    float bin_dt;             /* BIN_DT (tracer.rs:12) */
    float max_time;           /* MAX_TIME (tracer.rs:13) */
    float floor_rel;          /* relative energy floor, 1e-7 (tracer.rs:163, :179) */
    float radius;             /* RECEIVER_RADIUS 0.5 (tracer.rs:15) */
    float energy[OMG_MAX_BANDS];  /* source energy per band */
} omg_trace_params;

/* Per-ray parity record (optional output). */
typedef struct {
    uint32_t iters;           /* bounce-loop iterations = bounce-rays */
    uint32_t term;            /* 0 escape, 1 time, 2 dead after transmit, 3 dead after reflect, 4 max bounces */
    uint32_t hit_hash;        /* FNV-1a over the hit triangle ids */
    uint32_t flags;           /* bit0 near-edge hit, bit1 exact-t tie */
} omg_ray_record;

/* Traversal counters (optional output): AABB tests, triangle tests,
 * bounce-rays -- the inputs of the algorithmic-bytes roofline. */
typedef struct {
    uint64_t aabb_tests;
    uint64_t tri_tests;
    uint64_t bounce_rays;
} omg_counters;

typedef struct {
    uint32_t n_tris, n_prims, n_nodes, max_depth;
    uint64_t device_bytes;
} omg_scene_info;

/* ------------------------------------------------------------ lifecycle */

/* GpuTracer::new / GpuMeshTracer::new (omg-gpu/src/lib.rs:33-47): returns
 * OMG_ERR_NO_DEVICE when no GPU is usable, OMG_ERR_VERSION when the caller
 * was built against another layout. */
int omg_init(int device, uint32_t layout_version, omg_ctx **out);
void omg_shutdown(omg_ctx *ctx);
/* Thread-local message for the last failing call on this thread. */
const char *omg_last_error(void);
int omg_device_name(omg_ctx *ctx, char *buf, uint32_t len);
/* Tracer launch shape (no reference counterpart): BVH nodes of the top
 * levels staged in LDS per block (in 64 B paired-node units; a 4-wide node
 * takes two), traversal-stack entries per ray kept in LDS (16 or 32;
 * deeper entries spill to scratch), lanes per ray (1, 4 = a DPP quad per
 * ray, 16 = a DPP row per ray over the 16-wide nodes, 0 = by launch size)
 * and the BVH width walked (2 = the reference's binary tree, 4 = its
 * 4-wide regrouping, 16 = its 16-wide regrouping; a scene without usable
 * wider nodes keeps the next narrower width).  Every width visits the
 * binary walk's leaves in its order and gives its nearest hits bit for
 * bit, ties and their flags included; width 2 also reproduces the
 * reference's traversal counts.  Defaults: 512, 16, 0, 4.  A call that
 * asks for traversal counters walks width 2 (the counters are the
 * reference traversal's). */
int omg_set_launch_config(omg_ctx *ctx, uint32_t node_cache, uint32_t stack_lds, uint32_t lanes,
                          uint32_t width);
/* Plain launches (omg_trace_world, omg_trace_world_async, omg_trace_box
 * excluded) of at least min_rays rays (rays per source x sources) run on
 * the persistent regenerating kernel of the batched launch instead of the
 * single-job kernel: the same rays and per-ray records, echograms equal up
 * to the order of their f32 sums.  Default 1048576 (the size from which a
 * plain launch would trace one lane per ray); 0 = never.  No
 * reference counterpart (the reference has one kernel per call). */
int omg_set_route_rays(omg_ctx *ctx, uint64_t min_rays);
/* Plain launches that trace a quad per ray over the 4-wide tree (below the
 * routing threshold, with the whole stack in LDS) hand every ray still
 * alive after `bounce` bounces to a second, persistent kernel that
 * continues it on a 16-lane row over the 16-wide nodes: the first bounces
 * keep all rays in flight four lanes each, the long-lived rays then walk
 * four binary levels per node fetch.  node_cache16: 16-wide nodes of the
 * top levels staged in LDS per block of that kernel (512 B each);
 * waves_per_simd: resident waves per SIMD its grid is sized for (1..8).
 * The same rays and per-ray records; echograms equal up to the order of
 * their f32 sums.  Defaults 4, 32, 6; bounce 0 = off.  No reference
 * counterpart. */
int omg_set_handoff(omg_ctx *ctx, uint32_t bounce, uint32_t node_cache16, uint32_t waves_per_simd);
/* Lanes per ray of that second kernel: 16 (one 16-wide node slot per
 * lane, four rays per wave), 8 (two slots per lane, eight rays per wave:
 * fewer instructions per ray and step, a longer step) or 0 (the default:
 * 8 for launches of at least 262144 rays, where the kernel is bound by
 * instruction issue, else 16).  The same rays and per-ray records either
 * way.  No reference counterpart. */
int omg_set_handoff_lanes(omg_ctx *ctx, uint32_t lanes);

/* ------------------------------------------------------------ scenes */

/* Mesh::with_surfaces_edges + build_bvh (mesh.rs:198-405) and the BVH
 * upload of GpuMeshTracer::new (omg-gpu/src/lib.rs:274-287): positions
 * [nv][3], indices [nt][3], per-triangle material / surface id /
 * tessellation limit (tri_surface and tri_max_edge may be NULL: 0 and
 * 4 m), materials [nm].  Tessellation and SAH reproduce the reference's
 * primitives bit for bit. */
int omg_scene_create(omg_ctx *ctx, const float *positions, uint32_t nv, const uint32_t *indices,
                     uint32_t nt, const uint16_t *tri_material, const uint16_t *tri_surface,
                     const float *tri_max_edge, const omg_material *materials, uint32_t nm,
                     uint32_t nbands, omg_scene **out);
/* MeshBuilder::tri + build (mesh.rs:108-165): a triangle soup [nt][3][3]
 * with the reference's 0.1 mm vertex dedup applied first. */
int omg_scene_create_soup(omg_ctx *ctx, const float *tri_verts, uint32_t nt,
                          const uint16_t *tri_material, const uint16_t *tri_surface,
                          const float *tri_max_edge, const omg_material *materials, uint32_t nm,
                          uint32_t nbands, omg_scene **out);
int omg_scene_info_get(omg_scene *scene, omg_scene_info *info);
/* Host-only inspection of the scene build (no device needed): the
 * MeshBuilder dedup (mesh.rs:108-114) and the packed primitives / paired
 * nodes exactly as they would be uploaded.  prims_out rows are 12 words
 * (a[3], tri, e1[3], mat, e2[3], surf); nodes_out rows are 16 words
 * (lmin[3], lref, lmax[3], rref, rmin[3], 0, rmax[3], 0); either may be
 * NULL to query the counts.  A leaf reference has bit 31 set, its count in
 * bits 28..30 and its first primitive in bits 0..27. */
int omg_dedup_soup(const float *tri_verts, uint32_t nt, float *positions_out,
                   uint32_t *indices_out, uint32_t *nv_out);
int omg_bvh_build_host(const float *positions, uint32_t nv, const uint32_t *indices, uint32_t nt,
                       const uint16_t *tri_material, const uint16_t *tri_surface,
                       const float *tri_max_edge, float *prims_out, uint32_t prim_cap,
                       uint32_t *n_prims, float *nodes_out, uint32_t node_cap, uint32_t *n_nodes,
                       uint32_t *root_ref, uint32_t *max_depth);

/* Host-only inspection of the 4-wide collapse of the same tree: nodes4_out
 * rows are 32 words (per child: min[3], ref, max[3], 0; ref 0xFFFFFFFF =
 * unused slot, bit 31 = leaf as above, else a row index); NULL queries the
 * counts.  *stack4 is the worst-case traversal stack; *usable = 1 when it
 * fits the 64-entry stack (the scene then walks width 4 on request). */
int omg_bvh4_build_host(const float *positions, uint32_t nv, const uint32_t *indices, uint32_t nt,
                        const uint16_t *tri_material, const uint16_t *tri_surface,
                        const float *tri_max_edge, float *nodes4_out, uint32_t node_cap,
                        uint32_t *n_nodes4, uint32_t *root4, uint32_t *stack4, uint32_t *usable);

/* The same scene built on the device: Mesh::with_surfaces_edges +
 * build_bvh (mesh.rs:198-405) with the tessellation, the 16-bin SAH tree,
 * its breadth-first flatten and the 4-wide collapse computed by kernels,
 * level by level.  Same arguments and errors as omg_scene_create (a soup
 * goes through omg_dedup_soup first), and the same scene: primitives,
 * paired nodes and 4-wide nodes equal the host build's (box coordinates
 * equal as values; a zero may carry the other sign).  In addition it
 * refuses non-finite vertex coordinates (OMG_ERR_INVALID) and patches
 * more than 60 splits deep (OMG_ERR_CAP), and keeps its inputs on the
 * device for omg_scene_refit. */
int omg_scene_create_device(omg_ctx *ctx, const float *positions, uint32_t nv,
                            const uint32_t *indices, uint32_t nt, const uint16_t *tri_material,
                            const uint16_t *tri_surface, const float *tri_max_edge,
                            const omg_material *materials, uint32_t nm, uint32_t nbands,
                            omg_scene **out);

typedef struct {
    uint32_t n_prims, n_nodes, n_nodes4;
    uint32_t root_ref, root4, stack4, has4, max_depth, empty, pad;
    float root_min[3], root_max[3];
    /* the 16-wide nodes derived from the 4-wide ones (16-lane walks): their
     * count, worst-case stack and whether the stack fits (has16) */
    uint32_t n_nodes16, stack16, has16, pad16;
} omg_scene_tree;

/* The resident tree of any scene, in the row layouts of omg_bvh_build_host
 * and omg_bvh4_build_host (n_nodes4 is 0 when the scene walks width 2
 * only); each output may be NULL. */
int omg_scene_export(omg_scene *scene, float *prims_out, uint32_t prim_cap, float *nodes_out,
                     uint32_t node_cap, float *nodes4_out, uint32_t node4_cap,
                     omg_scene_tree *tree);

/* Refit for moving geometry (no reference counterpart: the reference
 * rebuilds, and keeps door leaves as box overlays to avoid it,
 * scene.rs:84-89).  New positions [nv][3] for a scene built by
 * omg_scene_create_device, same nv and triangles: the patches are
 * re-derived from the moved vertices and every box of both trees is
 * recomputed bottom-up; the tree topology is kept.  Refused
 * (OMG_ERR_INVALID, scene unchanged) when a triangle's patch count would
 * change -- build the scene again then.  Nearest hits are the same as
 * with a fresh build of the moved mesh except among hits at exactly equal
 * distance. */
int omg_scene_refit(omg_scene *scene, const float *positions, uint32_t nv);

void omg_scene_destroy(omg_scene *scene);

/* ------------------------------------------------------------ late-field tracer */

/* WorldLateBackend::trace_world (crates/omg-scene/src/late.rs:81-91;
 * omg-gpu/src/lib.rs:845-873), multi-source and multi-listener: every
 * source traces params->n_rays rays, each scored against every listener
 * (a ray's path does not depend on the listener).  Host buffers:
 * sources [n_src][3], listeners [n_lis][3], overlays [n_ovl] (<= 64),
 * bins [n_src][n_lis][nbins][nbands], dirs [n_src][n_lis][nbins][3].
 * records [n_src][n_rays] and counters are optional (NULL). */
int omg_trace_world(omg_scene *scene, const omg_trace_params *params, const float *sources,
                    uint32_t n_src, const float *listeners, uint32_t n_lis,
                    const omg_overlay *overlays, uint32_t n_ovl, float *bins, float *dirs,
                    omg_ray_record *records, omg_counters *counters);

/* omg_trace_world with the per-bounce hit record of every ray (parity
 * checking; no reference counterpart): hit_ids and hit_t are
 * [n_src * n_rays][max_bounces], the triangle id (0xFFFF0000 | i for
 * overlay i, 0xFFFFFFFF = none) and distance of each bounce's hit. */
int omg_trace_world_hits(omg_scene *scene, const omg_trace_params *params, const float *sources,
                         uint32_t n_src, const float *listeners, uint32_t n_lis,
                         const omg_overlay *overlays, uint32_t n_ovl, float *bins, float *dirs,
                         omg_ray_record *records, uint32_t *hit_ids, float *hit_t);

/* LateBackend::trace (late.rs:28-40; omg-gpu/src/lib.rs:1051-1067) over the
 * analytic Shoebox (scene.rs:25-82): size[3], walls[6] in the order
 * x-, x+, y-, y+, z-, z+.  Same buffers as omg_trace_world. */
int omg_trace_box(omg_ctx *ctx, const float size[3], const omg_material walls[6],
                  uint32_t nbands, const omg_trace_params *params, const float *sources,
                  uint32_t n_src, const float *listeners, uint32_t n_lis, float *bins,
                  float *dirs, omg_ray_record *records, omg_counters *counters);

/* Device-resident variant of omg_trace_world: sources, listeners and the
 * outputs are device pointers, the launch is enqueued on `stream`
 * (a hipStream_t, NULL = the context's stream) and the call returns
 * without waiting.  Overlays are host memory, copied before the launch. */
int omg_trace_world_async(omg_scene *scene, const omg_trace_params *params,
                          const float *d_sources, uint32_t n_src, const float *d_listeners,
                          uint32_t n_lis, const omg_overlay *overlays, uint32_t n_ovl,
                          float *d_bins, float *d_dirs, omg_ray_record *d_records,
                          void *stream);

/* Allocation-free variant of omg_trace_world_async, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments, kernels and results; the call allocates
 * nothing: the job's material and overlay tables and the 16-lane hand-off
 * queue live in a fixed device array (about 24 MiB per GPU, held until the
 * process ends) instead of the context's buffers.  Limits: at most 262144
 * rays (rays per source x sources) and 4096 materials (scene materials +
 * overlays) per call, refused with OMG_ERR_CAP above.  The tables are
 * copied (the call then waits for `stream`) only when they differ from the
 * last call's.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other contexts) race. */
int omg_trace_world_async_noalloc(omg_scene *scene, const omg_trace_params *params,
                                  const float *d_sources, uint32_t n_src,
                                  const float *d_listeners, uint32_t n_lis,
                                  const omg_overlay *overlays, uint32_t n_ovl, float *d_bins,
                                  float *d_dirs, omg_ray_record *d_records, void *stream);

/* ------------------------------------------------------------ batched persistent launch */

/* One query of a batch: the arguments of one omg_trace_world call
 * (late.rs:81-91), host memory.  No reference counterpart for the batch
 * itself: the reference runs one job per call. */
typedef struct {
    omg_scene *scene;
    omg_trace_params params;
    const float *sources;
    uint32_t n_src;
    const float *listeners;
    uint32_t n_lis;
    const omg_overlay *overlays;
    uint32_t n_ovl;
} omg_query;

/* Launch shape of the persistent kernel: lanes per ray (1, or 4 = a DPP
 * quad), traversal-stack entries per ray kept in LDS (16 or 32; 8, 12 or
 * 16 with regen), resident waves per SIMD the grid is sized for (1..8),
 * and regen = 1 to refill lanes whose ray has ended with new rays between
 * bounces, or regen = 2 (one lane per ray, width 4, 12 or 16 entries) to
 * shade a lane's segment and start its next walk - of the same ray or of a
 * new one - as soon as most of its wave waits, instead of after the wave's
 * longest walk.  The valid combinations are listed by the error message
 * of omg_batch_set_shape.  ms: the sampled time the autotuner measured for
 * it (0 when set by hand). */
typedef struct {
    uint32_t lanes;
    uint32_t stack_lds;
    uint32_t waves_per_simd;
    uint32_t regen;
    uint32_t width;             /* BVH width: 2, or 4 when every scene has it (16: 16-lane rows) */
    float ms;
} omg_batch_shape;

#define OMG_BATCH_RECORDS 1u   /* keep per-ray records (omg_batch_result) */
#define OMG_BATCH_COUNTERS 2u  /* count traversal work (omg_batch_counters); walks width 2 */

typedef struct omg_batch omg_batch;

/* Uploads every query's inputs, lays the echograms of all queries out in
 * one device arena, splits the rays into work items of 64 consecutive
 * rays of one source and captures clear + persistent launch in a
 * hipGraph.  Each query is validated as omg_trace_world validates it.
 * The scenes must stay alive until omg_batch_destroy. */
int omg_batch_create(omg_ctx *ctx, const omg_query *queries, uint32_t n_queries, uint32_t flags,
                     omg_batch **out);
/* Replays the graph on `stream` (NULL = the context's stream) without waiting. */
int omg_batch_launch(omg_batch *batch, void *stream);
int omg_batch_sync(omg_batch *batch);
/* Copies query q's echograms (and records, with OMG_BATCH_RECORDS) to host
 * buffers laid out as omg_trace_world's; waits for the last launch. */
int omg_batch_result(omg_batch *batch, uint32_t q, float *bins, float *dirs,
                     omg_ray_record *records);
/* Device addresses of query q's echograms (valid until omg_batch_destroy). */
int omg_batch_device_output(omg_batch *batch, uint32_t q, float **d_bins, float **d_dirs);
int omg_batch_counters(omg_batch *batch, omg_counters *counters);
/* Total rays and work items of the batch. */
int omg_batch_size(omg_batch *batch, uint64_t *rays, uint32_t *items);
int omg_batch_get_shape(omg_batch *batch, omg_batch_shape *shape);
int omg_batch_set_shape(omg_batch *batch, const omg_batch_shape *shape);
/* Autotuner: times every launch shape on a strided sample of the batch's
 * work items (every `stride`-th; the stride is lowered so that the sample
 * keeps at least 256 items per compute unit), keeps the fastest, re-captures the graph
 * and caches the choice on the context for batches of the same class
 * (widest band count, total rays and mean BVH size, each to a power of 2).
 * out may be NULL.  Outputs of the sampled launches are overwritten by
 * the next omg_batch_launch. */
int omg_batch_tune(omg_batch *batch, uint32_t stride, omg_batch_shape *out);
void omg_batch_destroy(omg_batch *batch);

/* ------------------------------------------------------------ echogram post-processing */

/* estimate_reverb (tracer.rs:205-257): rt60[nbands], level[nbands]. */
int omg_estimate_reverb(const float *bins, uint32_t nbins, uint32_t nbands, float bin_dt,
                        float *rt60, float *level);
/* Echogram::late_direction (tracer.rs:62-77). */
int omg_late_direction(const float *bins, const float *dirs, uint32_t nbins, uint32_t nbands,
                       float bin_dt, float from_s, float dir_out[3], float *aniso);

/* ------------------------------------------------------------ ambient dome */

/* Dome panel (dome.rs:41-53): a rectangle in a vertical plane. axis 0 =
 * plane of constant x, 1 = constant y.  glass != 0 transmits bounce rays
 * with trans^2; a leaf (glass == 0) blocks them.  The straight-through ray
 * pays trans^2 for either kind (dome.rs:407-420). */
typedef struct {
    uint32_t axis;
    float plane;
    float lat0, lat1, z0, z1;
    uint32_t glass;
    float trans[OMG_MAX_BANDS];
} omg_dome_panel;

typedef struct {
    uint32_t n_fan;           /* N_RAYS, the golden-spiral fan size (dome.rs:34) */
    uint32_t n_rays;          /* rays traced, <= n_fan, even stride (dome.rs:432-434) */
    uint32_t max_events;      /* MAX_EVENTS 6 (dome.rs:35) */
    uint32_t nbands;
    float horizon;            /* HORIZON 80 m (dome.rs:36); the sky-sphere radius */
    float through_len;        /* straight-through segment, 50 m (dome.rs:392) */
} omg_dome_params;

typedef struct {
    uint32_t escaped;         /* the bounce path reached the sky */
    uint32_t events;          /* events used by the bounce path */
    uint32_t used_through;    /* the straight-through ray was traced */
    uint32_t bin;             /* 0..7 azimuth sector, 8 zenith (dome.rs:304-310) */
    float tp[OMG_MAX_BANDS];  /* per-band throughput delivered */
    float dir[3];             /* departure direction */
} omg_dome_ray;

/* DomeProbe::sample without its EMA (dome.rs:425-459) for n_eyes listener
 * positions: energy [n_eyes][9][nbands] and dsum [n_eyes][9][3] are the
 * per-bin sums before the EMA; rays [n_eyes][n_rays] optional. */
int omg_dome_sample(omg_scene *scene, const omg_dome_params *params, const float *eyes,
                    uint32_t n_eyes, const omg_dome_panel *panels, uint32_t n_panels,
                    float *energy, float *dsum, omg_dome_ray *rays, omg_counters *counters);

/* Device-resident variant: d_eyes, d_energy, d_dsum and d_rays (scratch of
 * n_eyes * n_rays omg_dome_ray, required) are device pointers; the launch
 * is enqueued on `stream` (NULL = the context's stream) without waiting.
 * Panels are host memory, copied before the launch.  A crossing-list
 * overflow is reported by the next omg_dome_status call. */
int omg_dome_sample_async(omg_scene *scene, const omg_dome_params *params, const float *d_eyes,
                          uint32_t n_eyes, const omg_dome_panel *panels, uint32_t n_panels,
                          float *d_energy, float *d_dsum, omg_dome_ray *d_rays, void *stream);
/* Waits for the context's dome work and returns OMG_ERR_OVERFLOW if any
 * asynchronous dome launch since the last call overflowed, else OMG_OK. */
int omg_dome_status(omg_scene *scene);

/* Stateful probe = DomeProbe (dome.rs:293-302, :460-477): applies the 0.35
 * EMA (first call primes) and the bin-direction normalisation.  out is
 * [9][3 + nbands] rows of (dir[3], energy[nbands]). */
typedef struct omg_dome_probe omg_dome_probe;
int omg_dome_probe_create(omg_scene *scene, const omg_dome_params *params, omg_dome_probe **out);
int omg_dome_probe_sample(omg_dome_probe *probe, const float eye[3],
                          const omg_dome_panel *panels, uint32_t n_panels, float *out);
void omg_dome_probe_destroy(omg_dome_probe *probe);

/* ------------------------------------------------------------ image sources */

/* ism::image_source_taps (crates/omg-core/src/ism.rs:15-84): taps as rows
 * of (key, delay, dir[3], gains[nbands]).  *n_out receives the count. */
int omg_ism_box(const float size[3], const omg_material walls[6], uint32_t nbands,
                const float src[3], const float lis[3], uint32_t max_order, float *taps,
                uint32_t cap, uint32_t *n_out);

/* Overlay face table row for the solve (pt_mesh.rs:29-41) */
typedef struct {
    float n[3];
    float d;
    float refl[OMG_MAX_BANDS];
    float trans[OMG_MAX_BANDS];
    uint32_t has_rect;
    float rmin[3], rmax[3];
} omg_surface;

/* Transient blocker for leg transmission (pt.rs:31-35, Aabb) */
typedef struct {
    float min[3], max[3];
    float trans[OMG_MAX_BANDS];
} omg_blocker;

/* Solved path (pt_mesh.rs:111-120).  valid = 0: no record. */
typedef struct {
    uint32_t valid;
    float delay;
    float dir[3];
    float gains[OMG_MAX_BANDS];
} omg_mesh_record;

/* SurfaceTable::build (pt_mesh.rs:51-80): surfaces of the scene's mesh;
 * *n_out receives the count (the first overlay-face id). */
int omg_surface_table(omg_scene *scene, omg_surface *out, uint32_t cap, uint32_t *n_out);
/* SurfaceTable::append_box (pt_mesh.rs:87-106), host only: the six faces of
 * an overlay box (furniture) as rect-bounded reflectors with its material,
 * out6[0..5] = x min, x max, y min, y max, z min, z max -- appended after
 * the table, the first one's index is the `base` of omg_discover. */
int omg_surface_append_box(const float mn[3], const float mx[3], const omg_material *mat,
                           uint32_t nbands, omg_surface *out6);

/* WorldDiscovery::discover (crates/omg-scene/src/early_world.rs:24-26;
 * mesh_chains, pt_mesh.rs:375-457; GpuWorldDiscovery passes the furniture
 * boxes when furniture is on, crates/omg-gpu/src/lib.rs:832-843): the
 * rotating fan over the mesh and the overlay boxes [n_boxes][6] (min[3],
 * max[3]; at most 64, refused above), whose faces reflect too with surface
 * ids base + box * 6 + face (face = axis * 2 + 0 for the min side, 1 for
 * the max side, as SurfaceTable::append_box numbers them).  Chains as [3]
 * u16 + order, deduplicated in first-seen order. */
int omg_discover(omg_scene *scene, const float *boxes, uint32_t n_boxes, uint16_t base,
                 const float listener[3], uint32_t n_rays, uint32_t rot, uint16_t *chains,
                 uint8_t *orders, uint32_t cap, uint32_t *n_out);

/* WorldSolve::solve_batch (early_world.rs:38-52; omg-gpu/src/lib.rs:510-609):
 * records[si * n_chains + ci] for every (source, chain).  surfaces: the
 * table of omg_surface_table plus any appended overlay faces. */
int omg_solve_batch(omg_scene *scene, const omg_surface *surfaces, uint32_t n_surf,
                    const float *sources, uint32_t n_src, const uint16_t *chains,
                    const uint8_t *orders, uint32_t n_chains, const float listener[3],
                    const omg_blocker *extras, uint32_t n_extras, const float *air_per_m,
                    omg_mesh_record *records);

/* ------------------------------------------------------------ serving loop */

/* A propagation request (no reference counterpart: the reference traces
 * one listener's late field per sim tick, late.rs:81-91, and samples the
 * dome per tick, dome.rs:425-478).  Host memory, copied at submit. */
typedef struct {
    omg_scene *scene;           /* one of the server's resident scenes */
    omg_trace_params params;
    const float *sources;       /* [n_src][3] */
    uint32_t n_src;
    const float *listeners;     /* [n_lis][3] */
    uint32_t n_lis;
    uint32_t dome;              /* 1: also sample the dome fan at every listener */
    uint32_t flags;             /* OMG_REQ_* */
} omg_request;

#define OMG_REQ_KEEP 1u         /* keep the results for omg_server_result */
#define OMG_REQ_RECORDS 2u      /* also keep per-ray records (with OMG_REQ_KEEP) */

typedef struct {
    uint64_t hbm_budget;        /* bytes: resident scenes + in-flight work arenas */
    uint32_t slots;             /* batches in flight at once, 1..16.  Each slot runs its tracer
                                   and its dome fans on two streams while 2 * slots + 2 fit
                                   GPU_MAX_HW_QUEUES (HIP's hardware queues per process, 4 by
                                   default; always with 1-2 slots), else on one stream */
    uint32_t max_batch;         /* requests per batch, 1..4096 */
    uint32_t latency_rays;      /* batches of at most this many rays run a DPP quad per ray
                                   with regeneration over 16-ray items (shorter per-ray chains,
                                   lanes refilled as rays end); 0 = never */
    omg_batch_shape shape;      /* persistent launch shape of the other batches */
    omg_dome_params dome;       /* dome fan of requests with dome = 1 */
} omg_server_config;

typedef struct {
    uint64_t submitted, completed, refused, batches;
    uint64_t committed_bytes;   /* resident scenes + work arenas now */
    uint64_t peak_committed_bytes;
    uint64_t rays;              /* tracer rays launched so far */
    double mean_batch;          /* requests per batch */
    uint32_t max_in_flight;     /* most batches on the device at once */
} omg_server_stats;

typedef struct {
    double queue_s;             /* submit -> launch */
    double latency_s;           /* arrival (submit, or the replay's scheduled time) -> results on the host */
    uint32_t batch_size;        /* requests in the batch it ran in */
    uint32_t slot;
} omg_request_stats;

typedef struct omg_server omg_server;

/* Starts the scheduler thread over resident scenes.  The scenes' device
 * bytes count against the budget (OMG_ERR_BUDGET if they alone exceed
 * it).  While the server runs, its thread is the only user of the
 * context's device work; the caller's thread only submits and collects. */
int omg_server_create(omg_ctx *ctx, const omg_server_config *cfg, omg_scene *const *scenes,
                      uint32_t n_scenes, omg_server **out);
/* Queues a request; *ticket identifies it.  Refused with OMG_ERR_BUDGET
 * when its work footprint could never fit beside the resident scenes.
 * Requests are admitted to a batch in arrival order as a slot and budget
 * free up; batches launch as soon as a slot is free (continuous batching:
 * requests arriving while others run go into the next batch, which
 * overlaps them on the device). */
int omg_server_submit(omg_server *srv, const omg_request *req, uint64_t *ticket);
/* Results of a finished OMG_REQ_KEEP request, laid out as omg_trace_world's
 * (bins, dirs, records) and omg_dome_sample's (dome_energy [n_lis][9][nbands],
 * dome_dsum [n_lis][9][3]); any output may be NULL.  OMG_ERR_NOT_READY while
 * it runs.  The results are released after a successful call. */
int omg_server_result(omg_server *srv, uint64_t ticket, float *bins, float *dirs,
                      omg_ray_record *records, float *dome_energy, float *dome_dsum,
                      omg_request_stats *st);
/* Blocks until the request has finished (then as omg_server_result). */
int omg_server_wait(omg_server *srv, uint64_t ticket, float *bins, float *dirs,
                    omg_ray_record *records, float *dome_energy, float *dome_dsum,
                    omg_request_stats *st);
/* Open-loop load: submits reqs[i] at arrival_s[i] seconds after the call
 * (ascending), waits for all, and returns each request's latency from its
 * scheduled arrival to its results on the host, and its wait for a launch
 * (queue_s, may be NULL).  A refused request reads -1.  tickets may be NULL. */
int omg_server_replay(omg_server *srv, const omg_request *reqs, const double *arrival_s,
                      uint32_t n, double *latency_s, double *queue_s, uint64_t *tickets);
int omg_server_stats_get(omg_server *srv, omg_server_stats *st);
/* Drains the in-flight work, stops the thread, frees the arenas. */
void omg_server_destroy(omg_server *srv);

/* ------------------------------------------------------------ late-field state (sim side)
 *
 * The consumer of the tracer inside the reference's Sim: TraceGate
 * (crates/omg-scene/src/sim.rs:38-76) decides whether a slot's late field
 * is re-traced this tick; a fresh echogram is folded into the slot's
 * running average with Echogram::ema (crates/omg-core/src/tracer.rs:46-57,
 * alpha 1.0 on a slot's first update, 0.3 after, sim.rs:330-333 and
 * :374-377); the average yields ReverbParams via estimate_reverb
 * (tracer.rs:205-257), the late direction and anisotropy via
 * late_direction(0.08) (tracer.rs:62-77), and the level is scaled for
 * the source-receiver distance change since the last trace (sim.rs:382-395).
 * Sim::reverb_world (sim.rs:349-397) is that sequence.  Here the averages of
 * all slots stay resident on the device, one call updates any number of
 * slots, and only the parameters come back to the host. */
typedef struct omg_late_field omg_late_field;

/* TraceGate (sim.rs:38-76); caller-owned state, no device needed. */
typedef struct omg_trace_gate {
    float last_src[3];
    float last_lis[3];
    float last_energy[OMG_MAX_BANDS];
    uint32_t age;
} omg_trace_gate;
/* TraceGate::new(phase) (sim.rs:46-53) */
void omg_trace_gate_init(omg_trace_gate *g, uint32_t phase);
/* TraceGate::should_trace (sim.rs:55-75): *fire = 1 when the slot should be
 * re-traced (motion > 0.25 m, a band's energy moved by > 10 %, or max_age
 * ticks since the last trace - the quality tier's gate_max_age,
 * quality.rs:133-139, :177-179). */
int omg_trace_gate_step(omg_trace_gate *g, const float src[3], const float lis[3],
                        const float *energy, uint32_t nbands, uint32_t max_age, int *fire);

/* n_slots averages of [nbins][nbands] bins + [nbins][3] dirs, zero at start;
 * slot s's gate starts at phase s % 8 (Sim::new, sim.rs:99-103).
 * 2 <= nbands <= OMG_MAX_BANDS (late_direction reads band 1). */
int omg_late_field_create(omg_ctx *ctx, uint32_t n_slots, uint32_t nbins, uint32_t nbands,
                          float bin_dt, omg_late_field **out);
/* The slot's TraceGate::should_trace. */
int omg_late_field_gate(omg_late_field *lf, uint32_t slot, const float src[3], const float lis[3],
                        const float *energy, uint32_t max_age, int *fire);

typedef struct omg_late_update {
    uint32_t slot;
    uint32_t traced;          /* 1: fold cur_bins/cur_dirs into the slot's average */
    const float *cur_bins;    /* device [nbins][nbands] (a trace's echogram), or NULL */
    const float *cur_dirs;    /* device [nbins][3], or NULL */
    float src[3];             /* source and receiver of this tick */
    float receiver[3];
} omg_late_update;

typedef struct omg_late_out {
    float rt60[OMG_MAX_BANDS];   /* ReverbParams (params.rs:19-29) */
    float level[OMG_MAX_BANDS];  /* after the distance compensation */
    float dir[3];                /* late_direction(0.08) */
    float aniso;
    uint32_t version;            /* the slot's update count after this call */
    uint32_t traced;
} omg_late_out;

/* Updates n distinct slots (sim.rs:365-397 after the trace decision):
 * EMA for the traced ones, then reverb parameters, late direction and the
 * distance compensation for every one.  The fresh echograms must be
 * complete on `stream` (a hipStream_t; NULL = the field's stream, then
 * the caller has synchronized them).  Blocks until out[] is written. */
int omg_late_field_update(omg_late_field *lf, const omg_late_update *ups, uint32_t n,
                          omg_late_out *out, void *stream);
typedef struct omg_late_trace {
    uint32_t slot;
    omg_trace_params params;  /* nbands and nbins(max_time, bin_dt) must be the field's */
    float src[3];
    float receiver[3];
} omg_late_trace;
/* Traces each job's late field on the scene (one source, one listener:
 * WorldLateBackend::trace_world, late.rs:81-91) into the field's own device
 * staging, then updates the slots as omg_late_field_update with traced = 1
 * (sim.rs:366-377).  Slots must be distinct.  Blocks until out[] is written. */
int omg_late_field_trace_world(omg_late_field *lf, omg_scene *scene, const omg_late_trace *jobs,
                               uint32_t n, const omg_overlay *overlays, uint32_t n_ovl,
                               omg_late_out *out);
/* Copies a slot's average to host buffers ([nbins][nbands], [nbins][3]). */
int omg_late_field_read(omg_late_field *lf, uint32_t slot, float *bins, float *dirs);
void omg_late_field_destroy(omg_late_field *lf);

/* ------------------------------------------------------------ flat wire formats
 *
 * The reference's web bridge queues traces as flat f32 jobs and takes
 * echograms back as flat f32 buffers (crates/omg-web/src/lib.rs:151-161,
 * :367-378, web/gpu.js:395-440); ParamBlock crosses thread and worker
 * boundaries as a flat f32 layout (crates/omg-core/src/params.rs:62-143).
 * These entry points take and give exactly those layouts, so a remote or
 * asynchronous driver can hand its job queue to this library unchanged. */
#define OMG_WIRE_BOX_JOB_F32S 57        /* GPU_JOB_F32S (lib.rs:157) */
#define OMG_WIRE_BOX_JOB_VERSION 2      /* GPU_JOB_VERSION (lib.rs:158) */
#define OMG_WIRE_WORLD_MAX_PANELS 64    /* WLATE_MAX_PANELS (lib.rs:373) */
#define OMG_WIRE_WORLD_JOB_F32S (10 + OMG_WIRE_WORLD_MAX_PANELS * 16)  /* WLATE_JOB_F32S */
#define OMG_WIRE_WORLD_JOB_VERSION 3    /* WLATE_VERSION (lib.rs:378) */
#define OMG_WIRE_ECHO_BINS 300          /* MAX_TIME / BIN_DT (tracer.rs:12-14) */
#define OMG_WIRE_ECHO_F32S (OMG_WIRE_ECHO_BINS * 3 + OMG_WIRE_ECHO_BINS * 3)  /* GPU_ECHO_F32S */
#define OMG_WIRE_PARAM_HEADER 24        /* FLAT_HEADER (params.rs:74) */
#define OMG_WIRE_PARAM_PER_TAP 9        /* FLAT_PER_TAP (params.rs:75) */

/* One decoded world job (lib.rs:394-410 writes it; gpu.js:420-440 reads
 * n_rays and seed as u32 and uses unit source energy). */
typedef struct {
    uint32_t id;
    uint32_t n_rays;
    uint32_t seed;
    float src[3];
    float lis[3];
    uint32_t n_panels;       /* min(job[9], 64) */
    omg_overlay panels[OMG_WIRE_WORLD_MAX_PANELS];
} omg_wire_world_job;
/* Decodes one 1034-f32 world job; OMG_ERR_VERSION unless version == 3. */
int omg_wire_decode_world_job(const float *job, uint32_t version, omg_wire_world_job *out);

/* Runs the flat box jobs (57 f32 each: id, n_rays, seed, room size, source,
 * listener, band energy, 6 walls x (absorption x3, scattering,
 * transmission x3), lib.rs:151-156, :168-185) on the analytic Shoebox with
 * the reference tracer's constants (10 ms x 3 s, 64 bounces, floor 1e-7,
 * receiver 0.5 m) and writes each job's echogram in the injection layout
 * (300 x 3 bins then 300 x 3 dirs, lib.rs:159-160) to echo_out[k * 1800],
 * its sim id to ids_out[k].  n_f32 must be a multiple of 57. */
int omg_wire_trace_box_jobs(omg_ctx *ctx, const float *jobs, uint32_t n_f32, uint32_t version,
                            float *echo_out, uint32_t *ids_out, uint32_t *n_jobs_out);
/* The same for flat world jobs (1034 f32 each) on a 3-band mesh scene,
 * panels as overlays. */
int omg_wire_trace_world_jobs(omg_scene *scene, const float *jobs, uint32_t n_f32,
                              uint32_t version, float *echo_out, uint32_t *ids_out,
                              uint32_t *n_jobs_out);

/* ParamBlock (params.rs:62-70) with its taps (params.rs:8-16) and the
 * optional RemoteReverb (params.rs:46-55). */
typedef struct {
    uint32_t key;
    float delay_s;
    float dir[3];
    float gains[3];
} omg_tap;
typedef struct {
    uint64_t version;
    float rt60[3];
    float level[3];
    uint32_t has_remote;
    float remote_rt60[3];
    float remote_send[3];
    float remote_sh[9];
} omg_param_block;
/* ParamBlock::write_flat (params.rs:82-107): 24 + 9 * n_taps f32. */
int omg_param_block_write_flat(const omg_param_block *pb, const omg_tap *taps, uint32_t n_taps,
                               float *out, uint32_t cap, uint32_t *n_out);
/* ParamBlock::read_flat (params.rs:109-142); *n_taps is the flat block's
 * tap count, taps receives at most cap of them. */
int omg_param_block_read_flat(const float *data, uint32_t n_f32, omg_param_block *pb,
                              omg_tap *taps, uint32_t cap, uint32_t *n_taps);

/* ------------------------------------------------------------ knife-edge paths
 *
 * AutoPaths (crates/omg-core/src/paths.rs:62-332): the direct segment
 * with emergent wall thickness, plus, when its mid band is blocked, bent
 * paths over the mesh's diffraction edges priced by the Kurze-Anderson
 * knife edge (crates/omg-core/src/diffraction.rs:27-47).  The edge graph is
 * extracted on the host once (extract_edges + budget_edges,
 * crates/omg-core/src/mesh.rs:639-682, 40 degree dihedral threshold) and
 * kept on the device; queries run many (source, listener) pairs in one
 * launch.  Three bands at 125, 790 and 5600 Hz (3-band scenes only). */
#define OMG_PATH_MAX_POINTS 25          /* the silhouette hull: 24 drops + the listener */
#define OMG_PATH_MAX_EDGES 4096         /* edge graph cap: refused above */
#define OMG_PATH_MAX_CANDIDATES 256     /* edge_candidates cap: refused above */
#define OMG_PATH_MAX_PAIR_EDGES 64      /* pair_edges cap: refused above */

/* DiffractionEdge (mesh.rs:626-634): endpoints and static importance
 * (length x sharpness). */
typedef struct {
    float a[3];
    float importance;
    float b[3];
    uint32_t reserved;
} omg_path_edge;
/* PathBudget (paths.rs:20-45); the reference's default is {24, 8, 4, 0}. */
typedef struct {
    uint32_t edge_candidates;
    uint32_t pair_edges;
    uint32_t max_paths;
    uint32_t taps;              /* strict tap semantics: legs priced past the apex */
} omg_path_budget;
/* FoundPath (paths.rs:48-60): key 0 = direct, 1 + edge = single bend,
 * 100000 + 4096 * e1 + e2 = double bend, 90000 = over the silhouette. */
typedef struct {
    uint32_t key;
    uint32_t n_points;
    float length;
    float gains[3];
    float points[OMG_PATH_MAX_POINTS][3];
} omg_path;

typedef struct omg_paths omg_paths;

/* extract_edges + budget_edges (mesh.rs:639-682) on the host, no device
 * needed: writes at most cap edges, importance-descending (equal
 * importance in ascending vertex-pair order), *n_out = the budgeted count. */
int omg_paths_extract_edges(const float *positions, uint32_t nv, const uint32_t *indices,
                            uint32_t nt, float min_dihedral_deg, uint32_t budget,
                            omg_path_edge *out, uint32_t cap, uint32_t *n_out);
/* AutoPaths::new (paths.rs:70-77): the scene's budgeted edge graph on the
 * device.  OMG_ERR_CAP above OMG_PATH_MAX_EDGES edges.  Queries walk the
 * scene's BVH: the scene must stay alive while the object is queried (it
 * may be destroyed before omg_paths_destroy). */
int omg_paths_create(omg_scene *scene, uint32_t edge_budget, omg_paths **out);
/* AutoPaths::edge_count (paths.rs:79-81) and the edges themselves. */
int omg_paths_edges(omg_paths *paths, omg_path_edge *out, uint32_t cap, uint32_t *n_out);
/* AutoPaths::transmission (paths.rs:84-113) for n segments a[i] -> b[i]:
 * gains [n][3]. */
int omg_paths_transmission(omg_paths *paths, const float *a, const float *b, uint32_t n,
                           float *gains);
/* AutoPaths::find (paths.rs:117-332) for n queries: out [n][1 + max_paths]
 * (the direct path first, then the strongest bent paths; unused slots are
 * zero), n_paths [n].  Blocks until written.  OMG_ERR_OVERFLOW when a
 * segment has more than 64 triangle crossings, counted before crossings
 * at the same position merge (segment_hits' dedup, mesh.rs:589-603). */
int omg_paths_find(omg_paths *paths, const omg_path_budget *budget, const float *src,
                   const float *lis, uint32_t n, omg_path *out, uint32_t *n_paths);
/* Device-resident variant: d_src, d_lis [n][3], d_out, d_n_paths and
 * d_work (n uint32 of scratch) are device pointers; enqueued on `stream`
 * (NULL = the context's stream) without waiting.  The object's own
 * per-query plan buffer and counters are used too, so calls on one object
 * must not overlap (a growing plan buffer first waits for the device).  An
 * overflow is reported by the next omg_paths_status. */
int omg_paths_find_async(omg_paths *paths, const omg_path_budget *budget, const float *d_src,
                         const float *d_lis, uint32_t n, omg_path *d_out, uint32_t *d_n_paths,
                         uint32_t *d_work, void *stream);
/* Waits for the object's queued work; OMG_ERR_OVERFLOW if any launch since
 * the last call overflowed, else OMG_OK. */
int omg_paths_status(omg_paths *paths);
/* AABB and triangle tests of the traversals of the last omg_paths_find
 * call made with counting on (omg_paths_set_counting). */
int omg_paths_set_counting(omg_paths *paths, uint32_t on);
int omg_paths_counters(omg_paths *paths, omg_counters *counters);
/* The shared numerics, host-evaluated (the device runs the same code):
 * knife_edge_amp (diffraction.rs:27-42) and bend_point (paths.rs:411-436). */
float omg_knife_edge_amp(float detour_m, float freq_hz);
void omg_bend_point(const float a[3], const float b[3], const float src[3], const float lis[3],
                    float p[3]);
void omg_paths_destroy(omg_paths *paths);

#ifdef __cplusplus
}
#endif
#endif
