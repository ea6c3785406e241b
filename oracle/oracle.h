/*
 * CPU oracle for the omg-audio stochastic energy path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference Rust implementation, used by tests/ (as the checker) and by
 * bench.py's cpu_baseline leg (as the timed CPU port).  No product code
 * may link, load or call it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the reference checkout).  Numerics follow the reference
 * exactly: f32 everywhere, no FMA contraction (built with
 * -ffp-contract=off), IEEE division/sqrt, fmaxf/fminf NaN-ignoring
 * semantics of Rust's f32::max/min, Rust's expression order.
 *
 * Two RNG modes:
 *   OR_RNG_XORSHIFT  the reference's sequential xorshift64* stream
 *                    (crates/omg-core/src/rng.rs:2-30), libm sin/cos;
 *                    used for the reference's statistical goldens.
 *   OR_RNG_COUNTER   a counter-based per-ray stream (Philox4x32-10,
 *                    Salmon et al., SC'11) keyed by the job seed, with a
 *                    fixed draw slot per bounce, and the project's
 *                    Cody-Waite + Taylor sin/cos.  This is the mode the GPU
 *                    product implements; parity is bit-exact per ray.
 */
#ifndef OMG_ORACLE_H
#define OMG_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXB 8           /* max frequency bands */
#define OR_RNG_XORSHIFT 0
#define OR_RNG_COUNTER 1

/* material.rs:7-17 generalised to nbands <= OR_MAXB */
typedef struct {
    float absorption[OR_MAXB];
    float scattering;
    float transmission[OR_MAXB];
} or_material;

/* ---------------------------------------------------------------- rng */
uint64_t or_xorshift_next(uint64_t *state);                 /* rng.rs:10-16 */
float or_xorshift_f32(uint64_t *state);                     /* rng.rs:19-21 */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_sincosf(float x, float *s, float *c);                /* project sin/cos */

/* ---------------------------------------------------------------- mesh */
typedef struct or_mesh or_mesh;

/* Mesh::with_surfaces_edges + build_bvh (mesh.rs:198-405) */
or_mesh *or_mesh_build(const float *positions, uint32_t nv, const uint32_t *indices, uint32_t nt,
                       const uint16_t *tri_material, const uint16_t *tri_surface,
                       const float *tri_max_edge, const or_material *materials, uint32_t nm,
                       uint32_t nbands);
void or_mesh_free(or_mesh *m);
void or_mesh_sizes(const or_mesh *m, uint32_t *n_nodes, uint32_t *n_prims);
/* visit_bvh (mesh.rs:470-481): nodes as {bmin[3], bmax[3], a, b} (8 words),
 * prims as {a[3], e1[3], e2[3], tri, material} (11 words). */
void or_mesh_export(const or_mesh *m, float *nodes8, float *prims11);

/* MeshBuilder vertex dedup (mesh.rs:108-126): turns a triangle soup
 * (nt x 3 x 3 floats) into positions/indices.  Returns nv. */
uint32_t or_dedup_tris(const float *tri_verts, uint32_t nt, float *positions_out,
                       uint32_t *indices_out);

/* Mesh::raycast (mesh.rs:484-541).  Returns 1 on hit.  stats (optional):
 * [0] += AABB tests, [1] += triangle tests.  flags (optional): bit0 hit
 * within 1e-6 barycentric of a prim edge, bit1 exact-t tie between
 * distinct original triangles among tested prims. */
int or_mesh_raycast(const or_mesh *m, const float o[3], const float d[3], float *t, uint32_t *tri,
                    uint64_t *stats, uint32_t *flags);
/* Mesh::segment_hits (mesh.rs:546-604).  Returns the count (<= cap). */
int or_mesh_segment_hits(const or_mesh *m, const float a[3], const float b[3], float *t_out,
                         uint32_t *tri_out, uint16_t *mat_out, int cap, uint64_t *stats);
void or_mesh_tri_normal(const or_mesh *m, uint32_t tri, float n[3]);   /* mesh.rs:458-461 */

/* ---------------------------------------------------------------- geometry */
#define OR_GEOM_BOX 0
#define OR_GEOM_MESH 1
typedef struct {
    int kind;
    float size[3];              /* Shoebox (scene.rs:25-28) */
    or_material walls[6];
    const or_mesh *mesh;
    int n_panels;               /* WithPanels (scene.rs:90-138) */
    const float *panels;        /* n_panels x {min[3], max[3]} */
    const or_material *panel_mats;
} or_geom;

typedef struct {
    float t;
    float normal[3];
    int material_ref;           /* <0: panel -(i+1); box: wall; mesh: material index */
    uint32_t tri;               /* mesh tri id, 0xFFFFFFFF for box walls / panels */
    uint32_t flags;
} or_hit;

int or_raycast_hit(const or_geom *g, const float p[3], const float d[3], or_hit *h, uint64_t *stats);

/* ---------------------------------------------------------------- tracer */
typedef struct {
    uint32_t n_rays;
    uint32_t nbands;
    uint32_t max_bounces;       /* 64 in tracer.rs:103 */
    uint32_t rng_mode;
    uint64_t seed;              /* xorshift seed, or Philox key */
    uint32_t src_index;         /* Philox counter word 2 */
    uint32_t dir_band;          /* band weighting the arrival direction (tracer.rs:123) */
    float bin_dt;               /* BIN_DT, tracer.rs:12 */
    float max_time;             /* MAX_TIME, tracer.rs:13 */
    float floor_rel;            /* 1e-7 relative energy floor, tracer.rs:163 */
    float radius;               /* RECEIVER_RADIUS, tracer.rs:15 */
    float energy[OR_MAXB];
} or_trace_params;

typedef struct {
    uint32_t iters;             /* bounce-loop iterations (bounce-rays) */
    uint32_t term;              /* OR_TERM_* */
    uint32_t hit_hash;          /* FNV-1a over hit tri ids */
    uint32_t flags;             /* OR of raycast flags over the ray's hits */
} or_ray_rec;

#define OR_TERM_ESCAPE 0
#define OR_TERM_TIME 1
#define OR_TERM_DEAD_TRANSMIT 2
#define OR_TERM_DEAD_REFLECT 3
#define OR_TERM_MAX_BOUNCES 4

uint32_t or_nbins(float max_time, float bin_dt);  /* tracer.rs:14 */

/* tracer::trace (tracer.rs:80-202), scoring n_lis listeners in one pass.
 * bins: [n_lis][nbins][nbands] f32, dirs: [n_lis][nbins][3] f32 (both
 * cleared first).  bins64/dirs64 (optional) are the same sums in f64.
 * rng_state (xorshift mode) is read and advanced.  recs: [n_rays],
 * hit_ids: [n_rays][max_bounces] (0xFFFFFFFF = none), stats: [aabb, tri,
 * bounce-rays] (all optional). */
int or_trace(const or_geom *g, const float src[3], const float *lis, uint32_t n_lis,
             const or_trace_params *p, uint64_t *rng_state, float *bins, float *dirs,
             double *bins64, double *dirs64, or_ray_rec *recs, uint32_t *hit_ids,
             uint64_t *stats);

/* Threaded batch of independent traces (CPU baseline): job i traces source
 * srcs[i] with its own params[i] and writes bins/dirs at the given offsets. */
typedef struct {
    const or_geom *g;
    const float *src;
    const float *lis;
    uint32_t n_lis;
    or_trace_params p;
    float *bins;
    float *dirs;
    uint64_t stats[3];
} or_trace_job;
void or_trace_jobs(or_trace_job *jobs, uint32_t n_jobs, uint32_t n_threads);

/* Echogram::ema (tracer.rs:46-57) */
void or_ema(float *bins, float *dirs, const float *obins, const float *odirs, uint32_t nbins,
            uint32_t nbands, float alpha);
/* Echogram::late_direction (tracer.rs:62-77) */
void or_late_direction(const float *bins, const float *dirs, uint32_t nbins, uint32_t nbands,
                       float bin_dt, float from_s, float dir_out[3], float *aniso);
/* estimate_reverb (tracer.rs:205-257) */
void or_estimate_reverb(const float *bins, uint32_t nbins, uint32_t nbands, float bin_dt,
                        float *rt60, float *level);

/* ---------------------------------------------------------------- Sim late path */
/* TraceGate (crates/omg-scene/src/sim.rs:38-76) */
typedef struct {
    float last_src[3], last_lis[3], last_energy[OR_MAXB];
    uint32_t age;
} or_gate;
/* TraceGate::new (sim.rs:46-53) */
void or_gate_init(or_gate *g, uint32_t phase);
/* TraceGate::should_trace (sim.rs:55-75); returns 1 to trace */
int or_gate_step(or_gate *g, const float src[3], const float lis[3], const float *energy,
                 uint32_t nbands, uint32_t max_age);
/* The late state of one Sim: echo_avg, version, rp_dist (sim.rs:79-112) */
typedef struct {
    float *bins, *dirs;
    uint32_t version;
    float rp_dist;
} or_late_slot;
void or_late_slot_init(or_late_slot *s, float *bins, float *dirs);
/* Sim::reverb_world after the trace decision (sim.rs:374-397): EMA of the
 * fresh echogram when traced, late_direction(0.08), estimate_reverb, the
 * distance compensation of the level. */
void or_late_step(or_late_slot *s, uint32_t nbins, uint32_t nbands, float bin_dt, int traced,
                  const float *cur_bins, const float *cur_dirs, const float src[3],
                  const float receiver[3], float *rt60, float *level, float dir[3], float *aniso);

/* ---------------------------------------------------------------- ISM */
/* ism::image_source_taps (ism.rs:15-84); taps as {key, delay, dir[3],
 * gains[nbands]} rows of (5 + nbands) floats.  Returns count. */
int or_image_source_taps(const float size[3], const or_material walls[6], uint32_t nbands,
                         const float src[3], const float lis[3], uint32_t max_order,
                         const float *air_per_m, float *taps, int cap);

/* ---------------------------------------------------------------- dome */
typedef struct {
    int axis;                   /* 0: plane x = const, 1: plane y = const */
    float plane;
    float lat0, lat1, z0, z1;
    int glass;                  /* glass transmits on bounce rays; leaves block */
    float trans[OR_MAXB];       /* amplitude transmission used by the through ray */
} or_panel;

typedef struct {
    uint32_t n_fan;             /* N_RAYS (dome.rs:34) */
    uint32_t n_rays;            /* rays traced (<= n_fan, dome.rs:432-434) */
    uint32_t max_events;        /* MAX_EVENTS (dome.rs:35) */
    uint32_t nbands;
    float horizon;              /* HORIZON (dome.rs:36) */
    float through_len;          /* 50 m (dome.rs:392) */
    int use_project_trig;       /* 1: project sin/cos for the fan, 0: libm */
} or_dome_params;

typedef struct {
    uint32_t escaped;           /* bounce path reached the sky */
    uint32_t events;            /* events consumed by trace_escape */
    uint32_t used_through;      /* trace_through was evaluated */
    uint32_t bin;
    float tp[OR_MAXB];          /* combined per-ray throughput (dome.rs:439-450) */
} or_dome_ray;

void or_dome_fan(uint32_t n_fan, int use_project_trig, float *dirs);     /* dome.rs:316-324 */
uint32_t or_dome_bin_of(const float d[3]);                               /* dome.rs:304-310 */
/* One DomeProbe::sample pass without the EMA (dome.rs:425-459): e[9][nbands],
 * dsum[9][3]; per-ray results optional. */
void or_dome_sample(const or_mesh *m, const or_dome_params *p, const float eye[3],
                    const or_panel *panels, int n_panels, float *e, float *dsum,
                    or_dome_ray *rays, uint64_t *stats);

/* ---------------------------------------------------------------- pt_mesh */
typedef struct {
    float n[3];
    float d;
    float refl[OR_MAXB];
    float trans[OR_MAXB];
    int has_rect;
    float rmin[3], rmax[3];
} or_surface;

/* SurfaceTable::build (pt_mesh.rs:51-80); returns the count written (<= cap) */
int or_surface_table(const or_mesh *m, or_surface *out, int cap);
/* SurfaceTable::append_box (pt_mesh.rs:87-106) */
void or_surface_append_box(or_surface *out6, const float mn[3], const float mx[3],
                           const or_material *mat, uint32_t nbands);

typedef struct {
    float min[3], max[3];
    float trans[OR_MAXB];
} or_aabb;

typedef struct {
    float delay;
    float dir[3];
    float gains[OR_MAXB];
} or_record;

/* mesh_record (pt_mesh.rs:185-324): returns 1 when a record is produced. */
int or_mesh_record(const or_mesh *m, const or_surface *table, int n_surf, const uint16_t *chain,
                   int order, const float src[3], const float lis[3], const or_aabb *extras,
                   int n_extras, float air0, float air1, float air2, or_record *rec);
/* mesh_chains (pt_mesh.rs:375-457): the rotating fan over the mesh and the
 * overlay boxes [n_boxes][6] (min, max; face ids base + box*6 + face):
 * chains as [3] u16 + order, deduplicated in first-seen order.  Returns
 * the count. */
int or_mesh_chains(const or_mesh *m, const float *boxes, int n_boxes, uint16_t base,
                   const float lis[3], uint32_t n_rays, uint32_t rot, int use_project_trig,
                   uint16_t *chains, uint8_t *orders, int cap);

/* ---------------------------------------------------------------- knife-edge paths */
/* DiffractionEdge (mesh.rs:626-634) */
typedef struct {
    float a[3], b[3];
    float importance;
} or_edge;
/* PathBudget (paths.rs:22-45) */
typedef struct {
    uint32_t edge_candidates, pair_edges, max_paths, taps;
} or_path_budget;
#define OR_PATH_MAX_POINTS 25
/* FoundPath (paths.rs:48-60) with at most 25 points (the silhouette's hull) */
typedef struct {
    uint32_t key, n_points;
    float length;
    float gains[3];
    float points[OR_PATH_MAX_POINTS][3];
} or_path;
/* extract_edges + budget_edges (mesh.rs:639-682); edges of equal
 * importance in ascending vertex-pair order */
int or_paths_edges(const or_mesh *m, float min_dihedral_deg, uint32_t budget, or_edge *out, int cap);
void or_paths_transmission(const or_mesh *m, const float a[3], const float b[3], float g[3]);
float or_knife_edge_amp(float detour_m, float freq_hz);                  /* diffraction.rs:27-42 */
void or_knife_edge_bands(float detour_m, float out[3]);                  /* diffraction.rs:45-47 */
void or_bend_point(const float a[3], const float b[3], const float src[3], const float lis[3],
                   float p[3]);                                          /* paths.rs:411-436 */
int or_paths_find(const or_mesh *m, const or_edge *edges, int ne, const float src[3],
                  const float lis[3], const or_path_budget *b, or_path *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
