/*
 * Oracle restatement of the tracer (crates/omg-core/src/tracer.rs), the
 * geometry seams (scene.rs) and shoebox image sources (ism.rs).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "or_mesh_internal.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define WALL_EPS 1e-3f   /* tracer.rs:16 */

/* ------------------------------------------------------------ geometry */

/* Shoebox::raycast (scene.rs:41-61) */
static void box_raycast(const float size[3], v3 p, v3 d, float *t_out, int *w_out)
{
    float best_t = FLT_MAX;
    int best_w = 0;
    for (int axis = 0; axis < 3; axis++) {
        float di = vget(d, axis);
        if (di > 1e-9f) {
            float t = (size[axis] - vget(p, axis)) / di;
            if (t < best_t) {
                best_t = t;
                best_w = 2 * axis + 1;
            }
        } else if (di < -1e-9f) {
            float t = -vget(p, axis) / di;
            if (t < best_t) {
                best_t = t;
                best_w = 2 * axis;
            }
        }
    }
    *t_out = best_t;
    *w_out = best_w;
}

/* AcousticGeometry::raycast_hit for Shoebox (scene.rs:64-82), Mesh
 * (mesh.rs:607-620) and WithPanels (scene.rs:96-138). */
int or_raycast_hit(const or_geom *g, const float pf[3], const float df[3], or_hit *h,
                   uint64_t *stats)
{
    v3 p = vload(pf), d = vload(df);
    int found = 0;
    h->flags = 0;
    if (g->kind == OR_GEOM_BOX) {
        const float *s = g->size;
        int inside = p.x > 0.0f && p.x < s[0] && p.y > 0.0f && p.y < s[1] && p.z > 0.0f &&
                     p.z < s[2];
        if (inside) {
            float t;
            int wall;
            box_raycast(s, p, d, &t, &wall);
            if (isfinite(t) && !(t <= 0.0f)) {
                v3 n = V(0.0f, 0.0f, 0.0f);
                vset(&n, wall / 2, (wall % 2 == 0) ? 1.0f : -1.0f);
                h->t = t;
                vstore(h->normal, n);
                h->material_ref = wall;
                h->tri = (uint32_t)wall;
                found = 1;
            }
        }
    } else {
        float t;
        uint32_t tri, fl = 0;
        if (or_mesh_raycast(g->mesh, pf, df, &t, &tri, stats, &fl)) {
            v3 n = g->mesh->normals[tri];
            if (vdot(n, d) > 0.0f)
                n = vmul(n, -1.0f);
            h->t = t;
            vstore(h->normal, n);
            h->material_ref = g->mesh->tri_material[tri];
            h->tri = tri;
            h->flags = fl;
            found = 1;
        }
    }
    for (int i = 0; i < g->n_panels; i++) {
        const float *mn = g->panels + 6 * i, *mx = mn + 3;
        float t0 = 0.0f, t1 = FLT_MAX;
        int axis = 3;
        for (int a = 0; a < 3; a++) {
            float da = vget(d, a), pa = vget(p, a), lo = mn[a], hi = mx[a];
            if (fabsf(da) < 1e-9f) {
                if (pa < lo || pa > hi) {
                    t0 = FLT_MAX;
                    break;
                }
            } else {
                float ta = (lo - pa) / da, tb = (hi - pa) / da;
                if (ta > tb) {
                    float tmp = ta;
                    ta = tb;
                    tb = tmp;
                }
                if (ta > t0) {
                    t0 = ta;
                    axis = a;
                }
                t1 = fminf(t1, tb);
                if (t0 > t1) {
                    t0 = FLT_MAX;
                    break;
                }
            }
        }
        if (t0 > 1e-4f && t0 < FLT_MAX && axis < 3 && (!found || t0 < h->t)) {
            v3 n = V(0.0f, 0.0f, 0.0f);
            vset(&n, axis, vget(d, axis) > 0.0f ? -1.0f : 1.0f);
            h->t = t0;
            vstore(h->normal, n);
            h->material_ref = -(i + 1);
            h->tri = 0xFFFF0000u | (uint32_t)i;
            h->flags = 0;
            found = 1;
        }
    }
    return found;
}

static const or_material *hit_material(const or_geom *g, const or_hit *h)
{
    if (h->material_ref < 0)
        return &g->panel_mats[-h->material_ref - 1];
    if (g->kind == OR_GEOM_BOX)
        return &g->walls[h->material_ref];
    return &g->mesh->materials[h->material_ref];
}

/* ------------------------------------------------------------ tracer */

uint32_t or_nbins(float max_time, float bin_dt)
{
    return (uint32_t)sat_usize(max_time / bin_dt);   /* tracer.rs:14 */
}

typedef struct {
    int mode;
    uint64_t *xs;               /* xorshift state */
    uint32_t key[2];
    uint32_t ray, src;
    uint32_t words[4];
} rng_t;

static inline float u24(uint32_t w) { return (float)(w >> 8) / 16777216.0f; }

static void counter_block(rng_t *r, uint32_t slot)
{
    uint32_t ctr[4] = {r->ray, slot, r->src, 0u};
    or_philox4x32_10(ctr, r->key, r->words);
}

/* Rng::unit_sphere (rng.rs:24-29) from two uniforms */
static v3 unit_sphere_from(float u1, float u2, int project_trig)
{
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    float phi = OR_TAU * u2;
    float s, c;
    if (project_trig) {
        or_sincosf(phi, &s, &c);
    } else {
        s = sinf(phi);
        c = cosf(phi);
    }
    return V(r * c, r * s, z);
}

static inline uint32_t fnv1a(uint32_t h, uint32_t v)
{
    for (int i = 0; i < 4; i++) {
        h ^= (v >> (8 * i)) & 0xFFu;
        h *= 16777619u;
    }
    return h;
}

/* tracer::trace (tracer.rs:80-202) */
int or_trace(const or_geom *g, const float srcf[3], const float *lis, uint32_t n_lis,
             const or_trace_params *P, uint64_t *rng_state, float *bins, float *dirs,
             double *bins64, double *dirs64, or_ray_rec *recs, uint32_t *hit_ids,
             uint64_t *stats)
{
    const uint32_t NB = P->nbands;
    const uint32_t nbins = or_nbins(P->max_time, P->bin_dt);
    if (NB == 0 || NB > OR_MAXB || P->dir_band >= NB)
        return -1;
    size_t eg = (size_t)nbins * NB, dg = (size_t)nbins * 3;
    memset(bins, 0, n_lis * eg * sizeof(float));
    memset(dirs, 0, n_lis * dg * sizeof(float));
    if (bins64)
        memset(bins64, 0, n_lis * eg * sizeof(double));
    if (dirs64)
        memset(dirs64, 0, n_lis * dg * sizeof(double));
    const float per_ray = 1.0f / (float)P->n_rays;
    const v3 source = vload(srcf);
    rng_t rng;
    rng.mode = P->rng_mode;
    rng.xs = rng_state;
    rng.key[0] = (uint32_t)P->seed;
    rng.key[1] = (uint32_t)(P->seed >> 32);
    rng.src = P->src_index;
    uint64_t local_stats[2] = {0, 0};
    uint64_t bounce_rays = 0;
    const int counter = P->rng_mode == OR_RNG_COUNTER;

    for (uint32_t ray = 0; ray < P->n_rays; ray++) {
        rng.ray = ray;
        v3 pos = source, dir;
        if (counter) {
            counter_block(&rng, 0);
            dir = unit_sphere_from(u24(rng.words[0]), u24(rng.words[1]), 1);
        } else {
            float u1 = or_xorshift_f32(rng.xs);
            float u2 = or_xorshift_f32(rng.xs);
            dir = unit_sphere_from(u1, u2, 0);
        }
        float energy[OR_MAXB];
        for (uint32_t b = 0; b < NB; b++)
            energy[b] = per_ray * P->energy[b];
        float dist_total = 0.0f;
        uint32_t iters = 0, term = OR_TERM_MAX_BOUNCES, hash = 2166136261u, rflags = 0;
        if (hit_ids)
            for (uint32_t k = 0; k < P->max_bounces; k++)
                hit_ids[(size_t)ray * P->max_bounces + k] = 0xFFFFFFFFu;

        for (uint32_t bounce = 0; bounce < P->max_bounces; bounce++) {
            iters++;
            float pf[3], df[3];
            vstore(pf, pos);
            vstore(df, dir);
            or_hit hit;
            int has = or_raycast_hit(g, pf, df, &hit, local_stats);
            float t_hit = has ? hit.t : FLT_MAX;

            /* receiver sphere crossing (tracer.rs:110-128) */
            for (uint32_t li = 0; li < n_lis; li++) {
                v3 L = vload(lis + 3 * li);
                v3 to_l = vsub(L, pos);
                float s = vdot(to_l, dir);
                if (s > 0.0f && s < t_hit) {
                    v3 closest = vadd(pos, vmul(dir, s));
                    if (vlen(vsub(closest, L)) < P->radius) {
                        float arrival = (dist_total + s) / OR_SPEED_OF_SOUND;
                        uint64_t bin = sat_usize(arrival / P->bin_dt);
                        if (bin < nbins) {
                            float *ob = bins + li * eg + bin * NB;
                            float *od = dirs + li * dg + bin * 3;
                            for (uint32_t b = 0; b < NB; b++)
                                ob[b] += energy[b];
                            float ed = energy[P->dir_band];
                            od[0] -= dir.x * ed;
                            od[1] -= dir.y * ed;
                            od[2] -= dir.z * ed;
                            if (bins64) {
                                double *qb = bins64 + li * eg + bin * NB;
                                for (uint32_t b = 0; b < NB; b++)
                                    qb[b] += (double)energy[b];
                            }
                            if (dirs64) {
                                double *qd = dirs64 + li * dg + bin * 3;
                                qd[0] -= (double)(dir.x * ed);
                                qd[1] -= (double)(dir.y * ed);
                                qd[2] -= (double)(dir.z * ed);
                            }
                        }
                    }
                }
            }
            if (!has) {
                term = OR_TERM_ESCAPE;
                break;
            }
            hash = fnv1a(hash, hit.tri);
            rflags |= hit.flags;
            if (hit_ids)
                hit_ids[(size_t)ray * P->max_bounces + bounce] = hit.tri;
            pos = vadd(pos, vmul(dir, t_hit));
            dist_total += t_hit;
            if (dist_total / OR_SPEED_OF_SOUND > P->max_time) {
                term = OR_TERM_TIME;
                break;
            }
            const or_material *mat = hit_material(g, &hit);
            if (counter)
                counter_block(&rng, bounce + 1);
            float p_t = 0.0f;
            for (uint32_t b = 0; b < NB; b++)
                p_t = fmaxf(p_t, mat->transmission[b] * mat->transmission[b]);
            float p = p_t > 1e-5f ? fminf(fmaxf(p_t, 0.02f), 0.5f) : 0.0f;
            int transmitted = 0;
            if (p > 0.0f) {
                float u = counter ? u24(rng.words[0]) : or_xorshift_f32(rng.xs);
                transmitted = u < p;
            }
            const float floor_e = P->floor_rel * per_ray;
            int alive = 0;
            if (transmitted) {
                for (uint32_t b = 0; b < NB; b++) {
                    energy[b] *= (mat->transmission[b] * mat->transmission[b]) / p;
                    if (energy[b] > floor_e)
                        alive = 1;
                }
                if (!alive) {
                    term = OR_TERM_DEAD_TRANSMIT;
                    break;
                }
                pos = vadd(pos, vmul(dir, 2.0f * WALL_EPS));
                continue;
            }
            for (uint32_t b = 0; b < NB; b++) {
                float keep = fmaxf(
                    1.0f - mat->absorption[b] - mat->transmission[b] * mat->transmission[b], 0.0f);
                energy[b] *= keep / (1.0f - p);
                if (energy[b] > floor_e)
                    alive = 1;
            }
            if (!alive) {
                term = OR_TERM_DEAD_REFLECT;
                break;
            }
            v3 normal = vload(hit.normal);
            float us = counter ? u24(rng.words[1]) : or_xorshift_f32(rng.xs);
            if (us < mat->scattering) {
                v3 sp;
                if (counter) {
                    sp = unit_sphere_from(u24(rng.words[2]), u24(rng.words[3]), 1);
                } else {
                    float u1 = or_xorshift_f32(rng.xs);
                    float u2 = or_xorshift_f32(rng.xs);
                    sp = unit_sphere_from(u1, u2, 0);
                }
                dir = vnorm(vadd(normal, sp));
                if (vdot(dir, normal) < 1e-3f)
                    dir = normal;
            } else {
                dir = vsub(dir, vmul(normal, 2.0f * vdot(dir, normal)));
            }
            pos = vadd(pos, vmul(normal, WALL_EPS));
        }
        bounce_rays += iters;
        if (recs) {
            recs[ray].iters = iters;
            recs[ray].term = term;
            recs[ray].hit_hash = hash;
            recs[ray].flags = rflags;
        }
    }
    if (stats) {
        stats[0] += local_stats[0];
        stats[1] += local_stats[1];
        stats[2] += bounce_rays;
    }
    return 0;
}

/* ------------------------------------------------------------ CPU baseline pool */

typedef struct {
    or_trace_job *jobs;
    uint32_t n;
    uint32_t next;
    pthread_mutex_t mu;
} pool_t;

static void *pool_worker(void *arg)
{
    pool_t *pl = (pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&pl->mu);
        uint32_t i = pl->next++;
        pthread_mutex_unlock(&pl->mu);
        if (i >= pl->n)
            break;
        or_trace_job *j = &pl->jobs[i];
        /* xorshift mode: one sequential stream per job, seeded from the seed
         * and the source index (never zero); the counter mode ignores it */
        uint64_t st = (j->p.seed ? j->p.seed : 1u) ^ ((uint64_t)j->p.src_index * 0x9E3779B97F4A7C15ull);
        if (!st)
            st = 1;
        memset(j->stats, 0, sizeof j->stats);
        or_trace(j->g, j->src, j->lis, j->n_lis, &j->p, &st, j->bins, j->dirs, NULL, NULL, NULL,
                 NULL, j->stats);
    }
    return NULL;
}

void or_trace_jobs(or_trace_job *jobs, uint32_t n_jobs, uint32_t n_threads)
{
    pool_t pl;
    pl.jobs = jobs;
    pl.n = n_jobs;
    pl.next = 0;
    pthread_mutex_init(&pl.mu, NULL);
    if (n_threads < 1)
        n_threads = 1;
    pthread_t *th = (pthread_t *)malloc(n_threads * sizeof(pthread_t));
    for (uint32_t i = 0; i < n_threads; i++)
        pthread_create(&th[i], NULL, pool_worker, &pl);
    for (uint32_t i = 0; i < n_threads; i++)
        pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&pl.mu);
}

/* ------------------------------------------------------------ echogram post-processing */

void or_ema(float *bins, float *dirs, const float *obins, const float *odirs, uint32_t nbins,
            uint32_t nbands, float alpha)
{
    for (size_t i = 0; i < (size_t)nbins * nbands; i++)
        bins[i] += alpha * (obins[i] - bins[i]);
    for (size_t i = 0; i < (size_t)nbins * 3; i++)
        dirs[i] += alpha * (odirs[i] - dirs[i]);
}

void or_late_direction(const float *bins, const float *dirs, uint32_t nbins, uint32_t nbands,
                       float bin_dt, float from_s, float dir_out[3], float *aniso)
{
    uint64_t start = sat_usize(from_s / bin_dt);
    if (start > nbins)
        start = nbins;
    float v[3] = {0.0f, 0.0f, 0.0f}, e = 0.0f;
    for (uint64_t i = start; i < nbins; i++) {
        for (int k = 0; k < 3; k++)
            v[k] += dirs[i * 3 + k];
        e += bins[i * nbands + 1];
    }
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len < 1e-12f || e < 1e-12f) {
        dir_out[0] = 1.0f;
        dir_out[1] = 0.0f;
        dir_out[2] = 0.0f;
        *aniso = 0.0f;
        return;
    }
    dir_out[0] = v[0] / len;
    dir_out[1] = v[1] / len;
    dir_out[2] = v[2] / len;
    *aniso = fminf(len / e, 1.0f);
}

void or_estimate_reverb(const float *bins, uint32_t nbins, uint32_t nbands, float bin_dt,
                        float *rt60, float *level)
{
    double *edc = (double *)malloc((nbins ? nbins : 1) * sizeof(double));
    for (uint32_t b = 0; b < nbands; b++) {
        rt60[b] = 0.5f;
        double total = 0.0;
        for (int64_t i = (int64_t)nbins - 1; i >= 0; i--) {
            total += (double)bins[i * nbands + b];
            edc[i] = total;
        }
        if (total <= 0.0)
            continue;
        int64_t t5 = -1, t25 = -1;
        for (uint32_t i = 0; i < nbins; i++) {
            double r = edc[i] / total;
            double dd = 10.0 * log10(r > 1e-12 ? r : 1e-12);
            if (t5 < 0 && dd <= -5.0)
                t5 = i;
            if (t25 < 0 && dd <= -25.0) {
                t25 = i;
                break;
            }
        }
        if (t5 >= 0 && t25 >= 0 && t25 > t5) {
            double r5 = edc[t5] / total, r25 = edc[t25] / total;
            double db5 = 10.0 * log10(r5 > 1e-12 ? r5 : 1e-12);
            double db25 = 10.0 * log10(r25 > 1e-12 ? r25 : 1e-12);
            float dt = (float)(t25 - t5) * bin_dt;
            float ddb = (float)(db25 - db5);
            float v = -60.0f * dt / ddb;
            rt60[b] = fminf(fmaxf(v, 0.1f), 5.0f);
        }
    }
    free(edc);
    uint64_t late_start = sat_usize(0.080f / bin_dt);
    for (uint32_t b = 0; b < nbands; b++) {
        float late_e = 0.0f;
        for (uint64_t i = late_start; i < nbins; i++)
            late_e += bins[i * nbands + b];
        level[b] = fminf(sqrtf(late_e), 0.7f);
    }
}

/* ------------------------------------------------------------ Sim late path (sim.rs) */

void or_gate_init(or_gate *g, uint32_t phase)
{
    /* last_src / last_lis = Vec3::new(f32::MAX, 0, 0), last_energy = 0 (sim.rs:47-52) */
    for (int k = 0; k < 3; k++) {
        g->last_src[k] = k == 0 ? FLT_MAX : 0.0f;
        g->last_lis[k] = k == 0 ? FLT_MAX : 0.0f;
    }
    for (int b = 0; b < OR_MAXB; b++)
        g->last_energy[b] = 0.0f;
    g->age = phase;
}

int or_gate_step(or_gate *g, const float src[3], const float lis[3], const float *energy,
                 uint32_t nbands, uint32_t max_age)
{
    g->age += 1;
    int moved = 0;
    for (uint32_t b = 0; b < nbands; b++) {
        float e0 = g->last_energy[b], e1 = energy[b];
        if (fabsf(e1 - e0) > 0.1f * fmaxf(fmaxf(e0, e1), 1e-6f))
            moved = 1;
    }
    float ds = vlen(vsub(vload(src), vload(g->last_src)));
    float dl = vlen(vsub(vload(lis), vload(g->last_lis)));
    if (g->age >= max_age || ds > 0.25f || dl > 0.25f || moved) {
        for (int k = 0; k < 3; k++) {
            g->last_src[k] = src[k];
            g->last_lis[k] = lis[k];
        }
        for (uint32_t b = 0; b < nbands; b++)
            g->last_energy[b] = energy[b];
        g->age = 0;
        return 1;
    }
    return 0;
}

void or_late_slot_init(or_late_slot *s, float *bins, float *dirs)
{
    s->bins = bins;
    s->dirs = dirs;
    s->version = 0;
    s->rp_dist = FLT_MAX;
}

void or_late_step(or_late_slot *s, uint32_t nbins, uint32_t nbands, float bin_dt, int traced,
                  const float *cur_bins, const float *cur_dirs, const float src[3],
                  const float receiver[3], float *rt60, float *level, float dir[3], float *aniso)
{
    v3 d = vsub(vload(src), vload(receiver));
    if (traced) {
        float alpha = s->version == 0 ? 1.0f : 0.3f;
        or_ema(s->bins, s->dirs, cur_bins, cur_dirs, nbins, nbands, alpha);
        s->rp_dist = vlen(d);
    }
    s->version += 1;
    or_late_direction(s->bins, s->dirs, nbins, nbands, bin_dt, 0.08f, dir, aniso);
    or_estimate_reverb(s->bins, nbins, nbands, bin_dt, rt60, level);
    if (s->rp_dist < FLT_MAX) {
        float d_now = vlen(d);
        float k = fminf((s->rp_dist + 1.0f) / (d_now + 1.0f), 1.0f);
        for (uint32_t b = 0; b < nbands; b++)
            level[b] *= k;
    }
}

/* ------------------------------------------------------------ ISM (ism.rs:15-84) */

int or_image_source_taps(const float size[3], const or_material walls[6], uint32_t NB,
                         const float srcf[3], const float lisf[3], uint32_t max_order,
                         const float *air_per_m, float *taps, int cap)
{
    v3 source = vload(srcf), listener = vload(lisf);
    int n_max = ((int)max_order + 1) / 2 + 1;
    float refl[6][OR_MAXB];
    for (int w = 0; w < 6; w++)
        for (uint32_t b = 0; b < NB; b++)
            refl[w][b] = sqrtf(fmaxf(1.0f - walls[w].absorption[b], 0.0f));
    int n = 0;
    const int row = 5 + (int)NB;
    for (int px = 0; px < 2; px++)
        for (int py = 0; py < 2; py++)
            for (int pz = 0; pz < 2; pz++)
                for (int nx = -n_max; nx <= n_max; nx++)
                    for (int ny = -n_max; ny <= n_max; ny++)
                        for (int nz = -n_max; nz <= n_max; nz++) {
                            unsigned hits[6] = {(unsigned)abs(nx - px), (unsigned)abs(nx),
                                                (unsigned)abs(ny - py), (unsigned)abs(ny),
                                                (unsigned)abs(nz - pz), (unsigned)abs(nz)};
                            unsigned order = 0;
                            for (int k = 0; k < 6; k++)
                                order += hits[k];
                            if (order > max_order)
                                continue;
                            v3 img = V((float)(1 - 2 * px) * source.x + 2.0f * (float)nx * size[0],
                                       (float)(1 - 2 * py) * source.y + 2.0f * (float)ny * size[1],
                                       (float)(1 - 2 * pz) * source.z + 2.0f * (float)nz * size[2]);
                            v3 to_img = vsub(img, listener);
                            float dist = fmaxf(vlen(to_img), 0.3f);
                            v3 dir = vnorm(to_img);
                            if (n < cap) {
                                float *o = taps + (size_t)n * row;
                                o[0] = (float)n;
                                o[1] = dist / OR_SPEED_OF_SOUND;
                                o[2] = dir.x;
                                o[3] = dir.y;
                                o[4] = dir.z;
                                for (uint32_t b = 0; b < NB; b++) {
                                    float g = expf(-air_per_m[b] * dist) / dist;
                                    for (int w = 0; w < 6; w++)
                                        for (unsigned k = 0; k < hits[w]; k++)
                                            g *= refl[w][b];
                                    o[5 + b] = g;
                                }
                            }
                            n++;
                        }
    return n < cap ? n : cap;
}
