/*
 * Oracle restatement of the ambient dome (crates/omg-scene/src/dome.rs)
 * and the mesh image-source solver (crates/omg-core/src/pt_mesh.rs,
 * pt.rs:31-69 for Aabb::clip).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "or_mesh_internal.h"
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ dome */

/* DomeProbe::new fan (dome.rs:316-324) */
void or_dome_fan(uint32_t n_fan, int use_project_trig, float *dirs)
{
    float ga = OR_PI * (3.0f - sqrtf(5.0f));
    for (uint32_t i = 0; i < n_fan; i++) {
        float z = 1.0f - 2.0f * ((float)i + 0.5f) / (float)n_fan;
        float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
        float phi = ga * (float)i;
        float s, c;
        if (use_project_trig) {
            or_sincosf(phi, &s, &c);
        } else {
            s = sinf(phi);
            c = cosf(phi);
        }
        dirs[3 * i] = r * c;
        dirs[3 * i + 1] = r * s;
        dirs[3 * i + 2] = z;
    }
}

/* bin_of (dome.rs:304-310) */
uint32_t or_dome_bin_of(const float d[3])
{
    if (d[2] > 0.87f)
        return 8;
    float az = atan2f(d[1], d[0]) + OR_PI;
    uint64_t k = sat_usize(az / (OR_TAU / 8.0f));
    return k < 7 ? (uint32_t)k : 7u;
}

/* Panel::hit (dome.rs:56-68) */
static int panel_hit(const or_panel *pl, v3 o, v3 d, float t_max, float *t_out)
{
    float po = pl->axis == 0 ? o.x : o.y;
    float pd = pl->axis == 0 ? d.x : d.y;
    if (fabsf(pd) < 1e-8f)
        return 0;
    float t = (pl->plane - po) / pd;
    if (t < 1e-4f || t >= t_max)
        return 0;
    float lat = pl->axis == 0 ? o.y + t * d.y : o.x + t * d.x;
    float z = o.z + t * d.z;
    if (lat > pl->lat0 && lat < pl->lat1 && z > pl->z0 && z < pl->z1) {
        *t_out = t;
        return 1;
    }
    return 0;
}

/* DomeProbe::trace_escape (dome.rs:330-382).  Returns 1 when the ray
 * escapes with throughput tp. */
static int trace_escape(const or_mesh *m, const or_dome_params *P, v3 pos, v3 dir,
                        const or_panel *panels, int n_panels, float *tp, uint32_t *events,
                        uint64_t *stats)
{
    const uint32_t NB = P->nbands;
    for (uint32_t b = 0; b < NB; b++)
        tp[b] = 1.0f;
    uint32_t ev = 0;
    int result = 0;
    for (uint32_t e = 0; e < P->max_events; e++) {
        ev++;
        float pf[3], df[3], t_mesh = FLT_MAX, tm;
        uint32_t tri = 0;
        vstore(pf, pos);
        vstore(df, dir);
        int mesh_hit = or_mesh_raycast(m, pf, df, &tm, &tri, stats, NULL);
        if (mesh_hit)
            t_mesh = tm;
        float t_pan = t_mesh;
        int pan = -1;
        for (int i = 0; i < n_panels; i++) {
            float t;
            if (panel_hit(&panels[i], pos, dir, t_pan, &t)) {
                t_pan = t;
                pan = i;
            }
        }
        if (pan >= 0) {
            if (!panels[pan].glass) {
                result = 0;
                goto done;
            }
            for (uint32_t b = 0; b < NB; b++)
                tp[b] *= panels[pan].trans[b] * panels[pan].trans[b];
            pos = vadd(pos, vmul(dir, t_pan + 1e-3f));
            continue;
        }
        int esc = pos.z > -0.05f || dir.z > 0.05f;
        if (!mesh_hit || tm > P->horizon) {
            result = esc;
            goto done;
        }
        const or_material *mat = &m->materials[m->tri_material[tri]];
        for (uint32_t b = 0; b < NB; b++)
            tp[b] *= 1.0f - mat->absorption[b];
        if (tp[1] < 2e-3f) {
            result = 0;
            goto done;
        }
        v3 n = m->normals[tri];
        v3 n_o = vdot(dir, n) > 0.0f ? vmul(n, -1.0f) : n;
        v3 hit = vadd(pos, vmul(dir, tm));
        dir = vnorm(vsub(dir, vmul(n, 2.0f * vdot(dir, n))));
        pos = vadd(hit, vmul(n_o, 1e-3f));
    }
    result = 0;
done:
    *events = ev;
    return result;
}

/* DomeProbe::trace_through (dome.rs:390-422) */
static void trace_through(const or_mesh *m, const or_dome_params *P, v3 eye, v3 dir,
                          const or_panel *panels, int n_panels, float *tp, uint64_t *stats)
{
    const uint32_t NB = P->nbands;
    v3 end = vadd(eye, vmul(dir, P->through_len));
    for (uint32_t b = 0; b < NB; b++)
        tp[b] = 0.0f;
    if (end.z < -0.05f)
        return;
    for (uint32_t b = 0; b < NB; b++)
        tp[b] = 1.0f;
    float a[3], e[3];
    vstore(a, eye);
    vstore(e, end);
    int cap = 4096;
    uint16_t *mats = (uint16_t *)malloc(cap * sizeof(uint16_t));
    int n = or_mesh_segment_hits(m, a, e, NULL, NULL, mats, cap, stats);
    for (int i = 0; i < n; i++) {
        const or_material *mt = &m->materials[mats[i]];
        for (uint32_t b = 0; b < NB; b++)
            tp[b] *= mt->transmission[b] * mt->transmission[b];
        if (tp[0] < 1e-7f) {
            for (uint32_t b = 0; b < NB; b++)
                tp[b] = 0.0f;
            free(mats);
            return;
        }
    }
    free(mats);
    for (int i = 0; i < n_panels; i++) {
        float t;
        if (panel_hit(&panels[i], eye, dir, P->horizon, &t)) {
            for (uint32_t b = 0; b < NB; b++)
                tp[b] *= panels[i].trans[b] * panels[i].trans[b];
        }
    }
}

/* DomeProbe::sample (dome.rs:425-459), without the EMA stage */
void or_dome_sample(const or_mesh *m, const or_dome_params *P, const float eyef[3],
                    const or_panel *panels, int n_panels, float *e, float *dsum,
                    or_dome_ray *rays, uint64_t *stats)
{
    const uint32_t NB = P->nbands;
    float *fan = (float *)malloc(3 * (size_t)P->n_fan * sizeof(float));
    or_dome_fan(P->n_fan, P->use_project_trig, fan);
    memset(e, 0, 9 * NB * sizeof(float));
    memset(dsum, 0, 9 * 3 * sizeof(float));
    v3 eye = vload(eyef);
    for (uint32_t j = 0; j < P->n_rays; j++) {
        uint32_t i = (uint32_t)(((uint64_t)j * P->n_fan) / P->n_rays);
        v3 d0 = vload(fan + 3 * i);
        float bp[OR_MAXB], tp[OR_MAXB];
        uint32_t events = 0;
        int esc = trace_escape(m, P, eye, d0, panels, n_panels, bp, &events, stats);
        int through = 0;
        if (esc && bp[0] > 0.5f) {
            memcpy(tp, bp, NB * sizeof(float));
        } else {
            through = 1;
            trace_through(m, P, eye, d0, panels, n_panels, tp, stats);
            if (esc)
                for (uint32_t b = 0; b < NB; b++)
                    tp[b] = fmaxf(tp[b], bp[b]);
        }
        uint32_t bi = or_dome_bin_of(fan + 3 * i);
        if (tp[0] > 1e-9f) {
            for (uint32_t b = 0; b < NB; b++)
                e[bi * NB + b] += tp[b] / (float)P->n_rays;
            float w = fmaxf(tp[1], tp[0] * 0.2f);
            dsum[bi * 3] += d0.x * w;
            dsum[bi * 3 + 1] += d0.y * w;
            dsum[bi * 3 + 2] += d0.z * w;
        }
        if (rays) {
            rays[j].escaped = (uint32_t)esc;
            rays[j].events = events;
            rays[j].used_through = (uint32_t)through;
            rays[j].bin = bi;
            for (uint32_t b = 0; b < OR_MAXB; b++)
                rays[j].tp[b] = b < NB ? tp[b] : 0.0f;
        }
    }
    free(fan);
}

/* ------------------------------------------------------------ pt_mesh */

int or_surface_table(const or_mesh *m, or_surface *out, int cap)
{
    int n = 0;
    for (uint32_t t = 0; t < m->nt; t++) {
        int sid = m->tri_surface[t];
        while (n <= sid && n < cap) {
            memset(&out[n], 0, sizeof(or_surface));
            out[n].n[2] = 1.0f;
            out[n].d = FLT_MAX;
            n++;
        }
        if (sid >= cap)
            continue;
        or_surface *s = &out[sid];
        if (s->d == FLT_MAX) {
            v3 nn = m->normals[t];
            v3 p = m->positions[m->indices[3 * (size_t)t]];
            const or_material *mt = &m->materials[m->tri_material[t]];
            vstore(s->n, nn);
            s->d = vdot(nn, p);
            for (uint32_t b = 0; b < m->nbands; b++) {
                s->refl[b] = sqrtf(fmaxf(1.0f - mt->absorption[b], 0.0f));
                s->trans[b] = mt->transmission[b];
            }
            s->has_rect = 0;
        }
    }
    return n;
}

void or_surface_append_box(or_surface *out6, const float mnf[3], const float mxf[3],
                           const or_material *mat, uint32_t NB)
{
    v3 mn = vload(mnf), mx = vload(mxf);
    int k = 0;
    for (int axis = 0; axis < 3; axis++)
        for (int side = 0; side < 2; side++) {
            v3 n = V(0.0f, 0.0f, 0.0f);
            vset(&n, axis, side == 0 ? -1.0f : 1.0f);
            float plane = side == 0 ? vget(mn, axis) : vget(mx, axis);
            v3 fmin = mn, fmax = mx;
            vset(&fmin, axis, plane);
            vset(&fmax, axis, plane);
            or_surface *s = &out6[k++];
            memset(s, 0, sizeof *s);
            vstore(s->n, n);
            s->d = vdot(n, side == 0 ? mn : mx);
            for (uint32_t b = 0; b < NB; b++) {
                s->refl[b] = sqrtf(fmaxf(1.0f - mat->absorption[b], 0.0f));
                s->trans[b] = mat->transmission[b];
            }
            s->has_rect = 1;
            vstore(s->rmin, fmin);
            vstore(s->rmax, fmax);
        }
}

/* Aabb::clip (pt.rs:42-69) */
static int aabb_clip(const or_aabb *x, v3 a, v3 b)
{
    v3 d = vsub(b, a);
    float t0 = 0.0f, t1 = 1.0f;
    for (int axis = 0; axis < 3; axis++) {
        float da = vget(d, axis), aa = vget(a, axis), mn = x->min[axis], mx = x->max[axis];
        if (fabsf(da) < 1e-9f) {
            if (aa < mn || aa > mx)
                return 0;
        } else {
            float ta = (mn - aa) / da, tb = (mx - aa) / da;
            if (ta > tb) {
                float tmp = ta;
                ta = tb;
                tb = tmp;
            }
            t0 = fmaxf(t0, ta);
            t1 = fminf(t1, tb);
            if (t0 > t1)
                return 0;
        }
    }
    return t1 > 1e-4f && t0 < 1.0f - 1e-4f;
}

#define SEG_CAP 4096

static int all_below(const float *x, uint32_t nb, float lim)
{
    for (uint32_t b = 0; b < nb; b++)
        if (!(x[b] < lim))
            return 0;
    return 1;
}

/* leg_transmission (pt_mesh.rs:145-181) */
static int leg_transmission(const or_mesh *m, v3 a, v3 b, int has_skip, uint16_t skip_sid,
                            float skip_t, const or_aabb *extras, int n_extras, float *trans)
{
    const uint32_t NB = m->nbands;
    float af[3], bf[3];
    vstore(af, a);
    vstore(bf, b);
    static __thread float ts[SEG_CAP];
    static __thread uint32_t tris[SEG_CAP];
    static __thread uint16_t mats[SEG_CAP];
    int n = or_mesh_segment_hits(m, af, bf, ts, tris, mats, SEG_CAP, NULL);
    float seg_len = fmaxf(vlen(vsub(b, a)), 1e-6f);
    for (int i = 0; i < n; i++) {
        if (has_skip && m->tri_surface[tris[i]] == skip_sid &&
            fabsf(ts[i] - skip_t) * seg_len < 0.05f)
            continue;
        const or_material *mt = &m->materials[mats[i]];
        for (uint32_t k = 0; k < NB; k++)
            trans[k] *= mt->transmission[k];
        if (all_below(trans, NB, 1e-6f))
            return 0;
    }
    for (int i = 0; i < n_extras; i++) {
        if (aabb_clip(&extras[i], a, b)) {
            for (uint32_t k = 0; k < NB; k++)
                trans[k] *= extras[i].trans[k];
            if (all_below(trans, NB, 1e-6f))
                return 0;
        }
    }
    return 1;
}

static v3 mirror(v3 p, const or_surface *s)
{
    v3 n = vload(s->n);
    return vsub(p, vmul(n, 2.0f * (vdot(n, p) - s->d)));
}

/* mesh_record (pt_mesh.rs:185-324); air_per_m holds nbands coefficients
 * packed as (air0, air1, air2) for the reference's three bands. */
int or_mesh_record(const or_mesh *m, const or_surface *table, int n_surf, const uint16_t *chain,
                   int order, const float srcf[3], const float lisf[3], const or_aabb *extras,
                   int n_extras, float air0, float air1, float air2, or_record *rec)
{
    const uint32_t NB = m->nbands;
    const float air_per_m[3] = {air0, air1, air2};
    v3 src_pos = vload(srcf), listener = vload(lisf);
    v3 img = src_pos;
    for (int k = order - 1; k >= 0; k--) {
        if (chain[k] >= n_surf)
            return 0;
        img = mirror(img, &table[chain[k]]);
    }
    v3 to_img = vsub(img, listener);
    float dist = vlen(to_img);
    if (dist < 1e-4f)
        return 0;
    v3 dir = vmul(to_img, 1.0f / dist);
    float gm[OR_MAXB];
    for (uint32_t b = 0; b < NB; b++)
        gm[b] = 1.0f;
    v3 pos = listener, d = dir;
    for (int k = 0; k < order; k++) {
        uint16_t sid = chain[k];
        if (sid >= n_surf)
            return 0;
        const or_surface *s = &table[sid];
        v3 sn = vload(s->n);
        float denom = vdot(sn, d);
        if (fabsf(denom) < 1e-9f)
            return 0;
        float t = (s->d - vdot(sn, pos)) / denom;
        if (t <= 1e-4f)
            return 0;
        v3 hit = vadd(pos, vmul(d, t));
        float of[3], df[3];
        vstore(of, vadd(pos, vmul(d, 1e-4f)));
        vstore(df, d);
        float rt;
        uint32_t rtri;
        if (s->has_rect) {
            const float eps = 1e-3f;
            for (int a = 0; a < 3; a++)
                if (vget(hit, a) < s->rmin[a] - eps || vget(hit, a) > s->rmax[a] + eps)
                    return 0;
            if (or_mesh_raycast(m, of, df, &rt, &rtri, NULL, NULL)) {
                if (rt < t - 1e-3f) {
                    const or_material *mt = &m->materials[m->tri_material[rtri]];
                    if (all_below(mt->transmission, NB, 1e-6f))
                        return 0;
                }
            }
        } else {
            if (!or_mesh_raycast(m, of, df, &rt, &rtri, NULL, NULL))
                return 0;
            int on_surface = m->tri_surface[rtri] == sid && fabsf(rt - t) < 2e-2f;
            if (!on_surface) {
                if (rt < t - 1e-3f) {
                    const or_material *mt = &m->materials[m->tri_material[rtri]];
                    if (all_below(mt->transmission, NB, 1e-4f))
                        return 0;
                    float af[3], bf[3];
                    vstore(af, pos);
                    vstore(bf, vadd(hit, vmul(d, 0.05f)));
                    static __thread float ts[SEG_CAP];
                    static __thread uint32_t tris[SEG_CAP];
                    int n = or_mesh_segment_hits(m, af, bf, ts, tris, NULL, SEG_CAP, NULL);
                    int any = 0;
                    for (int i = 0; i < n; i++)
                        if (m->tri_surface[tris[i]] == sid && fabsf(ts[i] * (t + 0.05f) - t) < 0.1f) {
                            any = 1;
                            break;
                        }
                    if (!any)
                        return 0;
                } else {
                    return 0;
                }
            }
        }
        if (!leg_transmission(m, pos, hit, 1, sid, 1.0f, extras, n_extras, gm))
            return 0;
        for (uint32_t b = 0; b < NB; b++)
            gm[b] *= s->refl[b];
        pos = vadd(hit, vmul(sn, vdot(sn, d) < 0.0f ? 1e-4f : -1e-4f));
        d = vsub(d, vmul(sn, 2.0f * vdot(sn, d)));
    }
    if (order > 0) {
        const or_surface *s = &table[chain[order - 1]];
        v3 sn = vload(s->n);
        if ((vdot(sn, pos) - s->d) * (vdot(sn, src_pos) - s->d) < 0.0f)
            return 0;
    }
    if (!leg_transmission(m, pos, src_pos, 0, 0, 0.0f, extras, n_extras, gm))
        return 0;
    if (order > 0) {
        v3 to_src = vsub(src_pos, pos);
        float leg = vlen(to_src);
        if (leg > 1e-4f && vdot(vmul(to_src, 1.0f / leg), d) < 0.995f)
            return 0;
    }
    float dtot = fmaxf(dist, 0.3f);
    float gains[OR_MAXB];
    for (uint32_t b = 0; b < NB; b++) {
        float air = expf(-air_per_m[b < 3 ? b : 2] * dtot);
        gains[b] = air / dtot * gm[b];
    }
    if (all_below(gains, NB, 2e-5f))
        return 0;
    rec->delay = dist / OR_SPEED_OF_SOUND;
    vstore(rec->dir, dir);
    for (uint32_t b = 0; b < OR_MAXB; b++)
        rec->gains[b] = b < NB ? gains[b] : 0.0f;
    return 1;
}

/* mesh_chains (pt_mesh.rs:375-457), no overlay boxes */
typedef struct {
    uint64_t *keys;
    uint64_t cap, n;
} u64set;

static int set_insert(u64set *s, uint64_t k)
{
    if ((s->n + 1) * 2 > s->cap) {
        uint64_t ncap = s->cap ? s->cap * 2 : 1024;
        uint64_t *nk = (uint64_t *)calloc(ncap, sizeof(uint64_t));
        for (uint64_t i = 0; i < s->cap; i++)
            if (s->keys[i]) {
                uint64_t h = (s->keys[i] * 0x9E3779B97F4A7C15ULL) & (ncap - 1);
                while (nk[h])
                    h = (h + 1) & (ncap - 1);
                nk[h] = s->keys[i];
            }
        free(s->keys);
        s->keys = nk;
        s->cap = ncap;
    }
    uint64_t h = (k * 0x9E3779B97F4A7C15ULL) & (s->cap - 1);
    while (s->keys[h]) {
        if (s->keys[h] == k)
            return 0;
        h = (h + 1) & (s->cap - 1);
    }
    s->keys[h] = k;
    s->n++;
    return 1;
}

int or_mesh_chains(const or_mesh *m, const float *boxes, int n_boxes, uint16_t base,
                   const float lisf[3], uint32_t n_rays, uint32_t rot, int use_project_trig,
                   uint16_t *chains, uint8_t *orders, int cap)
{
    u64set seen = {0, 0, 0};
    float ga = OR_PI * (3.0f - sqrtf(5.0f));
    uint64_t js = (0xC6Bull ^ (uint64_t)rot) | 1ull;
    int n = 0;
    for (uint32_t i = 0; i < n_rays; i++) {
        float z = 1.0f - 2.0f * ((float)i + 0.5f) / (float)n_rays;
        float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
        float phi = ga * (float)i + (float)rot * 0.61803398875f * OR_TAU +
                    or_xorshift_f32(&js) * 0.02f;
        float s, c;
        if (use_project_trig) {
            or_sincosf(phi, &s, &c);
        } else {
            s = sinf(phi);
            c = cosf(phi);
        }
        v3 dir = V(r * c, r * s, z);
        v3 pos = vload(lisf);
        uint16_t chain[3] = {0xFFFF, 0xFFFF, 0xFFFF};
        for (int k = 0; k < 3; k++) {
            float pf[3], df[3], t;
            uint32_t tri;
            vstore(pf, pos);
            vstore(df, dir);
            int has = or_mesh_raycast(m, pf, df, &t, &tri, NULL, NULL);
            if (!has) t = FLT_MAX;
            uint16_t sid = has ? m->tri_surface[tri] : 0;
            v3 normal = has ? m->normals[tri] : V(0.0f, 0.0f, 1.0f);
            /* overlay boxes: slab-entry test from 1e-4, nearest wins (pt_mesh.rs:391-425) */
            for (int bi = 0; bi < n_boxes; bi++) {
                const float *mn = boxes + 6 * bi, *mx = boxes + 6 * bi + 3;
                float t0 = 1e-4f, t1 = t;
                int axis = 3;
                for (int a = 0; a < 3; a++) {
                    float da = df[a];
                    if (fabsf(da) < 1e-9f) {
                        if (pf[a] < mn[a] || pf[a] > mx[a]) {
                            t0 = FLT_MAX;
                            break;
                        }
                    } else {
                        float ta = (mn[a] - pf[a]) / da, tb = (mx[a] - pf[a]) / da;
                        if (ta > tb) {
                            float sw = ta;
                            ta = tb;
                            tb = sw;
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
                if (t0 < t && axis < 3) {
                    t = t0;
                    int face = axis * 2 + (df[axis] > 0.0f ? 0 : 1);
                    sid = (uint16_t)(base + (uint16_t)(bi * 6 + face));
                    float nv[3] = {0.0f, 0.0f, 0.0f};
                    nv[axis] = df[axis] > 0.0f ? -1.0f : 1.0f;
                    normal = V(nv[0], nv[1], nv[2]);
                    has = 1;
                }
            }
            if (!has)
                break;
            if (t <= 1e-4f || t > 200.0f)
                break;
            pos = vadd(pos, vmul(dir, t));
            chain[k] = sid;
            uint64_t key = 1;
            for (int q = 0; q <= k; q++)
                key = key * 65536u + chain[q] + 1u;
            if (set_insert(&seen, key)) {
                if (n < cap) {
                    memcpy(chains + 3 * n, chain, sizeof chain);
                    orders[n] = (uint8_t)(k + 1);
                }
                n++;
            }
            v3 nn = vdot(normal, dir) > 0.0f ? vmul(normal, -1.0f) : normal;
            dir = vsub(dir, vmul(nn, 2.0f * vdot(dir, nn)));
            pos = vadd(pos, vmul(nn, 1e-4f));
        }
    }
    free(seen.keys);
    return n < cap ? n : cap;
}
