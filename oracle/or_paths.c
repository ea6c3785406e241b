/*
 * AutoPaths knife-edge path search (crates/omg-core/src/paths.rs,
 * diffraction.rs, mesh.rs:622-682, material.rs:59-62), restated in C.
 * TEST INFRASTRUCTURE ONLY: the checker of the device path search.
 *
 * Float expressions keep the reference's evaluation order.  acosf (edge
 * extraction) is the C library's, as Rust's f32::acos calls the platform
 * libm.  The knife-edge kernel's tanh, tan, log10 and pow are evaluated
 * in f64 and rounded once to f32: a correctly rounded f32 result in all
 * but double-rounding cases (about 1 in 2^28 arguments), within an ulp of
 * the platform's f32 functions, and reproducible by the device, whose f64
 * functions are accurate to a few f64 ulps.  Like the tracer's shared
 * sine/cosine this replaces the platform libm on both sides.
 *
 * One deliberate difference: extract_edges iterates a HashMap, so the
 * reference's order among edges of equal importance is unspecified; here
 * edges are produced in ascending (min vertex, max vertex) order before
 * the stable importance sort.
 */
#include <stdlib.h>
#include <string.h>

#include "or_mesh_internal.h"

#define NB 3
static const float kBandHz[NB] = {125.0f, 790.0f, 5600.0f};   /* diffraction.rs:15 */
static const float kPi = 3.14159274101257324219f;
static const float kSpeed = 343.0f;

/* f32::total_cmp as an integer key (larger key = larger value) */
static int32_t total_key(float f)
{
    int32_t k;
    memcpy(&k, &f, 4);
    return k ^ (int32_t)((uint32_t)(k >> 31) >> 1);
}

typedef struct {
    float v;
    int idx;   /* payload */
    int ord;   /* original position: ties keep it (stable sort) */
} rank_t;

static int rank_desc(const void *a, const void *b)
{
    const rank_t *x = a, *y = b;
    int32_t kx = total_key(x->v), ky = total_key(y->v);
    if (kx != ky)
        return kx > ky ? -1 : 1;
    return x->ord - y->ord;
}

/* slice.sort_by(|x, y| y.0.total_cmp(&x.0)) then truncate(n) */
static int sort_truncate(rank_t *r, int n, int keep)
{
    for (int i = 0; i < n; i++)
        r[i].ord = i;
    qsort(r, n, sizeof *r, rank_desc);
    return n < keep ? n : keep;
}

/* Material::transmission_at (material.rs:59-62) */
static void transmission_at(const or_material *m, float thickness, float out[NB])
{
    float k = fminf(0.15f / fmaxf(thickness, 0.02f), 2.0f);
    for (int b = 0; b < NB; b++)
        out[b] = fminf(m->transmission[b] * k, 0.9f);
}

/* AutoPaths::transmission (paths.rs:84-113) */
void or_paths_transmission(const or_mesh *m, const float af[3], const float bf[3], float g[NB])
{
    int cap = 4096;
    float *t = malloc(cap * sizeof(float));
    uint32_t *tri = malloc(cap * sizeof(uint32_t));
    uint16_t *mat = malloc(cap * sizeof(uint16_t));
    int n = or_mesh_segment_hits(m, af, bf, t, tri, mat, cap, NULL);
    float len = vlen(vsub(vload(bf), vload(af)));
    for (int b = 0; b < NB; b++)
        g[b] = 1.0f;
    int i = 0;
    while (i < n) {
        int enter = i;
        const or_material *mm = &m->materials[mat[enter]];
        float thickness;
        if (i + 1 < n && mat[i + 1] == mat[enter]) {
            i += 1;
            thickness = fmaxf((t[i] - t[enter]) * len, 0.01f);
        } else {
            thickness = 0.1f;
        }
        float tr[NB];
        transmission_at(mm, thickness, tr);
        for (int b = 0; b < NB; b++)
            g[b] *= tr[b];
        i += 1;
    }
    free(t);
    free(tri);
    free(mat);
}

/* knife_edge_amp (diffraction.rs:27-42) */
float or_knife_edge_amp(float detour_m, float freq_hz)
{
    float n = 2.0f * detour_m * freq_hz / kSpeed;
    if (n <= -0.1916f)
        return 1.0f;
    float x = sqrtf(2.0f * kPi * fabsf(n));
    float ratio;
    if (x < 1e-3f)
        ratio = 1.0f;
    else if (n >= 0.0f)
        ratio = x / (float)tanh((double)x);
    else
        ratio = x / (float)tan((double)x);
    float a_db = 5.0f + 20.0f * (float)log10((double)ratio);
    return (float)pow(10.0, (double)(-a_db / 20.0f));
}

void or_knife_edge_bands(float detour_m, float out[NB])
{
    for (int b = 0; b < NB; b++)
        out[b] = or_knife_edge_amp(detour_m, kBandHz[b]);
}

/* bend_point (paths.rs:411-436): golden section on t in [0, 1] */
static v3 bend_point(v3 a, v3 b, v3 src, v3 lis)
{
    v3 ab = vsub(b, a);
#define COST(tt) (vlen(vsub(vadd(a, vmul(ab, (tt))), src)) + vlen(vsub(lis, vadd(a, vmul(ab, (tt))))))
    const float phi = 0.618034f;
    float lo = 0.0f, hi = 1.0f;
    float m1 = hi - phi * (hi - lo), m2 = lo + phi * (hi - lo);
    float c1 = COST(m1), c2 = COST(m2);
    for (int k = 0; k < 24; k++) {
        if (c1 <= c2) {
            hi = m2;
            m2 = m1;
            c2 = c1;
            m1 = hi - phi * (hi - lo);
            c1 = COST(m1);
        } else {
            lo = m1;
            m1 = m2;
            c1 = c2;
            m2 = lo + phi * (hi - lo);
            c2 = COST(m2);
        }
    }
#undef COST
    return vadd(a, vmul(ab, 0.5f * (lo + hi)));
}

void or_bend_point(const float a[3], const float b[3], const float src[3], const float lis[3],
                   float p[3])
{
    vstore(p, bend_point(vload(a), vload(b), vload(src), vload(lis)));
}

/* ------------------------------------------------------------ edges */

typedef struct {
    uint64_t key;
    uint32_t tri;
    uint32_t ord;
} half_t;

static int half_cmp(const void *a, const void *b)
{
    const half_t *x = a, *y = b;
    if (x->key != y->key)
        return x->key < y->key ? -1 : 1;
    return (int)x->ord - (int)y->ord;
}

/* extract_edges (mesh.rs:639-676) then budget_edges (mesh.rs:678-682).
 * Returns the count written (<= cap). */
int or_paths_edges(const or_mesh *m, float min_dihedral_deg, uint32_t budget, or_edge *out, int cap)
{
    uint32_t nh = 3 * m->nt;
    half_t *h = malloc((nh ? nh : 1) * sizeof *h);
    for (uint32_t t = 0; t < m->nt; t++)
        for (int k = 0; k < 3; k++) {
            uint32_t i = m->indices[3 * t + k], j = m->indices[3 * t + (k + 1) % 3];
            uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
            h[3 * t + k].key = (uint64_t)lo << 32 | hi;
            h[3 * t + k].tri = t;
            h[3 * t + k].ord = 3 * t + k;
        }
    qsort(h, nh, sizeof *h, half_cmp);
    or_edge *all = malloc((nh ? nh : 1) * sizeof *all);
    int ne = 0;
    for (uint32_t s = 0; s < nh;) {
        uint32_t e = s;
        while (e < nh && h[e].key == h[s].key)
            e++;
        uint32_t i = (uint32_t)(h[s].key >> 32), j = (uint32_t)h[s].key;
        v3 a = m->positions[i], b = m->positions[j];
        float len = vlen(vsub(b, a));
        int keep = len >= 1e-4f;
        float sharp = 0.0f;
        if (keep) {
            if (e - s == 1) {
                sharp = 1.0f;
            } else if (e - s == 2) {
                v3 n0 = m->normals[h[s].tri], n1 = m->normals[h[s + 1].tri];
                float c = vdot(n0, n1);
                c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
                float angle = acosf(c) * 57.2957795130823208767981548141051703f;
                if (angle < min_dihedral_deg)
                    keep = 0;
                else
                    sharp = angle / 180.0f;
            } else {
                keep = 0;
            }
        }
        if (keep) {
            vstore(all[ne].a, a);
            vstore(all[ne].b, b);
            all[ne].importance = len * sharp;
            ne++;
        }
        s = e;
    }
    rank_t *r = malloc((ne ? ne : 1) * sizeof *r);
    for (int i = 0; i < ne; i++) {
        r[i].v = all[i].importance;
        r[i].idx = i;
    }
    int nk = sort_truncate(r, ne, (int)budget);
    int w = nk < cap ? nk : cap;
    for (int i = 0; i < w; i++)
        out[i] = all[r[i].idx];
    free(r);
    free(all);
    free(h);
    return w;
}

/* ------------------------------------------------------------ find */

static void push_path(or_path *out, int *n, int cap, const v3 *pts, int np, float length,
                      const float g[NB], uint32_t key)
{
    if (*n >= cap)
        return;
    or_path *p = &out[(*n)++];
    memset(p, 0, sizeof *p);
    p->key = key;
    p->n_points = (uint32_t)np;
    p->length = length;
    for (int b = 0; b < NB; b++)
        p->gains[b] = g[b];
    for (int k = 0; k < np && k < OR_PATH_MAX_POINTS; k++)
        vstore(p->points[k], pts[k]);
}

/* rank_corridor (paths.rs:143-163) */
static int rank_corridor(const or_edge *edges, int ne, v3 src, v3 seg, float seg_len2, float lo_t,
                         float hi_t, int n, int *out)
{
    rank_t *r = malloc((ne ? ne : 1) * sizeof *r);
    int k = 0;
    for (int i = 0; i < ne; i++) {
        v3 c = vmul(vadd(vload(edges[i].a), vload(edges[i].b)), 0.5f);
        float t = vdot(vsub(c, src), seg) / seg_len2;
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        if (t < lo_t || t > hi_t)
            continue;
        v3 on = vadd(src, vmul(seg, t));
        float d2 = vdot(vsub(c, on), vsub(c, on));
        r[k].v = edges[i].importance / (d2 + 0.5f);
        r[k].idx = i;
        k++;
    }
    int nk = sort_truncate(r, k, n);
    for (int i = 0; i < nk; i++)
        out[i] = r[i].idx;
    free(r);
    return nk;
}

/* over_silhouette (paths.rs:337-405); returns 1 and the path if found */
static int over_silhouette(const or_mesh *m, v3 src, v3 lis, or_path *p)
{
    enum { K = 24 };
    v3 flat = V(lis.x - src.x, lis.y - src.y, 0.0f);
    float run = vlen(flat);
    if (run < 1.0f)
        return 0;
    const float down[3] = {0.0f, 0.0f, -1.0f};
    float ps[K + 1], pz[K + 1];
    int np = 0;
    ps[np] = 0.0f;
    pz[np] = src.z;
    np++;
    for (int i = 1; i < K; i++) {
        float t = (float)i / (float)K;
        float x = src.x + flat.x * t;
        float y = src.y + flat.y * t;
        float line_z = src.z + (lis.z - src.z) * t;
        float o[3] = {x, y, 60.0f}, th;
        uint32_t tri;
        if (or_mesh_raycast(m, o, down, &th, &tri, NULL, NULL)) {
            float top = 60.0f - th;
            if (top > line_z + 0.05f) {
                ps[np] = run * t;
                pz[np] = top + 0.03f;
                np++;
            }
        }
    }
    if (np == 1)
        return 0;
    ps[np] = run;
    pz[np] = lis.z;
    np++;
    float hs[K + 1], hz[K + 1];
    int nh = 0;
    for (int k = 0; k < np; k++) {
        while (nh >= 2) {
            float ax = hs[nh - 2], az = hz[nh - 2], bx = hs[nh - 1], bz = hz[nh - 1];
            if ((bx - ax) * (pz[k] - az) - (bz - az) * (ps[k] - ax) >= 0.0f)
                nh--;
            else
                break;
        }
        hs[nh] = ps[k];
        hz[nh] = pz[k];
        nh++;
    }
    if (nh < 3)
        return 0;
#define D2(i, j) sqrtf((hs[i] - hs[j]) * (hs[i] - hs[j]) + (hz[i] - hz[j]) * (hz[i] - hz[j]))
    float g[NB] = {1.0f, 1.0f, 1.0f};
    float length = 0.0f;
    for (int i = 0; i + 1 < nh; i++)
        length += D2(i, i + 1);
    for (int i = 1; i < nh - 1; i++) {
        float detour = D2(i - 1, i) + D2(i, i + 1) - D2(i - 1, i + 1);
        float ke[NB];
        or_knife_edge_bands(fmaxf(detour, 1e-4f), ke);
        for (int b = 0; b < NB; b++)
            g[b] *= ke[b];
    }
#undef D2
    v3 pts[K + 1];
    for (int i = 0; i < nh; i++) {
        float t = hs[i] / run;
        pts[i] = V(src.x + flat.x * t, src.y + flat.y * t, hz[i]);
    }
    int n = 0;
    push_path(p, &n, 1, pts, nh, length, g, 90000u);
    return 1;
}

typedef struct {
    float bound;
    v3 p1, p2;
    int ei, ej;
} pair_t;

/* AutoPaths::find (paths.rs:117-332).  Returns the number of paths
 * written (the direct path first, then the strongest bent paths). */
int or_paths_find(const or_mesh *m, const or_edge *edges, int ne, const float srcf[3],
                  const float lisf[3], const or_path_budget *bud, or_path *out, int cap)
{
    v3 src = vload(srcf), lis = vload(lisf);
    int n_out = 0;
    float direct[NB];
    or_paths_transmission(m, srcf, lisf, direct);
    int blocked = direct[1] < 0.25f;
    v3 dpts[2] = {src, lis};
    push_path(out, &n_out, cap, dpts, 2, vlen(vsub(lis, src)), direct, 0u);
    if (!blocked)
        return n_out;
    v3 seg = vsub(lis, src);
    float seg_len2 = fmaxf(vdot(seg, seg), 1e-6f);
    float direct_len = vlen(vsub(lis, src));

    int nb_cap = (int)bud->edge_candidates + (int)(bud->pair_edges * bud->pair_edges) + 1;
    or_path *bent = calloc(nb_cap, sizeof *bent);
    int nbent = 0;

    /* single bends ranked by their knife-edge bound (edges <= 1024: no prefilter) */
    int *pre = malloc((ne ? ne : 1) * sizeof(int));
    int npre;
    if (ne > 1024) {
        npre = rank_corridor(edges, ne, src, seg, seg_len2, 0.0f, 1.0f, 1024, pre);
    } else {
        npre = ne;
        for (int i = 0; i < ne; i++)
            pre[i] = i;
    }
    rank_t *rk = malloc((npre ? npre : 1) * sizeof *rk);
    int nr = 0;
    for (int q = 0; q < npre; q++) {
        int ei = pre[q];
        v3 p = bend_point(vload(edges[ei].a), vload(edges[ei].b), src, lis);
        float d1 = vlen(vsub(p, src)), d2 = vlen(vsub(lis, p));
        if (d1 < 0.2f || d2 < 0.2f)
            continue;
        float detour = d1 + d2 - direct_len;
        float ke[NB];
        or_knife_edge_bands(fmaxf(detour, 1e-4f), ke);
        rk[nr].v = ke[1];
        rk[nr].idx = ei;
        nr++;
    }
    nr = sort_truncate(rk, nr, (int)bud->edge_candidates);
    for (int q = 0; q < nr; q++) {
        int ei = rk[q].idx;
        v3 p = bend_point(vload(edges[ei].a), vload(edges[ei].b), src, lis);
        float d1 = vlen(vsub(p, src)), d2 = vlen(vsub(lis, p));
        if (d1 < 0.2f || d2 < 0.2f)
            continue;
        v3 e1 = p, e2 = p;
        if (bud->taps) {
            e1 = vadd(p, vmul(vnorm(vsub(p, src)), 0.06f));
            e2 = vadd(p, vmul(vnorm(vsub(p, lis)), 0.06f));
        }
        float l1[NB], l2[NB], f1[3], f2[3];
        vstore(f1, e1);
        vstore(f2, e2);
        or_paths_transmission(m, srcf, f1, l1);
        or_paths_transmission(m, f2, lisf, l2);
        if (l1[0] * l2[0] < 1e-4f)
            continue;
        float detour = d1 + d2 - direct_len;
        float ke[NB], g[NB];
        or_knife_edge_bands(fmaxf(detour, 1e-4f), ke);
        for (int b = 0; b < NB; b++)
            g[b] = ke[b] * l1[b] * l2[b];
        v3 pts[3] = {src, p, lis};
        push_path(bent, &nbent, nb_cap, pts, 3, d1 + d2, g, 1u + (uint32_t)ei);
    }
    free(rk);
    free(pre);

    /* double bends */
    int *near_src = malloc((bud->pair_edges + 1) * sizeof(int));
    int *near_lis = malloc((bud->pair_edges + 1) * sizeof(int));
    int ns = rank_corridor(edges, ne, src, seg, seg_len2, 0.0f, 0.65f, (int)bud->pair_edges, near_src);
    int nl = rank_corridor(edges, ne, src, seg, seg_len2, 0.35f, 1.0f, (int)bud->pair_edges, near_lis);
    pair_t *cand = malloc((ns * nl + 1) * sizeof *cand);
    int nc = 0;
    for (int a = 0; a < ns; a++)
        for (int c = 0; c < nl; c++) {
            int ei = near_src[a], ej = near_lis[c];
            if (ei == ej)
                continue;
            v3 e1a = vload(edges[ei].a), e1b = vload(edges[ei].b);
            v3 e2a = vload(edges[ej].a), e2b = vload(edges[ej].b);
            v3 p2 = vmul(vadd(e2a, e2b), 0.5f);
            v3 p1 = bend_point(e1a, e1b, src, p2);
            for (int it = 0; it < 3; it++) {
                p2 = bend_point(e2a, e2b, p1, lis);
                p1 = bend_point(e1a, e1b, src, p2);
            }
            float d1 = vlen(vsub(p1, src)), dm = vlen(vsub(p2, p1)), d2 = vlen(vsub(lis, p2));
            if (d1 < 0.2f || dm < 0.2f || d2 < 0.2f)
                continue;
            float det1 = d1 + dm - vlen(vsub(p2, src));
            float det2 = dm + d2 - vlen(vsub(lis, p1));
            float k1[NB], k2[NB];
            or_knife_edge_bands(fmaxf(det1, 1e-4f), k1);
            or_knife_edge_bands(fmaxf(det2, 1e-4f), k2);
            cand[nc].bound = k1[1] * k2[1];
            cand[nc].p1 = p1;
            cand[nc].p2 = p2;
            cand[nc].ei = ei;
            cand[nc].ej = ej;
            nc++;
        }
    rank_t *rc = malloc((nc ? nc : 1) * sizeof *rc);
    for (int i = 0; i < nc; i++) {
        rc[i].v = cand[i].bound;
        rc[i].idx = i;
    }
    int ncs = sort_truncate(rc, nc, 64);
    for (int q = 0; q < ncs; q++) {
        const pair_t *c = &cand[rc[q].idx];
        v3 p1 = c->p1, p2 = c->p2;
        float d1 = vlen(vsub(p1, src)), dm = vlen(vsub(p2, p1)), d2 = vlen(vsub(lis, p2));
        v3 x1 = p1, m1 = p1, m2 = p2, x2 = p2;
        if (bud->taps) {
            x1 = vadd(p1, vmul(vnorm(vsub(p1, src)), 0.06f));
            m1 = vadd(p1, vmul(vnorm(vsub(p1, p2)), 0.06f));
            m2 = vadd(p2, vmul(vnorm(vsub(p2, p1)), 0.06f));
            x2 = vadd(p2, vmul(vnorm(vsub(p2, lis)), 0.06f));
        }
        float fx1[3], fm1[3], fm2[3], fx2[3], l1[NB], lm[NB], l2[NB];
        vstore(fx1, x1);
        vstore(fm1, m1);
        vstore(fm2, m2);
        vstore(fx2, x2);
        or_paths_transmission(m, srcf, fx1, l1);
        or_paths_transmission(m, fm1, fm2, lm);
        or_paths_transmission(m, fx2, lisf, l2);
        if (l1[0] * lm[0] * l2[0] < 1e-4f)
            continue;
        float det1 = d1 + dm - vlen(vsub(p2, src));
        float det2 = dm + d2 - vlen(vsub(lis, p1));
        float k1[NB], k2[NB], g[NB];
        or_knife_edge_bands(fmaxf(det1, 1e-4f), k1);
        or_knife_edge_bands(fmaxf(det2, 1e-4f), k2);
        for (int b = 0; b < NB; b++)
            g[b] = k1[b] * k2[b] * l1[b] * lm[b] * l2[b];
        v3 pts[4] = {src, p1, p2, lis};
        push_path(bent, &nbent, nb_cap, pts, 4, d1 + dm + d2, g,
                  100000u + (uint32_t)(c->ei * 4096 + c->ej));
    }
    free(rc);
    free(cand);
    free(near_src);
    free(near_lis);

    /* over the silhouette */
    or_path sil;
    if (over_silhouette(m, src, lis, &sil) && nbent < nb_cap)
        bent[nbent++] = sil;

    /* strongest first, capped */
    rank_t *rb = malloc((nbent ? nbent : 1) * sizeof *rb);
    for (int i = 0; i < nbent; i++) {
        rb[i].v = bent[i].gains[1];
        rb[i].idx = i;
    }
    int keep = sort_truncate(rb, nbent, (int)bud->max_paths);
    for (int i = 0; i < keep && n_out < cap; i++)
        out[n_out++] = bent[rb[i].idx];
    free(rb);
    free(bent);
    return n_out;
}
