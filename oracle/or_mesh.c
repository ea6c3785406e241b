/*
 * Oracle restatement of crates/omg-core/src/mesh.rs: vertex dedup,
 * tessellation, binned-SAH BVH, ordered raycast, segment crossings.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include "or_mesh_internal.h"
#include <stdlib.h>
#include <string.h>

#define LEAF_SIZE 4u              /* mesh.rs:16 */
#define LEAF_BIT 0x80000000u      /* mesh.rs:17 */
#define STACK_CAP 256             /* the reference's stack is 64 (mesh.rs:492) */

/* ------------------------------------------------------------ builder dedup */

/* MeshBuilder::vertex (mesh.rs:108-114): key = (v * 1e4).round() as i64
 * per axis; the first position with a key wins. */
static int64_t qkey(float v)
{
    float r = roundf(v * 1e4f);
    if (r != r)
        return 0;
    if (r >= 9223372036854775808.0f)
        return INT64_MAX;
    if (r <= -9223372036854775808.0f)
        return INT64_MIN;
    return (int64_t)r;
}

typedef struct {
    int64_t k[3];
    uint32_t idx;
    uint32_t used;
} vslot;

static uint64_t hash3(const int64_t k[3])
{
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < 3; i++) {
        h ^= (uint64_t)k[i];
        h *= 1099511628211ULL;
        h ^= h >> 29;
    }
    return h;
}

uint32_t or_dedup_tris(const float *tv, uint32_t nt, float *pos_out, uint32_t *idx_out)
{
    uint64_t cap = 16;
    while (cap < (uint64_t)nt * 3 * 2)
        cap <<= 1;
    vslot *tab = (vslot *)calloc(cap, sizeof(vslot));
    uint32_t nv = 0;
    for (uint64_t i = 0; i < (uint64_t)nt * 3; i++) {
        const float *p = tv + i * 3;
        int64_t k[3] = {qkey(p[0]), qkey(p[1]), qkey(p[2])};
        uint64_t h = hash3(k) & (cap - 1);
        for (;;) {
            vslot *s = &tab[h];
            if (!s->used) {
                s->used = 1;
                memcpy(s->k, k, sizeof k);
                s->idx = nv;
                memcpy(pos_out + (uint64_t)nv * 3, p, 3 * sizeof(float));
                nv++;
                idx_out[i] = s->idx;
                break;
            }
            if (s->k[0] == k[0] && s->k[1] == k[1] && s->k[2] == k[2]) {
                idx_out[i] = s->idx;
                break;
            }
            h = (h + 1) & (cap - 1);
        }
    }
    free(tab);
    return nv;
}

/* ------------------------------------------------------------ BVH build */

typedef struct {
    v3 a, b, c;
    uint32_t t;
} work_item;

static void pbounds(const or_prim *p, v3 *lo, v3 *hi)
{
    v3 a = p->a, b = vadd(p->a, p->e1), c = vadd(p->a, p->e2);
    *lo = vmin(vmin(a, b), c);
    *hi = vmax(vmax(a, b), c);
}
static v3 pcentroid(const or_prim *p)
{
    return vadd(p->a, vmul(vadd(p->e1, p->e2), 1.0f / 3.0f));
}
static float area(v3 lo, v3 hi)
{
    v3 e = vmax(vsub(hi, lo), V(0.0f, 0.0f, 0.0f));
    return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
}

static uint32_t push_node(or_mesh *m, v3 bmin, v3 bmax)
{
    if (m->n_nodes == m->cap_nodes) {
        m->cap_nodes = m->cap_nodes ? m->cap_nodes * 2 : 64;
        m->nodes = (or_node *)realloc(m->nodes, m->cap_nodes * sizeof(or_node));
    }
    or_node *n = &m->nodes[m->n_nodes];
    n->bmin = bmin;
    n->bmax = bmax;
    n->a = 0;
    n->b = 0;
    return m->n_nodes++;
}

static size_t bin_of(const or_prim *p, int axis, v3 cmin, float cwidth)
{
    float f = (vget(pcentroid(p), axis) - vget(cmin, axis)) / cwidth * 16.0f;
    return (size_t)sat_usize(fminf(f, 15.0f));
}

/* Mesh::build_node (mesh.rs:292-405) */
static uint32_t build_node(or_mesh *m, or_prim *prims, size_t start, size_t end, uint32_t depth)
{
    v3 bmin = V(FLT_MAX, FLT_MAX, FLT_MAX), bmax = V(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    v3 cmin = bmin, cmax = bmax;
    for (size_t i = start; i < end; i++) {
        v3 lo, hi;
        pbounds(&prims[i], &lo, &hi);
        bmin = vmin(bmin, lo);
        bmax = vmax(bmax, hi);
        v3 c = pcentroid(&prims[i]);
        cmin = vmin(cmin, c);
        cmax = vmax(cmax, c);
    }
    uint32_t idx = push_node(m, bmin, bmax);
    if (depth > m->max_depth)
        m->max_depth = depth;
    size_t count = end - start;
    if (count <= LEAF_SIZE) {
        m->nodes[idx].a = LEAF_BIT | (uint32_t)start;
        m->nodes[idx].b = (uint32_t)count;
        return idx;
    }
    v3 cext = vsub(cmax, cmin);
    int axis = (cext.x >= cext.y && cext.x >= cext.z) ? 0 : (cext.y >= cext.z ? 1 : 2);
    float cwidth = vget(cext, axis);
    size_t mid;
    if (cwidth < 1e-6f) {
        mid = start + count / 2;
    } else {
        v3 bnmin[16], bnmax[16];
        uint32_t bnn[16];
        for (int k = 0; k < 16; k++) {
            bnmin[k] = V(FLT_MAX, FLT_MAX, FLT_MAX);
            bnmax[k] = V(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            bnn[k] = 0;
        }
        for (size_t i = start; i < end; i++) {
            size_t k = bin_of(&prims[i], axis, cmin, cwidth);
            v3 lo, hi;
            pbounds(&prims[i], &lo, &hi);
            bnmin[k] = vmin(bnmin[k], lo);
            bnmax[k] = vmax(bnmax[k], hi);
            bnn[k] += 1;
        }
        float right_cost[16] = {0};
        v3 amin = V(FLT_MAX, FLT_MAX, FLT_MAX), amax = V(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        uint32_t an = 0;
        for (int k = 15; k >= 1; k--) {
            amin = vmin(amin, bnmin[k]);
            amax = vmax(amax, bnmax[k]);
            an += bnn[k];
            right_cost[k - 1] = an > 0 ? area(amin, amax) * (float)an : 0.0f;
        }
        int best_cut = -1;
        float best_cost = FLT_MAX;
        amin = V(FLT_MAX, FLT_MAX, FLT_MAX);
        amax = V(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        an = 0;
        for (int k = 0; k < 15; k++) {
            amin = vmin(amin, bnmin[k]);
            amax = vmax(amax, bnmax[k]);
            an += bnn[k];
            if (an == 0 || an == (uint32_t)count)
                continue;
            float cost = area(amin, amax) * (float)an + right_cost[k];
            if (cost < best_cost) {
                best_cost = cost;
                best_cut = k;
            }
        }
        if (best_cut < 0) {
            mid = start + count / 2;
        } else {
            /* stable two-pointer partition by bin (mesh.rs:383-395) */
            or_prim *s = prims + start;
            size_t i = 0, j = count;
            while (i < j) {
                if (bin_of(&s[i], axis, cmin, cwidth) <= (size_t)best_cut) {
                    i++;
                } else {
                    j--;
                    or_prim tmp = s[i];
                    s[i] = s[j];
                    s[j] = tmp;
                }
            }
            mid = start + i;
        }
    }
    if (mid < start + 1)
        mid = start + 1;
    if (mid > end - 1)
        mid = end - 1;
    uint32_t left = build_node(m, prims, start, mid, depth + 1);
    uint32_t right = build_node(m, prims, mid, end, depth + 1);
    m->nodes[idx].a = left;
    m->nodes[idx].b = right;
    return idx;
}

/* Mesh::with_surfaces_edges + build_bvh (mesh.rs:198-287) */
or_mesh *or_mesh_build(const float *positions, uint32_t nv, const uint32_t *indices, uint32_t nt,
                       const uint16_t *tri_material, const uint16_t *tri_surface,
                       const float *tri_max_edge, const or_material *materials, uint32_t nm,
                       uint32_t nbands)
{
    or_mesh *m = (or_mesh *)calloc(1, sizeof(or_mesh));
    m->nv = nv;
    m->nt = nt;
    m->nm = nm;
    m->nbands = nbands;
    m->positions = (v3 *)malloc((nv ? nv : 1) * sizeof(v3));
    for (uint32_t i = 0; i < nv; i++)
        m->positions[i] = vload(positions + 3 * (size_t)i);
    m->indices = (uint32_t *)malloc((nt ? nt : 1) * 3 * sizeof(uint32_t));
    memcpy(m->indices, indices, (size_t)nt * 3 * sizeof(uint32_t));
    m->tri_material = (uint16_t *)malloc((nt ? nt : 1) * sizeof(uint16_t));
    memcpy(m->tri_material, tri_material, nt * sizeof(uint16_t));
    m->tri_surface = (uint16_t *)calloc(nt ? nt : 1, sizeof(uint16_t));
    if (tri_surface)
        memcpy(m->tri_surface, tri_surface, nt * sizeof(uint16_t));
    m->materials = (or_material *)malloc((nm ? nm : 1) * sizeof(or_material));
    memcpy(m->materials, materials, nm * sizeof(or_material));
    m->normals = (v3 *)malloc((nt ? nt : 1) * sizeof(v3));
    for (uint32_t t = 0; t < nt; t++) {
        v3 a = m->positions[indices[3 * (size_t)t]];
        v3 b = m->positions[indices[3 * (size_t)t + 1]];
        v3 c = m->positions[indices[3 * (size_t)t + 2]];
        m->normals[t] = vnorm(vcross(vsub(b, a), vsub(c, a)));   /* mesh.rs:458-461 */
    }

    /* tessellation: LIFO work list, longest-edge midpoint split (mesh.rs:244-279) */
    size_t wcap = (size_t)nt + 64, wn = 0;
    work_item *work = (work_item *)malloc(wcap * sizeof(work_item));
    for (uint32_t t = 0; t < nt; t++) {
        work[wn].a = m->positions[indices[3 * (size_t)t]];
        work[wn].b = m->positions[indices[3 * (size_t)t + 1]];
        work[wn].c = m->positions[indices[3 * (size_t)t + 2]];
        work[wn].t = t;
        wn++;
    }
    size_t pcap = (size_t)nt + 16, pn = 0;
    or_prim *prims = (or_prim *)malloc(pcap * sizeof(or_prim));
    while (wn > 0) {
        work_item w = work[--wn];
        v3 a = w.a, b = w.b, c = w.c;
        float lab = vlen(vsub(b, a)), lbc = vlen(vsub(c, b)), lca = vlen(vsub(a, c));
        float longest = fmaxf(fmaxf(lab, lbc), lca);
        float lim = tri_max_edge ? tri_max_edge[w.t] : 4.0f;
        if (longest > lim) {
            if (wn + 2 > wcap) {
                wcap *= 2;
                work = (work_item *)realloc(work, wcap * sizeof(work_item));
            }
            if (lab >= lbc && lab >= lca) {
                v3 mm = vmul(vadd(a, b), 0.5f);
                work[wn++] = (work_item){a, mm, c, w.t};
                work[wn++] = (work_item){mm, b, c, w.t};
            } else if (lbc >= lca) {
                v3 mm = vmul(vadd(b, c), 0.5f);
                work[wn++] = (work_item){a, b, mm, w.t};
                work[wn++] = (work_item){a, mm, c, w.t};
            } else {
                v3 mm = vmul(vadd(c, a), 0.5f);
                work[wn++] = (work_item){a, b, mm, w.t};
                work[wn++] = (work_item){mm, b, c, w.t};
            }
            continue;
        }
        if (pn == pcap) {
            pcap *= 2;
            prims = (or_prim *)realloc(prims, pcap * sizeof(or_prim));
        }
        prims[pn].a = a;
        prims[pn].e1 = vsub(b, a);
        prims[pn].e2 = vsub(c, a);
        prims[pn].tri = w.t;
        prims[pn].material = tri_material[w.t];
        pn++;
    }
    free(work);
    m->packed = prims;
    m->n_prims = (uint32_t)pn;
    if (pn > 0)
        build_node(m, prims, 0, pn, 1);
    return m;
}

void or_mesh_free(or_mesh *m)
{
    if (!m)
        return;
    free(m->positions);
    free(m->indices);
    free(m->tri_material);
    free(m->tri_surface);
    free(m->materials);
    free(m->normals);
    free(m->nodes);
    free(m->packed);
    free(m);
}

void or_mesh_sizes(const or_mesh *m, uint32_t *n_nodes, uint32_t *n_prims)
{
    *n_nodes = m->n_nodes;
    *n_prims = m->n_prims;
}

void or_mesh_export(const or_mesh *m, float *nodes8, float *prims11)
{
    for (uint32_t i = 0; i < m->n_nodes; i++) {
        const or_node *n = &m->nodes[i];
        float *o = nodes8 + 8 * (size_t)i;
        vstore(o, n->bmin);
        vstore(o + 3, n->bmax);
        memcpy(o + 6, &n->a, 4);
        memcpy(o + 7, &n->b, 4);
    }
    for (uint32_t i = 0; i < m->n_prims; i++) {
        const or_prim *p = &m->packed[i];
        float *o = prims11 + 11 * (size_t)i;
        vstore(o, p->a);
        vstore(o + 3, p->e1);
        vstore(o + 6, p->e2);
        uint32_t mat = p->material;
        memcpy(o + 9, &p->tri, 4);
        memcpy(o + 10, &mat, 4);
    }
}

void or_mesh_tri_normal(const or_mesh *m, uint32_t tri, float n[3])
{
    vstore(n, m->normals[tri]);
}

/* ------------------------------------------------------------ queries */

/* Mesh::ray_aabb (mesh.rs:409-426) */
static inline int ray_aabb(v3 bmin, v3 bmax, v3 o, v3 inv_d, float tmax, float *t_enter)
{
    float t0 = 0.0f, t1 = tmax;
    for (int axis = 0; axis < 3; axis++) {
        float inv = vget(inv_d, axis);
        float ta = (vget(bmin, axis) - vget(o, axis)) * inv;
        float tb = (vget(bmax, axis) - vget(o, axis)) * inv;
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
    *t_enter = t0;
    return 1;
}

/* Mesh::ray_tri_packed (mesh.rs:437-456); also returns u, v */
int or_ray_tri(v3 a, v3 e1, v3 e2, v3 o, v3 d, float *t_out, float *u_out, float *v_out)
{
    v3 p = vcross(d, e2);
    float det = vdot(e1, p);
    if (fabsf(det) < 1e-9f)
        return 0;
    float inv = 1.0f / det;
    v3 s = vsub(o, a);
    float u = vdot(s, p) * inv;
    if (!(u >= -1e-6f && u <= 1.0f + 1e-6f))
        return 0;
    v3 q = vcross(s, e1);
    float v = vdot(d, q) * inv;
    if (v < -1e-6f || u + v > 1.0f + 1e-6f)
        return 0;
    float t = vdot(e2, q) * inv;
    if (!(t > 1e-5f))
        return 0;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return 1;
}

static inline int near_edge(float u, float v)
{
    return u < 1e-6f || v < 1e-6f || u + v > 1.0f - 1e-6f;
}

/* Mesh::raycast (mesh.rs:484-541) */
int or_mesh_raycast(const or_mesh *m, const float of[3], const float df[3], float *t_out,
                    uint32_t *tri_out, uint64_t *stats, uint32_t *flags)
{
    if (m->n_nodes == 0)
        return 0;
    v3 o = vload(of), d = vload(df);
    v3 inv_d = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int found = 0;
    float best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;
    uint32_t best_tri = 0, tie = 0;
    uint32_t st_node[STACK_CAP];
    float st_t[STACK_CAP];
    int sp = 1, max_sp = 1;
    st_node[0] = 0;
    st_t[0] = 0.0f;
    uint64_t n_aabb = 0, n_tri = 0;
    while (sp > 0) {
        sp--;
        uint32_t ni = st_node[sp];
        float te = st_t[sp];
        if (found && te >= best_t)
            continue;
        const or_node *node = &m->nodes[ni];
        if (node->a & LEAF_BIT) {
            uint32_t start = node->a & ~LEAF_BIT;
            for (uint32_t k = start; k < start + node->b; k++) {
                const or_prim *pt = &m->packed[k];
                float t, u, v;
                n_tri++;
                if (or_ray_tri(pt->a, pt->e1, pt->e2, o, d, &t, &u, &v)) {
                    if (!found || t < best_t) {
                        found = 1;
                        best_t = t;
                        best_tri = pt->tri;
                        best_u = u;
                        best_v = v;
                        tie = 0;
                    } else if (t == best_t && pt->tri != best_tri) {
                        tie = 1;
                    }
                }
            }
        } else {
            float tmax = found ? best_t : FLT_MAX;
            const or_node *l = &m->nodes[node->a];
            const or_node *r = &m->nodes[node->b];
            float a = 0.0f, b = 0.0f;
            n_aabb += 2;
            int hl = ray_aabb(l->bmin, l->bmax, o, inv_d, tmax, &a);
            int hr = ray_aabb(r->bmin, r->bmax, o, inv_d, tmax, &b);
            if (hl && hr) {
                if (a <= b) {
                    st_node[sp] = node->b;
                    st_t[sp] = b;
                    st_node[sp + 1] = node->a;
                    st_t[sp + 1] = a;
                } else {
                    st_node[sp] = node->a;
                    st_t[sp] = a;
                    st_node[sp + 1] = node->b;
                    st_t[sp + 1] = b;
                }
                sp += 2;
            } else if (hl) {
                st_node[sp] = node->a;
                st_t[sp] = a;
                sp += 1;
            } else if (hr) {
                st_node[sp] = node->b;
                st_t[sp] = b;
                sp += 1;
            }
            if (sp > max_sp)
                max_sp = sp;
            if (sp + 2 > STACK_CAP)
                abort();
        }
    }
    if (stats) {
        stats[0] += n_aabb;
        stats[1] += n_tri;
    }
    if (flags) {
        *flags = 0;
        if (found && near_edge(best_u, best_v))
            *flags |= 1u;
        if (found && tie)
            *flags |= 2u;
        if (max_sp > 64)
            *flags |= 4u;   /* the reference's 64-entry stack would overflow */
    }
    if (!found)
        return 0;
    *t_out = best_t;
    *tri_out = best_tri;
    return 1;
}

typedef struct {
    float t;
    uint32_t tri;
    uint16_t material;
} seghit;

/* Mesh::segment_hits (mesh.rs:546-604) */
int or_mesh_segment_hits(const or_mesh *m, const float af[3], const float bf[3], float *t_out,
                         uint32_t *tri_out, uint16_t *mat_out, int cap, uint64_t *stats)
{
    if (m->n_nodes == 0)
        return 0;
    v3 a = vload(af), b = vload(bf);
    v3 d = vsub(b, a);
    float len = vlen(d);
    if (len < 1e-6f)
        return 0;
    v3 dn = vmul(d, 1.0f / len);
    v3 inv_d = V(1.0f / dn.x, 1.0f / dn.y, 1.0f / dn.z);
    uint32_t stack[STACK_CAP];
    int sp = 1;
    stack[0] = 0;
    int hcap = 64, n = 0;
    seghit *out = (seghit *)malloc(hcap * sizeof(seghit));
    uint64_t n_aabb = 0, n_tri = 0;
    while (sp > 0) {
        sp--;
        const or_node *node = &m->nodes[stack[sp]];
        float te;
        n_aabb++;
        if (!ray_aabb(node->bmin, node->bmax, a, inv_d, len, &te))
            continue;
        if (node->a & LEAF_BIT) {
            uint32_t start = node->a & ~LEAF_BIT;
            for (uint32_t k = start; k < start + node->b; k++) {
                const or_prim *pt = &m->packed[k];
                float t, u, v;
                n_tri++;
                if (or_ray_tri(pt->a, pt->e1, pt->e2, a, dn, &t, &u, &v)) {
                    float tt = t / len;
                    if (tt > 1e-4f && tt < 1.0f - 1e-4f) {
                        if (n == hcap) {
                            hcap *= 2;
                            out = (seghit *)realloc(out, hcap * sizeof(seghit));
                        }
                        out[n].t = tt;
                        out[n].tri = pt->tri;
                        out[n].material = pt->material;
                        n++;
                    }
                }
            }
        } else {
            stack[sp] = node->a;
            stack[sp + 1] = node->b;
            sp += 2;
            if (sp + 2 > STACK_CAP)
                abort();
        }
    }
    if (stats) {
        stats[0] += n_aabb;
        stats[1] += n_tri;
    }
    /* stable sort by t (slice::sort_by is stable) */
    for (int i = 1; i < n; i++) {
        seghit x = out[i];
        int j = i - 1;
        while (j >= 0 && out[j].t > x.t) {
            out[j + 1] = out[j];
            j--;
        }
        out[j + 1] = x;
    }
    /* same-t dedup, lowest surface id wins (mesh.rs:589-603) */
    float tol = 1e-4f / fmaxf(len, 1.0f);
    int w = 0;
    for (int i = 0; i < n; i++) {
        if (w > 0 && fabsf(out[i].t - out[w - 1].t) < tol) {
            if (m->tri_surface[out[i].tri] < m->tri_surface[out[w - 1].tri]) {
                out[w - 1].tri = out[i].tri;
                out[w - 1].material = out[i].material;
            }
        } else {
            out[w++] = out[i];
        }
    }
    int k = w < cap ? w : cap;
    for (int i = 0; i < k; i++) {
        if (t_out)
            t_out[i] = out[i].t;
        if (tri_out)
            tri_out[i] = out[i].tri;
        if (mat_out)
            mat_out[i] = out[i].material;
    }
    free(out);
    return k;
}
