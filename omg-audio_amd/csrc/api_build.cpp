// Scenes built on the device (omg_scene_create_device), their export for
// inspection (omg_scene_export) and refits of moved vertices
// (omg_scene_refit).
//
// The build runs the tessellation and the tree level by level on the
// context's stream (build_kernels.hip); the host only reads back a few
// words per level to size the next launches.  The result is the host
// build's scene (omg_scene_create) exactly: same primitives in the same
// order, same paired nodes in breadth-first order, same 4-wide collapse.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "build_kernels.h"
#include "omg_numerics.h"
#include "runtime.h"
#include "scene_build.h"

using namespace omg;

namespace {

// Bump allocation of sub-buffers (256 B aligned) out of one device block:
// a first pass with base == nullptr measures, the second hands out pointers.
struct Carve {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    T* take(size_t count)
    {
        off = (off + 255) & ~(size_t)255;
        T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
};

// Working memory of one build: the level loop and the collapse each carve
// their buffers out of a single allocation.
struct Scratch {
    DevBuf meta, cnt, pos, tris, lscan_refit;
    DevBuf big, big_off, first_of, items[2], oc, opos;
    DevBuf arena, arena4;
    // level loop (arena)
    BRec *rec_a = nullptr, *rec_b = nullptr;
    Prim* prims_t = nullptr;
    uint32_t *flags = nullptr, *lscan = nullptr, *a_r = nullptr, *b_l = nullptr;
    Node2* nodes = nullptr;
    uint32_t *seg_start[2] = {}, *seg_count[2] = {}, *seg_int[2] = {}, *seg_irank[2] = {};
    uint32_t* seg_box = nullptr;
    ISplit* split = nullptr;
    uint32_t* bins = nullptr;
    // collapse (arena4)
    Node4* nodes4 = nullptr;
    uint32_t *src4 = nullptr, *queue[2] = {}, *path[2] = {}, *nint = nullptr, *iscan = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    ~Scratch()
    {
        for (DevBuf* b : {&meta, &cnt, &pos, &tris, &lscan_refit, &big, &big_off, &first_of,
                          &items[0], &items[1], &oc, &opos, &arena, &arena4})
            b->release();
        if (scan_tmp) (void)hipFree(scan_tmp);
    }
    void carve_levels(Carve& C, size_t n)
    {
        size_t n_int_max = n / 5 + 1;   // a splitting segment holds at least 5 primitives
        rec_a = C.take<BRec>(n);
        rec_b = C.take<BRec>(n);
        prims_t = C.take<Prim>(n);
        flags = C.take<uint32_t>(n + 1);
        lscan = C.take<uint32_t>(n + 1);
        a_r = C.take<uint32_t>(n);
        b_l = C.take<uint32_t>(n);
        nodes = C.take<Node2>(n);
        for (int k = 0; k < 2; k++) {
            seg_start[k] = C.take<uint32_t>(n + 2);
            seg_count[k] = C.take<uint32_t>(n + 2);
            seg_int[k] = C.take<uint32_t>(n + 2);
            seg_irank[k] = C.take<uint32_t>(n + 2);
        }
        seg_box = C.take<uint32_t>(12 * (n + 1));
        split = C.take<ISplit>(n_int_max);
        bins = C.take<uint32_t>(128 * n_int_max);
    }
    void carve_collapse(Carve& C, size_t base)
    {
        nodes4 = C.take<Node4>(base);
        src4 = C.take<uint32_t>(4 * base);
        for (int k = 0; k < 2; k++) {
            queue[k] = C.take<uint32_t>(base + 1);
            path[k] = C.take<uint32_t>(base + 1);
        }
        nint = C.take<uint32_t>(base + 2);
        iscan = C.take<uint32_t>(base + 2);
    }
};

// grow-only with doubling, for buffers resized round by round
hipError_t grow(DevBuf& b, size_t bytes)
{
    if (bytes <= b.cap) return hipSuccess;
    return b.ensure_exact(std::max(bytes, 2 * b.cap));
}

int alloc_levels(Scratch& X, size_t n)
{
    Carve M;
    X.carve_levels(M, n);
    OMG_HIP_CHECK(X.arena.ensure_exact(M.off));
    Carve C;
    C.base = X.arena.as<char>();
    X.carve_levels(C, n);
    return OMG_OK;
}

int alloc_collapse(Scratch& X, size_t base)
{
    Carve M;
    X.carve_collapse(M, base);
    OMG_HIP_CHECK(X.arena4.ensure_exact(M.off));
    Carve C;
    C.base = X.arena4.as<char>();
    X.carve_collapse(C, base);
    return OMG_OK;
}

void scene_free(omg_scene* s)
{
    for (DevBuf* b : {&s->nodes, &s->prims, &s->tris, &s->mats, &s->mats_raw, &s->tri_surf,
                      &s->nodes4, &s->nodes16, &s->b_pos, &s->b_idx, &s->b_mat, &s->b_surf,
                      &s->b_maxe, &s->b_cnt, &s->b_off, &s->leaf_of, &s->src4})
        b->release();
    delete s;
}

int d2h(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    OMG_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    OMG_HIP_CHECK(hipStreamSynchronize(st));
    return OMG_OK;
}

int h2d_exact(DevBuf& b, const void* src, size_t bytes, hipStream_t st)
{
    OMG_HIP_CHECK(b.ensure_exact(bytes ? bytes : 16));
    if (bytes) OMG_HIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st));
    return OMG_OK;
}

int build_error(uint32_t e)
{
    if (e & kBuildErrIndex) return fail(OMG_ERR_INVALID, "triangle index out of range");
    if (e & kBuildErrMaterial) return fail(OMG_ERR_INVALID, "tri_material out of range");
    if (e & kBuildErrFinite) return fail(OMG_ERR_INVALID, "non-finite vertex coordinate");
    if (e & kBuildErrDepth) return fail(OMG_ERR_CAP, "tessellation deeper than 60 splits");
    if (e & kBuildErrRefit)
        return fail(OMG_ERR_INVALID, "the moved vertices change the tessellation: build the scene again");
    return OMG_OK;
}

// Patch counts per triangle (T.cnt_rev, scanned into T.off_rev): the
// triangles that split are tessellated in breadth-first rounds over a work
// list.  With `expect` (a refit) every count must equal the build's.  On
// return *items / *m are the final work list for launch_tess_emit.
int tessellate(Scratch& X, TessArgs& T, uint32_t* off_rev, const uint32_t* expect, hipStream_t st,
               const TessItem** items, uint32_t* m_out)
{
    const uint32_t nt = T.nt;
    OMG_HIP_CHECK(X.big.ensure(((size_t)nt + 1) * 4));
    OMG_HIP_CHECK(X.big_off.ensure(((size_t)nt + 1) * 4));
    OMG_HIP_CHECK(X.first_of.ensure(((size_t)nt + 1) * 4));
    T.big = X.big.as<uint32_t>();
    T.big_off = X.big_off.as<uint32_t>();
    T.first_of = X.first_of.as<uint32_t>();
    OMG_HIP_CHECK(launch_tess_count(T, st));
    OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, T.big, X.big_off.as<uint32_t>(), (size_t)nt + 1, st));
    BuildMeta meta{};
    if (int rc = d2h(&meta, T.meta, sizeof meta, st)) return rc;
    if (int rc = build_error(meta.err)) return rc;
    uint32_t m = 0;
    if (int rc = d2h(&m, X.big_off.as<uint32_t>() + nt, 4, st)) return rc;
    int cur = 0;
    if (m) {
        OMG_HIP_CHECK(grow(X.items[0], (size_t)m * sizeof(TessItem)));
        OMG_HIP_CHECK(launch_tess_seed(T, X.items[0].as<TessItem>(), st));
        for (uint32_t round = 1;; round++) {
            OMG_HIP_CHECK(grow(X.oc, ((size_t)m + 1) * 4));
            OMG_HIP_CHECK(grow(X.opos, ((size_t)m + 1) * 4));
            TessRound R{};
            R.in = X.items[cur].as<TessItem>();
            R.m = m;
            R.oc = X.oc.as<uint32_t>();
            R.pos = X.opos.as<uint32_t>();
            R.tri_max_edge = T.tri_max_edge;
            R.nt = nt;
            OMG_HIP_CHECK(launch_tess_round_count(R, st));
            OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, R.oc, X.opos.as<uint32_t>(), (size_t)m + 1, st));
            uint32_t next = 0;
            if (int rc = d2h(&next, X.opos.as<uint32_t>() + m, 4, st)) return rc;
            if (next == m) break;   // nothing split: every item is final
            if (round > kTessMaxDepth) return build_error(kBuildErrDepth);
            if (next >= kLeafStartMask) return fail(OMG_ERR_INVALID, "too many primitives for the leaf encoding");
            OMG_HIP_CHECK(grow(X.items[cur ^ 1], (size_t)next * sizeof(TessItem)));
            R.out = X.items[cur ^ 1].as<TessItem>();
            OMG_HIP_CHECK(launch_tess_round_write(R, st));
            cur ^= 1;
            m = next;
        }
        OMG_HIP_CHECK(launch_tess_leaves(T, X.items[cur].as<TessItem>(), m, st));
    }
    if (expect) {
        OMG_HIP_CHECK(launch_tess_check(T.cnt_rev, expect, nt, T.meta, st));
        if (int rc = d2h(&meta, T.meta, sizeof meta, st)) return rc;
        if (int rc = build_error(meta.err)) return rc;
    }
    OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, T.cnt_rev, off_rev, (size_t)nt + 1, st));
    T.off_rev = off_rev;
    *items = m ? X.items[cur].as<TessItem>() : nullptr;
    *m_out = m;
    return OMG_OK;
}

// Tree levels over the n records in X.rec_a, then the 4-wide collapse.
int build_tree(omg_scene* s, Scratch& X, uint32_t n, hipStream_t st)
{
    // level 0: one segment of every primitive
    uint32_t root[2] = {0, n};
    uint32_t root_rank[2] = {0, n > 4 ? 1u : 0u};
    OMG_HIP_CHECK(hipMemcpyAsync(X.seg_start[0], &root[0], 4, hipMemcpyHostToDevice, st));
    OMG_HIP_CHECK(hipMemcpyAsync(X.seg_count[0], &root[1], 4, hipMemcpyHostToDevice, st));
    OMG_HIP_CHECK(hipMemcpyAsync(X.seg_irank[0], &root_rank[0], 8, hipMemcpyHostToDevice, st));
    uint32_t n_seg = 1, n_int = root_rank[1], base = 0, base_prev = 0, level = 0;
    int cur = 0;
    s->level_base.clear();
    for (;;) {
        if (level + 1 + 2 > (uint32_t)kStackCap)
            return fail(OMG_ERR_CAP, "BVH deeper than the traversal stack");
        LevelArgs A{};
        A.n = n;
        A.level = level;
        A.n_seg = n_seg;
        A.n_int = n_int;
        A.base = base;
        A.base_prev = base_prev;
        A.seg_start = X.seg_start[cur];
        A.seg_count = X.seg_count[cur];
        A.seg_irank = X.seg_irank[cur];
        A.seg_box = X.seg_box;
        A.cur = (level & 1) ? X.rec_b : X.rec_a;
        A.out = (level & 1) ? X.rec_a : X.rec_b;
        A.flags = X.flags;
        A.lscan = X.lscan;
        A.aR = X.a_r;
        A.bL = X.b_l;
        A.nodes = X.nodes;
        A.meta = X.meta.as<BuildMeta>();
        A.split = X.split;
        A.bins = X.bins;
        A.nseg_start = X.seg_start[cur ^ 1];
        A.nseg_count = X.seg_count[cur ^ 1];
        A.nseg_int = X.seg_int[cur ^ 1];
        OMG_HIP_CHECK(launch_box_init(A, st));
        OMG_HIP_CHECK(launch_box(A, st));
        OMG_HIP_CHECK(launch_small(A, st));
        OMG_HIP_CHECK(launch_seg(A, st));
        if (!n_int) break;
        s->level_base.push_back(base);
        OMG_HIP_CHECK(launch_bin(A, st));
        OMG_HIP_CHECK(launch_split(A, st));
        OMG_HIP_CHECK(launch_flag(A, st));
        OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, A.flags, X.lscan, (size_t)n + 1, st));
        OMG_HIP_CHECK(launch_ptable(A, st));
        OMG_HIP_CHECK(launch_permute(A, st));
        size_t nn = 2 * (size_t)n_int + 1;
        uint32_t* nrank = X.seg_irank[cur ^ 1];
        OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, A.nseg_int, nrank, nn, st));
        uint32_t next_int = 0;
        if (int rc = d2h(&next_int, nrank + (nn - 1), 4, st)) return rc;
        base_prev = base;
        base += n_int;
        n_seg = 2 * n_int;
        n_int = next_int;
        cur ^= 1;
        level++;
    }
    const BRec* final_recs = (level & 1) ? X.rec_b : X.rec_a;
    s->max_depth = level + 1;
    s->n_nodes = base;
    OMG_HIP_CHECK(s->prims.ensure_exact((size_t)n * sizeof(Prim)));
    OMG_HIP_CHECK(s->leaf_of.ensure_exact((size_t)n * 4));
    OMG_HIP_CHECK(launch_gather_prims(X.prims_t, final_recs, n, s->prims.as<Prim>(),
                                      s->leaf_of.as<uint32_t>(), st));
    OMG_HIP_CHECK(s->nodes.ensure_exact(base ? (size_t)base * sizeof(Node2) : 16));
    if (base)
        OMG_HIP_CHECK(hipMemcpyAsync(s->nodes.ptr, X.nodes, (size_t)base * sizeof(Node2),
                                     hipMemcpyDeviceToDevice, st));

    // 4-wide collapse, one level of 4-wide nodes at a time
    uint32_t n4 = 0;
    if (base) {
        if (int rc = alloc_collapse(X, base)) return rc;
        uint32_t zero2[2] = {0, 0};
        OMG_HIP_CHECK(hipMemcpyAsync(X.queue[0], &zero2[0], 4, hipMemcpyHostToDevice, st));
        OMG_HIP_CHECK(hipMemcpyAsync(X.path[0], &zero2[1], 4, hipMemcpyHostToDevice, st));
        uint32_t q = 1;
        int qc = 0;
        while (q) {
            Collapse4Args C{};
            C.nodes = s->nodes.as<Node2>();
            C.nodes4 = X.nodes4;
            C.src4 = X.src4;
            C.q = q;
            C.base4 = n4;
            C.base4_next = n4 + q;
            C.queue = X.queue[qc];
            C.queue_next = X.queue[qc ^ 1];
            C.path = X.path[qc];
            C.path_next = X.path[qc ^ 1];
            C.n_int = X.nint;
            C.int_scan = X.iscan;
            C.meta = X.meta.as<BuildMeta>();
            OMG_HIP_CHECK(launch_collapse_expand(C, st));
            OMG_HIP_CHECK(scan_u32(X.scan_tmp, X.scan_bytes, C.n_int, X.iscan,
                                   (size_t)q + 1, st));
            OMG_HIP_CHECK(launch_collapse_link(C, st));
            uint32_t next = 0;
            if (int rc = d2h(&next, X.iscan + q, 4, st)) return rc;
            n4 += q;
            q = next;
            qc ^= 1;
        }
    }
    BuildMeta meta{};
    if (int rc = d2h(&meta, X.meta.ptr, sizeof meta, st)) return rc;
    s->root_ref = meta.root_ref;
    for (int k = 0; k < 3; k++) {
        s->root_min[k] = meta.root_min[k];
        s->root_max[k] = meta.root_max[k];
    }
    if (base) {
        s->root4 = 0;
        s->stack4 = meta.stack4;
        s->has4 = s->stack4 + 1 <= (uint32_t)kStackCap;
    } else {
        s->root4 = s->root_ref;
        s->stack4 = 0;
        s->has4 = true;
    }
    s->n_nodes4 = s->has4 ? n4 : 0;
    if (s->has4 && n4) {
        OMG_HIP_CHECK(s->nodes4.ensure_exact((size_t)n4 * sizeof(Node4)));
        OMG_HIP_CHECK(s->src4.ensure_exact((size_t)n4 * 16));
        OMG_HIP_CHECK(hipMemcpyAsync(s->nodes4.ptr, X.nodes4, (size_t)n4 * sizeof(Node4),
                                     hipMemcpyDeviceToDevice, st));
        OMG_HIP_CHECK(hipMemcpyAsync(s->src4.ptr, X.src4, (size_t)n4 * 16, hipMemcpyDeviceToDevice, st));
    }
    return OMG_OK;
}

void tri_normals(omg_scene* s)
{
    s->tris_host.resize(s->n_tris);
    for (uint32_t t = 0; t < s->n_tris; t++) {
        auto P = [&](uint32_t v) {
            return mk(s->positions[3 * (size_t)v], s->positions[3 * (size_t)v + 1],
                      s->positions[3 * (size_t)v + 2]);
        };
        f3 a = P(s->indices[3 * (size_t)t]), b = P(s->indices[3 * (size_t)t + 1]),
           c = P(s->indices[3 * (size_t)t + 2]);
        f3 n = normalize(cross(sub(b, a), sub(c, a)));
        s->tris_host[t] = TriInfo{{n.x, n.y, n.z}, s->tri_material[t]};
    }
}

}  // namespace

int omg::ensure_host_mesh(omg_scene* s)
{
    const size_t nt = s->n_tris;
    if (s->device_built && s->positions.size() != (size_t)s->nv * 3) {
        hipStream_t st = s->ctx->stream;
        OMG_HIP_CHECK(hipSetDevice(s->ctx->device));
        if (int rc = quiesce(s->ctx)) return rc;
        s->positions.resize((size_t)s->nv * 3);
        s->indices.resize(nt * 3);
        s->tri_material.resize(nt);
        s->tri_surface.assign(nt, 0);
        if (s->nv) OMG_HIP_CHECK(hipMemcpyAsync(s->positions.data(), s->b_pos.ptr, (size_t)s->nv * 12, hipMemcpyDeviceToHost, st));
        if (nt) {
            OMG_HIP_CHECK(hipMemcpyAsync(s->indices.data(), s->b_idx.ptr, nt * 12, hipMemcpyDeviceToHost, st));
            OMG_HIP_CHECK(hipMemcpyAsync(s->tri_material.data(), s->b_mat.ptr, nt * 2, hipMemcpyDeviceToHost, st));
            if (s->b_surf.ptr)
                OMG_HIP_CHECK(hipMemcpyAsync(s->tri_surface.data(), s->b_surf.ptr, nt * 2, hipMemcpyDeviceToHost, st));
        }
        OMG_HIP_CHECK(hipStreamSynchronize(st));
        s->tris_host.clear();
    }
    if (s->tris_host.size() != nt) tri_normals(s);
    return OMG_OK;
}

extern "C" int omg_scene_create_device(omg_ctx* c, const float* positions, uint32_t nv,
                                       const uint32_t* indices, uint32_t nt,
                                       const uint16_t* tri_material, const uint16_t* tri_surface,
                                       const float* tri_max_edge, const omg_material* materials,
                                       uint32_t nm, uint32_t nbands, omg_scene** out)
{
    if (!c || !out || (nt && (!positions || !indices || !tri_material)) || (nm && !materials))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (nbands < 1 || nbands > OMG_MAX_BANDS) return fail(OMG_ERR_INVALID, "nbands out of range");
    *out = nullptr;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    omg_scene* s = new omg_scene();
    s->ctx = c;
    s->nbands = nbands;
    s->n_tris = nt;
    s->nv = nv;
    s->device_built = true;
    s->materials.assign(materials, materials + nm);
    s->mat_params.resize(nm);
    for (uint32_t m = 0; m < nm; m++)
        mat_params(materials[m].absorption, materials[m].scattering, materials[m].transmission,
                   nbands, s->mat_params[m]);
    // the host copies of the mesh (for the surface table) are downloaded on
    // first use: omg::ensure_host_mesh
    Scratch X;
    auto run = [&]() -> int {
        if (int rc = h2d_exact(s->b_pos, positions, (size_t)nv * 12, st)) return rc;
        if (int rc = h2d_exact(s->b_idx, indices, (size_t)nt * 12, st)) return rc;
        if (int rc = h2d_exact(s->b_mat, tri_material, (size_t)nt * 2, st)) return rc;
        if (tri_surface)
            if (int rc = h2d_exact(s->b_surf, tri_surface, (size_t)nt * 2, st)) return rc;
        if (tri_max_edge)
            if (int rc = h2d_exact(s->b_maxe, tri_max_edge, (size_t)nt * 4, st)) return rc;
        OMG_HIP_CHECK(X.meta.ensure(sizeof(BuildMeta)));
        OMG_HIP_CHECK(hipMemsetAsync(X.meta.ptr, 0, sizeof(BuildMeta), st));
        OMG_HIP_CHECK(s->b_cnt.ensure_exact(((size_t)nt + 1) * 4));
        OMG_HIP_CHECK(s->b_off.ensure_exact(((size_t)nt + 1) * 4));
        OMG_HIP_CHECK(s->tris.ensure_exact(nt ? (size_t)nt * sizeof(TriInfo) : 16));
        OMG_HIP_CHECK(s->tri_surf.ensure_exact(nt ? (size_t)nt * 4 : 16));
        TessArgs T{};
        T.pos = s->b_pos.as<float>();
        T.nv = nv;
        T.idx = s->b_idx.as<uint32_t>();
        T.nt = nt;
        T.tri_mat = s->b_mat.as<uint16_t>();
        T.tri_surf = tri_surface ? s->b_surf.as<uint16_t>() : nullptr;
        T.tri_max_edge = tri_max_edge ? s->b_maxe.as<float>() : nullptr;
        T.nm = nm;
        T.cnt_rev = s->b_cnt.as<uint32_t>();
        T.tris = s->tris.as<TriInfo>();
        T.tri_surf32 = s->tri_surf.as<uint32_t>();
        T.meta = X.meta.as<BuildMeta>();
        const TessItem* items = nullptr;
        uint32_t m = 0, n = 0;
        if (int rc = tessellate(X, T, s->b_off.as<uint32_t>(), nullptr, st, &items, &m)) return rc;
        if (int rc = d2h(&n, s->b_off.as<uint32_t>() + nt, 4, st)) return rc;
        if (n >= kLeafStartMask) return fail(OMG_ERR_INVALID, "too many primitives for the leaf encoding");
        s->n_prims = n;
        if (n == 0) {
            s->empty = true;
            s->has4 = true;
            return OMG_OK;
        }
        if (int rc = alloc_levels(X, n)) return rc;
        T.prims = X.prims_t;
        T.recs = X.rec_a;
        OMG_HIP_CHECK(launch_tess_emit(T, items, m, st));
        return build_tree(s, X, n, st);
    };
    int rc = run();
    if (!rc) rc = derive16(s, st);
    if (!rc) rc = upload(s->mats, s->mat_params.data(), s->mat_params.size() * sizeof(MatParams), st);
    if (!rc) rc = upload(s->mats_raw, s->materials.data(), s->materials.size() * sizeof(omg_material), st);
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "device scene build");
    if (rc) {
        (void)hipStreamSynchronize(st);
        scene_free(s);
        return rc;
    }
    s->device_bytes = 0;
    for (DevBuf* b : {&s->nodes, &s->prims, &s->tris, &s->mats, &s->mats_raw, &s->tri_surf,
                      &s->nodes4, &s->nodes16, &s->b_pos, &s->b_idx, &s->b_mat, &s->b_surf,
                      &s->b_maxe, &s->b_cnt, &s->b_off, &s->leaf_of, &s->src4})
        s->device_bytes += b->cap;
    c->live_scenes++;
    *out = s;
    return OMG_OK;
}

extern "C" int omg_scene_export(omg_scene* s, float* prims_out, uint32_t prim_cap,
                                float* nodes_out, uint32_t node_cap, float* nodes4_out,
                                uint32_t node4_cap, omg_scene_tree* tree)
{
    if (!s || !tree) return fail(OMG_ERR_INVALID, "bad argument");
    std::memset(tree, 0, sizeof *tree);
    tree->n_prims = s->n_prims;
    tree->n_nodes = s->n_nodes;
    tree->n_nodes4 = s->has4 ? s->n_nodes4 : 0;
    tree->root_ref = s->root_ref;
    tree->root4 = s->root4;
    tree->stack4 = s->stack4;
    tree->has4 = s->has4 ? 1u : 0u;
    tree->max_depth = s->max_depth;
    tree->empty = s->empty ? 1u : 0u;
    for (int k = 0; k < 3; k++) {
        tree->root_min[k] = s->root_min[k];
        tree->root_max[k] = s->root_max[k];
    }
    tree->n_nodes16 = s->n_nodes16;
    tree->stack16 = s->stack16;
    tree->has16 = s->has16 ? 1u : 0u;
    hipStream_t st = s->ctx->stream;
    OMG_HIP_CHECK(hipSetDevice(s->ctx->device));
    if (int rc = quiesce(s->ctx)) return rc;
    if (prims_out && s->n_prims) {
        if (prim_cap < s->n_prims) return fail(OMG_ERR_CAP, "prim buffer too small");
        if (int rc = d2h(prims_out, s->prims.ptr, (size_t)s->n_prims * sizeof(Prim), st)) return rc;
    }
    if (nodes_out && s->n_nodes) {
        if (node_cap < s->n_nodes) return fail(OMG_ERR_CAP, "node buffer too small");
        if (int rc = d2h(nodes_out, s->nodes.ptr, (size_t)s->n_nodes * sizeof(Node2), st)) return rc;
    }
    if (nodes4_out && tree->n_nodes4) {
        if (node4_cap < tree->n_nodes4) return fail(OMG_ERR_CAP, "4-wide node buffer too small");
        if (int rc = d2h(nodes4_out, s->nodes4.ptr, (size_t)tree->n_nodes4 * sizeof(Node4), st)) return rc;
    }
    return OMG_OK;
}

extern "C" int omg_scene_refit(omg_scene* s, const float* positions, uint32_t nv)
{
    if (!s || (nv && !positions)) return fail(OMG_ERR_INVALID, "bad argument");
    if (!s->device_built) return fail(OMG_ERR_INVALID, "refit needs a scene built by omg_scene_create_device");
    if (nv != s->nv) return fail(OMG_ERR_INVALID, "refit keeps the vertex count");
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    if (int rc = quiesce(c)) return rc;
    hipStream_t st = c->stream;
    const uint32_t nt = s->n_tris;
    Scratch X;
    OMG_HIP_CHECK(X.meta.ensure(sizeof(BuildMeta)));
    OMG_HIP_CHECK(hipMemsetAsync(X.meta.ptr, 0, sizeof(BuildMeta), st));
    if (int rc = h2d_exact(X.pos, positions, (size_t)nv * 12, st)) return rc;
    OMG_HIP_CHECK(X.cnt.ensure(((size_t)nt + 1) * 4));
    OMG_HIP_CHECK(X.tris.ensure(nt ? (size_t)nt * sizeof(TriInfo) : 16));
    TessArgs T{};
    T.pos = X.pos.as<float>();
    T.nv = nv;
    T.idx = s->b_idx.as<uint32_t>();
    T.nt = nt;
    T.tri_mat = s->b_mat.as<uint16_t>();
    T.tri_surf = s->b_surf.ptr ? s->b_surf.as<uint16_t>() : nullptr;
    T.tri_max_edge = s->b_maxe.ptr ? s->b_maxe.as<float>() : nullptr;
    T.nm = (uint32_t)s->materials.size();
    T.cnt_rev = X.cnt.as<uint32_t>();
    T.tris = X.tris.as<TriInfo>();
    T.meta = X.meta.as<BuildMeta>();
    const TessItem* items = nullptr;
    uint32_t m = 0;
    // the offsets of an accepted refit equal the build's: scan into scratch
    OMG_HIP_CHECK(X.lscan_refit.ensure(((size_t)nt + 1) * 4));
    if (int rc = tessellate(X, T, X.lscan_refit.as<uint32_t>(), s->b_cnt.as<uint32_t>(), st, &items, &m))
        return rc;
    // the moved mesh keeps every triangle's patches: commit it
    if (nv) OMG_HIP_CHECK(hipMemcpyAsync(s->b_pos.ptr, X.pos.ptr, (size_t)nv * 12, hipMemcpyDeviceToDevice, st));
    if (nt) OMG_HIP_CHECK(hipMemcpyAsync(s->tris.ptr, X.tris.ptr, (size_t)nt * sizeof(TriInfo), hipMemcpyDeviceToDevice, st));
    s->positions.clear();
    s->tris_host.clear();
    if (s->n_prims == 0) return hipStreamSynchronize(st) == hipSuccess ? OMG_OK : fail(OMG_ERR_DEVICE, "refit");
    T.pos = s->b_pos.as<float>();
    T.prims = s->prims.as<Prim>();
    T.leaf_of = s->leaf_of.as<uint32_t>();
    T.recs = nullptr;
    OMG_HIP_CHECK(launch_tess_emit(T, items, m, st));
    // paired nodes deepest level first, then the root box and the 4-wide copies
    for (size_t L = s->level_base.size(); L-- > 0;) {
        uint32_t first = s->level_base[L];
        uint32_t end = L + 1 < s->level_base.size() ? s->level_base[L + 1] : s->n_nodes;
        OMG_HIP_CHECK(launch_refit_level(s->nodes.as<Node2>(), s->prims.as<Prim>(), first, end - first, st));
    }
    OMG_HIP_CHECK(launch_refit_root(s->nodes.as<Node2>(), s->prims.as<Prim>(), s->root_ref,
                                    X.meta.as<BuildMeta>(), st));
    if (s->has4 && s->n_nodes4)
        OMG_HIP_CHECK(launch_refit4(s->nodes.as<Node2>(), s->src4.as<uint32_t>(), s->nodes4.as<Node4>(),
                                    s->n_nodes4, st));
    if (int rc = derive16(s, st)) return rc;
    BuildMeta meta{};
    if (int rc = d2h(&meta, X.meta.ptr, sizeof meta, st)) return rc;
    for (int k = 0; k < 3; k++) {
        s->root_min[k] = meta.root_min[k];
        s->root_max[k] = meta.root_max[k];
    }
    return OMG_OK;
}

// The 16-wide nodes of a scene, derived from its resident 4-wide nodes
// level by level (the 4-wide collapse's breadth-first scheme: expand, scan
// of the internal children, link).  A 16-wide node takes a 4-wide node and
// expands each of its internal slots once more, so its internal children
// are the 4-wide nodes two levels down; their count per level is read back
// to size the next level.  Stream-ordered; the working buffers are freed
// before returning.
int omg::derive16(omg_scene* s, hipStream_t st)
{
    s->nodes16.release();
    s->n_nodes16 = 0;
    s->stack16 = 0;
    s->has16 = false;
    if (!s->has4) return OMG_OK;
    if (s->empty || (s->root4 & kLeafBit) || s->n_nodes4 == 0) {
        // no internal node: the walk tests the root leaf (or nothing)
        s->root16 = s->root4;
        s->has16 = true;
        return OMG_OK;
    }
    const size_t n4 = s->n_nodes4;
    DevBuf out, work;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(st);
        out.release();
        work.release();
        if (scan_tmp) (void)hipFree(scan_tmp);
        return rc;
    };
    uint32_t *queue[2], *path[2], *nint, *iscan, *stack16;
    Carve M;
    for (int k = 0; k < 2; k++) {
        queue[k] = M.take<uint32_t>(n4 + 1);
        path[k] = M.take<uint32_t>(n4 + 1);
    }
    nint = M.take<uint32_t>(n4 + 2);
    iscan = M.take<uint32_t>(n4 + 2);
    stack16 = M.take<uint32_t>(1);
    if (out.ensure_exact(n4 * sizeof(Node16)) != hipSuccess || work.ensure_exact(M.off) != hipSuccess)
        return done(fail(OMG_ERR_DEVICE, "16-wide node memory"));
    Carve C;
    C.base = work.as<char>();
    for (int k = 0; k < 2; k++) {
        queue[k] = C.take<uint32_t>(n4 + 1);
        path[k] = C.take<uint32_t>(n4 + 1);
    }
    nint = C.take<uint32_t>(n4 + 2);
    iscan = C.take<uint32_t>(n4 + 2);
    stack16 = C.take<uint32_t>(1);
    const uint32_t zero = 0;
    if (hipMemcpyAsync(queue[0], &zero, 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(path[0], &zero, 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(stack16, &zero, 4, hipMemcpyHostToDevice, st) != hipSuccess)
        return done(fail(OMG_ERR_DEVICE, "16-wide derivation"));
    uint32_t q = 1, n16 = 0;
    int qc = 0;
    while (q) {
        Derive16Args D{};
        D.nodes4 = s->nodes4.as<Node4>();
        D.nodes16 = out.as<Node16>();
        D.q = q;
        D.base16 = n16;
        D.base16_next = n16 + q;
        D.queue = queue[qc];
        D.queue_next = queue[qc ^ 1];
        D.path = path[qc];
        D.path_next = path[qc ^ 1];
        D.n_int = nint;
        D.int_scan = iscan;
        D.stack16 = stack16;
        if (launch_derive16_expand(D, st) != hipSuccess ||
            scan_u32(scan_tmp, scan_bytes, nint, iscan, (size_t)q + 1, st) != hipSuccess ||
            launch_derive16_link(D, st) != hipSuccess)
            return done(fail(OMG_ERR_DEVICE, "16-wide derivation"));
        uint32_t next = 0;
        if (int rc = d2h(&next, iscan + q, 4, st)) return done(rc);
        n16 += q;
        q = next;
        qc ^= 1;
    }
    uint32_t sk = 0;
    if (int rc = d2h(&sk, stack16, 4, st)) return done(rc);
    if (s->nodes16.ensure_exact((size_t)n16 * sizeof(Node16)) != hipSuccess ||
        hipMemcpyAsync(s->nodes16.ptr, out.ptr, (size_t)n16 * sizeof(Node16), hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return done(fail(OMG_ERR_DEVICE, "16-wide node memory"));
    s->n_nodes16 = n16;
    s->root16 = 0;
    s->stack16 = sk;
    s->has16 = sk <= kStack16;
    return done(OMG_OK);
}
