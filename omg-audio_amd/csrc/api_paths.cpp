// C ABI: the knife-edge path search (AutoPaths, crates/omg-core/src/paths.rs)
// over a scene's BVH.  The edge graph is extracted on the host once and
// kept on the device; queries run on the device (paths_kernels.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "paths_kernels.h"
#include "runtime.h"

#pragma clang fp contract(off)

using namespace omg;

static_assert(sizeof(omg_path_edge) == sizeof(PathEdgeDev), "edge layout");
static_assert(sizeof(omg_path) == sizeof(PathDev), "path layout");

struct omg_paths {
    omg_scene* scene = nullptr;   // must outlive the object's queries
    omg_ctx* ctx = nullptr;
    std::vector<omg_path_edge> edges;
    DevBuf d_edges, d_src, d_lis, d_out, d_n, d_work, d_cnt, d_trav, d_plan;
    bool counting = false;
    omg_counters last{};
    // an async search ran since omg_paths_status last read the overflow
    // word, and an overflow it flagged that a blocking call has since cleared
    bool async_unread = false;
    bool pending_overflow = false;
};

namespace {

// f32::total_cmp as a signed integer key (larger key = larger value)
int32_t total_key(float f)
{
    int32_t k;
    std::memcpy(&k, &f, 4);
    return k ^ (int32_t)((uint32_t)(k >> 31) >> 1);
}

// Before a blocking call clears the counter words: an overflow flagged by
// an earlier async search stays pending for omg_paths_status.
int keep_async_overflow(omg_paths* p, omg_ctx* c)
{
    if (!p->async_unread) return OMG_OK;
    uint32_t flag = 0;
    OMG_HIP_CHECK(hipMemcpyAsync(&flag, p->d_cnt.as<uint32_t>() + 1, sizeof flag,
                                 hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    p->pending_overflow |= flag != 0;
    p->async_unread = false;
    return OMG_OK;
}

int check_budget(const omg_paths* p, const omg_path_budget* b)
{
    if (!p || !b) return fail(OMG_ERR_INVALID, "bad argument");
    if (b->edge_candidates > OMG_PATH_MAX_CANDIDATES)
        return fail(OMG_ERR_CAP, "edge_candidates above OMG_PATH_MAX_CANDIDATES");
    if (b->pair_edges > OMG_PATH_MAX_PAIR_EDGES)
        return fail(OMG_ERR_CAP, "pair_edges above OMG_PATH_MAX_PAIR_EDGES");
    return OMG_OK;
}

// The per-query plans of the ranking pass (grow-only, kept by the object).
hipError_t ensure_plan(omg_paths* p, const omg_path_budget* b, uint32_t n)
{
    const PlanLayout pl = plan_layout((uint32_t)p->edges.size(), b->edge_candidates, b->pair_edges);
    return p->d_plan.ensure((size_t)n * pl.words * sizeof(float));
}

void fill_args(omg_paths* p, const omg_path_budget* b, PathsArgs& A)
{
    omg_scene* s = p->scene;
    fill_geom(s, A.g);
    A.mats = s->mats_raw.as<omg_material_dev>();
    A.edges = p->d_edges.as<PathEdgeDev>();
    A.ne = (uint32_t)p->edges.size();
    A.edge_candidates = b->edge_candidates;
    A.pair_edges = b->pair_edges;
    A.max_paths = b->max_paths;
    A.taps = b->taps ? 1u : 0u;
    A.plan = p->d_plan.as<float>();
}

}  // namespace

// extract_edges (mesh.rs:639-676): every undirected edge of the ORIGINAL
// triangles; an edge of one triangle is a boundary (sharpness 1), of two a
// crease kept when their normals differ by at least min_dihedral_deg
// (sharpness angle / 180), of more non-manifold and dropped; edges shorter
// than 1e-4 m are dropped.  Then budget_edges (mesh.rs:678-682): stable
// sort by importance descending, truncate.  The reference iterates a hash
// map, so its order among equal importances is unspecified; here it is
// ascending (min vertex, max vertex).
namespace {

// The budgeted edge graph; false on a vertex index out of range.
bool extract_edges(const float* positions, uint32_t nv, const uint32_t* indices, uint32_t nt,
                   float min_dihedral_deg, uint32_t budget, std::vector<omg_path_edge>& all)
{
    all.clear();
    struct Half {
        uint64_t key;
        uint32_t tri;
    };
    std::vector<Half> h;
    h.reserve((size_t)nt * 3);
    for (uint32_t t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) {
            uint32_t i = indices[3 * (size_t)t + k], j = indices[3 * (size_t)t + (k + 1) % 3];
            if (i >= nv || j >= nv) return false;
            uint32_t lo = std::min(i, j), hi = std::max(i, j);
            h.push_back(Half{(uint64_t)lo << 32 | hi, t});
        }
    std::stable_sort(h.begin(), h.end(), [](const Half& x, const Half& y) { return x.key < y.key; });
    auto vert = [&](uint32_t i) {
        return mk(positions[3 * (size_t)i], positions[3 * (size_t)i + 1], positions[3 * (size_t)i + 2]);
    };
    auto normal = [&](uint32_t t) {   // mesh.rs:458-461
        f3 a = vert(indices[3 * (size_t)t]), b = vert(indices[3 * (size_t)t + 1]);
        f3 c = vert(indices[3 * (size_t)t + 2]);
        return normalize(cross(sub(b, a), sub(c, a)));
    };
    for (size_t s = 0; s < h.size();) {
        size_t e = s;
        while (e < h.size() && h[e].key == h[s].key) e++;
        f3 a = vert((uint32_t)(h[s].key >> 32)), b = vert((uint32_t)h[s].key);
        float len = length(sub(b, a));
        bool keep = len >= 1e-4f;
        float sharp = 0.0f;
        if (keep) {
            if (e - s == 1) {
                sharp = 1.0f;
            } else if (e - s == 2) {
                float c = dot(normal(h[s].tri), normal(h[s + 1].tri));
                c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
                float angle = acosf(c) * 57.2957795130823208767981548141051703f;   // f32::to_degrees
                if (angle < min_dihedral_deg)
                    keep = false;
                else
                    sharp = angle / 180.0f;
            } else {
                keep = false;
            }
        }
        if (keep) {
            omg_path_edge d{};
            d.a[0] = a.x;
            d.a[1] = a.y;
            d.a[2] = a.z;
            d.b[0] = b.x;
            d.b[1] = b.y;
            d.b[2] = b.z;
            d.importance = len * sharp;
            all.push_back(d);
        }
        s = e;
    }
    std::stable_sort(all.begin(), all.end(), [](const omg_path_edge& x, const omg_path_edge& y) {
        return total_key(x.importance) > total_key(y.importance);
    });
    if (all.size() > budget) all.resize(budget);
    return true;
}

}  // namespace

extern "C" int omg_paths_extract_edges(const float* positions, uint32_t nv, const uint32_t* indices,
                                       uint32_t nt, float min_dihedral_deg, uint32_t budget,
                                       omg_path_edge* out, uint32_t cap, uint32_t* n_out)
{
    if (!n_out || (nt && (!positions || !indices)) || (cap && !out))
        return fail(OMG_ERR_INVALID, "bad argument");
    *n_out = 0;
    std::vector<omg_path_edge> all;
    if (!extract_edges(positions, nv, indices, nt, min_dihedral_deg, budget, all))
        return fail(OMG_ERR_INVALID, "vertex index out of range");
    uint32_t n = (uint32_t)all.size();
    *n_out = n;
    if (out && cap) std::memcpy(out, all.data(), std::min(n, cap) * sizeof(omg_path_edge));
    return n > cap ? fail(OMG_ERR_CAP, "edge buffer too small") : OMG_OK;
}

extern "C" int omg_paths_create(omg_scene* s, uint32_t edge_budget, omg_paths** out)
{
    if (!s || !out) return fail(OMG_ERR_INVALID, "bad argument");
    *out = nullptr;
    if (s->nbands != 3) return fail(OMG_ERR_INVALID, "the path search prices 3 bands");
    if (int rc = ensure_host_mesh(s)) return rc;
    uint32_t nv = (uint32_t)(s->positions.size() / 3);
    std::vector<omg_path_edge> edges;
    // AutoPaths::new's graph: a 40 degree crease threshold (paths.rs:72-77)
    if (!extract_edges(s->positions.data(), nv, s->indices.data(), s->n_tris, 40.0f, edge_budget, edges))
        return fail(OMG_ERR_INVALID, "vertex index out of range");
    if (edges.size() > OMG_PATH_MAX_EDGES) return fail(OMG_ERR_CAP, "edge graph above OMG_PATH_MAX_EDGES");
    const uint32_t n = (uint32_t)edges.size();
    omg_paths* p = new omg_paths();
    p->scene = s;
    p->ctx = s->ctx;
    p->edges = std::move(edges);
    int rc = OMG_OK;
    omg_ctx* c = s->ctx;
    if (!rc) {
        hipError_t e = hipSetDevice(c->device);
        if (e != hipSuccess) rc = fail(OMG_ERR_DEVICE, hipGetErrorString(e));
    }
    if (!rc && n) rc = upload(p->d_edges, p->edges.data(), (size_t)n * sizeof(omg_path_edge), c->stream);
    if (!rc && !n && p->d_edges.ensure(sizeof(omg_path_edge)) != hipSuccess)
        rc = fail(OMG_ERR_DEVICE, "paths: edge buffer");
    if (!rc) {
        hipError_t e = p->d_cnt.ensure(2 * sizeof(uint32_t));
        if (e == hipSuccess) e = p->d_trav.ensure(2 * sizeof(unsigned long long));
        if (e == hipSuccess) e = launch_fill_zero_u64(p->d_trav.as<unsigned long long>(), 2, c->stream);
        if (e == hipSuccess) e = launch_fill_zero(p->d_cnt.as<float>(), 2, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(OMG_ERR_DEVICE, std::string("paths: ") + hipGetErrorString(e));
    }
    if (rc) {
        for (DevBuf* d : {&p->d_edges, &p->d_cnt, &p->d_trav}) d->release();
        delete p;
        return rc;
    }
    c->live_scenes++;
    *out = p;
    return OMG_OK;
}

extern "C" int omg_paths_edges(omg_paths* p, omg_path_edge* out, uint32_t cap, uint32_t* n_out)
{
    if (!p || !n_out || (cap && !out)) return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t n = (uint32_t)p->edges.size();
    *n_out = n;
    if (out && cap) std::memcpy(out, p->edges.data(), std::min(n, cap) * sizeof(omg_path_edge));
    return n > cap && out ? fail(OMG_ERR_CAP, "edge buffer too small") : OMG_OK;
}

extern "C" int omg_paths_transmission(omg_paths* p, const float* a, const float* b, uint32_t n,
                                      float* gains)
{
    if (!p || (n && (!a || !b || !gains))) return fail(OMG_ERR_INVALID, "bad argument");
    if (n == 0) return OMG_OK;
    omg_ctx* c = p->scene->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    int rc = quiesce(c);
    if (!rc) rc = keep_async_overflow(p, c);
    if (rc) return rc;
    rc = upload(p->d_src, a, (size_t)n * 3 * sizeof(float), c->stream);
    if (!rc) rc = upload(p->d_lis, b, (size_t)n * 3 * sizeof(float), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(p->d_out.ensure((size_t)n * 3 * sizeof(float)));
    OMG_HIP_CHECK(launch_fill_zero(p->d_cnt.as<float>(), 2, c->stream));
    PathTransArgs T{};
    fill_geom(p->scene, T.g);
    T.mats = p->scene->mats_raw.as<omg_material_dev>();
    T.a = p->d_src.as<float>();
    T.b = p->d_lis.as<float>();
    T.n = n;
    T.gains = p->d_out.as<float>();
    T.counters = p->d_cnt.as<uint32_t>();
    OMG_HIP_CHECK(launch_path_transmission(T, c->stream));
    uint32_t cnt[2];
    OMG_HIP_CHECK(hipMemcpyAsync(gains, T.gains, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(cnt, T.counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    return cnt[1] ? fail(OMG_ERR_OVERFLOW, "segment crossing list overflowed") : OMG_OK;
}

extern "C" int omg_paths_find(omg_paths* p, const omg_path_budget* b, const float* src,
                              const float* lis, uint32_t n, omg_path* out, uint32_t* n_paths)
{
    int rc = check_budget(p, b);
    if (rc) return rc;
    if (n && (!src || !lis || !out || !n_paths)) return fail(OMG_ERR_INVALID, "bad argument");
    if (n == 0) return OMG_OK;
    omg_ctx* c = p->scene->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    rc = quiesce(c);
    if (!rc) rc = keep_async_overflow(p, c);
    if (rc) return rc;
    const size_t stride = 1 + (size_t)b->max_paths;
    rc = upload(p->d_src, src, (size_t)n * 3 * sizeof(float), c->stream);
    if (!rc) rc = upload(p->d_lis, lis, (size_t)n * 3 * sizeof(float), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(p->d_out.ensure((size_t)n * stride * sizeof(PathDev)));
    OMG_HIP_CHECK(p->d_n.ensure((size_t)n * sizeof(uint32_t)));
    OMG_HIP_CHECK(p->d_work.ensure((size_t)n * sizeof(uint32_t)));
    OMG_HIP_CHECK(ensure_plan(p, b, n));
    OMG_HIP_CHECK(launch_fill_zero(p->d_cnt.as<float>(), 2, c->stream));
    if (p->counting) OMG_HIP_CHECK(launch_fill_zero_u64(p->d_trav.as<unsigned long long>(), 2, c->stream));
    PathsArgs A{};
    fill_args(p, b, A);
    A.src = p->d_src.as<float>();
    A.lis = p->d_lis.as<float>();
    A.n = n;
    A.out = p->d_out.as<PathDev>();
    A.n_out = p->d_n.as<uint32_t>();
    A.work = p->d_work.as<uint32_t>();
    A.counters = p->d_cnt.as<uint32_t>();
    A.trav = p->d_trav.as<unsigned long long>();
    OMG_HIP_CHECK(launch_paths(A, p->counting, c->stream));
    uint32_t cnt[2];
    unsigned long long trav[2] = {0, 0};
    OMG_HIP_CHECK(hipMemcpyAsync(out, A.out, (size_t)n * stride * sizeof(PathDev), hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(n_paths, A.n_out, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(cnt, A.counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (p->counting)
        OMG_HIP_CHECK(hipMemcpyAsync(trav, A.trav, sizeof trav, hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    p->last.aabb_tests = trav[0];
    p->last.tri_tests = trav[1];
    p->last.bounce_rays = 0;
    return cnt[1] ? fail(OMG_ERR_OVERFLOW, "segment crossing list overflowed") : OMG_OK;
}

extern "C" int omg_paths_find_async(omg_paths* p, const omg_path_budget* b, const float* d_src,
                                    const float* d_lis, uint32_t n, omg_path* d_out,
                                    uint32_t* d_n_paths, uint32_t* d_work, void* stream)
{
    int rc = check_budget(p, b);
    if (rc) return rc;
    if (n && (!d_src || !d_lis || !d_out || !d_n_paths || !d_work))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (n == 0) return OMG_OK;
    omg_ctx* c = p->scene->ctx;
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    if (p->d_plan.cap < (size_t)n * plan_layout((uint32_t)p->edges.size(), b->edge_candidates, b->pair_edges).words * sizeof(float)) {
        // growing the plan buffer frees the old one: wait for work that may read it
        OMG_HIP_CHECK(hipDeviceSynchronize());
        OMG_HIP_CHECK(ensure_plan(p, b, n));
    }
    // the blocked count is per launch; the overflow flag stays set until
    // omg_paths_status reads it (the count word is cleared by a fill kernel)
    OMG_HIP_CHECK(launch_fill_zero(p->d_cnt.as<float>(), 1, st));
    PathsArgs A{};
    fill_args(p, b, A);
    A.src = d_src;
    A.lis = d_lis;
    A.n = n;
    A.out = reinterpret_cast<PathDev*>(d_out);
    A.n_out = d_n_paths;
    A.work = d_work;
    A.counters = p->d_cnt.as<uint32_t>();
    A.trav = p->d_trav.as<unsigned long long>();
    // blocking calls on this context wait for it before reusing its buffers
    c->async_inflight = true;
    p->async_unread = true;
    OMG_HIP_CHECK(launch_paths(A, false, st));
    return OMG_OK;
}

extern "C" int omg_paths_status(omg_paths* p)
{
    if (!p) return fail(OMG_ERR_INVALID, "bad argument");
    omg_ctx* c = p->scene->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    OMG_HIP_CHECK(hipDeviceSynchronize());
    uint32_t cnt[2];
    OMG_HIP_CHECK(hipMemcpy(cnt, p->d_cnt.ptr, sizeof cnt, hipMemcpyDeviceToHost));
    OMG_HIP_CHECK(launch_fill_zero(p->d_cnt.as<float>(), 2, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->async_inflight = false;
    const bool over = cnt[1] != 0 || p->pending_overflow;
    p->pending_overflow = false;
    p->async_unread = false;
    return over ? fail(OMG_ERR_OVERFLOW, "segment crossing list overflowed") : OMG_OK;
}

extern "C" int omg_paths_set_counting(omg_paths* p, uint32_t on)
{
    if (!p) return fail(OMG_ERR_INVALID, "bad argument");
    p->counting = on != 0;
    return OMG_OK;
}

extern "C" int omg_paths_counters(omg_paths* p, omg_counters* counters)
{
    if (!p || !counters) return fail(OMG_ERR_INVALID, "bad argument");
    *counters = p->last;
    return OMG_OK;
}

extern "C" float omg_knife_edge_amp(float detour_m, float freq_hz)
{
    return knife_edge_amp(detour_m, freq_hz);
}

extern "C" void omg_bend_point(const float a[3], const float b[3], const float src[3],
                               const float lis[3], float p[3])
{
    f3 r = bend_point(mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]), mk(src[0], src[1], src[2]),
                      mk(lis[0], lis[1], lis[2]));
    p[0] = r.x;
    p[1] = r.y;
    p[2] = r.z;
}

extern "C" void omg_paths_destroy(omg_paths* p)
{
    if (!p) return;
    omg_ctx* c = p->ctx;   // the scene may already be gone
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf* d : {&p->d_edges, &p->d_src, &p->d_lis, &p->d_out, &p->d_n, &p->d_work, &p->d_cnt,
                      &p->d_trav, &p->d_plan})
        d->release();
    delete p;
    ctx_release(c);
}
