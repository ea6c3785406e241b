// C ABI: the batched persistent launch.
//
// Many trace_world queries (different scenes, sources, listeners, ray and
// band counts) become one list of work items of up to 64 consecutive rays
// of one source, consumed by a persistent work-stealing kernel
// (k_trace_batch).  The clear of the echogram arena and the launch are
// captured once in a hipGraph (fill kernels only, no memset nodes) and
// replayed per omg_batch_launch.  The launch shape comes from a sampled
// autotuner whose choice is cached on the context per batch class.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

#include "runtime.h"

#pragma clang fp contract(off)

using namespace omg;

namespace {

struct QueryOut {
    omg_scene* scene;
    uint32_t n_src, n_lis, n_rays, nbins, nbands;
    size_t bins_off, dirs_off, rec_off;  // elements of the output arenas
};

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

uint32_t log2_ceil(uint64_t x)
{
    uint32_t l = 0;
    while ((1ull << l) < x && l < 63) l++;
    return l;
}

}  // namespace

struct omg_batch {
    omg_ctx* ctx = nullptr;
    uint32_t flags = 0;
    std::vector<QueryOut> q;
    std::vector<BatchItem> items_host;
    DevBuf in, jobs, items, sample, ctl, bins, dirs, rec;
    size_t n_bins_f = 0, n_dirs_f = 0, n_rec = 0;
    uint64_t rays = 0;
    uint32_t n_items = 0, nbm = 0;
    uint64_t class_key = 0;
    BatchShape shape{1, 16, 4, 1, 2};
    bool all4 = true;              // every scene has its 4-wide tree
    bool deep = true;              // ... and its stack fits kDeepStack entries
    bool all16 = true;             // every scene has a usable 16-wide tree
    float shape_ms = 0.0f;
    hipStream_t stream = nullptr;  // capture stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t last = nullptr;    // stream of the last launch
};

namespace {

BatchArgs batch_args(omg_batch* b, const BatchItem* items, uint32_t n_items)
{
    BatchArgs B{};
    B.jobs = b->jobs.as<JobView>();
    B.items = items;
    B.n_items = n_items;
    B.next_item = b->ctl.as<uint32_t>();
    B.counters = (b->flags & OMG_BATCH_COUNTERS) ? b->ctl.as<unsigned long long>() + 1 : nullptr;
    return B;
}

void drop_graph(omg_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Records the cursor/counter clear, the echogram clear and the persistent
// launch into a graph.
int capture(omg_batch* b)
{
    drop_graph(b);
    omg_ctx* c = b->ctx;
    BatchArgs B = batch_args(b, b->items.as<BatchItem>(), b->n_items);
    OMG_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch_fill_zero_u64(b->ctl.as<unsigned long long>(), 4, b->stream);
    if (e == hipSuccess) e = launch_fill_zero(b->bins.as<float>(), b->n_bins_f, b->stream);
    if (e == hipSuccess) e = launch_fill_zero(b->dirs.as<float>(), b->n_dirs_f, b->stream);
    // traversal counters are the reference walk's: a counting batch walks the binary tree
    BatchShape sh = b->shape;
    if (b->flags & OMG_BATCH_COUNTERS) {
        sh.width = 2;
        if (sh.regen > 1) sh.regen = 1;
        if (sh.stack_lds == kDeepStack) sh.stack_lds = 16;
        if (sh.lanes == 16) {
            sh.lanes = 4;
            sh.stack_lds = 16;
        }
    }
    if (e == hipSuccess)
        e = launch_trace_batch(B, b->nbm, sh, c->n_cu, (b->flags & OMG_BATCH_COUNTERS) != 0,
                               b->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(OMG_ERR_DEVICE, std::string("batch capture: ") +
                                        hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    e = hipGraphInstantiateWithFlags(&b->exec, g, 0);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return fail(OMG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return OMG_OK;
}

void batch_free(omg_batch* b)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->last) (void)hipStreamSynchronize(b->last);
    drop_graph(b);
    for (DevBuf* d : {&b->in, &b->jobs, &b->items, &b->sample, &b->ctl, &b->bins, &b->dirs, &b->rec})
        d->release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

}  // namespace

extern "C" int omg_batch_create(omg_ctx* c, const omg_query* queries, uint32_t n, uint32_t flags,
                                omg_batch** out)
{
    if (!c || !queries || !out || n == 0) return fail(OMG_ERR_INVALID, "bad argument");
    if (flags & ~(OMG_BATCH_RECORDS | OMG_BATCH_COUNTERS)) return fail(OMG_ERR_INVALID, "unknown flags");
    *out = nullptr;
    // validate every query first: nothing is allocated for a refused batch
    std::vector<uint32_t> nbins(n);
    uint32_t nbm = 1;
    uint64_t rays = 0, items = 0;
    double lg_prims = 0.0;
    bool all4 = true, deep = true, all16 = true;
    for (uint32_t i = 0; i < n; i++) {
        const omg_query& Q = queries[i];
        if (!Q.scene || Q.scene->ctx != c || !Q.sources || !Q.listeners || (Q.n_ovl && !Q.overlays))
            return fail(OMG_ERR_INVALID, "query " + std::to_string(i) + ": bad argument");
        int rc = validate_params(&Q.params, Q.scene->nbands, Q.n_src, Q.n_lis, Q.n_ovl, &nbins[i]);
        if (rc) return fail(rc, "query " + std::to_string(i) + ": " + omg_last_error());
        nbm = std::max(nbm, Q.scene->nbands);
        all4 = all4 && Q.scene->has4;
        all16 = all16 && Q.scene->has16;
        deep = deep && Q.scene->has4 && Q.scene->stack4 <= kDeepStack;
        uint64_t r = (uint64_t)Q.params.n_rays * Q.n_src;
        rays += r;
        items += (uint64_t)Q.n_src * ((Q.params.n_rays + kBatchItemRays - 1) / kBatchItemRays);
        lg_prims += (double)r * std::log2((double)std::max(Q.scene->n_prims, 1u));
    }
    if (items >= (1ull << 32)) return fail(OMG_ERR_CAP, "too many work items");
    OMG_HIP_CHECK(hipSetDevice(c->device));

    omg_batch* b = new omg_batch();
    b->ctx = c;
    b->flags = flags;
    b->nbm = nbm;
    b->rays = rays;
    b->all4 = all4;
    b->all16 = all16;
    b->deep = deep;
    b->q.resize(n);
    // host staging of every query's tables: mats | overlays | sources | listeners
    std::vector<size_t> off_mats(n), off_ovl(n), off_src(n), off_lis(n);
    std::vector<std::vector<MatParams>> mats(n);
    std::vector<std::vector<Overlay>> ovl(n);
    size_t in_bytes = 0, bf = 0, df = 0, rec = 0;
    for (uint32_t i = 0; i < n; i++) {
        const omg_query& Q = queries[i];
        job_tables(Q.scene->mat_params, Q.overlays, Q.n_ovl, Q.scene->nbands, mats[i], ovl[i]);
        off_mats[i] = in_bytes;
        in_bytes = align16(in_bytes + mats[i].size() * sizeof(MatParams));
        off_ovl[i] = in_bytes;
        in_bytes = align16(in_bytes + ovl[i].size() * sizeof(Overlay));
        off_src[i] = in_bytes;
        in_bytes = align16(in_bytes + (size_t)Q.n_src * 3 * sizeof(float));
        off_lis[i] = in_bytes;
        in_bytes = align16(in_bytes + (size_t)Q.n_lis * 3 * sizeof(float));
        QueryOut& o = b->q[i];
        o.scene = Q.scene;
        o.n_src = Q.n_src;
        o.n_lis = Q.n_lis;
        o.n_rays = Q.params.n_rays;
        o.nbins = nbins[i];
        o.nbands = Q.scene->nbands;
        o.bins_off = bf;
        o.dirs_off = df;
        o.rec_off = rec;
        bf += (size_t)Q.n_src * Q.n_lis * nbins[i] * o.nbands;
        df += (size_t)Q.n_src * Q.n_lis * nbins[i] * 3;
        rec += (size_t)Q.n_src * Q.params.n_rays;
        // keep every query's arena slice 16 B aligned
        bf = (bf + 3) & ~size_t(3);
        df = (df + 3) & ~size_t(3);
    }
    std::vector<uint8_t> in(in_bytes ? in_bytes : 16);
    for (uint32_t i = 0; i < n; i++) {
        const omg_query& Q = queries[i];
        std::memcpy(in.data() + off_mats[i], mats[i].data(), mats[i].size() * sizeof(MatParams));
        if (!ovl[i].empty())
            std::memcpy(in.data() + off_ovl[i], ovl[i].data(), ovl[i].size() * sizeof(Overlay));
        std::memcpy(in.data() + off_src[i], Q.sources, (size_t)Q.n_src * 3 * sizeof(float));
        std::memcpy(in.data() + off_lis[i], Q.listeners, (size_t)Q.n_lis * 3 * sizeof(float));
    }
    b->n_bins_f = bf;
    b->n_dirs_f = df;
    b->n_rec = rec;
    int rc = OMG_OK;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        b->stream = nullptr;
        rc = fail(OMG_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (!rc) rc = upload(b->in, in.data(), in.size(), b->stream);
    if (!rc && b->bins.ensure(std::max<size_t>(bf, 4) * sizeof(float)) != hipSuccess)
        rc = fail(OMG_ERR_DEVICE, "echogram arena allocation");
    if (!rc && b->dirs.ensure(std::max<size_t>(df, 4) * sizeof(float)) != hipSuccess)
        rc = fail(OMG_ERR_DEVICE, "direction arena allocation");
    if (!rc && (flags & OMG_BATCH_RECORDS) && b->rec.ensure(rec * sizeof(omg_ray_record)) != hipSuccess)
        rc = fail(OMG_ERR_DEVICE, "record arena allocation");
    if (!rc && b->ctl.ensure(4 * sizeof(unsigned long long)) != hipSuccess)
        rc = fail(OMG_ERR_DEVICE, "control words allocation");
    std::vector<JobView> jobs(n);
    if (!rc) {
        uint8_t* base = b->in.as<uint8_t>();
        for (uint32_t i = 0; i < n; i++) {
            const omg_query& Q = queries[i];
            const QueryOut& o = b->q[i];
            JobView& J = jobs[i];
            std::memset(&J, 0, sizeof J);
            fill_geom(Q.scene, J.g);
            J.mats = reinterpret_cast<const MatParams*>(base + off_mats[i]);
            J.overlays = reinterpret_cast<const Overlay*>(base + off_ovl[i]);
            J.sources = reinterpret_cast<const float*>(base + off_src[i]);
            J.listeners = reinterpret_cast<const float*>(base + off_lis[i]);
            J.n_src = Q.n_src;
            J.n_lis = Q.n_lis;
            job_scalars(&Q.params, o.nbands, Q.n_ovl, J);
            J.nbins = o.nbins;
            J.bins = b->bins.as<float>() + o.bins_off;
            J.dirs = b->dirs.as<float>() + o.dirs_off;
            J.records = (flags & OMG_BATCH_RECORDS)
                            ? reinterpret_cast<uint32_t*>(b->rec.as<omg_ray_record>() + o.rec_off)
                            : nullptr;
        }
        rc = upload(b->jobs, jobs.data(), jobs.size() * sizeof(JobView), b->stream);
    }
    if (!rc) {
        // queries of one scene are consecutive in the work list, so at any
        // moment the device walks few BVHs and they stay in L2 / Infinity
        // Cache.  Larger scenes come first, and in a scene the queries whose
        // rays may live longest: the launch ends with the rays taken last,
        // so the longest work is taken early
        std::vector<uint32_t> order(n);
        for (uint32_t i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            const omg_query& qx = queries[x];
            const omg_query& qy = queries[y];
            if (qx.scene != qy.scene) {
                if (qx.scene->n_prims != qy.scene->n_prims) return qx.scene->n_prims > qy.scene->n_prims;
                return std::less<const omg_scene*>()(qx.scene, qy.scene);
            }
            if (qx.params.max_bounces != qy.params.max_bounces)
                return qx.params.max_bounces > qy.params.max_bounces;
            return qx.params.max_time > qy.params.max_time;
        });
        b->items_host.reserve(items);
        for (uint32_t i : order)
            for (uint32_t s = 0; s < b->q[i].n_src; s++)
                for (uint32_t r = 0; r < b->q[i].n_rays; r += kBatchItemRays)
                    b->items_host.push_back(
                        BatchItem{i, s, r, std::min(kBatchItemRays, b->q[i].n_rays - r)});
        b->n_items = (uint32_t)b->items_host.size();
        rc = upload(b->items, b->items_host.data(), b->items_host.size() * sizeof(BatchItem), b->stream);
    }
    // the staging vectors die with this frame
    if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "batch upload");
    if (!rc) {
        uint32_t lg = rays ? (uint32_t)std::lround(lg_prims / (double)rays) : 0;
        b->class_key = ((uint64_t)(nbm <= 3 ? 3 : nbm <= 6 ? 6 : 8) << 16) |
                       ((uint64_t)log2_ceil(rays) << 8) | lg;
        // untuned default: the regenerating walk (4-wide, rays replaced at
        // walk steps, where every scene has the 4-wide tree)
        b->shape = all4 ? BatchShape{1, 12, 6, 2, 4} : BatchShape{1, 8, 8, 1, 2};
        auto it = c->tuned.find(b->class_key);
        if (it != c->tuned.end() && (it->second.width == 2 || (it->second.width == 4 && all4) ||
                                     (it->second.width == 16 && all16)) &&
            (it->second.stack_lds != kDeepStack || deep))
            b->shape = it->second;
        rc = capture(b);
    }
    if (rc) {
        batch_free(b);
        delete b;
        return rc;
    }
    c->live_scenes++;
    *out = b;
    return OMG_OK;
}

extern "C" int omg_batch_launch(omg_batch* b, void* stream)
{
    if (!b || !b->exec) return fail(OMG_ERR_INVALID, "bad argument");
    OMG_HIP_CHECK(hipSetDevice(b->ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : b->ctx->stream;
    OMG_HIP_CHECK(hipGraphLaunch(b->exec, st));
    b->last = st;
    return OMG_OK;
}

extern "C" int omg_batch_sync(omg_batch* b)
{
    if (!b) return fail(OMG_ERR_INVALID, "bad argument");
    OMG_HIP_CHECK(hipSetDevice(b->ctx->device));
    if (b->last) OMG_HIP_CHECK(hipStreamSynchronize(b->last));
    return OMG_OK;
}

extern "C" int omg_batch_result(omg_batch* b, uint32_t qi, float* bins, float* dirs,
                                omg_ray_record* records)
{
    if (!b || qi >= b->q.size() || !bins || !dirs) return fail(OMG_ERR_INVALID, "bad argument");
    if (records && !(b->flags & OMG_BATCH_RECORDS))
        return fail(OMG_ERR_INVALID, "batch was created without OMG_BATCH_RECORDS");
    int rc = omg_batch_sync(b);
    if (rc) return rc;
    const QueryOut& o = b->q[qi];
    size_t nb_f = (size_t)o.n_src * o.n_lis * o.nbins * o.nbands;
    size_t nd_f = (size_t)o.n_src * o.n_lis * o.nbins * 3;
    OMG_HIP_CHECK(hipMemcpy(bins, b->bins.as<float>() + o.bins_off, nb_f * sizeof(float),
                            hipMemcpyDeviceToHost));
    OMG_HIP_CHECK(hipMemcpy(dirs, b->dirs.as<float>() + o.dirs_off, nd_f * sizeof(float),
                            hipMemcpyDeviceToHost));
    if (records)
        OMG_HIP_CHECK(hipMemcpy(records, b->rec.as<omg_ray_record>() + o.rec_off,
                                (size_t)o.n_src * o.n_rays * sizeof(omg_ray_record),
                                hipMemcpyDeviceToHost));
    return OMG_OK;
}

extern "C" int omg_batch_device_output(omg_batch* b, uint32_t qi, float** d_bins, float** d_dirs)
{
    if (!b || qi >= b->q.size() || !d_bins || !d_dirs) return fail(OMG_ERR_INVALID, "bad argument");
    *d_bins = b->bins.as<float>() + b->q[qi].bins_off;
    *d_dirs = b->dirs.as<float>() + b->q[qi].dirs_off;
    return OMG_OK;
}

extern "C" int omg_batch_counters(omg_batch* b, omg_counters* out)
{
    if (!b || !out) return fail(OMG_ERR_INVALID, "bad argument");
    if (!(b->flags & OMG_BATCH_COUNTERS))
        return fail(OMG_ERR_INVALID, "batch was created without OMG_BATCH_COUNTERS");
    int rc = omg_batch_sync(b);
    if (rc) return rc;
    unsigned long long w[4];
    OMG_HIP_CHECK(hipMemcpy(w, b->ctl.ptr, sizeof w, hipMemcpyDeviceToHost));
    out->aabb_tests = w[1];
    out->tri_tests = w[2];
    out->bounce_rays = w[3];
    return OMG_OK;
}

extern "C" int omg_batch_size(omg_batch* b, uint64_t* rays, uint32_t* items)
{
    if (!b || !rays || !items) return fail(OMG_ERR_INVALID, "bad argument");
    *rays = b->rays;
    *items = b->n_items;
    return OMG_OK;
}

extern "C" int omg_batch_get_shape(omg_batch* b, omg_batch_shape* s)
{
    if (!b || !s) return fail(OMG_ERR_INVALID, "bad argument");
    s->lanes = b->shape.lanes;
    s->stack_lds = b->shape.stack_lds;
    s->waves_per_simd = b->shape.waves_per_simd;
    s->regen = b->shape.regen;
    s->width = b->shape.width;
    s->ms = b->shape_ms;
    return OMG_OK;
}

extern "C" int omg_batch_set_shape(omg_batch* b, const omg_batch_shape* s)
{
    if (!b || !s) return fail(OMG_ERR_INVALID, "bad argument");
    BatchShape sh{s->lanes, s->stack_lds, s->waves_per_simd, s->regen, s->width};
    if (!valid_batch_shape(sh))
        return fail(OMG_ERR_INVALID,
                    "shape: lanes 1|4 with stack_lds 16|32 (width 2) or 16 (width 4); regen 1 "
                    "with lanes 1, stack_lds 8|12|16 and width 2|4, or lanes 4, stack_lds 16 and "
                    "width 2; lanes 4 with stack_lds 40 and width 4; regen 1 with lanes 16, "
                    "stack_lds 96 and width 16; regen 2 with lanes 1, stack_lds 12|16 and "
                    "width 4; waves_per_simd 1..8");
    if (sh.width == 4 && !b->all4) return fail(OMG_ERR_INVALID, "a scene of the batch has no 4-wide tree");
    if (sh.width == 16 && !b->all16)
        return fail(OMG_ERR_INVALID, "a scene of the batch has no usable 16-wide tree");
    if (sh.stack_lds == kDeepStack && !b->deep)
        return fail(OMG_ERR_INVALID, "a scene of the batch needs more than 40 stack entries");
    OMG_HIP_CHECK(hipSetDevice(b->ctx->device));
    int rc = omg_batch_sync(b);
    if (rc) return rc;
    b->shape = sh;
    b->shape_ms = 0.0f;
    return capture(b);
}

extern "C" int omg_batch_tune(omg_batch* b, uint32_t stride, omg_batch_shape* out)
{
    if (!b) return fail(OMG_ERR_INVALID, "bad argument");
    if (stride == 0) stride = 1;
    omg_ctx* c = b->ctx;
    // a sample too small to keep every lane busy for several rays judges
    // occupancy by the launch's tail: keep at least 256 items per CU
    const uint64_t min_items = (uint64_t)c->n_cu * 256u;
    const uint64_t cap = std::max<uint64_t>(1, b->items_host.size() / min_items);
    if (stride > cap) stride = (uint32_t)cap;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    int rc = omg_batch_sync(b);
    if (rc) return rc;
    std::vector<BatchItem> smp;
    for (size_t i = 0; i < b->items_host.size(); i += stride) smp.push_back(b->items_host[i]);
    rc = upload(b->sample, smp.data(), smp.size() * sizeof(BatchItem), b->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(hipStreamSynchronize(b->stream));
    BatchArgs B = batch_args(b, b->sample.as<BatchItem>(), (uint32_t)smp.size());
    B.counters = nullptr;
    hipEvent_t e0, e1;
    OMG_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(OMG_ERR_DEVICE, "hipEventCreate");
    }
    hipError_t e = hipSuccess;
    // best of two sampled launches of one shape
    auto time_shape = [&](const BatchShape& sh) {
        float ms = 1e30f;
        for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {
            e = hipEventRecord(e0, b->stream);
            if (e == hipSuccess) e = launch_fill_zero_u64(b->ctl.as<unsigned long long>(), 1, b->stream);
            if (e == hipSuccess) e = launch_trace_batch(B, b->nbm, sh, c->n_cu, false, b->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e == hipSuccess) ms = std::min(ms, t);
        }
        return ms;
    };
    // the current shape is kept unless a candidate is 3 % faster, so that
    // timing noise does not flip between near-equal shapes from run to run
    BatchShape best = b->shape;
    float best_ms = time_shape(best);
    // (lanes, stack_lds, regen, width): a lane or a quad per ray, with and
    // without regeneration (between bounces, or at walk steps), over the
    // binary or the 4-wide tree; each at 2..8 resident waves per SIMD
    const uint32_t modes[18][4] = {{1u, 8u, 1u, 2u},  {1u, 12u, 1u, 2u}, {1u, 16u, 1u, 2u},
                                   {1u, 8u, 1u, 4u},  {1u, 12u, 1u, 4u}, {1u, 16u, 1u, 4u},
                                   {1u, 12u, 2u, 4u}, {1u, 16u, 2u, 4u},
                                   {4u, 16u, 1u, 2u}, {4u, 40u, 1u, 4u}, {1u, 16u, 0u, 2u},
                                   {1u, 32u, 0u, 2u}, {4u, 16u, 0u, 2u}, {4u, 32u, 0u, 2u},
                                   {1u, 16u, 0u, 4u}, {4u, 16u, 0u, 4u}, {4u, 40u, 0u, 4u},
                                   {16u, 96u, 1u, 16u}};
    static_assert(kDeepStack == 40 && kStack16 == 96, "tuner modes name the deep stacks");
    for (const auto& m : modes) {
        if (m[3] == 4 && !b->all4) continue;
        if (m[3] == 16 && !b->all16) continue;
        if (m[1] == kDeepStack && !b->deep) continue;
        for (uint32_t wps : {2u, 4u, 6u, 8u}) {
            BatchShape sh{m[0], m[1], wps, m[2], m[3]};
            float ms = time_shape(sh);
            if (e != hipSuccess) break;
            if (ms < 0.97f * best_ms) {
                best_ms = ms;
                best = sh;
            }
        }
        if (e != hipSuccess) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(OMG_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
    b->shape = best;
    b->shape_ms = best_ms;
    c->tuned[b->class_key] = best;
    rc = capture(b);
    if (rc) return rc;
    if (out) omg_batch_get_shape(b, out);
    return OMG_OK;
}

extern "C" void omg_batch_destroy(omg_batch* b)
{
    if (!b) return;
    omg_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    batch_free(b);
    delete b;
    ctx_release(c);
}
