// C ABI: the serving loop.
//
// A scheduler thread owns `slots` streams.  Submitted requests wait in a
// FIFO; whenever a slot is free the thread packs the waiting requests that
// fit (the HBM budget and max_batch) into one batch: every request's
// tables, job view and work items go to the device in one pinned copy, the
// persistent regenerating tracer runs over all of them, the dome fans run
// per scene group, and the outputs come back in one pinned copy, all on
// the slot's stream.  Batches on different slots overlap on the device, so
// requests that arrive while others run are admitted as soon as a slot and
// memory free up (continuous batching).  Device memory is the resident
// scenes plus the slots' grow-only work arenas; an arena only grows while
// the total stays within the budget.
#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "dome_kernels.h"
#include "runtime.h"

#pragma clang fp contract(off)

using namespace omg;

namespace {

using Clock = std::chrono::steady_clock;

size_t a16(size_t x) { return (x + 15) & ~size_t(15); }

struct Req {
    uint64_t ticket = 0;
    omg_scene* scene = nullptr;
    omg_trace_params p{};
    std::vector<float> src, lis;
    uint32_t n_src = 0, n_lis = 0, dome = 0, flags = 0, nbins = 0, nbands = 0;
    size_t footprint = 0;
    double t_arrival = 0, t_launch = 0, t_done = 0;
    uint32_t batch_size = 0, slot = 0;
    int status = OMG_OK;
    bool done = false;
    std::vector<float> bins, dirs, dome_e, dome_d;
    std::vector<omg_ray_record> rec;
    // placement in the slot's arena for the batch it runs in (bytes)
    size_t o_bins = 0, o_dirs = 0, o_dome_e = 0, o_dome_d = 0, o_rec = 0;

    size_t n_bins_f() const { return (size_t)n_src * n_lis * nbins * nbands; }
    size_t n_dirs_f() const { return (size_t)n_src * n_lis * nbins * 3; }
};
using ReqP = std::shared_ptr<Req>;

// rays per work item of a latency-mode batch: one pass of a wave of quads
constexpr uint32_t kLatencyItemRays = 16;
// First sizes of a slot's device arena and pinned staging.  Freeing device
// or pinned memory waits for the whole device, so a slot that grew its
// buffers a step at a time would stall every other slot's batch each time;
// starting at these sizes (a batch of a few requests of up to 32 x 16384 rays needs a
// few MB) makes growth rare.
constexpr size_t kMinArenaBytes = 16u << 20;
constexpr size_t kMinStagingBytes = 4u << 20;
// Sizes given to every slot at omg_server_create when the budget leaves
// each slot at least kMinArenaBytes: a batch of 64 requests of up to 32
// sources x 16384 rays fits without growing, so a fresh server serves its
// first requests at its steady-state latency.
constexpr size_t kPresizeArenaBytes = 64u << 20;
constexpr size_t kPresizeStagingBytes = 8u << 20;

struct Slot {
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;    // dome fans, overlapped with the tracer
    hipEvent_t ev = nullptr, fork = nullptr, join = nullptr;
    DevBuf arena;
    // pinned staging, mapped into the device's address space: the batch's
    // copy kernels read host_in and write host_out through *_dev
    void* host_in = nullptr;
    void* host_in_dev = nullptr;
    size_t host_in_cap = 0;
    void* host_out = nullptr;
    void* host_out_dev = nullptr;
    size_t host_out_cap = 0;
    std::vector<ReqP> reqs;
    bool busy = false;
    size_t out_off = 0, out_bytes = 0;   // [ctl | outputs | records] copied back
};

// Coherent (fine-grained) pinned memory: the device reads and writes it
// over the bus without caching it, so a copy kernel's reads see the host's
// latest tables and its writes are visible once the batch's event fires.
int host_grow(void*& p, void*& dev, size_t& cap, size_t want)
{
    if (want <= cap) return OMG_OK;
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    cap = 0;
    size_t n = std::max(want + want / 4 + 4096, kMinStagingBytes);
    if (hipHostMalloc(&p, n, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
        p = nullptr;
        return fail(OMG_ERR_DEVICE, "hipHostMalloc");
    }
    if (hipHostGetDevicePointer(&dev, p, 0) != hipSuccess || !dev) {
        (void)hipHostFree(p);
        p = dev = nullptr;
        return fail(OMG_ERR_DEVICE, "hipHostGetDevicePointer");
    }
    cap = n;
    return OMG_OK;
}

}  // namespace

struct omg_server {
    omg_ctx* ctx = nullptr;
    omg_server_config cfg{};
    BatchShape shape{};
    bool deep = false;  // every scene's 4-wide stack fits kDeepStack entries
    std::vector<omg_scene*> scenes;
    uint64_t scenes_bytes = 0;
    std::vector<Slot> slots;
    DevBuf fan, fan_bin, no_panels;
    Clock::time_point t0;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqP> pending;
    std::unordered_map<uint64_t, ReqP> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    omg_server_stats st{};
    uint64_t batched_reqs = 0;
    std::thread th;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    uint64_t committed() const
    {
        uint64_t b = scenes_bytes;
        for (const Slot& s : slots) b += s.arena.cap;
        return b;
    }

    // Device bytes a request occupies in a work arena (inputs, job view,
    // work items, outputs, records, dome scratch), with alignment slack.
    size_t footprint(const Req& r) const
    {
        size_t b = a16(r.scene->mat_params.size() * sizeof(MatParams)) + a16(r.n_src * 12) +
                   a16(r.n_lis * 12) + a16(sizeof(JobView)) +
                   a16((size_t)r.n_src * ((r.p.n_rays + kLatencyItemRays - 1) / kLatencyItemRays) *
                       sizeof(BatchItem)) +
                   a16(r.n_bins_f() * 4) + a16(r.n_dirs_f() * 4);
        if (r.flags & OMG_REQ_RECORDS) b += a16((size_t)r.n_src * r.p.n_rays * sizeof(omg_ray_record));
        if (r.dome)
            b += a16(r.n_lis * 12) + a16((size_t)r.n_lis * kDomeBins * r.nbands * 4) +
                 a16((size_t)r.n_lis * kDomeBins * 3 * 4) +
                 a16((size_t)r.n_lis * cfg.dome.n_rays * sizeof(DomeRayDev));
        return b + 256;
    }

    void loop();
    int launch(uint32_t si, std::vector<ReqP>& batch, uint64_t room);
    void finish(uint32_t si, int status);
};

// Packs one batch into slot si and enqueues it.  Layout of the arena:
// [inputs: per request mats | sources | listeners, dome eyes per group,
//  job views, work items][ctl 32 B | per request bins | dirs, dome energy |
//  dsum per group | records][dome ray scratch].  The second part is cleared
// on the device and copied back after the kernels.
int omg_server::launch(uint32_t si, std::vector<ReqP>& batch, uint64_t room)
{
    Slot& S = slots[si];
    const uint32_t n = (uint32_t)batch.size();
    // dome groups: requests with a dome, by scene, in batch order
    std::vector<omg_scene*> gscene;
    std::vector<std::vector<uint32_t>> gmem;
    for (uint32_t i = 0; i < n; i++) {
        if (!batch[i]->dome) continue;
        uint32_t g = 0;
        while (g < gscene.size() && gscene[g] != batch[i]->scene) g++;
        if (g == gscene.size()) {
            gscene.push_back(batch[i]->scene);
            gmem.emplace_back();
        }
        gmem[g].push_back(i);
    }
    const uint32_t ng = (uint32_t)gscene.size();
    std::vector<uint32_t> g_eyes(ng, 0);
    for (uint32_t g = 0; g < ng; g++)
        for (uint32_t i : gmem[g]) g_eyes[g] += batch[i]->n_lis;

    // small batches run in latency mode: a DPP quad per ray (the whole
    // traversal stack in LDS where the trees allow it) with regeneration
    // over 16-ray items; the grid is sized to the batch
    uint64_t rays = 0;
    for (uint32_t i = 0; i < n; i++) rays += (uint64_t)batch[i]->n_src * batch[i]->p.n_rays;
    const bool lat = cfg.latency_rays && rays <= cfg.latency_rays;
    const uint32_t item_rays = lat ? kLatencyItemRays : kBatchItemRays;
    const BatchShape sh = !lat                        ? shape
                          : deep && shape.width == 4 ? BatchShape{4, kDeepStack, 8, 1, 4}
                                                     : BatchShape{4, 16, 8, 1, shape.width};

    // inputs
    std::vector<size_t> o_mats(n), o_src(n), o_lis(n), g_eye_off(ng);
    size_t in = 0;
    uint64_t items = 0;
    uint32_t nbm = 1;
    for (uint32_t i = 0; i < n; i++) {
        const Req& r = *batch[i];
        o_mats[i] = in;
        in = a16(in + r.scene->mat_params.size() * sizeof(MatParams));
        o_src[i] = in;
        in = a16(in + r.n_src * 12);
        o_lis[i] = in;
        in = a16(in + r.n_lis * 12);
        items += (uint64_t)r.n_src * ((r.p.n_rays + item_rays - 1) / item_rays);
        nbm = std::max(nbm, r.nbands);
    }
    for (uint32_t g = 0; g < ng; g++) {
        g_eye_off[g] = in;
        in = a16(in + (size_t)g_eyes[g] * 12);
    }
    const size_t o_jobs = in;
    in = a16(in + n * sizeof(JobView));
    const size_t o_items = in;
    in = a16(in + items * sizeof(BatchItem));
    // outputs
    const size_t o_ctl = in;
    size_t out = in + 32;
    std::vector<size_t> g_e(ng), g_d(ng);
    for (uint32_t i = 0; i < n; i++) {
        Req& r = *batch[i];
        r.o_bins = out;
        out = a16(out + r.n_bins_f() * 4);
        r.o_dirs = out;
        out = a16(out + r.n_dirs_f() * 4);
    }
    for (uint32_t g = 0; g < ng; g++) {
        uint32_t nb = gscene[g]->nbands, e = 0;
        g_e[g] = out;
        out = a16(out + (size_t)g_eyes[g] * kDomeBins * nb * 4);
        g_d[g] = out;
        out = a16(out + (size_t)g_eyes[g] * kDomeBins * 3 * 4);
        for (uint32_t i : gmem[g]) {
            Req& r = *batch[i];
            r.o_dome_e = g_e[g] + (size_t)e * kDomeBins * nb * 4;
            r.o_dome_d = g_d[g] + (size_t)e * kDomeBins * 3 * 4;
            e += r.n_lis;
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        Req& r = *batch[i];
        if (!(r.flags & OMG_REQ_RECORDS)) continue;
        r.o_rec = out;
        out = a16(out + (size_t)r.n_src * r.p.n_rays * sizeof(omg_ray_record));
    }
    const size_t out_end = out;
    std::vector<size_t> g_rays(ng);
    for (uint32_t g = 0; g < ng; g++) {
        g_rays[g] = out;
        out = a16(out + (size_t)g_eyes[g] * cfg.dome.n_rays * sizeof(DomeRayDev));
    }
    const size_t total = out;

    int rc = OMG_OK;
    if (total > S.arena.cap) {
        // grow by doubling (from kMinArenaBytes) within the room the budget leaves this slot
        uint64_t want = std::max<uint64_t>({total, 2 * (uint64_t)S.arena.cap, kMinArenaBytes});
        want = std::min<uint64_t>(want, std::max<uint64_t>(room, total));
        if (S.arena.ensure_exact((size_t)want) != hipSuccess)
            return fail(OMG_ERR_DEVICE, "work arena allocation");
    }
    rc = host_grow(S.host_in, S.host_in_dev, S.host_in_cap, o_ctl);
    if (!rc) rc = host_grow(S.host_out, S.host_out_dev, S.host_out_cap, out_end - o_ctl);
    if (rc) return rc;
    uint8_t* H = static_cast<uint8_t*>(S.host_in);
    uint8_t* D = S.arena.as<uint8_t>();
    JobView* jobs = reinterpret_cast<JobView*>(H + o_jobs);
    BatchItem* it = reinterpret_cast<BatchItem*>(H + o_items);
    uint64_t k = 0;
    for (uint32_t i = 0; i < n; i++) {
        Req& r = *batch[i];
        std::memcpy(H + o_mats[i], r.scene->mat_params.data(), r.scene->mat_params.size() * sizeof(MatParams));
        std::memcpy(H + o_src[i], r.src.data(), r.n_src * 12);
        std::memcpy(H + o_lis[i], r.lis.data(), r.n_lis * 12);
        JobView& J = jobs[i];
        std::memset(&J, 0, sizeof J);
        fill_geom(r.scene, J.g);
        J.mats = reinterpret_cast<const MatParams*>(D + o_mats[i]);
        J.overlays = nullptr;
        J.sources = reinterpret_cast<const float*>(D + o_src[i]);
        J.listeners = reinterpret_cast<const float*>(D + o_lis[i]);
        J.n_src = r.n_src;
        J.n_lis = r.n_lis;
        job_scalars(&r.p, r.nbands, 0, J);
        J.nbins = r.nbins;
        J.bins = reinterpret_cast<float*>(D + r.o_bins);
        J.dirs = reinterpret_cast<float*>(D + r.o_dirs);
        J.records = (r.flags & OMG_REQ_RECORDS) ? reinterpret_cast<uint32_t*>(D + r.o_rec) : nullptr;
        for (uint32_t s = 0; s < r.n_src; s++)
            for (uint32_t ray = 0; ray < r.p.n_rays; ray += item_rays)
                it[k++] = BatchItem{i, s, ray, std::min(item_rays, r.p.n_rays - ray)};
    }
    for (uint32_t g = 0; g < ng; g++) {
        size_t e = 0;
        for (uint32_t i : gmem[g]) {
            std::memcpy(H + g_eye_off[g] + e * 12, batch[i]->lis.data(), batch[i]->n_lis * 12);
            e += batch[i]->n_lis;
        }
    }
    hipStream_t st = S.stream;
    // tables in and results out by copy kernels on the batch's stream: a
    // DMA copy's call can block the scheduler for milliseconds
    OMG_HIP_CHECK(launch_copy16(D, S.host_in_dev, o_ctl, st));
    OMG_HIP_CHECK(launch_fill_zero(reinterpret_cast<float*>(D + o_ctl), (out_end - o_ctl) / 4, st));
    BatchArgs B{};
    B.jobs = reinterpret_cast<const JobView*>(D + o_jobs);
    B.items = reinterpret_cast<const BatchItem*>(D + o_items);
    B.n_items = (uint32_t)items;
    B.next_item = reinterpret_cast<uint32_t*>(D + o_ctl);
    B.counters = nullptr;
    // with a side stream the dome fans overlap the tracer; without one they run first
    hipStream_t ds = S.side ? S.side : st;
    if (ng && S.side) {
        OMG_HIP_CHECK(hipEventRecord(S.fork, st));
        OMG_HIP_CHECK(hipStreamWaitEvent(S.side, S.fork, 0));
    }
    // a grid no larger than the batch needs leaves the rest of the chip to the other slots
    uint64_t lanes_needed = rays * sh.lanes;
    uint32_t max_blocks = (uint32_t)std::min<uint64_t>((lanes_needed + 255) / 256, 1u << 20);
    if (S.side) OMG_HIP_CHECK(launch_trace_batch(B, nbm, sh, ctx->n_cu, false, st, max_blocks));
    for (uint32_t g = 0; g < ng; g++) {
        DomeArgs A{};
        fill_geom(gscene[g], A.g);
        A.mats = gscene[g]->mats_raw.as<omg_material_dev>();
        A.panels = no_panels.as<DomePanelDev>();
        A.n_panels = 0;
        A.eyes = reinterpret_cast<const float*>(D + g_eye_off[g]);
        A.n_eyes = g_eyes[g];
        A.fan = fan.as<float>();
        A.fan_bin = fan_bin.as<uint32_t>();
        A.n_fan = cfg.dome.n_fan;
        A.n_rays = cfg.dome.n_rays;
        A.max_events = cfg.dome.max_events;
        A.nbands = gscene[g]->nbands;
        A.horizon = cfg.dome.horizon;
        A.through_len = cfg.dome.through_len;
        A.rays = reinterpret_cast<DomeRayDev*>(D + g_rays[g]);
        A.energy = reinterpret_cast<float*>(D + g_e[g]);
        A.dsum = reinterpret_cast<float*>(D + g_d[g]);
        A.status = reinterpret_cast<uint32_t*>(D + o_ctl + 8);
        A.counters = nullptr;
        OMG_HIP_CHECK(launch_dome(A, false, ds));
    }
    if (ng && S.side) {
        OMG_HIP_CHECK(hipEventRecord(S.join, S.side));
        OMG_HIP_CHECK(hipStreamWaitEvent(st, S.join, 0));
    }
    if (!S.side) OMG_HIP_CHECK(launch_trace_batch(B, nbm, sh, ctx->n_cu, false, st, max_blocks));
    OMG_HIP_CHECK(launch_copy16(S.host_out_dev, D + o_ctl, out_end - o_ctl, st));
    OMG_HIP_CHECK(hipEventRecord(S.ev, st));
    S.out_off = o_ctl;
    S.out_bytes = out_end - o_ctl;
    double t = now();
    for (uint32_t i = 0; i < n; i++) {
        batch[i]->t_launch = t;
        batch[i]->batch_size = n;
        batch[i]->slot = si;
    }
    S.reqs = batch;
    S.busy = true;
    return OMG_OK;
}

void omg_server::finish(uint32_t si, int status)
{
    Slot& S = slots[si];
    const uint8_t* O = static_cast<const uint8_t*>(S.host_out);
    bool overflow = status == OMG_OK && (reinterpret_cast<const uint32_t*>(O)[2] & 1u);
    double t = now();
    for (ReqP& rp : S.reqs) {
        Req& r = *rp;
        if (status == OMG_OK && (r.flags & OMG_REQ_KEEP)) {
            auto grab = [&](std::vector<float>& v, size_t off, size_t nf) {
                v.resize(nf);
                std::memcpy(v.data(), O + (off - S.out_off), nf * 4);
            };
            grab(r.bins, r.o_bins, r.n_bins_f());
            grab(r.dirs, r.o_dirs, r.n_dirs_f());
            if (r.dome) {
                grab(r.dome_e, r.o_dome_e, (size_t)r.n_lis * kDomeBins * r.nbands);
                grab(r.dome_d, r.o_dome_d, (size_t)r.n_lis * kDomeBins * 3);
            }
            if (r.flags & OMG_REQ_RECORDS) {
                r.rec.resize((size_t)r.n_src * r.p.n_rays);
                std::memcpy(r.rec.data(), O + (r.o_rec - S.out_off), r.rec.size() * sizeof(omg_ray_record));
            }
        }
        r.t_done = t;
    }
    {
        std::lock_guard<std::mutex> lk(m);
        for (ReqP& rp : S.reqs) {
            rp->status = status != OMG_OK ? status : (overflow && rp->dome) ? OMG_ERR_OVERFLOW : OMG_OK;
            rp->done = true;
            st.completed++;
        }
    }
    cv_done.notify_all();
    S.reqs.clear();
    S.busy = false;
}

void omg_server::loop()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        bool did = false;
        for (uint32_t si = 0; si < slots.size(); si++) {
            Slot& S = slots[si];
            if (!S.busy) continue;
            hipError_t q = hipEventQuery(S.ev);
            if (q == hipSuccess) {
                finish(si, OMG_OK);
                did = true;
            } else if (q != hipErrorNotReady) {
                if (S.side) (void)hipStreamSynchronize(S.side);
                (void)hipStreamSynchronize(S.stream);
                finish(si, OMG_ERR_DEVICE);
                did = true;
            }
        }
        for (uint32_t si = 0; si < slots.size(); si++) {
            Slot& S = slots[si];
            if (S.busy) continue;
            std::vector<ReqP> batch;
            uint64_t room_now = 0;
            {
                std::lock_guard<std::mutex> lk(m);
                if (pending.empty()) break;
                // room for this slot's arena: the budget less the scenes and the other arenas
                uint64_t others = scenes_bytes;
                for (uint32_t j = 0; j < slots.size(); j++)
                    if (j != si) others += slots[j].arena.cap;
                uint64_t room = cfg.hbm_budget > others ? cfg.hbm_budget - others : 0;
                size_t need = 4096;
                // arrival order: stop at the first request that does not fit
                while (!pending.empty() && batch.size() < cfg.max_batch) {
                    size_t f = pending.front()->footprint;
                    if (need + f > room) break;
                    need += f;
                    batch.push_back(pending.front());
                    pending.pop_front();
                }
                room_now = room;
            }
            if (batch.empty()) {
                // the head request needs memory that idle arenas hold: release them
                // (busy ones are released once their batch has finished)
                for (uint32_t j = 0; j < slots.size(); j++)
                    if (!slots[j].busy && slots[j].arena.cap && j != si) slots[j].arena.release();
                continue;
            }
            int rc = launch(si, batch, room_now);
            if (rc) {
                S.reqs = batch;
                S.busy = true;
                if (S.side) (void)hipStreamSynchronize(S.side);
                (void)hipStreamSynchronize(S.stream);
                finish(si, rc);
            }
            {
                std::lock_guard<std::mutex> lk(m);
                st.batches++;
                batched_reqs += batch.size();
                for (ReqP& r : batch) st.rays += (uint64_t)r->n_src * r->p.n_rays;
                uint64_t c = committed();
                st.committed_bytes = c;
                st.peak_committed_bytes = std::max(st.peak_committed_bytes, c);
                uint32_t busy = 0;
                for (const Slot& s2 : slots) busy += s2.busy ? 1u : 0u;
                st.max_in_flight = std::max(st.max_in_flight, busy);
            }
            did = true;
        }
        std::unique_lock<std::mutex> lk(m);
        bool any_busy = false;
        for (const Slot& s : slots) any_busy |= s.busy;
        if (stop && pending.empty() && !any_busy) break;
        if (!did) {
            if (any_busy) {
                lk.unlock();
                std::this_thread::yield();
            } else {
                cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

extern "C" int omg_server_create(omg_ctx* c, const omg_server_config* cfg, omg_scene* const* scenes,
                                 uint32_t n_scenes, omg_server** out)
{
    if (!c || !cfg || !out || (n_scenes && !scenes)) return fail(OMG_ERR_INVALID, "bad argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 16) return fail(OMG_ERR_INVALID, "slots must be 1..16");
    if (cfg->max_batch < 1 || cfg->max_batch > 4096) return fail(OMG_ERR_INVALID, "max_batch must be 1..4096");
    BatchShape sh{cfg->shape.lanes, cfg->shape.stack_lds, cfg->shape.waves_per_simd, cfg->shape.regen,
                  cfg->shape.width};
    if (!valid_batch_shape(sh)) return fail(OMG_ERR_INVALID, "launch shape");
    bool deep = n_scenes > 0;
    for (uint32_t i = 0; i < n_scenes; i++) {
        if (sh.width == 4 && scenes[i] && !scenes[i]->has4)
            return fail(OMG_ERR_INVALID, "a scene has no 4-wide tree");
        deep = deep && scenes[i] && scenes[i]->has4 && scenes[i]->stack4 <= kDeepStack;
    }
    if (sh.stack_lds == kDeepStack && !deep)
        return fail(OMG_ERR_INVALID, "a scene needs more than 40 stack entries");
    const omg_dome_params& dp = cfg->dome;
    if (dp.n_fan == 0 || dp.n_rays == 0 || dp.n_rays > dp.n_fan || dp.max_events == 0 || dp.max_events > 64)
        return fail(OMG_ERR_INVALID, "dome parameters");
    uint64_t sb = 0;
    for (uint32_t i = 0; i < n_scenes; i++) {
        if (!scenes[i] || scenes[i]->ctx != c) return fail(OMG_ERR_INVALID, "scene of another context");
        sb += scenes[i]->device_bytes;
    }
    if (sb > cfg->hbm_budget) return fail(OMG_ERR_BUDGET, "resident scenes exceed the HBM budget");
    OMG_HIP_CHECK(hipSetDevice(c->device));
    omg_server* s = new omg_server();
    s->ctx = c;
    s->cfg = *cfg;
    s->shape = sh;
    s->deep = deep;
    s->scenes.assign(scenes, scenes + n_scenes);
    s->scenes_bytes = sb;
    s->slots.resize(cfg->slots);
    int rc = OMG_OK;
    // two streams per slot (tracer, dome) while they fit the process's
    // hardware queues (GPU_MAX_HW_QUEUES, 4 by default, two kept for the
    // context's and the caller's streams); otherwise one stream per slot
    uint32_t hw_queues = 4;
    if (const char* q = std::getenv("GPU_MAX_HW_QUEUES")) {
        long v = std::strtol(q, nullptr, 10);
        if (v > 0 && v < 1024) hw_queues = (uint32_t)v;
    }
    const bool side = cfg->slots <= 2 || 2 * cfg->slots + 2 <= hw_queues;
    for (Slot& S : s->slots) {
        if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess ||
            (side && hipStreamCreateWithFlags(&S.side, hipStreamNonBlocking) != hipSuccess) ||
            hipEventCreateWithFlags(&S.ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&S.fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&S.join, hipEventDisableTiming) != hipSuccess) {
            rc = fail(OMG_ERR_DEVICE, "stream/event creation");
            break;
        }
    }
    std::vector<float> fan;
    std::vector<uint32_t> fan_bin;
    make_fan(dp.n_fan, fan, fan_bin);
    if (!rc) rc = upload(s->fan, fan.data(), fan.size() * 4, c->stream);
    if (!rc) rc = upload(s->fan_bin, fan_bin.data(), fan_bin.size() * 4, c->stream);
    if (!rc && s->no_panels.ensure(sizeof(DomePanelDev)) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "alloc");
    // Start-up work a first batch would otherwise pay for (tens of ms:
    // device and pinned allocations, which wait for the whole device, the
    // load of the kernels' code objects, each stream's first hardware
    // queue): every slot gets its arena and staging at their working size
    // (as far as the budget allows) and runs one tiny kernel of each
    // kernel file on each of its streams.
    const uint64_t share = (cfg->hbm_budget - sb) / cfg->slots;
    for (Slot& S : s->slots) {
        if (rc) break;
        if (share >= kMinArenaBytes &&
            S.arena.ensure_exact((size_t)std::min<uint64_t>(share, kPresizeArenaBytes)) != hipSuccess)
            rc = fail(OMG_ERR_DEVICE, "work arena allocation");
        if (!rc) rc = host_grow(S.host_in, S.host_in_dev, S.host_in_cap, kPresizeStagingBytes);
        if (!rc) rc = host_grow(S.host_out, S.host_out_dev, S.host_out_cap, kPresizeStagingBytes);
        for (hipStream_t q : {S.stream, S.side}) {
            if (rc || !q) continue;
            float* scratch = S.arena.cap ? S.arena.as<float>() : s->no_panels.as<float>();
            if (launch_fill_zero(scratch, 1, q) != hipSuccess || launch_dome_prime(scratch, q) != hipSuccess)
                rc = fail(OMG_ERR_DEVICE, "stream start-up");
        }
    }
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(OMG_ERR_DEVICE, "server start-up");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "fan upload");
    // then one 16-ray request per slot through the whole batch path (copies,
    // clear, tracer, dome fans, read-back), results discarded
    for (uint32_t si = 0; !rc && n_scenes && si < cfg->slots; si++) {
        auto r = std::make_shared<Req>();
        r->scene = scenes[0];
        omg_trace_params& p = r->p;
        p.n_rays = 16;
        p.nbands = scenes[0]->nbands;
        p.max_bounces = 8;
        p.dir_band = p.nbands > 1 ? 1 : 0;
        p.seed = 0;
        p.bin_dt = 0.01f;
        p.max_time = 0.05f;
        p.floor_rel = 1e-7f;
        p.radius = 0.5f;
        for (uint32_t b = 0; b < OMG_MAX_BANDS; b++) p.energy[b] = 1.0f;
        r->src.assign(3, 0.0f);
        r->lis.assign(3, 1.0f);
        r->n_src = r->n_lis = 1;
        r->dome = p.nbands >= 2 ? 1u : 0u;
        r->nbands = p.nbands;
        rc = validate_params(&p, p.nbands, 1, 1, 0, &r->nbins);
        if (rc) break;
        r->footprint = s->footprint(*r);
        Slot& S = s->slots[si];
        // within the budget: the room this slot's arena may take
        const uint64_t others = s->committed() - S.arena.cap;
        const uint64_t room = cfg->hbm_budget > others ? cfg->hbm_budget - others : 0;
        if (r->footprint + 4096 > room) continue;
        std::vector<ReqP> one{r};
        rc = s->launch(si, one, room);
        if (!rc && S.side && hipStreamSynchronize(S.side) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "start-up");
        if (!rc && hipStreamSynchronize(S.stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "start-up");
        S.reqs.clear();
        S.busy = false;
    }
    if (rc) {
        for (Slot& S : s->slots) {
            for (hipEvent_t e : {S.ev, S.fork, S.join})
                if (e) (void)hipEventDestroy(e);
            for (hipStream_t q : {S.stream, S.side})
                if (q) (void)hipStreamDestroy(q);
        }
        for (Slot& S : s->slots) {
            S.arena.release();
            for (void* h : {S.host_in, S.host_out})
                if (h) (void)hipHostFree(h);
        }
        s->fan.release();
        s->fan_bin.release();
        s->no_panels.release();
        delete s;
        return rc;
    }
    s->st.committed_bytes = s->st.peak_committed_bytes = s->committed();
    s->t0 = Clock::now();
    c->live_scenes++;
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return OMG_OK;
}

namespace {

int submit_at(omg_server* s, const omg_request* q, double t_arrival, uint64_t* ticket)
{
    if (!s || !q || !ticket) return fail(OMG_ERR_INVALID, "bad argument");
    if (std::find(s->scenes.begin(), s->scenes.end(), q->scene) == s->scenes.end())
        return fail(OMG_ERR_INVALID, "scene is not resident on this server");
    if (!q->sources || !q->listeners) return fail(OMG_ERR_INVALID, "bad argument");
    if (q->flags & ~(OMG_REQ_KEEP | OMG_REQ_RECORDS)) return fail(OMG_ERR_INVALID, "unknown flags");
    uint32_t nbins;
    int rc = validate_params(&q->params, q->scene->nbands, q->n_src, q->n_lis, 0, &nbins);
    if (rc) return rc;
    if (q->dome && q->scene->nbands < 2) return fail(OMG_ERR_INVALID, "dome needs >= 2 bands");
    auto r = std::make_shared<Req>();
    r->scene = q->scene;
    r->p = q->params;
    r->src.assign(q->sources, q->sources + (size_t)q->n_src * 3);
    r->lis.assign(q->listeners, q->listeners + (size_t)q->n_lis * 3);
    r->n_src = q->n_src;
    r->n_lis = q->n_lis;
    r->dome = q->dome ? 1u : 0u;
    r->flags = q->flags;
    r->nbins = nbins;
    r->nbands = q->scene->nbands;
    r->footprint = s->footprint(*r);
    r->t_arrival = t_arrival;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->st.submitted++;
        if (r->footprint + 4096 + s->scenes_bytes > s->cfg.hbm_budget) {
            s->st.refused++;
            return fail(OMG_ERR_BUDGET, "request does not fit the HBM budget beside the resident scenes");
        }
        r->ticket = s->next_ticket++;
        s->reqs[r->ticket] = r;
        s->pending.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return OMG_OK;
}

int collect(omg_server* s, uint64_t ticket, bool block, float* bins, float* dirs,
            omg_ray_record* records, float* dome_e, float* dome_d, omg_request_stats* st)
{
    if (!s) return fail(OMG_ERR_INVALID, "bad argument");
    ReqP r;
    {
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return fail(OMG_ERR_INVALID, "unknown ticket");
        r = it->second;
        if (!r->done) {
            if (!block) return fail(OMG_ERR_NOT_READY, "request still running");
            s->cv_done.wait(lk, [&] { return r->done; });
        }
        s->reqs.erase(ticket);
    }
    if (st) {
        st->queue_s = r->t_launch - r->t_arrival;
        st->latency_s = r->t_done - r->t_arrival;
        st->batch_size = r->batch_size;
        st->slot = r->slot;
    }
    if (r->status != OMG_OK) return fail(r->status, "request failed on the device");
    bool keep = (r->flags & OMG_REQ_KEEP) != 0;
    if ((bins || dirs || records || dome_e || dome_d) && !keep)
        return fail(OMG_ERR_INVALID, "request was submitted without OMG_REQ_KEEP");
    if (bins) std::memcpy(bins, r->bins.data(), r->bins.size() * 4);
    if (dirs) std::memcpy(dirs, r->dirs.data(), r->dirs.size() * 4);
    if (records) {
        if (!(r->flags & OMG_REQ_RECORDS)) return fail(OMG_ERR_INVALID, "request kept no records");
        std::memcpy(records, r->rec.data(), r->rec.size() * sizeof(omg_ray_record));
    }
    if (dome_e || dome_d) {
        if (!r->dome) return fail(OMG_ERR_INVALID, "request sampled no dome");
        if (dome_e) std::memcpy(dome_e, r->dome_e.data(), r->dome_e.size() * 4);
        if (dome_d) std::memcpy(dome_d, r->dome_d.data(), r->dome_d.size() * 4);
    }
    return OMG_OK;
}

}  // namespace

extern "C" int omg_server_submit(omg_server* s, const omg_request* q, uint64_t* ticket)
{
    if (!s) return fail(OMG_ERR_INVALID, "bad argument");
    return submit_at(s, q, s->now(), ticket);
}

extern "C" int omg_server_result(omg_server* s, uint64_t ticket, float* bins, float* dirs,
                                 omg_ray_record* records, float* dome_e, float* dome_d,
                                 omg_request_stats* st)
{
    return collect(s, ticket, false, bins, dirs, records, dome_e, dome_d, st);
}

extern "C" int omg_server_wait(omg_server* s, uint64_t ticket, float* bins, float* dirs,
                               omg_ray_record* records, float* dome_e, float* dome_d,
                               omg_request_stats* st)
{
    return collect(s, ticket, true, bins, dirs, records, dome_e, dome_d, st);
}

extern "C" int omg_server_replay(omg_server* s, const omg_request* reqs, const double* arrival_s,
                                 uint32_t n, double* latency_s, double* queue_s, uint64_t* tickets)
{
    if (!s || (n && (!reqs || !arrival_s || !latency_s))) return fail(OMG_ERR_INVALID, "bad argument");
    for (uint32_t i = 1; i < n; i++)
        if (arrival_s[i] < arrival_s[i - 1]) return fail(OMG_ERR_INVALID, "arrivals must ascend");
    std::vector<uint64_t> tk(n, 0);
    const double base = s->now();
    for (uint32_t i = 0; i < n; i++) {
        double due = base + arrival_s[i];
        // sleep until shortly before the arrival, then spin to it
        for (;;) {
            double left = due - s->now();
            if (left <= 0) break;
            if (left > 300e-6)
                std::this_thread::sleep_for(std::chrono::duration<double>(left - 200e-6));
            else
                std::this_thread::yield();
        }
        if (submit_at(s, &reqs[i], due, &tk[i]) != OMG_OK) tk[i] = 0;
    }
    int rc = OMG_OK;
    for (uint32_t i = 0; i < n; i++) {
        latency_s[i] = -1.0;
        if (queue_s) queue_s[i] = -1.0;
        if (!tk[i]) continue;
        if (reqs[i].flags & OMG_REQ_KEEP) {
            // keep the results for the caller: wait without collecting
            std::unique_lock<std::mutex> lk(s->m);
            ReqP r = s->reqs[tk[i]];
            s->cv_done.wait(lk, [&] { return r->done; });
            latency_s[i] = r->t_done - r->t_arrival;
            if (queue_s) queue_s[i] = r->t_launch - r->t_arrival;
            if (r->status != OMG_OK) rc = r->status;
        } else {
            omg_request_stats st;
            int e = omg_server_wait(s, tk[i], nullptr, nullptr, nullptr, nullptr, nullptr, &st);
            if (e == OMG_OK) {
                latency_s[i] = st.latency_s;
                if (queue_s) queue_s[i] = st.queue_s;
            } else {
                rc = e;
            }
        }
    }
    if (tickets) std::memcpy(tickets, tk.data(), n * sizeof(uint64_t));
    return rc == OMG_OK ? OMG_OK : fail(rc, "a request failed on the device");
}

extern "C" int omg_server_stats_get(omg_server* s, omg_server_stats* out)
{
    if (!s || !out) return fail(OMG_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->st;
    out->mean_batch = s->st.batches ? (double)s->batched_reqs / (double)s->st.batches : 0.0;
    return OMG_OK;
}

extern "C" void omg_server_destroy(omg_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->ctx->device);
    for (Slot& S : s->slots) {
        if (S.stream) (void)hipStreamSynchronize(S.stream);
        if (S.side) (void)hipStreamSynchronize(S.side);
        S.arena.release();
        if (S.host_in) (void)hipHostFree(S.host_in);
        if (S.host_out) (void)hipHostFree(S.host_out);
        for (hipEvent_t e : {S.ev, S.fork, S.join})
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t q : {S.stream, S.side})
            if (q) (void)hipStreamDestroy(q);
    }
    s->fan.release();
    s->fan_bin.release();
    s->no_panels.release();
    omg_ctx* c = s->ctx;
    delete s;
    ctx_release(c);
}
