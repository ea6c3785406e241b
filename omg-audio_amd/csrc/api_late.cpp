// C ABI of the late-field state: TraceGate on the host, the slots'
// averages on the device (late_kernels.hip), estimate_reverb finished on
// the host from the device reductions.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "late_kernels.h"
#include "omg_numerics.h"
#include "runtime.h"

using namespace omg;

struct omg_late_field {
    omg_ctx* ctx = nullptr;
    uint32_t n_slots = 0, nbins = 0, nbands = 0;
    float bin_dt = 0.0f;
    double r5 = 0.0, r25 = 0.0;
    uint32_t late_start = 0, dir_start = 0;
    DevBuf avg, jobs, red, cur, pts;  // averages; update tables; traced echograms and positions
    void* h_jobs = nullptr;  // pinned staging
    void* h_red = nullptr;
    void* h_pts = nullptr;
    size_t h_jobs_cap = 0, h_red_cap = 0, h_pts_cap = 0;
    std::vector<omg_trace_gate> gate;
    std::vector<uint32_t> version;
    std::vector<float> rp_dist;
    hipStream_t stream = nullptr;
};

namespace {

float vlen(const float a[3], const float b[3])
{
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return sqrtf(x * x + y * y + z * z);
}

// 10 log10(max(r, 1e-12)) <= db  (estimate_reverb, tracer.rs:222-231)
bool at_or_below(double r, double db)
{
    return 10.0 * std::log10(r > 1e-12 ? r : 1e-12) <= db;
}

// The largest ratio in [1e-12, 1] whose level is at or below db: bisection
// over the ordered bit patterns of positive doubles, with the host's own
// log10 - the one the CPU reference and the finishing step use.  A device
// test `r <= threshold` is then the host's test.
double ratio_threshold(double db)
{
    uint64_t lo, hi;
    double a = 1e-12, b = 1.0;
    std::memcpy(&lo, &a, 8);
    std::memcpy(&hi, &b, 8);
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        double m;
        std::memcpy(&m, &mid, 8);
        if (at_or_below(m, db)) lo = mid;
        else hi = mid;
    }
    double r;
    std::memcpy(&r, &lo, 8);
    return r;
}

int grow_pinned(void*& p, size_t& cap, size_t want)
{
    if (want <= cap) return OMG_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return fail(OMG_ERR_DEVICE, "hipHostMalloc");
    }
    cap = want;
    return OMG_OK;
}

}  // namespace

extern "C" void omg_trace_gate_init(omg_trace_gate* g, uint32_t phase)
{
    if (!g) return;
    g->last_src[0] = FLT_MAX;
    g->last_src[1] = 0.0f;
    g->last_src[2] = 0.0f;
    g->last_lis[0] = FLT_MAX;
    g->last_lis[1] = 0.0f;
    g->last_lis[2] = 0.0f;
    for (float& e : g->last_energy) e = 0.0f;
    g->age = phase;
}

extern "C" int omg_trace_gate_step(omg_trace_gate* g, const float src[3], const float lis[3],
                                   const float* energy, uint32_t nbands, uint32_t max_age,
                                   int* fire)
{
    if (!g || !src || !lis || !energy || !fire || nbands == 0 || nbands > OMG_MAX_BANDS)
        return fail(OMG_ERR_INVALID, "bad argument");
    g->age += 1;
    bool moved = false;
    for (uint32_t b = 0; b < nbands; b++) {
        float e0 = g->last_energy[b], e1 = energy[b];
        if (fabsf(e1 - e0) > 0.1f * fmaxf(fmaxf(e0, e1), 1e-6f)) moved = true;
    }
    *fire = 0;
    if (g->age >= max_age || vlen(src, g->last_src) > 0.25f || vlen(lis, g->last_lis) > 0.25f ||
        moved) {
        for (int k = 0; k < 3; k++) {
            g->last_src[k] = src[k];
            g->last_lis[k] = lis[k];
        }
        for (uint32_t b = 0; b < nbands; b++) g->last_energy[b] = energy[b];
        g->age = 0;
        *fire = 1;
    }
    return OMG_OK;
}

extern "C" int omg_late_field_create(omg_ctx* c, uint32_t n_slots, uint32_t nbins, uint32_t nbands,
                                     float bin_dt, omg_late_field** out)
{
    if (!c || !out || n_slots == 0 || nbins == 0 || nbands < 2 || nbands > OMG_MAX_BANDS ||
        !(bin_dt > 0.0f))
        return fail(OMG_ERR_INVALID, "bad argument");
    *out = nullptr;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    omg_late_field* f = new omg_late_field();
    f->ctx = c;
    f->n_slots = n_slots;
    f->nbins = nbins;
    f->nbands = nbands;
    f->bin_dt = bin_dt;
    f->r5 = ratio_threshold(-5.0);
    f->r25 = ratio_threshold(-25.0);
    // (0.080 / BIN_DT) as usize, clamped to the echogram (tracer.rs:247, :63-66)
    f->late_start = std::min(sat_u32(0.080f / bin_dt), nbins);
    f->dir_start = std::min(sat_u32(0.08f / bin_dt), nbins);
    f->gate.resize(n_slots);
    for (uint32_t s = 0; s < n_slots; s++) omg_trace_gate_init(&f->gate[s], s % 8u);
    f->version.assign(n_slots, 0u);
    f->rp_dist.assign(n_slots, FLT_MAX);
    const size_t per = (size_t)nbins * (nbands + 3) * sizeof(float);
    hipError_t e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = f->avg.ensure_exact(per * n_slots);
    if (e == hipSuccess) e = hipMemsetAsync(f->avg.ptr, 0, per * n_slots, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) {
        f->avg.release();
        if (f->stream) (void)hipStreamDestroy(f->stream);
        delete f;
        return fail(OMG_ERR_DEVICE, std::string("late field: ") + hipGetErrorString(e));
    }
    c->live_scenes++;
    *out = f;
    return OMG_OK;
}

extern "C" int omg_late_field_gate(omg_late_field* f, uint32_t slot, const float src[3],
                                   const float lis[3], const float* energy, uint32_t max_age,
                                   int* fire)
{
    if (!f || slot >= f->n_slots) return fail(OMG_ERR_INVALID, "bad argument");
    return omg_trace_gate_step(&f->gate[slot], src, lis, energy, f->nbands, max_age, fire);
}

extern "C" int omg_late_field_update(omg_late_field* f, const omg_late_update* ups, uint32_t n,
                                     omg_late_out* out, void* stream)
{
    if (!f || (n && (!ups || !out))) return fail(OMG_ERR_INVALID, "bad argument");
    if (n == 0) return OMG_OK;
    std::vector<uint8_t> seen(f->n_slots, 0);
    for (uint32_t i = 0; i < n; i++) {
        const omg_late_update& u = ups[i];
        if (u.slot >= f->n_slots) return fail(OMG_ERR_INVALID, "update " + std::to_string(i) + ": slot");
        if (seen[u.slot]) return fail(OMG_ERR_INVALID, "update " + std::to_string(i) + ": slot repeated");
        seen[u.slot] = 1;
        if (u.traced > 1 || (u.traced && (!u.cur_bins || !u.cur_dirs)))
            return fail(OMG_ERR_INVALID, "update " + std::to_string(i) + ": traced echogram");
    }
    OMG_HIP_CHECK(hipSetDevice(f->ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : f->stream;
    const size_t per_f = (size_t)f->nbins * (f->nbands + 3);
    const size_t jb = (size_t)n * sizeof(LateJob), rb = (size_t)n * sizeof(LateRed);
    int rc = grow_pinned(f->h_jobs, f->h_jobs_cap, jb);
    if (!rc) rc = grow_pinned(f->h_red, f->h_red_cap, rb);
    if (rc) return rc;
    LateJob* J = reinterpret_cast<LateJob*>(f->h_jobs);
    float* base = f->avg.as<float>();
    for (uint32_t i = 0; i < n; i++) {
        const omg_late_update& u = ups[i];
        float* slot = base + (size_t)u.slot * per_f;
        J[i].avg_bins = slot;
        J[i].avg_dirs = slot + (size_t)f->nbins * f->nbands;
        J[i].cur_bins = u.traced ? u.cur_bins : nullptr;
        J[i].cur_dirs = u.traced ? u.cur_dirs : nullptr;
        J[i].alpha = f->version[u.slot] == 0 ? 1.0f : 0.3f;
        J[i].traced = u.traced;
    }
    OMG_HIP_CHECK(f->jobs.ensure(jb));
    OMG_HIP_CHECK(f->red.ensure(rb));
    OMG_HIP_CHECK(hipMemcpyAsync(f->jobs.ptr, J, jb, hipMemcpyHostToDevice, s));
    LateArgs A;
    A.jobs = f->jobs.as<LateJob>();
    A.n = n;
    A.nbins = f->nbins;
    A.nbands = f->nbands;
    A.late_start = f->late_start;
    A.dir_start = f->dir_start;
    A.r5 = f->r5;
    A.r25 = f->r25;
    A.out = f->red.as<LateRed>();
    OMG_HIP_CHECK(launch_late_ema(A, s));
    OMG_HIP_CHECK(launch_late_reduce(A, s));
    OMG_HIP_CHECK(hipMemcpyAsync(f->h_red, f->red.ptr, rb, hipMemcpyDeviceToHost, s));
    OMG_HIP_CHECK(hipStreamSynchronize(s));
    const LateRed* R = reinterpret_cast<const LateRed*>(f->h_red);
    for (uint32_t i = 0; i < n; i++) {
        const omg_late_update& u = ups[i];
        const LateRed& r = R[i];
        omg_late_out& o = out[i];
        std::memset(&o, 0, sizeof o);
        // estimate_reverb's RT60 from the -5 / -25 dB bins (tracer.rs:216-240)
        for (uint32_t b = 0; b < f->nbands; b++) {
            o.rt60[b] = 0.5f;
            if (!(r.total[b] > 0.0)) continue;
            if (r.t5[b] >= 0 && r.t25[b] >= 0 && r.t25[b] > r.t5[b]) {
                double r5 = r.edc5[b] / r.total[b], r25 = r.edc25[b] / r.total[b];
                double db5 = 10.0 * std::log10(r5 > 1e-12 ? r5 : 1e-12);
                double db25 = 10.0 * std::log10(r25 > 1e-12 ? r25 : 1e-12);
                float dt = (float)(r.t25[b] - r.t5[b]) * f->bin_dt;
                float ddb = (float)(db25 - db5);
                o.rt60[b] = fminf(fmaxf(-60.0f * dt / ddb, 0.1f), 5.0f);
            }
        }
        for (uint32_t b = 0; b < f->nbands; b++) o.level[b] = r.level[b];
        for (int k = 0; k < 3; k++) o.dir[k] = r.dir[k];
        o.aniso = r.aniso;
        // Sim::reverb_world bookkeeping and staleness compensation (sim.rs:374-395)
        if (u.traced) f->rp_dist[u.slot] = vlen(u.src, u.receiver);
        f->version[u.slot] += 1;
        if (f->rp_dist[u.slot] < FLT_MAX) {
            float d_now = vlen(u.src, u.receiver);
            float k = fminf((f->rp_dist[u.slot] + 1.0f) / (d_now + 1.0f), 1.0f);
            for (uint32_t b = 0; b < f->nbands; b++) o.level[b] *= k;
        }
        o.version = f->version[u.slot];
        o.traced = u.traced;
    }
    return OMG_OK;
}

extern "C" int omg_late_field_trace_world(omg_late_field* f, omg_scene* scene,
                                          const omg_late_trace* jobs, uint32_t n,
                                          const omg_overlay* overlays, uint32_t n_ovl,
                                          omg_late_out* out)
{
    if (!f || !scene || (n && (!jobs || !out)) || (n_ovl && !overlays))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (scene->ctx != f->ctx) return fail(OMG_ERR_INVALID, "scene of another context");
    if (n == 0) return OMG_OK;
    std::vector<uint8_t> seen(f->n_slots, 0);
    for (uint32_t i = 0; i < n; i++) {
        const omg_late_trace& j = jobs[i];
        if (j.slot >= f->n_slots || seen[j.slot])
            return fail(OMG_ERR_INVALID, "job " + std::to_string(i) + ": slot");
        seen[j.slot] = 1;
        uint32_t nb = 0;
        int rc = validate_params(&j.params, scene->nbands, 1, 1, n_ovl, &nb);
        if (rc) return fail(rc, "job " + std::to_string(i) + ": " + omg_last_error());
        if (j.params.nbands != f->nbands || nb != f->nbins)
            return fail(OMG_ERR_INVALID, "job " + std::to_string(i) + ": echogram shape of the field");
    }
    OMG_HIP_CHECK(hipSetDevice(f->ctx->device));
    const size_t per_f = (size_t)f->nbins * (f->nbands + 3);
    int rc = grow_pinned(f->h_pts, f->h_pts_cap, (size_t)n * 6 * sizeof(float));
    if (rc) return rc;
    float* P = reinterpret_cast<float*>(f->h_pts);
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            P[6 * i + k] = jobs[i].src[k];
            P[6 * i + 3 + k] = jobs[i].receiver[k];
        }
    OMG_HIP_CHECK(f->pts.ensure((size_t)n * 6 * sizeof(float)));
    OMG_HIP_CHECK(f->cur.ensure((size_t)n * per_f * sizeof(float)));
    OMG_HIP_CHECK(hipMemcpyAsync(f->pts.ptr, P, (size_t)n * 6 * sizeof(float),
                                 hipMemcpyHostToDevice, f->stream));
    std::vector<omg_late_update> ups(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* dp = f->pts.as<float>() + 6 * i;
        float* cb = f->cur.as<float>() + i * per_f;
        float* cd = cb + (size_t)f->nbins * f->nbands;
        rc = omg_trace_world_async(scene, &jobs[i].params, dp, 1, dp + 3, 1, overlays, n_ovl, cb,
                                   cd, nullptr, f->stream);
        if (rc) return rc;
        omg_late_update& u = ups[i];
        u.slot = jobs[i].slot;
        u.traced = 1;
        u.cur_bins = cb;
        u.cur_dirs = cd;
        for (int k = 0; k < 3; k++) {
            u.src[k] = jobs[i].src[k];
            u.receiver[k] = jobs[i].receiver[k];
        }
    }
    return omg_late_field_update(f, ups.data(), n, out, f->stream);
}

extern "C" int omg_late_field_read(omg_late_field* f, uint32_t slot, float* bins, float* dirs)
{
    if (!f || slot >= f->n_slots || !bins || !dirs) return fail(OMG_ERR_INVALID, "bad argument");
    OMG_HIP_CHECK(hipSetDevice(f->ctx->device));
    const size_t per_f = (size_t)f->nbins * (f->nbands + 3);
    const float* p = f->avg.as<float>() + (size_t)slot * per_f;
    OMG_HIP_CHECK(hipMemcpyAsync(bins, p, (size_t)f->nbins * f->nbands * sizeof(float),
                                 hipMemcpyDeviceToHost, f->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(dirs, p + (size_t)f->nbins * f->nbands,
                                 (size_t)f->nbins * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                 f->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(f->stream));
    return OMG_OK;
}

extern "C" void omg_late_field_destroy(omg_late_field* f)
{
    if (!f) return;
    omg_ctx* c = f->ctx;
    (void)hipSetDevice(c->device);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    for (DevBuf* d : {&f->avg, &f->jobs, &f->red, &f->cur, &f->pts}) d->release();
    for (void* h : {f->h_jobs, f->h_red, f->h_pts})
        if (h) (void)hipHostFree(h);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    ctx_release(c);
}
