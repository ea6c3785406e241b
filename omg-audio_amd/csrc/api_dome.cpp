// C ABI: the ambient dome (crates/omg-scene/src/dome.rs:293-478).
#include <cmath>
#include <cstring>

#include "dome_kernels.h"
#include "runtime.h"

#pragma clang fp contract(off)

using namespace omg;

struct omg_dome_probe {
    omg_scene* scene = nullptr;
    omg_dome_params params{};
    float ema[kDomeBins][OMG_MAX_BANDS] = {};
    bool primed = false;
};

// DomeProbe::new fan (dome.rs:316-324) with the project sin/cos, and the
// sector of each direction (bin_of, dome.rs:304-310).
void omg::make_fan(uint32_t n_fan, std::vector<float>& fan, std::vector<uint32_t>& bins)
{
    fan.resize((size_t)n_fan * 3);
    bins.resize(n_fan);
    float ga = kPi * (3.0f - sqrtf(5.0f));
    for (uint32_t i = 0; i < n_fan; i++) {
        float z = 1.0f - 2.0f * ((float)i + 0.5f) / (float)n_fan;
        float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
        float s, c;
        sincos_cw(ga * (float)i, &s, &c);
        fan[3 * i] = r * c;
        fan[3 * i + 1] = r * s;
        fan[3 * i + 2] = z;
        if (z > 0.87f) {
            bins[i] = 8;
        } else {
            float az = atan2f(r * s, r * c) + kPi;
            uint32_t k = sat_u32(az / (kTau / 8.0f));
            bins[i] = k < 7 ? k : 7;
        }
    }
}

namespace {

int validate(const omg_scene* s, const omg_dome_params* p, uint32_t n_eyes, uint32_t n_panels)
{
    if (!s || !p) return fail(OMG_ERR_INVALID, "bad argument");
    if (p->nbands != s->nbands || p->nbands < 2) return fail(OMG_ERR_INVALID, "nbands");
    if (p->n_fan == 0 || p->n_rays == 0 || p->n_rays > p->n_fan) return fail(OMG_ERR_INVALID, "n_rays");
    if (p->max_events == 0 || p->max_events > 64) return fail(OMG_ERR_INVALID, "max_events");
    if (n_panels > OMG_MAX_OVERLAYS) return fail(OMG_ERR_CAP, "too many dome panels");
    if ((uint64_t)n_eyes * p->n_rays >= (1ull << 31)) return fail(OMG_ERR_CAP, "too many rays");
    return OMG_OK;
}

}  // namespace

extern "C" int omg_dome_sample(omg_scene* s, const omg_dome_params* p, const float* eyes,
                               uint32_t n_eyes, const omg_dome_panel* panels, uint32_t n_panels,
                               float* energy, float* dsum, omg_dome_ray* rays,
                               omg_counters* counters)
{
    int rc = validate(s, p, n_eyes, n_panels);
    if (rc) return rc;
    if (!eyes || !energy || !dsum || (n_panels && !panels)) return fail(OMG_ERR_INVALID, "bad argument");
    if (n_eyes == 0) return OMG_OK;
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    // the panels go to job_ovl, which an in-flight trace may read as overlays
    rc = quiesce(c);
    if (rc) return rc;
    c->tables_valid = false;
    std::vector<float> fan;
    std::vector<uint32_t> fan_bin;
    make_fan(p->n_fan, fan, fan_bin);
    static_assert(sizeof(omg_dome_panel) == sizeof(DomePanelDev), "panel layout");
    static_assert(sizeof(omg_dome_ray) == sizeof(DomeRayDev), "ray layout");
    static_assert(sizeof(omg_material) == sizeof(omg_material_dev), "material layout");
    size_t nr = (size_t)n_eyes * p->n_rays;
    size_t ne = (size_t)n_eyes * kDomeBins * p->nbands, nd = (size_t)n_eyes * kDomeBins * 3;
    rc = upload(c->aux0, fan.data(), fan.size() * sizeof(float), c->stream);
    if (!rc) rc = upload(c->aux1, fan_bin.data(), fan_bin.size() * sizeof(uint32_t), c->stream);
    if (!rc) rc = upload(c->job_ovl, panels, n_panels * sizeof(omg_dome_panel), c->stream);
    if (!rc) rc = upload(c->job_lis, eyes, (size_t)n_eyes * 3 * sizeof(float), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(c->aux2.ensure(nr * sizeof(DomeRayDev)));
    OMG_HIP_CHECK(c->job_bins.ensure(ne * sizeof(float)));
    OMG_HIP_CHECK(c->job_dirs.ensure(nd * sizeof(float)));
    OMG_HIP_CHECK(c->job_cnt.ensure(4 * sizeof(unsigned long long)));
    DomeArgs A{};
    fill_geom(s, A.g);
    A.mats = s->mats_raw.as<omg_material_dev>();
    A.panels = c->job_ovl.as<DomePanelDev>();
    A.n_panels = n_panels;
    A.eyes = c->job_lis.as<float>();
    A.n_eyes = n_eyes;
    A.fan = c->aux0.as<float>();
    A.fan_bin = c->aux1.as<uint32_t>();
    A.n_fan = p->n_fan;
    A.n_rays = p->n_rays;
    A.max_events = p->max_events;
    A.nbands = p->nbands;
    A.horizon = p->horizon;
    A.through_len = p->through_len;
    A.rays = c->aux2.as<DomeRayDev>();
    A.energy = c->job_bins.as<float>();
    A.dsum = c->job_dirs.as<float>();
    unsigned long long* cnt = c->job_cnt.as<unsigned long long>();
    A.counters = counters ? cnt : nullptr;
    A.status = reinterpret_cast<uint32_t*>(cnt + 3);
    OMG_HIP_CHECK(launch_fill_zero_u64(cnt, 4, c->stream));
    OMG_HIP_CHECK(launch_dome(A, counters != nullptr, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(energy, A.energy, ne * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(dsum, A.dsum, nd * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (rays)
        OMG_HIP_CHECK(hipMemcpyAsync(rays, A.rays, nr * sizeof(DomeRayDev), hipMemcpyDeviceToHost, c->stream));
    unsigned long long h[4];
    OMG_HIP_CHECK(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (counters) {
        counters->aabb_tests = h[0];
        counters->tri_tests = h[1];
        counters->bounce_rays = 0;
    }
    if (h[3] & 1ull) return fail(OMG_ERR_OVERFLOW, "a straight-through ray crossed too many surfaces");
    return OMG_OK;
}

extern "C" int omg_dome_sample_async(omg_scene* s, const omg_dome_params* p, const float* d_eyes,
                                     uint32_t n_eyes, const omg_dome_panel* panels,
                                     uint32_t n_panels, float* d_energy, float* d_dsum,
                                     omg_dome_ray* d_rays, void* stream)
{
    int rc = validate(s, p, n_eyes, n_panels);
    if (rc) return rc;
    if (!d_eyes || !d_energy || !d_dsum || !d_rays || (n_panels && !panels))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (n_eyes == 0) return OMG_OK;
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    if (c->fan_n != p->n_fan || !c->dome_status.ptr) {
        rc = quiesce(c);   // an earlier dome launch may still read the fan
        if (rc) return rc;
        std::vector<float> fan;
        std::vector<uint32_t> fan_bin;
        make_fan(p->n_fan, fan, fan_bin);
        rc = upload(c->fan_dirs, fan.data(), fan.size() * sizeof(float), c->stream);
        if (!rc) rc = upload(c->fan_bins, fan_bin.data(), fan_bin.size() * sizeof(uint32_t), c->stream);
        if (rc) return rc;
        if (!c->dome_status.ptr) {
            OMG_HIP_CHECK(c->dome_status.ensure(sizeof(uint32_t)));
            OMG_HIP_CHECK(launch_fill_zero((float*)c->dome_status.ptr, 1, c->stream));
        }
        OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
        c->fan_n = p->n_fan;
    }
    if (n_panels) {
        rc = quiesce(c);   // an earlier dome launch may still read the panels
        if (!rc) rc = upload(c->dome_panels, panels, n_panels * sizeof(omg_dome_panel), c->stream);
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "panels");
        if (rc) return rc;
    } else {
        OMG_HIP_CHECK(c->dome_panels.ensure(sizeof(omg_dome_panel)));
    }
    DomeArgs A{};
    fill_geom(s, A.g);
    A.mats = s->mats_raw.as<omg_material_dev>();
    A.panels = c->dome_panels.as<DomePanelDev>();
    A.n_panels = n_panels;
    A.eyes = d_eyes;
    A.n_eyes = n_eyes;
    A.fan = c->fan_dirs.as<float>();
    A.fan_bin = c->fan_bins.as<uint32_t>();
    A.n_fan = p->n_fan;
    A.n_rays = p->n_rays;
    A.max_events = p->max_events;
    A.nbands = p->nbands;
    A.horizon = p->horizon;
    A.through_len = p->through_len;
    A.rays = reinterpret_cast<DomeRayDev*>(d_rays);
    A.energy = d_energy;
    A.dsum = d_dsum;
    A.counters = nullptr;
    A.status = c->dome_status.as<uint32_t>();
    c->async_inflight = true;
    OMG_HIP_CHECK(launch_dome(A, false, st));
    return OMG_OK;
}

extern "C" int omg_dome_status(omg_scene* s)
{
    if (!s) return fail(OMG_ERR_INVALID, "bad argument");
    omg_ctx* c = s->ctx;
    if (!c->dome_status.ptr) return OMG_OK;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    OMG_HIP_CHECK(hipDeviceSynchronize());
    uint32_t v = 0;
    OMG_HIP_CHECK(hipMemcpy(&v, c->dome_status.ptr, sizeof v, hipMemcpyDeviceToHost));
    OMG_HIP_CHECK(launch_fill_zero((float*)c->dome_status.ptr, 1, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (v & 1u) return fail(OMG_ERR_OVERFLOW, "a straight-through ray crossed too many surfaces");
    return OMG_OK;
}

extern "C" int omg_dome_probe_create(omg_scene* s, const omg_dome_params* p, omg_dome_probe** out)
{
    if (!out) return fail(OMG_ERR_INVALID, "out is null");
    int rc = validate(s, p, 1, 0);
    if (rc) return rc;
    omg_dome_probe* pr = new omg_dome_probe();
    pr->scene = s;
    pr->params = *p;
    *out = pr;
    return OMG_OK;
}

// DomeProbe::sample (dome.rs:425-478): one GPU pass, then the EMA and the
// bin directions on the host.
extern "C" int omg_dome_probe_sample(omg_dome_probe* pr, const float eye[3],
                                     const omg_dome_panel* panels, uint32_t n_panels, float* out)
{
    if (!pr || !eye || !out) return fail(OMG_ERR_INVALID, "bad argument");
    const uint32_t nb = pr->params.nbands;
    float e[kDomeBins * OMG_MAX_BANDS], ds[kDomeBins * 3];
    int rc = omg_dome_sample(pr->scene, &pr->params, eye, 1, panels, n_panels, e, ds, nullptr, nullptr);
    if (rc) return rc;
    float alpha = pr->primed ? 0.35f : 1.0f;
    pr->primed = true;
    const uint32_t row = 3 + nb;
    for (uint32_t k = 0; k < kDomeBins; k++) {
        for (uint32_t b = 0; b < nb; b++) {
            pr->ema[k][b] += alpha * (e[k * nb + b] - pr->ema[k][b]);
            out[k * row + 3 + b] = pr->ema[k][b];
        }
        f3 d = mk(ds[3 * k], ds[3 * k + 1], ds[3 * k + 2]);
        if (length(d) > 1e-6f) {
            d = normalize(d);
        } else if (k == 8) {
            d = mk(0.0f, 0.0f, 1.0f);
        } else {
            float az = ((float)k + 0.5f) * kTau / 8.0f - kPi;
            d = mk(cosf(az), sinf(az), 0.0f);
        }
        out[k * row] = d.x;
        out[k * row + 1] = d.y;
        out[k * row + 2] = d.z;
    }
    return OMG_OK;
}

extern "C" void omg_dome_probe_destroy(omg_dome_probe* pr) { delete pr; }
