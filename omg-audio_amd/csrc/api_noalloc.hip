// C ABI: the device-resident plain launch for callers that cannot allocate
// (code under HIP graph capture, tight serving loops): the same kernels as
// omg_trace_world_async, with every byte of scratch the launch needs -- the
// job's material and overlay tables, the hand-off queue of the 16-lane walk
// and its two counters -- in one file-scope device array instead of the
// context's grow-on-demand buffers.
#include <cstring>
#include <mutex>
#include <vector>

#include "runtime.h"

using namespace omg;

namespace {

// Limits of one call.
constexpr uint32_t kNaMaxRays = 1u << 18;   // rays per call (rays per source x sources)
constexpr uint32_t kNaMaxMats = 4096;       // scene materials + overlays
constexpr int kNaMaxDevices = 64;

// Layout of `scratch`: [hand-off counters | materials | overlays | hand-off queue].
constexpr size_t kNaCtl = 0;
constexpr size_t kNaMats = ((size_t)kHandoffCtlWords * sizeof(uint32_t) + 255) / 256 * 256;
constexpr size_t kNaOvl = kNaMats + (size_t)kNaMaxMats * sizeof(MatParams);
constexpr size_t kNaQueue = kNaOvl + (size_t)OMG_MAX_OVERLAYS * sizeof(Overlay);
constexpr size_t kNaBytes = kNaQueue + (size_t)kNaMaxRays * kHandoffF4 * sizeof(float4);
static_assert(kNaOvl % 256 == 0 && kNaQueue % 256 == 0, "scratch regions are 256-byte aligned");

// Every GPU the library runs on holds its own copy (about 24 MiB) until the
// process ends.  Host code never names it in a launch or a copy: the
// device address of this GPU's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Per device: the address of its `scratch`, and the host image of the
// tables last copied there (sized once, at set-up).
struct DeviceScratch {
    unsigned char* ptr = nullptr;
    std::vector<unsigned char> staged, next;
    uint32_t staged_mats = 0, staged_ovl = 0;
    bool valid = false;
};
DeviceScratch g_dev[kNaMaxDevices];
std::mutex g_mu;

}  // namespace

// Called by omg_init with `device` current: resolves this device's copy of
// `scratch` once per device.  A device index past the table is left
// without one (the entry point then refuses it).
int omg::noalloc_setup(int device)
{
    if (device < 0 || device >= kNaMaxDevices) return OMG_OK;
    std::lock_guard<std::mutex> g(g_mu);
    DeviceScratch& d = g_dev[device];
    if (d.ptr) return OMG_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(OMG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d.staged.assign(kNaQueue - kNaMats, 0);
    d.next.assign(kNaQueue - kNaMats, 0);
    d.ptr = static_cast<unsigned char*>(p);
    return OMG_OK;
}

extern "C" int omg_trace_world_async_noalloc(omg_scene* s, const omg_trace_params* p,
                                             const float* d_sources, uint32_t n_src,
                                             const float* d_listeners, uint32_t n_lis,
                                             const omg_overlay* overlays, uint32_t n_ovl,
                                             float* d_bins, float* d_dirs,
                                             omg_ray_record* d_records, void* stream)
{
    if (!s || !d_sources || !d_listeners || !d_bins || !d_dirs || (n_ovl && !overlays))
        return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t nbins;
    int rc = validate_params(p, s->nbands, n_src, n_lis, n_ovl, &nbins);
    if (rc) return rc;
    if ((uint64_t)p->n_rays * n_src > kNaMaxRays)
        return fail(OMG_ERR_CAP, "omg_trace_world_async_noalloc: more than 262144 rays "
                                 "(rays per source x sources) in one call");
    const uint32_t n_base = (uint32_t)s->mat_params.size();
    if ((uint64_t)n_base + n_ovl > kNaMaxMats)
        return fail(OMG_ERR_CAP, "omg_trace_world_async_noalloc: more than 4096 materials "
                                 "(scene materials + overlays)");
    omg_ctx* c = s->ctx;
    if (c->device < 0 || c->device >= kNaMaxDevices || !g_dev[c->device].ptr)
        return fail(OMG_ERR_DEVICE, "omg_trace_world_async_noalloc: no scratch on this device");
    OMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    DeviceScratch& d = g_dev[c->device];
    {
        // the job's tables (scene materials, one per overlay, the overlay
        // boxes), copied only when they differ from the last call's
        std::lock_guard<std::mutex> g(g_mu);
        MatParams* mats = reinterpret_cast<MatParams*>(d.next.data());
        Overlay* ov = reinterpret_cast<Overlay*>(d.next.data() + (kNaOvl - kNaMats));
        if (n_base) std::memcpy(mats, s->mat_params.data(), n_base * sizeof(MatParams));
        for (uint32_t i = 0; i < n_ovl; i++) {
            mat_params(overlays[i].mat.absorption, overlays[i].mat.scattering,
                       overlays[i].mat.transmission, s->nbands, mats[n_base + i]);
            Overlay o{};
            for (int k = 0; k < 3; k++) {
                o.mn[k] = overlays[i].min[k];
                o.mx[k] = overlays[i].max[k];
            }
            o.mat = n_base + i;
            ov[i] = o;
        }
        const size_t mat_bytes = (size_t)(n_base + n_ovl) * sizeof(MatParams);
        const size_t ovl_bytes = (size_t)n_ovl * sizeof(Overlay);
        const bool same = d.valid && d.staged_mats == n_base + n_ovl && d.staged_ovl == n_ovl &&
                          std::memcmp(d.staged.data(), d.next.data(), mat_bytes) == 0 &&
                          std::memcmp(d.staged.data() + (kNaOvl - kNaMats),
                                      d.next.data() + (kNaOvl - kNaMats), ovl_bytes) == 0;
        if (!same) {
            d.valid = false;
            // launches on this stream that read the old tables finish first;
            // the copies have landed before the host image is rewritten
            OMG_HIP_CHECK(hipStreamSynchronize(st));
            if (mat_bytes)
                OMG_HIP_CHECK(hipMemcpyAsync(d.ptr + kNaMats, d.next.data(), mat_bytes,
                                             hipMemcpyHostToDevice, st));
            if (ovl_bytes)
                OMG_HIP_CHECK(hipMemcpyAsync(d.ptr + kNaOvl, d.next.data() + (kNaOvl - kNaMats),
                                             ovl_bytes, hipMemcpyHostToDevice, st));
            OMG_HIP_CHECK(hipStreamSynchronize(st));
            d.staged.swap(d.next);
            d.staged_mats = n_base + n_ovl;
            d.staged_ovl = n_ovl;
            d.valid = true;
        }
    }
    TraceArgs A{};
    fill_geom(s, A.g);
    plain_shape(c, s, p, n_src, false, A);
    A.mats = reinterpret_cast<MatParams*>(d.ptr + kNaMats);
    A.overlays = reinterpret_cast<Overlay*>(d.ptr + kNaOvl);
    job_scalars(p, s->nbands, n_ovl, A);
    A.sources = d_sources;
    A.listeners = d_listeners;
    A.n_src = n_src;
    A.n_lis = n_lis;
    A.nbins = nbins;
    A.bins = d_bins;
    A.dirs = d_dirs;
    A.records = reinterpret_cast<uint32_t*>(d_records);
    A.counters = nullptr;
    A.ray_base = 0;
    A.ray_count = p->n_rays * n_src;
    if (A.handoff_bounce) {
        A.handoff = reinterpret_cast<float4*>(d.ptr + kNaQueue);
        A.handoff_ctl = reinterpret_cast<uint32_t*>(d.ptr + kNaCtl);
        OMG_HIP_CHECK(launch_fill_zero(reinterpret_cast<float*>(A.handoff_ctl), kHandoffCtlWords, st));
    }
    const size_t nb_f = (size_t)n_src * n_lis * nbins * s->nbands;
    const size_t nd_f = (size_t)n_src * n_lis * nbins * 3;
    OMG_HIP_CHECK(launch_fill_zero(d_bins, nb_f, st));
    OMG_HIP_CHECK(launch_fill_zero(d_dirs, nd_f, st));
    c->async_inflight = true;
    OMG_HIP_CHECK(launch_trace(A, true, false, st));
    return OMG_OK;
}
