// Flat wire formats of the reference's web bridge and ParamBlock: decode
// the flat jobs, run them through the tracer entry points, encode the
// echograms and parameter blocks in the reference's layouts.
#include <cmath>
#include <cstring>

#include "runtime.h"

using namespace omg;

namespace {

// JavaScript's ToUint32, which web/gpu.js applies when it stores a job's
// f32 words into a Uint32Array (n_rays, seed): truncate, then modulo 2^32.
uint32_t to_uint32(float f)
{
    if (!std::isfinite(f)) return 0u;
    double t = std::trunc((double)f);
    double m = std::fmod(t, 4294967296.0);
    if (m < 0.0) m += 4294967296.0;
    return (uint32_t)m;
}

// Rust's saturating `f32 as u32` (the wasm side's casts).
uint32_t as_u32(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

uint64_t as_u64(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)f;
}

// The reference tracer's constants (tracer.rs:12-16, :103, :163)
omg_trace_params wire_params(uint32_t n_rays, uint32_t seed, const float energy[3])
{
    omg_trace_params p;
    std::memset(&p, 0, sizeof p);
    p.n_rays = n_rays;
    p.nbands = 3;
    p.max_bounces = 64;
    p.dir_band = 1;
    p.seed = seed;
    p.bin_dt = 0.010f;
    p.max_time = 3.0f;
    p.floor_rel = 1e-7f;
    p.radius = 0.5f;
    for (int b = 0; b < 3; b++) p.energy[b] = energy[b];
    return p;
}

}  // namespace

extern "C" int omg_wire_decode_world_job(const float* job, uint32_t version, omg_wire_world_job* out)
{
    if (!job || !out) return fail(OMG_ERR_INVALID, "bad argument");
    if (version != OMG_WIRE_WORLD_JOB_VERSION) return fail(OMG_ERR_VERSION, "world job version");
    std::memset(out, 0, sizeof *out);
    out->id = as_u32(job[0]);
    out->n_rays = to_uint32(job[1]);
    out->seed = to_uint32(job[2]);
    for (int k = 0; k < 3; k++) {
        out->src[k] = job[3 + k];
        out->lis[k] = job[6 + k];
    }
    // Math.min(jobs[9], 64) (gpu.js:427)
    float np = job[9];
    uint32_t n = np > 0.0f ? (np >= 64.0f ? 64u : (uint32_t)np) : 0u;
    out->n_panels = n;
    // 16 f32 per panel, as the kernel's Panel (lib.rs:404-408, layout.rs:130-139)
    for (uint32_t i = 0; i < n; i++) {
        const float* q = job + 10 + 16 * i;
        omg_overlay& o = out->panels[i];
        for (int k = 0; k < 3; k++) {
            o.min[k] = q[k];
            o.max[k] = q[4 + k];
            o.mat.absorption[k] = q[8 + k];
            o.mat.transmission[k] = q[12 + k];
        }
        o.mat.scattering = q[3];
    }
    return OMG_OK;
}

extern "C" int omg_wire_trace_box_jobs(omg_ctx* c, const float* jobs, uint32_t n_f32, uint32_t version,
                                       float* echo_out, uint32_t* ids_out, uint32_t* n_jobs_out)
{
    if (!c || (n_f32 && (!jobs || !echo_out || !ids_out)) || !n_jobs_out)
        return fail(OMG_ERR_INVALID, "bad argument");
    if (version != OMG_WIRE_BOX_JOB_VERSION) return fail(OMG_ERR_VERSION, "box job version");
    if (n_f32 % OMG_WIRE_BOX_JOB_F32S) return fail(OMG_ERR_INVALID, "not a whole number of jobs");
    const uint32_t n = n_f32 / OMG_WIRE_BOX_JOB_F32S;
    *n_jobs_out = 0;
    for (uint32_t k = 0; k < n; k++) {
        const float* j = jobs + (size_t)k * OMG_WIRE_BOX_JOB_F32S;
        omg_material walls[6];
        std::memset(walls, 0, sizeof walls);
        for (int w = 0; w < 6; w++) {
            const float* q = j + 15 + 7 * w;
            for (int b = 0; b < 3; b++) {
                walls[w].absorption[b] = q[b];
                walls[w].transmission[b] = q[4 + b];
            }
            walls[w].scattering = q[3];
        }
        omg_trace_params p = wire_params(to_uint32(j[1]), to_uint32(j[2]), j + 12);
        float* e = echo_out + (size_t)k * OMG_WIRE_ECHO_F32S;
        int rc = omg_trace_box(c, j + 3, walls, 3, &p, j + 6, 1, j + 9, 1, e,
                               e + OMG_WIRE_ECHO_BINS * 3, nullptr, nullptr);
        if (rc) return fail(rc, "job " + std::to_string(k) + ": " + omg_last_error());
        ids_out[k] = as_u32(j[0]);
        *n_jobs_out = k + 1;
    }
    return OMG_OK;
}

extern "C" int omg_wire_trace_world_jobs(omg_scene* s, const float* jobs, uint32_t n_f32,
                                         uint32_t version, float* echo_out, uint32_t* ids_out,
                                         uint32_t* n_jobs_out)
{
    if (!s || (n_f32 && (!jobs || !echo_out || !ids_out)) || !n_jobs_out)
        return fail(OMG_ERR_INVALID, "bad argument");
    if (version != OMG_WIRE_WORLD_JOB_VERSION) return fail(OMG_ERR_VERSION, "world job version");
    if (n_f32 % OMG_WIRE_WORLD_JOB_F32S) return fail(OMG_ERR_INVALID, "not a whole number of jobs");
    const uint32_t n = n_f32 / OMG_WIRE_WORLD_JOB_F32S;
    *n_jobs_out = 0;
    static const float kUnit[3] = {1.0f, 1.0f, 1.0f};
    omg_wire_world_job J;
    for (uint32_t k = 0; k < n; k++) {
        int rc = omg_wire_decode_world_job(jobs + (size_t)k * OMG_WIRE_WORLD_JOB_F32S, version, &J);
        if (rc) return rc;
        omg_trace_params p = wire_params(J.n_rays, J.seed, kUnit);
        float* e = echo_out + (size_t)k * OMG_WIRE_ECHO_F32S;
        rc = omg_trace_world(s, &p, J.src, 1, J.lis, 1, J.panels, J.n_panels, e,
                             e + OMG_WIRE_ECHO_BINS * 3, nullptr, nullptr);
        if (rc) return fail(rc, "job " + std::to_string(k) + ": " + omg_last_error());
        ids_out[k] = J.id;
        *n_jobs_out = k + 1;
    }
    return OMG_OK;
}

extern "C" int omg_param_block_write_flat(const omg_param_block* pb, const omg_tap* taps,
                                          uint32_t n_taps, float* out, uint32_t cap,
                                          uint32_t* n_out)
{
    if (!pb || !n_out || (n_taps && !taps)) return fail(OMG_ERR_INVALID, "bad argument");
    const uint64_t len = OMG_WIRE_PARAM_HEADER + (uint64_t)n_taps * OMG_WIRE_PARAM_PER_TAP;
    *n_out = (uint32_t)len;
    if (!out || cap < len) return fail(OMG_ERR_CAP, "flat buffer too small");
    std::memset(out, 0, len * sizeof(float));
    out[0] = (float)pb->version;
    out[1] = (float)n_taps;
    for (int b = 0; b < 3; b++) {
        out[2 + b] = pb->rt60[b];
        out[5 + b] = pb->level[b];
    }
    if (pb->has_remote) {
        out[8] = 1.0f;
        for (int b = 0; b < 3; b++) {
            out[9 + b] = pb->remote_rt60[b];
            out[12 + b] = pb->remote_send[b];
        }
        for (int k = 0; k < 9; k++) out[15 + k] = pb->remote_sh[k];
    }
    for (uint32_t i = 0; i < n_taps; i++) {
        float* o = out + OMG_WIRE_PARAM_HEADER + (size_t)i * OMG_WIRE_PARAM_PER_TAP;
        o[0] = (float)taps[i].key;
        o[1] = taps[i].delay_s;
        for (int k = 0; k < 3; k++) {
            o[2 + k] = taps[i].dir[k];
            o[5 + k] = taps[i].gains[k];
        }
    }
    return OMG_OK;
}

extern "C" int omg_param_block_read_flat(const float* data, uint32_t n_f32, omg_param_block* pb,
                                         omg_tap* taps, uint32_t cap, uint32_t* n_taps)
{
    if (!data || !pb || !n_taps || (cap && !taps)) return fail(OMG_ERR_INVALID, "bad argument");
    if (n_f32 < OMG_WIRE_PARAM_HEADER) return fail(OMG_ERR_INVALID, "flat block shorter than its header");
    const uint32_t nt = as_u32(data[1]);
    if ((uint64_t)OMG_WIRE_PARAM_HEADER + (uint64_t)nt * OMG_WIRE_PARAM_PER_TAP > n_f32)
        return fail(OMG_ERR_INVALID, "flat block shorter than its taps");
    std::memset(pb, 0, sizeof *pb);
    pb->version = as_u64(data[0]);
    for (int b = 0; b < 3; b++) {
        pb->rt60[b] = data[2 + b];
        pb->level[b] = data[5 + b];
    }
    if (data[8] > 0.5f) {
        pb->has_remote = 1;
        for (int b = 0; b < 3; b++) {
            pb->remote_rt60[b] = data[9 + b];
            pb->remote_send[b] = data[12 + b];
        }
        for (int k = 0; k < 9; k++) pb->remote_sh[k] = data[15 + k];
    }
    *n_taps = nt;
    for (uint32_t i = 0; i < nt && i < cap; i++) {
        const float* o = data + OMG_WIRE_PARAM_HEADER + (size_t)i * OMG_WIRE_PARAM_PER_TAP;
        taps[i].key = as_u32(o[0]);
        taps[i].delay_s = o[1];
        for (int k = 0; k < 3; k++) {
            taps[i].dir[k] = o[2 + k];
            taps[i].gains[k] = o[5 + k];
        }
    }
    return OMG_OK;
}
