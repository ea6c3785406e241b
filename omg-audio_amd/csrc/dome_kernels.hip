// Ambient dome kernels (crates/omg-scene/src/dome.rs:330-459) for gfx950.
//
// k_dome_rays: one lane per (listener, fan ray): the bounce path to the sky
// (trace_escape) and, when it does not deliver more than half the low band,
// the straight-through segment (trace_through).  The fan is deterministic,
// so every per-ray result equals the CPU oracle's bit for bit.
// k_dome_reduce: one lane per (listener, bin) sums its rays in fan order,
// reproducing the reference's sequential f32 sums exactly.
#include "bvh_device.h"
#include "dome_kernels.h"

#pragma clang fp contract(off)

namespace omg {

// Panel::hit (dome.rs:56-68)
__device__ __forceinline__ bool panel_hit(const DomePanelDev& pl, f3 o, f3 d, float t_max, float& t_out)
{
    float po = pl.axis == 0 ? o.x : o.y;
    float pd = pl.axis == 0 ? d.x : d.y;
    if (fabsf(pd) < 1e-8f) return false;
    float t = (pl.plane - po) / pd;
    if (t < 1e-4f || t >= t_max) return false;
    float lat = pl.axis == 0 ? o.y + t * d.y : o.x + t * d.x;
    float z = o.z + t * d.z;
    if (lat > pl.lat0 && lat < pl.lat1 && z > pl.z0 && z < pl.z1) {
        t_out = t;
        return true;
    }
    return false;
}

template <int NBM, bool COUNT>
__global__ __launch_bounds__(128) void k_dome_rays(DomeArgs A)
{
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TravCount tc;
    bool active = gid < A.n_eyes * A.n_rays;
    uint32_t status = 0;
    if (active) {
        uint32_t eye_i = gid / A.n_rays;
        uint32_t j = gid - eye_i * A.n_rays;
        uint32_t i = (uint32_t)(((uint64_t)j * A.n_fan) / A.n_rays);
        f3 eye = mk(A.eyes[3 * eye_i], A.eyes[3 * eye_i + 1], A.eyes[3 * eye_i + 2]);
        f3 d0 = mk(A.fan[3 * i], A.fan[3 * i + 1], A.fan[3 * i + 2]);
        // ---- trace_escape (dome.rs:330-382)
        float bp[NBM];
#pragma unroll
        for (int b = 0; b < NBM; b++) bp[b] = 1.0f;
        f3 pos = eye, dir = d0;
        bool esc = false;
        uint32_t events = 0;
        for (uint32_t e = 0; e < A.max_events; e++) {
            events++;
            float tm;
            uint32_t tri, fl;
            bool mesh_hit = mesh_raycast<COUNT>(A.g, pos, dir, tm, tri, fl, tc);
            float t_pan = mesh_hit ? tm : FLT_MAX;
            int pan = -1;
            for (uint32_t k = 0; k < A.n_panels; k++) {
                float t;
                if (panel_hit(A.panels[k], pos, dir, t_pan, t)) {
                    t_pan = t;
                    pan = (int)k;
                }
            }
            if (pan >= 0) {
                const DomePanelDev& pl = A.panels[pan];
                if (!pl.glass) break;   // a door leaf blocks
#pragma unroll
                for (int b = 0; b < NBM; b++) bp[b] *= pl.trans[b] * pl.trans[b];
                pos = add(pos, mul(dir, t_pan + 1e-3f));
                continue;
            }
            bool sky = pos.z > -0.05f || dir.z > 0.05f;
            if (!mesh_hit || tm > A.horizon) {
                esc = sky;
                break;
            }
            float4 ti = *reinterpret_cast<const float4*>(A.g.tris + tri);
            const omg_material_dev& m = A.mats[__float_as_uint(ti.w)];
#pragma unroll
            for (int b = 0; b < NBM; b++) bp[b] *= 1.0f - m.absorption[b];
            if (bp[1 < NBM ? 1 : 0] < 2e-3f) break;
            f3 n = mk(ti.x, ti.y, ti.z);
            f3 n_o = dot(dir, n) > 0.0f ? mul(n, -1.0f) : n;
            f3 hit = add(pos, mul(dir, tm));
            dir = normalize(sub(dir, mul(n, 2.0f * dot(dir, n))));
            pos = add(hit, mul(n_o, 1e-3f));
        }
        // ---- sample: bounce first, straight-through when the bounce is weak (dome.rs:438-450)
        float tp[NBM];
        bool through = false;
        if (esc && bp[0] > 0.5f) {
#pragma unroll
            for (int b = 0; b < NBM; b++) tp[b] = bp[b];
        } else {
            through = true;
            // trace_through (dome.rs:390-422)
            f3 end = add(eye, mul(d0, A.through_len));
#pragma unroll
            for (int b = 0; b < NBM; b++) tp[b] = 0.0f;
            if (!(end.z < -0.05f)) {
#pragma unroll
                for (int b = 0; b < NBM; b++) tp[b] = 1.0f;
                SegHit hits[kDomeSegCap];
                int n = segment_hits<COUNT>(A.g, eye, end, hits, kDomeSegCap, tc);
                if (n < 0) {
                    status = 1;   // crossing list overflow
                    n = 0;
                }
                bool dead = false;
                for (int h = 0; h < n && !dead; h++) {
                    const omg_material_dev& m = A.mats[hits[h].mat];
#pragma unroll
                    for (int b = 0; b < NBM; b++) tp[b] *= m.transmission[b] * m.transmission[b];
                    if (tp[0] < 1e-7f) dead = true;
                }
                if (dead) {
#pragma unroll
                    for (int b = 0; b < NBM; b++) tp[b] = 0.0f;
                } else {
                    for (uint32_t k = 0; k < A.n_panels; k++) {
                        float t;
                        if (panel_hit(A.panels[k], eye, d0, A.horizon, t)) {
#pragma unroll
                            for (int b = 0; b < NBM; b++)
                                tp[b] *= A.panels[k].trans[b] * A.panels[k].trans[b];
                        }
                    }
                }
            }
            if (esc) {
#pragma unroll
                for (int b = 0; b < NBM; b++) tp[b] = fmaxf(tp[b], bp[b]);
            }
        }
        DomeRayDev* r = A.rays + gid;
        r->escaped = esc ? 1u : 0u;
        r->events = events;
        r->used_through = through ? 1u : 0u;
        r->bin = A.fan_bin[i];
#pragma unroll
        for (int b = 0; b < kMaxBands; b++) r->tp[b] = b < NBM ? tp[b] : 0.0f;
        r->dir[0] = d0.x;
        r->dir[1] = d0.y;
        r->dir[2] = d0.z;
    }
    if (A.status && status) atomicOr(A.status, status);
    if (COUNT && A.counters) {
        uint64_t a = tc.aabb, t = tc.tri;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off);
            t += __shfl_xor(t, off);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(A.counters + 0, (unsigned long long)a);
            atomicAdd(A.counters + 1, (unsigned long long)t);
        }
    }
}

// dome.rs:433-459 binning: e[bin][b] += tp[b] / n_rays and the
// departure-direction sum, in fan order.  One wave per eye: the wave loads
// 64 rays at a time (coalesced) and computes each ray's terms in parallel;
// lanes 0..8 then own one bin each and add the terms of the chunk's rays
// in fan order, which keeps the reference's sequential f32 sums exact.
template <int NBM>
__global__ __launch_bounds__(64) void k_dome_reduce(DomeArgs A)
{
    uint32_t eye_i = blockIdx.x, lane = threadIdx.x;
    float e[NBM];
#pragma unroll
    for (int b = 0; b < NBM; b++) e[b] = 0.0f;
    f3 ds = mk(0.0f, 0.0f, 0.0f);
    float nr = (float)A.n_rays;
    const DomeRayDev* rays = A.rays + (size_t)eye_i * A.n_rays;
    for (uint32_t base = 0; base < A.n_rays; base += 64) {
        uint32_t j = base + lane;
        uint32_t bin = 0xFFFFFFFFu;
        float c[NBM];
        f3 dc = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int b = 0; b < NBM; b++) c[b] = 0.0f;
        if (j < A.n_rays) {
            const DomeRayDev& r = rays[j];
            if (r.tp[0] > 1e-9f) {
                bin = r.bin;
#pragma unroll
                for (int b = 0; b < NBM; b++) c[b] = r.tp[b] / nr;
                float w = fmaxf(r.tp[1 < NBM ? 1 : 0], r.tp[0] * 0.2f);
                dc = mul(mk(r.dir[0], r.dir[1], r.dir[2]), w);
            }
        }
        uint32_t n = A.n_rays - base < 64u ? A.n_rays - base : 64u;
        for (uint32_t l = 0; l < n; l++) {
            uint32_t bl = __shfl(bin, (int)l);
            float cl[NBM];
#pragma unroll
            for (int b = 0; b < NBM; b++) cl[b] = __shfl(c[b], (int)l);
            f3 dl = mk(__shfl(dc.x, (int)l), __shfl(dc.y, (int)l), __shfl(dc.z, (int)l));
            if (bl == lane) {
#pragma unroll
                for (int b = 0; b < NBM; b++) e[b] += cl[b];
                ds = add(ds, dl);
            }
        }
    }
    if (lane < kDomeBins) {
        float* oe = A.energy + ((size_t)eye_i * kDomeBins + lane) * NBM;
#pragma unroll
        for (int b = 0; b < NBM; b++) oe[b] = e[b];
        float* od = A.dsum + ((size_t)eye_i * kDomeBins + lane) * 3;
        od[0] = ds.x;
        od[1] = ds.y;
        od[2] = ds.z;
    }
}

template <int NBM>
static hipError_t launch_dome_t(const DomeArgs& A, bool count, hipStream_t s)
{
    uint32_t n = A.n_eyes * A.n_rays;
    if (n) {
        if (count)
            hipLaunchKernelGGL((k_dome_rays<NBM, true>), dim3((n + 127) / 128), dim3(128), 0, s, A);
        else
            hipLaunchKernelGGL((k_dome_rays<NBM, false>), dim3((n + 127) / 128), dim3(128), 0, s, A);
    }
    if (A.n_eyes) hipLaunchKernelGGL((k_dome_reduce<NBM>), dim3(A.n_eyes), dim3(64), 0, s, A);
    return hipGetLastError();
}

__global__ void k_dome_prime(float* p)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = 0.0f;
}

hipError_t launch_dome_prime(float* p, hipStream_t s)
{
    hipLaunchKernelGGL(k_dome_prime, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_dome(const DomeArgs& A, bool count, hipStream_t s)
{
    switch (A.nbands) {
    case 1: return launch_dome_t<1>(A, count, s);
    case 2: return launch_dome_t<2>(A, count, s);
    case 3: return launch_dome_t<3>(A, count, s);
    case 4: return launch_dome_t<4>(A, count, s);
    case 5: return launch_dome_t<5>(A, count, s);
    case 6: return launch_dome_t<6>(A, count, s);
    case 7: return launch_dome_t<7>(A, count, s);
    case 8: return launch_dome_t<8>(A, count, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace omg
