// Image sources on the BVH (crates/omg-core/src/pt_mesh.rs) for gfx950.
//
// k_discover: one lane per listener fan ray (mesh_chains, pt_mesh.rs:375-457):
//   up to three specular bounces off the mesh and the overlay boxes,
//   recording the surface-id chain prefix.
// k_solve: one lane per (source, chain) (mesh_record, pt_mesh.rs:185-324):
//   mirror the source across the chain planes, validate every bounce point
//   by raycast, pay the transmission of every crossing of every leg.
#include "bvh_device.h"
#include "ism_kernels.h"

#pragma clang fp contract(off)

namespace omg {

__global__ __launch_bounds__(128) void k_discover(DiscoverArgs A)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_rays) return;
    float z = 1.0f - 2.0f * ((float)i + 0.5f) / (float)A.n_rays;
    float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    float phi = A.ga * (float)i + A.rot_phase + A.jitter[i] * 0.02f;
    float s, c;
    sincos_cw(phi, &s, &c);
    f3 dir = mk(r * c, r * s, z);
    f3 pos = mk(A.listener[0], A.listener[1], A.listener[2]);
    uint32_t chain[3] = {0xFFFFu, 0xFFFFu, 0xFFFFu};
    uint32_t levels = 0;
    TravCount tc;
    for (int k = 0; k < 3; k++) {
        float t;
        uint32_t tri, fl;
        bool has = mesh_raycast<false>(A.g, pos, dir, t, tri, fl, tc);
        uint32_t sid = 0;
        f3 nn = mk(0.0f, 0.0f, 1.0f);
        if (has) {
            sid = A.tri_surface[tri];
            float4 ti = *reinterpret_cast<const float4*>(A.g.tris + tri);
            nn = mk(ti.x, ti.y, ti.z);
        } else {
            t = FLT_MAX;
        }
        // overlay boxes: slab entry from 1e-4, nearest wins (pt_mesh.rs:391-425)
        for (uint32_t bi = 0; bi < A.n_boxes; bi++) {
            const float* mn = A.boxes + 6 * bi;
            const float* mx = mn + 3;
            float t0 = 1e-4f, t1 = t;
            int axis = 3;
            for (int a = 0; a < 3; a++) {
                const float da = get(dir, a), pa = get(pos, a);
                if (fabsf(da) < 1e-9f) {
                    if (pa < mn[a] || pa > mx[a]) {
                        t0 = FLT_MAX;
                        break;
                    }
                } else {
                    float ta = (mn[a] - pa) / da, tb = (mx[a] - pa) / da;
                    if (ta > tb) {
                        const float sw = ta;
                        ta = tb;
                        tb = sw;
                    }
                    if (ta > t0) {
                        t0 = ta;
                        axis = a;
                    }
                    t1 = fminf(t1, tb);
                    if (t0 > t1) {
                        t0 = FLT_MAX;
                        break;
                    }
                }
            }
            if (t0 < t && axis < 3) {
                t = t0;
                const float da = get(dir, axis);
                const uint32_t face = (uint32_t)axis * 2u + (da > 0.0f ? 0u : 1u);
                sid = (A.base + bi * 6u + face) & 0xFFFFu;
                const float sgn = da > 0.0f ? -1.0f : 1.0f;
                nn = mk(axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f);
                has = true;
            }
        }
        if (!has) break;
        if (t <= 1e-4f || t > 200.0f) break;
        pos = add(pos, mul(dir, t));
        chain[k] = sid;
        levels = (uint32_t)k + 1;
        f3 n = dot(nn, dir) > 0.0f ? mul(nn, -1.0f) : nn;
        dir = sub(dir, mul(n, 2.0f * dot(dir, n)));
        pos = add(pos, mul(n, 1e-4f));
    }
    reinterpret_cast<uint2*>(A.out)[i] =
        make_uint2(chain[0] | (chain[1] << 16), chain[2] | (levels << 16));
}

__device__ __forceinline__ bool all_below(const float* x, uint32_t nb, float lim)
{
    for (uint32_t b = 0; b < nb; b++)
        if (!(x[b] < lim)) return false;
    return true;
}

// Aabb::clip (crates/omg-core/src/pt.rs:42-69): interior crossing of a->b.
__device__ __forceinline__ bool aabb_clip(const BlockerDev& x, f3 a, f3 b)
{
    f3 d = sub(b, a);
    float t0 = 0.0f, t1 = 1.0f;
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        float da = get(d, axis), aa = get(a, axis), mn = x.min[axis], mx = x.max[axis];
        if (fabsf(da) < 1e-9f) {
            if (aa < mn || aa > mx) return false;
        } else {
            float ta = (mn - aa) / da, tb = (mx - aa) / da;
            if (ta > tb) { float s = ta; ta = tb; tb = s; }
            t0 = fmaxf(t0, ta);
            t1 = fminf(t1, tb);
            if (t0 > t1) return false;
        }
    }
    return t1 > 1e-4f && t0 < 1.0f - 1e-4f;
}

// leg_transmission (pt_mesh.rs:145-181).  Returns 0 dead, 1 alive, -1 overflow.
__device__ int leg_transmission(const SolveArgs& A, f3 a, f3 b, bool has_skip, uint32_t skip_sid,
                                float skip_t, float* trans, SegHit* buf)
{
    TravCount tc;
    int n = segment_hits<false>(A.g, a, b, buf, kSolveSegCap, tc);
    if (n < 0) return -1;
    float seg_len = fmaxf(length(sub(b, a)), 1e-6f);
    for (int i = 0; i < n; i++) {
        if (has_skip && buf[i].surf == skip_sid && fabsf(buf[i].t - skip_t) * seg_len < 0.05f)
            continue;
        const float* tr = A.mats[buf[i].mat].transmission;
        for (uint32_t k = 0; k < A.nbands; k++) trans[k] *= tr[k];
        if (all_below(trans, A.nbands, 1e-6f)) return 0;
    }
    for (uint32_t i = 0; i < A.n_extras; i++) {
        if (aabb_clip(A.extras[i], a, b)) {
            for (uint32_t k = 0; k < A.nbands; k++) trans[k] *= A.extras[i].trans[k];
            if (all_below(trans, A.nbands, 1e-6f)) return 0;
        }
    }
    return 1;
}

__device__ __forceinline__ f3 mirror(f3 p, const SurfaceDev& s)
{
    f3 n = mk(s.n[0], s.n[1], s.n[2]);
    return sub(p, mul(n, 2.0f * (dot(n, p) - s.d)));
}

// mesh_record (pt_mesh.rs:185-324); 0 = no record, 1 = record, -1 = overflow
__device__ int mesh_record(const SolveArgs& A, const uint32_t* chain, uint32_t order, f3 src_pos,
                           MeshRecordDev& rec)
{
    SegHit buf[kSolveSegCap];
    f3 listener = mk(A.listener[0], A.listener[1], A.listener[2]);
    f3 img = src_pos;
    for (int k = (int)order - 1; k >= 0; k--) {
        if (chain[k] >= A.n_surf) return 0;
        img = mirror(img, A.surfaces[chain[k]]);
    }
    f3 to_img = sub(img, listener);
    float dist = length(to_img);
    if (dist < 1e-4f) return 0;
    f3 dir = mul(to_img, 1.0f / dist);
    float gm[kMaxBands];
    for (uint32_t b = 0; b < A.nbands; b++) gm[b] = 1.0f;
    f3 pos = listener, d = dir;
    TravCount tc;
    for (uint32_t k = 0; k < order; k++) {
        uint32_t sid = chain[k];
        if (sid >= A.n_surf) return 0;
        const SurfaceDev& s = A.surfaces[sid];
        f3 sn = mk(s.n[0], s.n[1], s.n[2]);
        float denom = dot(sn, d);
        if (fabsf(denom) < 1e-9f) return 0;
        float t = (s.d - dot(sn, pos)) / denom;
        if (t <= 1e-4f) return 0;
        f3 hit = add(pos, mul(d, t));
        f3 o = add(pos, mul(d, 1e-4f));
        float rt;
        uint32_t rtri, fl;
        if (s.has_rect) {
            const float eps = 1e-3f;
            for (int a = 0; a < 3; a++)
                if (get(hit, a) < s.rmin[a] - eps || get(hit, a) > s.rmax[a] + eps) return 0;
            if (mesh_raycast<false>(A.g, o, d, rt, rtri, fl, tc)) {
                if (rt < t - 1e-3f) {
                    uint32_t m = A.g.tris[rtri].mat;
                    if (all_below(A.mats[m].transmission, A.nbands, 1e-6f)) return 0;
                }
            }
        } else {
            if (!mesh_raycast<false>(A.g, o, d, rt, rtri, fl, tc)) return 0;
            bool on_surface = A.tri_surface[rtri] == sid && fabsf(rt - t) < 2e-2f;
            if (!on_surface) {
                if (rt < t - 1e-3f) {
                    uint32_t m = A.g.tris[rtri].mat;
                    if (all_below(A.mats[m].transmission, A.nbands, 1e-4f)) return 0;
                    int n = segment_hits<false>(A.g, pos, add(hit, mul(d, 0.05f)), buf,
                                                kSolveSegCap, tc);
                    if (n < 0) return -1;
                    bool any = false;
                    for (int i = 0; i < n && !any; i++)
                        any = buf[i].surf == sid && fabsf(buf[i].t * (t + 0.05f) - t) < 0.1f;
                    if (!any) return 0;
                } else {
                    return 0;
                }
            }
        }
        int lt = leg_transmission(A, pos, hit, true, sid, 1.0f, gm, buf);
        if (lt <= 0) return lt;
        for (uint32_t b = 0; b < A.nbands; b++) gm[b] *= s.refl[b];
        pos = add(hit, mul(sn, dot(sn, d) < 0.0f ? 1e-4f : -1e-4f));
        d = sub(d, mul(sn, 2.0f * dot(sn, d)));
    }
    if (order > 0) {
        const SurfaceDev& s = A.surfaces[chain[order - 1]];
        f3 sn = mk(s.n[0], s.n[1], s.n[2]);
        if ((dot(sn, pos) - s.d) * (dot(sn, src_pos) - s.d) < 0.0f) return 0;
    }
    int lt = leg_transmission(A, pos, src_pos, false, 0, 0.0f, gm, buf);
    if (lt <= 0) return lt;
    if (order > 0) {
        f3 to_src = sub(src_pos, pos);
        float leg = length(to_src);
        if (leg > 1e-4f && dot(mul(to_src, 1.0f / leg), d) < 0.995f) return 0;
    }
    float dtot = fmaxf(dist, 0.3f);
    float gains[kMaxBands];
    for (uint32_t b = 0; b < A.nbands; b++) {
        float air = expf(-A.air[b] * dtot);
        gains[b] = air / dtot * gm[b];
    }
    if (all_below(gains, A.nbands, 2e-5f)) return 0;
    rec.valid = 1;
    rec.delay = dist / kSpeedOfSound;
    rec.dir[0] = dir.x;
    rec.dir[1] = dir.y;
    rec.dir[2] = dir.z;
    for (uint32_t b = 0; b < kMaxBands; b++) rec.gains[b] = b < A.nbands ? gains[b] : 0.0f;
    return 1;
}

__global__ __launch_bounds__(64) void k_solve(SolveArgs A)
{
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= A.n_src * A.n_chains) return;
    uint32_t si = gid / A.n_chains, ci = gid - si * A.n_chains;
    uint32_t chain[3] = {A.chains[3 * ci], A.chains[3 * ci + 1], A.chains[3 * ci + 2]};
    uint32_t order = A.orders[ci];
    f3 src = mk(A.sources[3 * si], A.sources[3 * si + 1], A.sources[3 * si + 2]);
    MeshRecordDev rec;
    rec.valid = 0;
    rec.delay = 0.0f;
    rec.dir[0] = rec.dir[1] = rec.dir[2] = 0.0f;
    for (int b = 0; b < kMaxBands; b++) rec.gains[b] = 0.0f;
    int r = order <= 3 ? mesh_record(A, chain, order, src, rec) : 0;
    if (r < 0) {
        atomicOr(A.status, 1u);
        rec.valid = 0;
    }
    A.records[gid] = rec;
}

hipError_t launch_discover(const DiscoverArgs& A, hipStream_t s)
{
    if (A.n_rays == 0) return hipSuccess;
    hipLaunchKernelGGL(k_discover, dim3((A.n_rays + 127) / 128), dim3(128), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_solve(const SolveArgs& A, hipStream_t s)
{
    uint32_t n = A.n_src * A.n_chains;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_solve, dim3((n + 63) / 64), dim3(64), 0, s, A);
    return hipGetLastError();
}

}  // namespace omg
