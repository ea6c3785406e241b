// Stochastic energy path tracer kernels for gfx950 (wave64).
//
// One lane owns one ray of one source for its whole life (tracer.rs:92-201).
// Traversal replays the reference's ordered binary-BVH walk (mesh.rs:484-541)
// over the paired 64 B nodes, so hit triangles, bounce counts and the f32
// bits of every per-ray quantity equal the CPU oracle's.  Randomness comes
// from a Philox4x32-10 block per (ray, bounce, source): no state crosses
// lanes, and a ray's path is the same whatever launch shape runs it.
#include "bvh_device.h"

#pragma clang fp contract(off)

namespace omg {

struct Hit {
    float t;
    f3 n;          // oriented against the ray
    uint32_t mat;  // index into the launch's material table
    uint32_t id;   // triangle id, wall index, or 0xFFFF0000 | overlay
    uint32_t flags;
};

// Shoebox::raycast_hit (scene.rs:41-82)
__device__ __forceinline__ bool box_raycast_hit(const GeomArgs& G, f3 p, f3 d, Hit& h)
{
    bool inside = p.x > 0.0f && p.x < G.box[0] && p.y > 0.0f && p.y < G.box[1] && p.z > 0.0f &&
                  p.z < G.box[2];
    if (!inside) return false;
    float best_t = FLT_MAX;
    int best_w = 0;
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        float di = get(d, axis), pi = get(p, axis);
        if (di > 1e-9f) {
            float t = (G.box[axis] - pi) / di;
            if (t < best_t) { best_t = t; best_w = 2 * axis + 1; }
        } else if (di < -1e-9f) {
            float t = -pi / di;
            if (t < best_t) { best_t = t; best_w = 2 * axis; }
        }
    }
    if (!isfinite(best_t) || best_t <= 0.0f) return false;
    float sgn = (best_w & 1) ? -1.0f : 1.0f;
    int ax = best_w >> 1;
    h.t = best_t;
    h.n = mk(ax == 0 ? sgn : 0.0f, ax == 1 ? sgn : 0.0f, ax == 2 ? sgn : 0.0f);
    h.mat = (uint32_t)best_w;
    h.id = (uint32_t)best_w;
    h.flags = 0;
    return true;
}

// WithPanels::raycast_hit overlay loop (scene.rs:99-135)
__device__ __forceinline__ void overlay_hits(const TraceArgs& A, f3 p, f3 d, bool& found, Hit& h)
{
    for (uint32_t i = 0; i < A.n_ovl; i++) {
        const Overlay ov = A.overlays[i];
        float t0 = 0.0f, t1 = FLT_MAX;
        int axis = 3;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (t0 == FLT_MAX) continue;
            float da = get(d, a), pa = get(p, a), lo = ov.mn[a], hi = ov.mx[a];
            if (fabsf(da) < 1e-9f) {
                if (pa < lo || pa > hi) t0 = FLT_MAX;
            } else {
                float ta = (lo - pa) / da, tb = (hi - pa) / da;
                if (ta > tb) { float s = ta; ta = tb; tb = s; }
                if (ta > t0) { t0 = ta; axis = a; }
                t1 = fminf(t1, tb);
                if (t0 > t1) t0 = FLT_MAX;
            }
        }
        if (t0 > 1e-4f && t0 < FLT_MAX && axis < 3 && (!found || t0 < h.t)) {
            float s = get(d, axis) > 0.0f ? -1.0f : 1.0f;
            h.t = t0;
            h.n = mk(axis == 0 ? s : 0.0f, axis == 1 ? s : 0.0f, axis == 2 ? s : 0.0f);
            h.mat = ov.mat;
            h.id = 0xFFFF0000u | i;
            h.flags = 0;
            found = true;
        }
    }
}

template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const LdsTrav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    return mesh_raycast_lds<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const QuadTrav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    return mesh_raycast_quad<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}

template <bool MESH, bool COUNT, int SK, class Trav>
__device__ __forceinline__ bool raycast_hit(const TraceArgs& A, const Trav& L, f3 p, f3 d,
                                            Hit& h, TravCount& c)
{
    bool found;
    if (MESH) {
        float t;
        uint32_t tri, fl;
        found = mesh_raycast_any<COUNT, SK>(A.g, L, p, d, t, tri, fl, c);
        if (found) {
            float4 ti = *reinterpret_cast<const float4*>(A.g.tris + tri);
            f3 n = mk(ti.x, ti.y, ti.z);
            if (dot(n, d) > 0.0f) n = mul(n, -1.0f);
            h.t = t;
            h.n = n;
            h.mat = __float_as_uint(ti.w);
            h.id = tri;
            h.flags = fl;
        }
    } else {
        found = box_raycast_hit(A.g, p, d, h);
    }
    if (A.n_ovl) overlay_hits(A, p, d, found, h);
    return found;
}

template <int LANES>
struct TravOf {
    using type = LdsTrav;
};
template <>
struct TravOf<4> {
    using type = QuadTrav;
};

__device__ __forceinline__ void init_trav(LdsTrav& L, lds_float4* lds, uint32_t nc, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.st_ref = reinterpret_cast<lds_u32*>(lds + 4 * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * blockDim.x);
    L.lane = threadIdx.x;
    L.bs = blockDim.x;
}
__device__ __forceinline__ void init_trav(QuadTrav& L, lds_float4* lds, uint32_t nc, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.nq = blockDim.x >> 2;
    L.st_ref = reinterpret_cast<lds_u32*>(lds + 4 * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * L.nq);
    L.quad = threadIdx.x >> 2;
    L.q = threadIdx.x & 3u;
}

// tracer::trace (tracer.rs:80-202).  LANES = 1: one lane per (source, ray);
// LANES = 4: a DPP quad per ray (see mesh_raycast_quad), the quad's lanes
// also splitting the listeners between them.
//
// Dynamic LDS: the first nc BVH nodes (breadth-first, so the top levels),
// then the SK-deep stack tops of the block's rays.  Blocks are dispatched
// round-robin over the 8 XCDs; the block index is remapped so that each
// XCD walks a contiguous range of rays (the same sources, hence the same
// parts of the tree in its own L2).
template <int NBM, bool MESH, bool COUNT, int SK, int LANES>
__global__ __launch_bounds__(256) void k_trace(TraceArgs A)
{
    extern __shared__ float4 lds[];
    typename TravOf<LANES>::type L;
    init_trav(L, (lds_float4*)lds, MESH ? A.node_cache : 0u, SK);
    if (MESH) {
        const float4* src4 = reinterpret_cast<const float4*>(A.g.nodes);
        for (uint32_t i = threadIdx.x; i < 4 * A.node_cache; i += blockDim.x) lds[i] = src4[i];
        __syncthreads();
    }
    uint32_t nb = gridDim.x, xcd = blockIdx.x & 7u, per = nb >> 3;
    uint32_t blk = (nb & 7u) == 0 ? xcd * per + (blockIdx.x >> 3) : blockIdx.x;
    uint32_t lid = (blk * blockDim.x + threadIdx.x) / LANES;
    const uint32_t q = LANES == 1 ? 0u : (threadIdx.x & (LANES - 1));
    bool active = lid < A.ray_count;
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    if (active) {
        uint32_t gid = A.ray_base + lid;
        uint32_t src = gid / A.n_rays;
        uint32_t ray = gid - src * A.n_rays;
        f3 pos = mk(A.sources[3 * src], A.sources[3 * src + 1], A.sources[3 * src + 2]);
        uint32_t w[4];
        philox4x32_10(ray, 0u, src, 0u, A.key0, A.key1, w);
        f3 dir = unit_sphere(u24(w[0]), u24(w[1]));
        float energy[NBM];
#pragma unroll
        for (int b = 0; b < NBM; b++) energy[b] = A.energy0[b];
        float dist_total = 0.0f;
        uint32_t iters = 0, term = 4u, hash = 2166136261u, rflags = 0;
        const size_t eg = (size_t)A.nbins * NBM, dg = (size_t)A.nbins * 3;
        float* ob_src = A.bins + (size_t)src * A.n_lis * eg;
        float* od_src = A.dirs + (size_t)src * A.n_lis * dg;
        TravCount tc;
        for (uint32_t bounce = 0; bounce < A.max_bounces; bounce++) {
            iters++;
            Hit h;
            bool has = raycast_hit<MESH, COUNT, SK>(A, L, pos, dir, h, tc);
            float t_hit = has ? h.t : FLT_MAX;
            // receiver sphere crossing along this segment (tracer.rs:110-128)
            for (uint32_t li = q; li < A.n_lis; li += LANES) {
                f3 Lp = mk(A.listeners[3 * li], A.listeners[3 * li + 1], A.listeners[3 * li + 2]);
                f3 to_l = sub(Lp, pos);
                float s = dot(to_l, dir);
                if (s > 0.0f && s < t_hit) {
                    f3 closest = add(pos, mul(dir, s));
                    if (length(sub(closest, Lp)) < A.radius) {
                        float arrival = (dist_total + s) / kSpeedOfSound;
                        uint32_t bin = sat_u32(arrival / A.bin_dt);
                        if (bin < A.nbins) {
                            float* ob = ob_src + li * eg + (size_t)bin * NBM;
                            float* od = od_src + li * dg + (size_t)bin * 3;
#pragma unroll
                            for (int b = 0; b < NBM; b++) unsafeAtomicAdd(ob + b, energy[b]);
                            float ed = energy[0];
#pragma unroll
                            for (int b = 1; b < NBM; b++)
                                if ((uint32_t)b == A.dir_band) ed = energy[b];
                            unsafeAtomicAdd(od + 0, -(dir.x * ed));
                            unsafeAtomicAdd(od + 1, -(dir.y * ed));
                            unsafeAtomicAdd(od + 2, -(dir.z * ed));
                        }
                    }
                }
            }
            if (A.hit_ids && q == 0) {
                size_t slot = (size_t)gid * A.max_bounces + bounce;
                A.hit_ids[slot] = has ? h.id : 0xFFFFFFFFu;
                A.hit_t[slot] = has ? h.t : 0.0f;
            }
            if (!has) { term = 0u; break; }
            hash = fnv1a_u32(hash, h.id);
            rflags |= h.flags;
            pos = add(pos, mul(dir, t_hit));
            dist_total += t_hit;
            if (dist_total / kSpeedOfSound > A.max_time) { term = 1u; break; }
            const MatParams* M = A.mats + h.mat;
            philox4x32_10(ray, bounce + 1u, src, 0u, A.key0, A.key1, w);
            float p = M->p;
            bool transmitted = p > 0.0f && u24(w[0]) < p;
            bool alive = false;
            if (transmitted) {
#pragma unroll
                for (int b = 0; b < NBM; b++) {
                    energy[b] *= M->tw[b];
                    if (energy[b] > A.floor_e) alive = true;
                }
                if (!alive) { term = 2u; break; }
                pos = add(pos, mul(dir, 2.0f * 1e-3f));
                continue;
            }
#pragma unroll
            for (int b = 0; b < NBM; b++) {
                energy[b] *= M->rw[b];
                if (energy[b] > A.floor_e) alive = true;
            }
            if (!alive) { term = 3u; break; }
            f3 n = h.n;
            if (u24(w[1]) < M->scattering) {
                dir = normalize(add(n, unit_sphere(u24(w[2]), u24(w[3]))));
                if (dot(dir, n) < 1e-3f) dir = n;
            } else {
                dir = sub(dir, mul(n, 2.0f * dot(dir, n)));
            }
            pos = add(pos, mul(n, 1e-3f));
        }
        if (A.records && q == 0) {
            uint4 r = make_uint4(iters, term, hash, rflags);
            reinterpret_cast<uint4*>(A.records)[gid] = r;
        }
        if (COUNT && q == 0) {
            aabb_acc = tc.aabb;
            tri_acc = tc.tri;
            iters_acc = iters;
        }
    }
    if (COUNT && A.counters) {
        // wave-level reduction, one atomic per wave and counter
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            aabb_acc += __shfl_xor(aabb_acc, off);
            tri_acc += __shfl_xor(tri_acc, off);
            iters_acc += __shfl_xor(iters_acc, off);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(A.counters + 0, (unsigned long long)aabb_acc);
            atomicAdd(A.counters + 1, (unsigned long long)tri_acc);
            atomicAdd(A.counters + 2, (unsigned long long)iters_acc);
        }
    }
}

__global__ void k_fill_zero(float* p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_fill_zero_u64(unsigned long long* p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ------------------------------------------------------------------ launchers

constexpr uint32_t kTraceBlock = 256;

template <int NBM, bool MESH, bool COUNT, int SK, int LANES>
static hipError_t launch_trace_l(const TraceArgs& A, hipStream_t s)
{
    uint64_t threads = (uint64_t)A.ray_count * LANES;
    uint32_t blocks = (uint32_t)((threads + kTraceBlock - 1) / kTraceBlock);
    if (blocks == 0) return hipSuccess;
    size_t lds = (MESH ? (size_t)A.node_cache * sizeof(Node2) : 0) +
                 (size_t)SK * (kTraceBlock / LANES) * 8;
    auto kern = k_trace<NBM, MESH, COUNT, SK, LANES>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kTraceBlock), lds, s, A);
    return hipGetLastError();
}

template <int NBM, bool MESH, bool COUNT, int SK>
static hipError_t launch_trace_t(const TraceArgs& A, hipStream_t s)
{
    if (MESH && A.lanes == 4) return launch_trace_l<NBM, MESH, COUNT, SK, 4>(A, s);
    return launch_trace_l<NBM, MESH, COUNT, SK, 1>(A, s);
}

template <int NBM, bool MESH, bool COUNT>
static hipError_t launch_trace_sk(const TraceArgs& A, hipStream_t s)
{
    if (!MESH) return launch_trace_t<NBM, MESH, COUNT, 1>(A, s);
    if (A.stack_lds >= 32) return launch_trace_t<NBM, MESH, COUNT, 32>(A, s);
    return launch_trace_t<NBM, MESH, COUNT, 16>(A, s);
}

template <bool MESH, bool COUNT>
static hipError_t launch_trace_nb(const TraceArgs& A, hipStream_t s)
{
    switch (A.nbands) {
    case 1: return launch_trace_sk<1, MESH, COUNT>(A, s);
    case 2: return launch_trace_sk<2, MESH, COUNT>(A, s);
    case 3: return launch_trace_sk<3, MESH, COUNT>(A, s);
    case 4: return launch_trace_sk<4, MESH, COUNT>(A, s);
    case 5: return launch_trace_sk<5, MESH, COUNT>(A, s);
    case 6: return launch_trace_sk<6, MESH, COUNT>(A, s);
    case 7: return launch_trace_sk<7, MESH, COUNT>(A, s);
    case 8: return launch_trace_sk<8, MESH, COUNT>(A, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace(const TraceArgs& A, bool mesh, bool count, hipStream_t s)
{
    if (mesh) return count ? launch_trace_nb<true, true>(A, s) : launch_trace_nb<true, false>(A, s);
    return count ? launch_trace_nb<false, true>(A, s) : launch_trace_nb<false, false>(A, s);
}

hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero_u64, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace omg
