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
__device__ __forceinline__ void overlay_hits(const JobView& A, f3 p, f3 d, bool& found, Hit& h)
{
    for (uint32_t i = 0; i < A.n_ovl; i++) {
        const Overlay ov = glb(A.overlays)[i];
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

template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const Lds4Trav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    return mesh_raycast4_lds<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const Row16Trav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    return mesh_raycast16<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const Row8Trav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    return mesh_raycast16x2<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_any(const GeomArgs& G, const Quad4Trav& L, f3 o, f3 d,
                                                 float& t, uint32_t& tri, uint32_t& fl,
                                                 TravCount& c)
{
    if constexpr (SK >= kDeepStack)
        return mesh_raycast4_quad_deep<COUNT, SK>(G, L, o, d, t, tri, fl, c);
    else
        return mesh_raycast4_quad<COUNT, SK>(G, L, o, d, t, tri, fl, c);
}

// The hit record of a mesh walk's nearest triangle (normal oriented
// against the ray), then the job's overlay boxes (WithPanels::raycast_hit).
__device__ __forceinline__ bool mesh_hit(const JobView& A, f3 p, f3 d, bool found, float t,
                                         uint32_t tri, uint32_t fl, Hit& h)
{
    if (found) {
        float4 ti = *reinterpret_cast<const OMG_GLOBAL float4*>(glb(A.g.tris) + tri);
        f3 n = mk(ti.x, ti.y, ti.z);
        if (dot(n, d) > 0.0f) n = mul(n, -1.0f);
        h.t = t;
        h.n = n;
        h.mat = __float_as_uint(ti.w);
        h.id = tri;
        h.flags = fl;
    }
    if (A.n_ovl) overlay_hits(A, p, d, found, h);
    return found;
}

template <bool MESH, bool COUNT, int SK, class Trav>
__device__ __forceinline__ bool raycast_hit(const JobView& A, const Trav& L, f3 p, f3 d,
                                            Hit& h, TravCount& c)
{
    if (MESH) {
        float t;
        uint32_t tri, fl;
        bool found = mesh_raycast_any<COUNT, SK>(A.g, L, p, d, t, tri, fl, c);
        return mesh_hit(A, p, d, found, t, tri, fl, h);
    }
    bool found = box_raycast_hit(A.g, p, d, h);
    if (A.n_ovl) overlay_hits(A, p, d, found, h);
    return found;
}

// Traversal state type by lanes per ray and BVH width.
template <int LANES, int W>
struct TravOf {
    using type = LdsTrav;
};
template <>
struct TravOf<4, 2> {
    using type = QuadTrav;
};
template <>
struct TravOf<1, 4> {
    using type = Lds4Trav;
};
template <>
struct TravOf<4, 4> {
    using type = Quad4Trav;
};
template <>
struct TravOf<16, 16> {
    using type = Row16Trav;
};
template <>
struct TravOf<8, 16> {
    using type = Row8Trav;
};

// float4s per node of width W (64 B paired node, 128 B 4-wide node, 512 B 16-wide node)
template <int W>
constexpr uint32_t node_f4() { return W == 16 ? 32u : W == 4 ? 8u : 4u; }

template <class T>
__device__ __forceinline__ void init_lane_trav(T& L, lds_float4* lds, uint32_t nc, uint32_t nf4, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.st_ref = reinterpret_cast<lds_u32*>(lds + nf4 * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * blockDim.x);
    L.lane = threadIdx.x;
    L.bs = blockDim.x;
}
template <class T>
__device__ __forceinline__ void init_quad_trav(T& L, lds_float4* lds, uint32_t nc, uint32_t nf4, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.nq = blockDim.x >> 2;
    L.nq_log2 = 29u - (uint32_t)__builtin_clz(blockDim.x);
    L.st_ref = reinterpret_cast<lds_u32*>(lds + nf4 * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * L.nq);
    L.quad = threadIdx.x >> 2;
    L.q = threadIdx.x & 3u;
}
__device__ __forceinline__ void init_trav(LdsTrav& L, lds_float4* lds, uint32_t nc, int sk)
{
    init_lane_trav(L, lds, nc, 4u, sk);
}
__device__ __forceinline__ void init_trav(Lds4Trav& L, lds_float4* lds, uint32_t nc, int sk)
{
    init_lane_trav(L, lds, nc, 8u, sk);
    L.bs_log2 = 31u - (uint32_t)__builtin_clz(blockDim.x);
}
__device__ __forceinline__ void init_trav(QuadTrav& L, lds_float4* lds, uint32_t nc, int sk)
{
    init_quad_trav(L, lds, nc, 4u, sk);
}
__device__ __forceinline__ void init_trav(Quad4Trav& L, lds_float4* lds, uint32_t nc, int sk)
{
    init_quad_trav(L, lds, nc, 8u, sk);
}
__device__ __forceinline__ void init_trav(Row16Trav& L, lds_float4* lds, uint32_t nc, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.nrow = blockDim.x >> 4;
    L.nrow_log2 = 27u - (uint32_t)__builtin_clz(blockDim.x);
    L.st_ref = reinterpret_cast<lds_u32*>(lds + 32u * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * L.nrow);
    L.row = threadIdx.x >> 4;
    L.r = threadIdx.x & 15u;
}
__device__ __forceinline__ void init_trav(Row8Trav& L, lds_float4* lds, uint32_t nc, int sk)
{
    L.nc = nc;
    L.nodes = lds;
    L.nrow = blockDim.x >> 3;
    L.nrow_log2 = 28u - (uint32_t)__builtin_clz(blockDim.x);
    L.st_ref = reinterpret_cast<lds_u32*>(lds + 32u * nc);
    L.st_t = reinterpret_cast<lds_f32*>(L.st_ref + sk * L.nrow);
    L.row = threadIdx.x >> 3;
    L.r = threadIdx.x & 7u;
}

// State of one ray between bounces (tracer.rs:92-102).
template <int NBM>
struct RayState {
    f3 pos, dir;
    float energy[NBM];
    float dist_total;
    uint32_t bounce;  // bounce-loop iterations so far
    uint32_t term, hash, rflags;
    uint32_t src, ray;
};

// Ray start (tracer.rs:92-102): source position, uniform direction from
// the Philox block (ray, 0, source), energy per band.
template <int NBM>
__device__ __forceinline__ void ray_begin(const JobView& A, uint32_t src, uint32_t ray,
                                          RayState<NBM>& S)
{
    S.src = src;
    S.ray = ray;
    const OMG_GLOBAL float* sp = glb(A.sources) + 3 * src;
    S.pos = mk(sp[0], sp[1], sp[2]);
    uint32_t w[4];
    philox4x32_10(ray, 0u, src, 0u, A.key0, A.key1, w);
    S.dir = unit_sphere(u24(w[0]), u24(w[1]));
#pragma unroll
    for (int b = 0; b < NBM; b++) S.energy[b] = A.energy0[b];
    S.dist_total = 0.0f;
    S.bounce = 0;
    S.term = 4u;
    S.hash = 2166136261u;
    S.rflags = 0;
}

// One iteration of the bounce loop (tracer.rs:103-200): raycast, score every
// listener the segment passes (tracer.rs:110-128) into the job's
// echograms, then escape / time cut-off / transmission / absorption /
// scattering.  Returns false when the ray has ended (S.term says why; 4 =
// the bounce limit).  NBM is the compile-time band capacity, nb the job's
// band count (<= NBM; energies beyond it are zero and never scored).
// LANES = 4 runs the raycast on a DPP quad (mesh_raycast_quad); the quad's
// lanes then split the listeners between them.
template <int NBM, int LANES>
__device__ __forceinline__ bool ray_shade(const JobView& A, uint32_t q, uint32_t nb,
                                          RayState<NBM>& S, bool has, const Hit& h,
                                          uint32_t* hit_ids, float* hit_t);

template <int NBM, bool MESH, bool COUNT, int SK, int LANES, class Trav>
__device__ __forceinline__ bool ray_step(const JobView& A, const Trav& L, uint32_t q, uint32_t nb,
                                         RayState<NBM>& S, uint32_t* hit_ids, float* hit_t,
                                         TravCount& tc)
{
    Hit h;
    bool has = raycast_hit<MESH, COUNT, SK>(A, L, S.pos, S.dir, h, tc);
    return ray_shade<NBM, LANES>(A, q, nb, S, has, h, hit_ids, hit_t);
}

// The rest of a bounce once the segment's hit is known (has, h): listener
// scoring, the per-bounce hit record, then the ray's fate.
template <int NBM, int LANES>
__device__ __forceinline__ bool ray_shade(const JobView& A, uint32_t q, uint32_t nb,
                                          RayState<NBM>& S, bool has, const Hit& h,
                                          uint32_t* hit_ids, float* hit_t)
{
    const uint32_t bounce = S.bounce;
    S.bounce = bounce + 1u;
    float t_hit = has ? h.t : FLT_MAX;
    const size_t eg = (size_t)A.nbins * nb, dg = (size_t)A.nbins * 3;
    for (uint32_t li = q; li < A.n_lis; li += LANES) {
        const OMG_GLOBAL float* lp = glb(A.listeners) + 3 * li;
        f3 Lp = mk(lp[0], lp[1], lp[2]);
        f3 to_l = sub(Lp, S.pos);
        float s = dot(to_l, S.dir);
        if (s > 0.0f && s < t_hit) {
            f3 closest = add(S.pos, mul(S.dir, s));
            if (length(sub(closest, Lp)) < A.radius) {
                float arrival = (S.dist_total + s) / kSpeedOfSound;
                uint32_t bin = sat_u32(arrival / A.bin_dt);
                if (bin < A.nbins) {
                    size_t pair = (size_t)S.src * A.n_lis + li;
                    float* ob = A.bins + pair * eg + (size_t)bin * nb;
                    float* od = A.dirs + pair * dg + (size_t)bin * 3;
#pragma unroll
                    for (int b = 0; b < NBM; b++)
                        if ((uint32_t)b < nb) atomic_add_f32(ob + b, S.energy[b]);
                    float ed = S.energy[0];
#pragma unroll
                    for (int b = 1; b < NBM; b++)
                        if ((uint32_t)b == A.dir_band) ed = S.energy[b];
                    atomic_add_f32(od + 0, -(S.dir.x * ed));
                    atomic_add_f32(od + 1, -(S.dir.y * ed));
                    atomic_add_f32(od + 2, -(S.dir.z * ed));
                }
            }
        }
    }
    if (hit_ids && q == 0) {
        size_t slot = ((size_t)S.src * A.n_rays + S.ray) * A.max_bounces + bounce;
        glb(hit_ids)[slot] = has ? h.id : 0xFFFFFFFFu;
        glb(hit_t)[slot] = has ? h.t : 0.0f;
    }
    if (!has) {
        S.term = 0u;
        return false;
    }
    S.hash = fnv1a_u32(S.hash, h.id);
    S.rflags |= h.flags;
    S.pos = add(S.pos, mul(S.dir, t_hit));
    S.dist_total += t_hit;
    if (S.dist_total / kSpeedOfSound > A.max_time) {
        S.term = 1u;
        return false;
    }
    const OMG_GLOBAL MatParams* M = glb(A.mats) + h.mat;
    uint32_t w[4];
    philox4x32_10(S.ray, bounce + 1u, S.src, 0u, A.key0, A.key1, w);
    float p = M->p;
    bool transmitted = p > 0.0f && u24(w[0]) < p;
    bool alive = false;
    if (transmitted) {
#pragma unroll
        for (int b = 0; b < NBM; b++) {
            S.energy[b] *= M->tw[b];
            if ((uint32_t)b < nb && S.energy[b] > A.floor_e) alive = true;
        }
        if (!alive) {
            S.term = 2u;
            return false;
        }
        S.pos = add(S.pos, mul(S.dir, 2.0f * 1e-3f));
        return S.bounce < A.max_bounces;
    }
#pragma unroll
    for (int b = 0; b < NBM; b++) {
        S.energy[b] *= M->rw[b];
        if ((uint32_t)b < nb && S.energy[b] > A.floor_e) alive = true;
    }
    if (!alive) {
        S.term = 3u;
        return false;
    }
    f3 n = h.n;
    if (u24(w[1]) < M->scattering) {
        S.dir = normalize(add(n, unit_sphere(u24(w[2]), u24(w[3]))));
        if (dot(S.dir, n) < 1e-3f) S.dir = n;
    } else {
        S.dir = sub(S.dir, mul(n, 2.0f * dot(S.dir, n)));
    }
    S.pos = add(S.pos, mul(n, 1e-3f));
    return S.bounce < A.max_bounces;
}

// Per-ray parity record: bounce-loop count, termination, FNV-1a of the hit
// ids, near-edge / tie flags.
template <int NBM>
__device__ __forceinline__ void ray_end(const JobView& A, const RayState<NBM>& S, uint32_t q)
{
    if (A.records && q == 0) {
        uint4 r = make_uint4(S.bounce, S.term, S.hash, S.rflags);
        reinterpret_cast<OMG_GLOBAL uint4*>(glb(A.records))[(size_t)S.src * A.n_rays + S.ray] = r;
    }
}

// Saves the state of a ray that goes on after the hand-off bounce (the
// group's first lane writes it) for k_trace_resume16.
template <int NBM>
__device__ __forceinline__ void handoff_save(float4* out, uint32_t* ctl, const RayState<NBM>& S,
                                             uint32_t q)
{
    if (q != 0) return;
    const uint32_t k = atomicAdd(ctl, 1u);
    float e[kMaxBands];
#pragma unroll
    for (int b = 0; b < kMaxBands; b++) e[b] = b < NBM ? S.energy[b] : 0.0f;
    OMG_GLOBAL float4* o = glb(out) + (size_t)kHandoffF4 * k;
    o[0] = make_float4(S.pos.x, S.pos.y, S.pos.z, S.dist_total);
    o[1] = make_float4(S.dir.x, S.dir.y, S.dir.z, __uint_as_float(S.bounce));
    o[2] = make_float4(__uint_as_float(S.term), __uint_as_float(S.hash), __uint_as_float(S.rflags),
                       __uint_as_float(S.src));
    o[3] = make_float4(__uint_as_float(S.ray), e[0], e[1], e[2]);
    o[4] = make_float4(e[3], e[4], e[5], e[6]);
    o[5] = make_float4(e[7], 0.0f, 0.0f, 0.0f);
}

template <int NBM>
__device__ __forceinline__ void handoff_load(const float4* in, uint32_t k, RayState<NBM>& S)
{
    const OMG_GLOBAL float4* o = glb(in) + (size_t)kHandoffF4 * k;
    const float4 a = o[0], b = o[1], c = o[2], d = o[3], e = o[4], f = o[5];
    S.pos = mk(a.x, a.y, a.z);
    S.dist_total = a.w;
    S.dir = mk(b.x, b.y, b.z);
    S.bounce = __float_as_uint(b.w);
    S.term = __float_as_uint(c.x);
    S.hash = __float_as_uint(c.y);
    S.rflags = __float_as_uint(c.z);
    S.src = __float_as_uint(c.w);
    S.ray = __float_as_uint(d.x);
    const float en[kMaxBands] = {d.y, d.z, d.w, e.x, e.y, e.z, e.w, f.x};
#pragma unroll
    for (int k2 = 0; k2 < NBM; k2++) S.energy[k2] = en[k2];
}

// tracer::trace (tracer.rs:80-202) for one ray of one source, start to end
// (max_bounces >= 1), or, with a hand-off bounce, until that many bounces
// (the state is then saved for the 16-lane walk).  Returns the bounce-loop
// count so far.
template <int NBM, bool MESH, bool COUNT, int SK, int LANES, class Trav>
__device__ __forceinline__ uint32_t trace_ray(const JobView& A, const Trav& L, uint32_t q,
                                              uint32_t src, uint32_t ray, uint32_t nb,
                                              uint32_t* hit_ids, float* hit_t, TravCount& tc,
                                              uint32_t handoff_bounce = 0u,
                                              float4* handoff = nullptr, uint32_t* ctl = nullptr)
{
    RayState<NBM> S;
    ray_begin(A, src, ray, S);
    bool cont;
    while ((cont = ray_step<NBM, MESH, COUNT, SK, LANES>(A, L, q, nb, S, hit_ids, hit_t, tc))) {
        if (S.bounce == handoff_bounce) break;
    }
    if (cont)
        handoff_save(handoff, ctl, S, q);
    else
        ray_end(A, S, q);
    return S.bounce;
}

// Wave-level reduction of the traversal counters, one atomic per wave and counter.
__device__ __forceinline__ void flush_counters(unsigned long long* counters, uint64_t aabb,
                                               uint64_t tri, uint64_t iters)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        aabb += __shfl_xor(aabb, off);
        tri += __shfl_xor(tri, off);
        iters += __shfl_xor(iters, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(counters + 0, (unsigned long long)aabb);
        atomicAdd(counters + 1, (unsigned long long)tri);
        atomicAdd(counters + 2, (unsigned long long)iters);
    }
}

// Single-job launch: thread -> (source, ray) in source-major order.
//
// Dynamic LDS: the first nc BVH nodes (breadth-first, so the top levels),
// then the SK-deep stack tops of the block's rays.  Blocks are dispatched
// round-robin over the 8 XCDs; the block index is remapped so that each
// XCD walks a contiguous range of rays (the same sources, hence the same
// parts of the tree in its own L2).
template <int NBM, bool MESH, bool COUNT, int SK, int LANES, int W>
__global__ __launch_bounds__(256) void k_trace(TraceArgs A)
{
    extern __shared__ float4 lds[];
    typename TravOf<LANES, W>::type L;
    init_trav(L, (lds_float4*)lds, MESH ? A.node_cache : 0u, SK);
    if (MESH) {
        const float4* src4 = W == 16 ? reinterpret_cast<const float4*>(A.g.nodes16)
                             : W == 4 ? reinterpret_cast<const float4*>(A.g.nodes4)
                                      : reinterpret_cast<const float4*>(A.g.nodes);
        for (uint32_t i = threadIdx.x; i < node_f4<W>() * A.node_cache; i += blockDim.x)
            lds[i] = src4[i];
        __syncthreads();
    }
    uint32_t nb = gridDim.x, xcd = blockIdx.x & 7u, per = nb >> 3;
    uint32_t blk = (nb & 7u) == 0 ? xcd * per + (blockIdx.x >> 3) : blockIdx.x;
    uint32_t lid = (blk * blockDim.x + threadIdx.x) / LANES;
    const uint32_t q = LANES == 1 ? 0u : (threadIdx.x & (LANES - 1));
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    if (lid < A.ray_count) {
        uint32_t gid = A.ray_base + lid;
        uint32_t src = gid / A.n_rays;
        uint32_t ray = gid - src * A.n_rays;
        TravCount tc;
        uint32_t iters = trace_ray<NBM, MESH, COUNT, SK, LANES>(A, L, q, src, ray, (uint32_t)NBM,
                                                                A.hit_ids, A.hit_t, tc,
                                                                A.handoff_bounce, A.handoff,
                                                                A.handoff_ctl);
        if (COUNT && q == 0) {
            aabb_acc = tc.aabb;
            tri_acc = tc.tri;
            iters_acc = iters;
        }
    }
    if (COUNT && A.counters) flush_counters(A.counters, aabb_acc, tri_acc, iters_acc);
}

// Second stage of a handed-off plain launch: the rays that k_trace saved
// after its hand-off bounce, each continued by a row of LANES lanes over
// the 16-wide nodes until it ends (a ray's later bounces are its own: the
// counter RNG and the saved state make them the same as in one launch).
// LANES = 16: one node slot per lane; LANES = 8: two slots per lane, eight
// rays per wave.  Persistent: the grid is sized to the resident wave slots
// and a row whose ray has ended takes the next saved ray from the cursor;
// dead rows of a wave are ranked by a ballot over their first lanes.  The
// saved rays are the long-lived ones, so a row per ray cuts each bounce's
// dependent chain of node fetches (four binary levels per fetch) where the
// early, crowded bounces use a quad per ray.  The LDS stack holds the
// scene's worst-case 16-wide stack (A.resume_stack entries per row).
//
// Admission: with few saved rays only some waves find work, and which ones
// reach the cursor first decides where they run; a launch whose working
// waves pile onto a few SIMDs (up to 6 on one, each waiting on the others'
// issue) takes half again as long as one whose waves spread out.  So a wave
// works only while its SIMD has fewer than ceil(saved rays / (SIMDs x rays
// per wave)) working waves, counted in the launch's own control words
// (cleared with the rest of them before every launch, graph replays
// included); a wave turned away exits.  The first wave on every SIMD is
// admitted (when there is a saved ray at all) and admitted waves take saved
// rays until none are left, so every saved ray is continued whatever the
// placement.
__device__ __forceinline__ bool resume_admit(const TraceArgs& A, uint32_t n, uint32_t rays_per_wave)
{
    if (A.resume_simds == 0) return true;
    const uint64_t per_round = (uint64_t)A.resume_simds * rays_per_wave;
    const uint32_t cap = (uint32_t)(((uint64_t)n + per_round - 1u) / per_round);
    if (cap >= 64u) return true;
    uint32_t ok = 1;
    if ((threadIdx.x & 63u) == 0u) {
        // this wave's SIMD: XCC, shader engine, shader array, CU, SIMD
        const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 4);     // HW_ID
        const uint32_t xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);    // XCC_ID
        const uint32_t key = ((xcc & 7u) << 10) | (((hw >> 13) & 7u) << 7) | (((hw >> 12) & 1u) << 6) |
                             (((hw >> 8) & 15u) << 2) | ((hw >> 4) & 3u);
        ok = atomicAdd(A.handoff_ctl + 2u + key, 1u) < cap ? 1u : 0u;
    }
    return __builtin_amdgcn_readfirstlane(ok) != 0u;
}

template <int NBM, int SK, int LANES>
__global__ __launch_bounds__(256) void k_trace_resume16(TraceArgs A)
{
    extern __shared__ float4 lds[];
    typename TravOf<LANES, 16>::type L;
    init_trav(L, (lds_float4*)lds, A.node_cache16, (int)A.resume_stack);
    for (uint32_t i = threadIdx.x; i < 32u * A.node_cache16; i += blockDim.x)
        lds[i] = reinterpret_cast<const float4*>(A.g.nodes16)[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = threadIdx.x & (LANES - 1u);
    const uint64_t before = (1ull << (lane & ~(LANES - 1u))) - 1ull;
    const uint32_t n = *reinterpret_cast<const volatile uint32_t*>(A.handoff_ctl);
    if (!resume_admit(A, n, 64u / LANES)) return;
    RayState<NBM> S;
    uint32_t live = 0, more = 1;
    for (;;) {
        // every lane of the wave is active here
        const uint64_t need = __ballot(live == 0u && r == 0u);
        if (need != 0 && more) {
            const uint32_t k = (uint32_t)__popcll(need);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(A.handoff_ctl + 1, k);
            base = __builtin_amdgcn_readfirstlane(base);
            const uint32_t rank = (uint32_t)__popcll(need & before);
            if (live == 0u && base + rank < n) {
                handoff_load(A.handoff, base + rank, S);
                live = 1u;
            }
            if (base + k >= n) more = 0;
        }
        if (__ballot(live != 0u) == 0) break;
        if (live) {
            TravCount tc;
            bool cont = ray_step<NBM, true, false, SK, LANES>(A, L, r, (uint32_t)NBM, S, A.hit_ids,
                                                             A.hit_t, tc);
            if (!cont) {
                ray_end(A, S, r);
                live = 0u;
            }
        }
    }
}

// Batched persistent launch: the grid is sized to the resident wave slots and
// every wave steals work items (kBatchItemRays consecutive rays of one source
// of one query) from a global cursor until the list is exhausted.  Queries
// of different scenes, sizes and band counts share the launch; each ray runs
// the same trace_ray as the single-job kernel, so its path and per-ray
// record are bit-identical to the query's own omg_trace_world.  No BVH
// nodes are staged in LDS (a wave moves between scenes); the LDS holds the
// SK-deep stack tops only.
template <int NBM, bool COUNT, int SK, int LANES, int W>
__global__ __launch_bounds__(256) void k_trace_batch(BatchArgs B)
{
    extern __shared__ float4 lds[];
    typename TravOf<LANES, W>::type L;
    init_trav(L, (lds_float4*)lds, 0u, SK);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t q = LANES == 1 ? 0u : (threadIdx.x & (LANES - 1));
    constexpr uint32_t kPerPass = 64u / LANES;
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    for (;;) {
        // every lane of the wave is active here: lane 0 is the first active lane
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(B.next_item, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= B.n_items) break;
        const BatchItem item = B.items[it];
        const JobView& J = B.jobs[item.query];
        for (uint32_t base = 0; base < item.count; base += kPerPass) {
            uint32_t lid = base + lane / LANES;
            if (lid < item.count) {
                TravCount tc;
                uint32_t iters = trace_ray<NBM, true, COUNT, SK, LANES>(
                    J, L, q, item.src, item.ray_start + lid, J.nbands, nullptr, nullptr, tc);
                if (COUNT && q == 0) {
                    aabb_acc += tc.aabb;
                    tri_acc += tc.tri;
                    iters_acc += iters;
                }
            }
        }
    }
    if (COUNT && B.counters) flush_counters(B.counters, aabb_acc, tri_acc, iters_acc);
}

// Batched persistent launch with per-lane ray regeneration: a lane whose
// ray has ended takes the next ray of its wave's current work item between
// bounces, so a wave's lanes stay busy while the rays' bounce counts differ.
// Dead lanes are ranked by a ballot (mbcnt over the dead-lane mask) and
// given consecutive rays; the wave takes a new item from the cursor when
// its item is used up.  Lanes of one wave may then hold rays of different
// queries: the job view is read per lane.  One lane per ray.
// Register budget: with 8 stack entries in LDS the kernel is held to 64
// VGPRs (8 waves per SIMD), otherwise to 6 waves per SIMD; the few values
// that no longer fit spill to scratch, which costs less than the lost
// occupancy of this latency-bound walk.
template <int NBM, bool COUNT, int SK, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SK <= 8 ? 8 : 6)))
void k_trace_batch_regen(BatchArgs B)
{
    extern __shared__ float4 lds[];
    typename TravOf<1, W>::type L;
    init_trav(L, (lds_float4*)lds, 0u, SK);
    const uint32_t lane = threadIdx.x & 63u;
    RayState<NBM> S;
    uint32_t live = 0, qi = 0;
    // the wave's current item: rays [it_next, it_end) of (it_q, it_src); uniform
    uint32_t it_q = 0, it_src = 0, it_next = 0, it_end = 0, more = 1;
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    for (;;) {
        // every lane of the wave is active here: lane 0 is the first active lane
        uint64_t need = __ballot(live == 0u);
        while (need != 0 && (it_next < it_end || more)) {
            if (it_next == it_end) {
                uint32_t it = 0;
                if (lane == 0) it = atomicAdd(B.next_item, 1u);
                it = __builtin_amdgcn_readfirstlane(it);
                if (it >= B.n_items) {
                    more = 0;
                    break;
                }
                const BatchItem I = B.items[it];
                it_q = I.query;
                it_src = I.src;
                it_next = I.ray_start;
                it_end = I.ray_start + I.count;
            }
            const uint32_t avail = it_end - it_next;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
            if (live == 0u && rank < avail) {
                ray_begin(B.jobs[it_q], it_src, it_next + rank, S);
                qi = it_q;
                live = 1u;
            }
            const uint32_t k = (uint32_t)__popcll(need);
            it_next += k < avail ? k : avail;
            need = __ballot(live == 0u);
        }
        if (__ballot(live != 0u) == 0) break;
        if (live) {
            const JobView& J = B.jobs[qi];
            TravCount tc;
            bool cont = ray_step<NBM, true, COUNT, SK, 1>(J, L, 0u, J.nbands, S, nullptr, nullptr, tc);
            if (COUNT) {
                aabb_acc += tc.aabb;
                tri_acc += tc.tri;
            }
            if (!cont) {
                ray_end(J, S, 0u);
                if (COUNT) iters_acc += S.bounce;
                live = 0u;
            }
        }
    }
    if (COUNT && B.counters) flush_counters(B.counters, aabb_acc, tri_acc, iters_acc);
}

// Regenerating launch with the bounce loop cut at walk steps: a lane whose
// walk has ended does not wait for the wave's longest walk.  Each iteration
// the wave takes one step of every lane's 4-wide walk; once at least T
// lanes of the wave have an ended walk or no ray (or no lane walks), those
// lanes shade their segment (ray_shade, the same code as ray_step's second
// half) and start their next walk, of the same ray or of a new one taken
// from the wave's item.  A lane's walk, hit and shading are those of
// k_trace_batch_regen, so every ray is bit-identical to it; only the
// interleaving of the lanes' work differs.  One lane per ray.
template <int NBM, bool COUNT, int SK, int T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
void k_trace_batch_regen_il(BatchArgs B)
{
    extern __shared__ float4 lds[];
    Lds4Trav L;
    init_trav(L, (lds_float4*)lds, 0u, SK);
    const uint32_t lane = threadIdx.x & 63u;
    RayState<NBM> S;
    Walk4 Wk;
    GeomArgs G{};  // the walking ray's geometry (node and primitive arrays)
    uint32_t ovf_ref[kStackCap - SK];
    float ovf_t[kStackCap - SK];
    // mode 0: no ray; 1: walking; 2: walk ended, segment not shaded yet
    uint32_t mode = 0, qi = 0;
    uint32_t it_q = 0, it_src = 0, it_next = 0, it_end = 0, more = 1;
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    for (;;) {
        // every lane of the wave is active here
        const uint64_t ended = __ballot(mode == 2u);
        const uint64_t idle = __ballot(mode == 0u);
        const uint64_t walking = __ballot(mode == 1u);
        const bool refill = more || it_next < it_end;
        const uint32_t waiting = (uint32_t)__popcll(ended) + (refill ? (uint32_t)__popcll(idle) : 0u);
        if (walking == 0 || waiting >= (uint32_t)T) {
            if (mode == 2u) {
                const JobView& J = B.jobs[qi];
                float t;
                uint32_t tri, fl;
                Hit h;
                const bool found = Wk.H.result(t, tri, fl);
                const bool has = mesh_hit(J, S.pos, S.dir, found, t, tri, fl, h);
                if (ray_shade<NBM, 1>(J, 0u, J.nbands, S, has, h, nullptr, nullptr)) {
                    mode = walk4_begin(J.g, L, Wk, S.dir) ? 1u : 2u;
                } else {
                    ray_end(J, S, 0u);
                    if (COUNT) iters_acc += S.bounce;
                    mode = 0u;
                }
            }
            uint64_t need = __ballot(mode == 0u);
            while (need != 0 && (it_next < it_end || more)) {
                if (it_next == it_end) {
                    uint32_t it = 0;
                    if (lane == 0) it = atomicAdd(B.next_item, 1u);
                    it = __builtin_amdgcn_readfirstlane(it);
                    if (it >= B.n_items) {
                        more = 0;
                        break;
                    }
                    const BatchItem I = B.items[it];
                    it_q = I.query;
                    it_src = I.src;
                    it_next = I.ray_start;
                    it_end = I.ray_start + I.count;
                }
                const uint32_t avail = it_end - it_next;
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
                if (mode == 0u && rank < avail) {
                    const JobView& J = B.jobs[it_q];
                    ray_begin(J, it_src, it_next + rank, S);
                    qi = it_q;
                    mode = walk4_begin(J.g, L, Wk, S.dir) ? 1u : 2u;
                }
                const uint32_t k = (uint32_t)__popcll(need);
                it_next += k < avail ? k : avail;
                need = __ballot(mode == 0u);
            }
            if (__ballot(mode != 0u) == 0) break;
            if (mode == 1u) {
                const JobView& J = B.jobs[qi];
                G.nodes4 = J.g.nodes4;
                G.prims = J.g.prims;
            }
        }
        if (mode == 1u) {
            TravCount tc;
            if (!walk4_step<COUNT, SK>(G, L, Wk, S.pos, S.dir, ovf_ref, ovf_t, tc)) mode = 2u;
            if (COUNT) {
                aabb_acc += tc.aabb;
                tri_acc += tc.tri;
            }
        }
    }
    if (COUNT && B.counters) flush_counters(B.counters, aabb_acc, tri_acc, iters_acc);
}

// Multi-lane variant of the regenerating launch: a group of LANES lanes
// runs each ray (a DPP quad on the 4-wide or binary tree: the 4-wide walk
// tests a node's four children on the quad's lanes, the binary walk splits
// the pair tests; a 16-lane row on the 16-wide tree, one slot per lane;
// the lanes share the listeners), and a group whose ray has ended takes
// the next ray of its wave's item.  Ranks go to the dead groups by the
// first lane of each group: a group's rank is the number of dead groups
// before it in the wave.  On the 4-wide and 16-wide trees the whole
// traversal stack (kDeepStack / kStack16 entries per group) lives in LDS.
template <int NBM, bool COUNT, int SK, int W, int LANES>
__global__ __launch_bounds__(256) void k_trace_batch_regenL(BatchArgs B)
{
    extern __shared__ float4 lds[];
    typename TravOf<LANES, W>::type L;
    init_trav(L, (lds_float4*)lds, 0u, SK);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t q = threadIdx.x & (LANES - 1u);
    const uint64_t before = (1ull << (lane & ~(LANES - 1u))) - 1ull;  // lanes ahead of this group
    RayState<NBM> S;
    uint32_t live = 0, qi = 0;
    uint32_t it_q = 0, it_src = 0, it_next = 0, it_end = 0, more = 1;
    uint64_t aabb_acc = 0, tri_acc = 0, iters_acc = 0;
    for (;;) {
        // every lane of the wave is active here: lane 0 is the first active lane
        uint64_t need = __ballot(live == 0u && q == 0u);
        while (need != 0 && (it_next < it_end || more)) {
            if (it_next == it_end) {
                uint32_t it = 0;
                if (lane == 0) it = atomicAdd(B.next_item, 1u);
                it = __builtin_amdgcn_readfirstlane(it);
                if (it >= B.n_items) {
                    more = 0;
                    break;
                }
                const BatchItem I = B.items[it];
                it_q = I.query;
                it_src = I.src;
                it_next = I.ray_start;
                it_end = I.ray_start + I.count;
            }
            const uint32_t avail = it_end - it_next;
            const uint32_t rank = (uint32_t)__popcll(need & before);
            if (live == 0u && rank < avail) {
                ray_begin(B.jobs[it_q], it_src, it_next + rank, S);
                qi = it_q;
                live = 1u;
            }
            const uint32_t k = (uint32_t)__popcll(need);
            it_next += k < avail ? k : avail;
            need = __ballot(live == 0u && q == 0u);
        }
        if (__ballot(live != 0u) == 0) break;
        if (live) {
            const JobView& J = B.jobs[qi];
            TravCount tc;
            bool cont = ray_step<NBM, true, COUNT, SK, LANES>(J, L, q, J.nbands, S, nullptr, nullptr, tc);
            if (COUNT && q == 0) {
                aabb_acc += tc.aabb;
                tri_acc += tc.tri;
            }
            if (!cont) {
                ray_end(J, S, q);
                if (COUNT && q == 0) iters_acc += S.bounce;
                live = 0u;
            }
        }
    }
    if (COUNT && B.counters) flush_counters(B.counters, aabb_acc, tri_acc, iters_acc);
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

// 16-byte copy, grid-stride: moves the serving loop's batch tables and
// results between mapped pinned host memory and the device on the batch's
// own stream (no DMA engine in the request's path)
__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

constexpr uint32_t kTraceBlock = 256;
// the walks' LDS stacks index depth x (groups per block) by a shift
static_assert((kTraceBlock & (kTraceBlock - 1u)) == 0u, "trace blocks are a power of two");

template <int NBM, bool MESH, bool COUNT, int SK, int LANES, int W>
static hipError_t launch_trace_l(const TraceArgs& A, hipStream_t s)
{
    uint64_t threads = (uint64_t)A.ray_count * LANES;
    uint32_t blocks = (uint32_t)((threads + kTraceBlock - 1) / kTraceBlock);
    if (blocks == 0) return hipSuccess;
    size_t lds = (MESH ? (size_t)A.node_cache * 16 * node_f4<W>() : 0) +
                 (size_t)SK * (kTraceBlock / LANES) * 8;
    auto kern = k_trace<NBM, MESH, COUNT, SK, LANES, W>;
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
    if (MESH && A.lanes == 4) return launch_trace_l<NBM, MESH, COUNT, SK, 4, 2>(A, s);
    return launch_trace_l<NBM, MESH, COUNT, SK, 1, 2>(A, s);
}

// 4-wide walks keep 16 stack entries per ray in LDS, or, on a quad, the
// whole worst-case stack (kDeepStack entries)
template <int NBM, bool COUNT>
static hipError_t launch_trace_w4(const TraceArgs& A, hipStream_t s)
{
    if (A.lanes == 4) {
        if (A.stack_lds == kDeepStack) return launch_trace_l<NBM, true, COUNT, kDeepStack, 4, 4>(A, s);
        return launch_trace_l<NBM, true, COUNT, 16, 4, 4>(A, s);
    }
    return launch_trace_l<NBM, true, COUNT, 16, 1, 4>(A, s);
}

template <int NBM, bool MESH, bool COUNT>
static hipError_t launch_trace_sk(const TraceArgs& A, hipStream_t s)
{
    if (!MESH) return launch_trace_t<NBM, MESH, COUNT, 1>(A, s);
    if (A.width == 16) {
        if (!A.g.nodes16 || A.lanes != 16) return hipErrorInvalidValue;
        return launch_trace_l<NBM, true, COUNT, kStack16, 16, 16>(A, s);
    }
    if (A.width == 4) {
        if (!A.g.nodes4) return hipErrorInvalidValue;
        return launch_trace_w4<NBM, COUNT>(A, s);
    }
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

template <int NBM>
static hipError_t launch_resume16_n(const TraceArgs& A, hipStream_t s)
{
    if (A.resume_stack == 0 || A.resume_stack > kStack16) return hipErrorInvalidValue;
    const uint32_t lanes = A.resume_lanes == 8 ? 8u : 16u;
    size_t lds = (size_t)A.node_cache16 * sizeof(Node16) +
                 (size_t)A.resume_stack * (kTraceBlock / lanes) * 8;
    if (lanes == 8)
        hipLaunchKernelGGL((k_trace_resume16<NBM, kStack16, 8>), dim3(A.resume_blocks),
                           dim3(kTraceBlock), lds, s, A);
    else
        hipLaunchKernelGGL((k_trace_resume16<NBM, kStack16, 16>), dim3(A.resume_blocks),
                           dim3(kTraceBlock), lds, s, A);
    return hipGetLastError();
}

static hipError_t launch_resume16(const TraceArgs& A, hipStream_t s)
{
    switch (A.nbands) {
    case 1: return launch_resume16_n<1>(A, s);
    case 2: return launch_resume16_n<2>(A, s);
    case 3: return launch_resume16_n<3>(A, s);
    case 4: return launch_resume16_n<4>(A, s);
    case 5: return launch_resume16_n<5>(A, s);
    case 6: return launch_resume16_n<6>(A, s);
    case 7: return launch_resume16_n<7>(A, s);
    case 8: return launch_resume16_n<8>(A, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace(const TraceArgs& A, bool mesh, bool count, hipStream_t s)
{
    hipError_t e;
    if (mesh) e = count ? launch_trace_nb<true, true>(A, s) : launch_trace_nb<true, false>(A, s);
    else e = count ? launch_trace_nb<false, true>(A, s) : launch_trace_nb<false, false>(A, s);
    if (e != hipSuccess || !mesh || count || A.handoff_bounce == 0) return e;
    if (!A.g.nodes16 || !A.handoff || !A.handoff_ctl || A.resume_blocks == 0) return hipErrorInvalidValue;
    return launch_resume16(A, s);
}

hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return hipSuccess;
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return hipErrorInvalidValue;
    size_t n = bytes / 16;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_copy16, dim3((uint32_t)blocks), dim3(256), 0, s, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), n);
    return hipGetLastError();
}

hipError_t launch_fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero_u64, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace omg

namespace omg {

template <int NBM, bool COUNT, int SK, int LANES, int W>
static hipError_t launch_batch_l(const BatchArgs& B, uint32_t blocks, hipStream_t s)
{
    size_t lds = (size_t)SK * (kTraceBlock / LANES) * 8;
    hipLaunchKernelGGL((k_trace_batch<NBM, COUNT, SK, LANES, W>), dim3(blocks), dim3(kTraceBlock),
                       lds, s, B);
    return hipGetLastError();
}

template <int NBM, bool COUNT, int SK>
static hipError_t launch_batch_regen(const BatchArgs& B, uint32_t width, uint32_t blocks,
                                     hipStream_t s)
{
    size_t lds = (size_t)SK * kTraceBlock * 8;
    if (width == 4)
        hipLaunchKernelGGL((k_trace_batch_regen<NBM, COUNT, SK, 4>), dim3(blocks), dim3(kTraceBlock),
                           lds, s, B);
    else
        hipLaunchKernelGGL((k_trace_batch_regen<NBM, COUNT, SK, 2>), dim3(blocks), dim3(kTraceBlock),
                           lds, s, B);
    return hipGetLastError();
}

// regen 2: the walk-step regenerating kernel, shading once 48 lanes of a
// wave wait (24, 32 and 40 measured within 1 % of it on the mixed 256-query
// workload)
template <int NBM, bool COUNT, int SK>
static hipError_t launch_batch_regen_il(const BatchArgs& B, uint32_t blocks, hipStream_t s)
{
    size_t lds = (size_t)SK * kTraceBlock * 8;
    hipLaunchKernelGGL((k_trace_batch_regen_il<NBM, COUNT, SK, 48>), dim3(blocks),
                       dim3(kTraceBlock), lds, s, B);
    return hipGetLastError();
}

template <int NBM, bool COUNT, int SK, int W, int LANES>
static hipError_t launch_batch_regenL(const BatchArgs& B, uint32_t blocks, hipStream_t s)
{
    size_t lds = (size_t)SK * (kTraceBlock / LANES) * 8;
    hipLaunchKernelGGL((k_trace_batch_regenL<NBM, COUNT, SK, W, LANES>), dim3(blocks),
                       dim3(kTraceBlock), lds, s, B);
    return hipGetLastError();
}

template <int NBM, bool COUNT>
static hipError_t launch_batch_nb(const BatchArgs& B, const BatchShape& sh, uint32_t blocks,
                                  hipStream_t s)
{
    if (sh.regen == 2)
        return sh.stack_lds == 16 ? launch_batch_regen_il<NBM, COUNT, 16>(B, blocks, s)
                                  : launch_batch_regen_il<NBM, COUNT, 12>(B, blocks, s);
    if (sh.regen && sh.lanes == 16) return launch_batch_regenL<NBM, COUNT, kStack16, 16, 16>(B, blocks, s);
    if (sh.regen && sh.lanes == 4)
        return sh.width == 4 ? launch_batch_regenL<NBM, COUNT, kDeepStack, 4, 4>(B, blocks, s)
                             : launch_batch_regenL<NBM, COUNT, 16, 2, 4>(B, blocks, s);
    if (sh.regen) {
        switch (sh.stack_lds) {
        case 8: return launch_batch_regen<NBM, COUNT, 8>(B, sh.width, blocks, s);
        case 12: return launch_batch_regen<NBM, COUNT, 12>(B, sh.width, blocks, s);
        case 16: return launch_batch_regen<NBM, COUNT, 16>(B, sh.width, blocks, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (sh.width == 4 && sh.lanes == 4)
        return sh.stack_lds == kDeepStack ? launch_batch_l<NBM, COUNT, kDeepStack, 4, 4>(B, blocks, s)
                                          : launch_batch_l<NBM, COUNT, 16, 4, 4>(B, blocks, s);
    if (sh.width == 4) return launch_batch_l<NBM, COUNT, 16, 1, 4>(B, blocks, s);
    if (sh.stack_lds >= 32)
        return sh.lanes == 4 ? launch_batch_l<NBM, COUNT, 32, 4, 2>(B, blocks, s)
                             : launch_batch_l<NBM, COUNT, 32, 1, 2>(B, blocks, s);
    return sh.lanes == 4 ? launch_batch_l<NBM, COUNT, 16, 4, 2>(B, blocks, s)
                         : launch_batch_l<NBM, COUNT, 16, 1, 2>(B, blocks, s);
}

// A plain job handed to the persistent kernel: its job view is written by
// a one-lane kernel (the view travels as a kernel argument, so no host
// staging outlives the call) and its work items are generated on the
// device, 64 consecutive rays of one source each, in source-major order.
__global__ void k_route_job(JobView J, JobView* out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = J;
}

__global__ void k_route_items(BatchItem* items, uint32_t n_items, uint32_t per_src, uint32_t n_rays)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    uint32_t src = i / per_src, first = (i - src * per_src) * kBatchItemRays;
    uint32_t left = n_rays - first;
    items[i] = BatchItem{0u, src, first, left < kBatchItemRays ? left : kBatchItemRays};
}

hipError_t launch_trace_routed(const TraceArgs& A, uint32_t nbm, uint32_t width, uint32_t n_cu,
                               JobView* d_job, BatchItem* d_items, uint32_t* d_cursor,
                               bool count, hipStream_t s)
{
    const uint32_t per_src = (A.n_rays + kBatchItemRays - 1) / kBatchItemRays;
    const uint64_t n_items = (uint64_t)per_src * A.n_src;
    if (n_items == 0) return hipSuccess;
    if (n_items > 0xFFFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_route_job, dim3(1), dim3(64), 0, s, static_cast<const JobView&>(A), d_job);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_route_items, dim3((uint32_t)((n_items + 255) / 256)), dim3(256), 0, s,
                       d_items, (uint32_t)n_items, per_src, A.n_rays);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_fill_zero(reinterpret_cast<float*>(d_cursor), 1, s)) != hipSuccess) return e;
    BatchArgs B{};
    B.jobs = d_job;
    B.items = d_items;
    B.n_items = (uint32_t)n_items;
    B.next_item = d_cursor;
    B.counters = A.counters;
    // the batched kernel's default shape: one lane per ray, 12 LDS stack
    // entries, 6 waves per SIMD, rays regenerated between bounces; a 16-lane
    // row per ray over the 16-wide tree when that is the width walked
    BatchShape sh = width == 16 ? BatchShape{16u, kStack16, 6u, 1u, 16u}
                    : width == 4 ? BatchShape{1u, 12u, 6u, 2u, 4u}
                                 : BatchShape{1u, 12u, 6u, 1u, width};
    return launch_trace_batch(B, nbm, sh, n_cu, count, s, 0);
}

// Band capacities compiled for batches: 3, 6 and 8 (a batch runs at the
// smallest one holding its widest query).
hipError_t launch_trace_batch(const BatchArgs& B, uint32_t nbm, const BatchShape& sh,
                              uint32_t n_cu, bool count, hipStream_t s, uint32_t max_blocks)
{
    if (B.n_items == 0) return hipSuccess;
    if (!valid_batch_shape(sh)) return hipErrorInvalidValue;
    // 4 SIMDs per CU and 4 waves per 256-lane block: one block per CU per wave slot,
    // fewer when the batch has less work than that (concurrent batches then co-run)
    uint32_t blocks = n_cu * sh.waves_per_simd;
    if (max_blocks && max_blocks < blocks) blocks = max_blocks;
    if (nbm <= 3) return count ? launch_batch_nb<3, true>(B, sh, blocks, s) : launch_batch_nb<3, false>(B, sh, blocks, s);
    if (nbm <= 6) return count ? launch_batch_nb<6, true>(B, sh, blocks, s) : launch_batch_nb<6, false>(B, sh, blocks, s);
    if (nbm <= 8) return count ? launch_batch_nb<8, true>(B, sh, blocks, s) : launch_batch_nb<8, false>(B, sh, blocks, s);
    return hipErrorInvalidValue;
}

}  // namespace omg
