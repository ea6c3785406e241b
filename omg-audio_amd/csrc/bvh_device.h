// Device-side BVH queries shared by the tracer, dome and image-source kernels.
//
// All traversals walk the paired 64 B node layout and replay the reference's
// visiting order exactly (crates/omg-core/src/mesh.rs:484-604), so every
// per-ray decision matches the CPU oracle bit for bit.
#pragma once
#include "trace_kernels.h"

#pragma clang fp contract(off)

namespace omg {

// Global memory pointers.  A pointer that a kernel reads out of memory (the
// geometry and tables of a batch's job views) is generic to the compiler,
// which then emits FLAT accesses: they wait on both the vector-memory and
// the LDS counters and cannot use the global addressing modes.  Every
// buffer behind these pointers is device memory, so the casts are exact.
// (The qualifier exists in the device pass only: the host pass parses these
// functions without a global address space to convert to.)
#if defined(__HIP_DEVICE_COMPILE__)
#define OMG_GLOBAL __attribute__((address_space(1)))
#else
#define OMG_GLOBAL
#endif
template <class T>
__device__ __forceinline__ const OMG_GLOBAL T* glb(const T* p)
{
    return (const OMG_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ OMG_GLOBAL T* glb(T* p)
{
    return (OMG_GLOBAL T*)p;
}
// Hardware f32 atomic add on device memory (no return value used).
__device__ __forceinline__ void atomic_add_f32(float* p, float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    (void)__builtin_amdgcn_global_atomic_fadd_f32(glb(p), v);
#endif
}

// One slab of the box test: narrows [t0, t1] to the box's extent along one axis.
__device__ __forceinline__ void ray_slab(float lo, float hi, float o, float inv, float& t0,
                                         float& t1)
{
    float ta = (lo - o) * inv, tb = (hi - o) * inv;
    if (ta > tb) { float s = ta; ta = tb; tb = s; }
    t0 = fmaxf(t0, ta);
    t1 = fminf(t1, tb);
}

// Mesh::ray_aabb (mesh.rs:409-426) on one box; t_enter is the entry
// distance.  The reference returns at the first axis where t0 > t1; here
// the three slabs are always narrowed and t0 > t1 is tested once.  Both
// give the same answer and the same t_enter: t0 starts at 0 and t1 at tmax
// (never NaN), fmaxf/fminf drop a NaN operand, so t0 never decreases and
// t1 never increases, and once t0 > t1 holds it holds to the end.
__device__ __forceinline__ bool ray_aabb(const float bmin[3], const float bmax[3], f3 o, f3 inv,
                                         float tmax, float& t_enter)
{
    float t0 = 0.0f, t1 = tmax;
    ray_slab(bmin[0], bmax[0], o.x, inv.x, t0, t1);
    ray_slab(bmin[1], bmax[1], o.y, inv.y, t0, t1);
    ray_slab(bmin[2], bmax[2], o.z, inv.z, t0, t1);
    t_enter = t0;
    return !(t0 > t1);
}

// Mesh::ray_tri_packed (mesh.rs:437-456): two-sided Moller-Trumbore.  The
// reference returns at the first failed test; here every quantity is
// computed with the same operations and the four tests are combined at the
// end, which gives the same answer and the same t, u, v without branches
// (t_out, u_out, v_out are written either way).
__device__ __forceinline__ bool ray_tri(const Prim& p, f3 o, f3 d, float& t_out, float& u_out,
                                        float& v_out)
{
    f3 e1 = mk(p.e1[0], p.e1[1], p.e1[2]);
    f3 e2 = mk(p.e2[0], p.e2[1], p.e2[2]);
    f3 pv = cross(d, e2);
    float det = dot(e1, pv);
    float inv = 1.0f / det;
    f3 s = sub(o, mk(p.a[0], p.a[1], p.a[2]));
    float u = dot(s, pv) * inv;
    f3 q = cross(s, e1);
    float v = dot(d, q) * inv;
    float t = dot(e2, q) * inv;
    const bool ok = !(fabsf(det) < 1e-9f) & (u >= -1e-6f) & (u <= 1.0f + 1e-6f) &
                    !((v < -1e-6f) | (u + v > 1.0f + 1e-6f)) & (t > 1e-5f);
    t_out = t;
    u_out = u;
    v_out = v;
    return ok;
}

__device__ __forceinline__ Prim load_prim(const Prim* prims, uint32_t k)
{
    const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(prims + k));
    float4 a = q[0], b = q[1], c = q[2];
    Prim p;
    p.a[0] = a.x; p.a[1] = a.y; p.a[2] = a.z; p.tri = __float_as_uint(a.w);
    p.e1[0] = b.x; p.e1[1] = b.y; p.e1[2] = b.z; p.mat = __float_as_uint(b.w);
    p.e2[0] = c.x; p.e2[1] = c.y; p.e2[2] = c.z; p.surf = __float_as_uint(c.w);
    return p;
}

struct TravCount {
    uint32_t aabb = 0, tri = 0;
};

constexpr uint32_t kNoHit = 0xFFFFFFFFu;

// Nearest-hit state of one raycast.  "No hit yet" is btri == kNoHit, an
// integer kept in a VGPR like the rest of the state: a bool here becomes a
// lane mask in SGPRs, which the compiler failed to reset for every lane
// between two raycasts of the quad traversal.
struct BestHit {
    float best = FLT_MAX;
    uint32_t btri = kNoHit;
    // the result flags of the winner: bit0 its hit lies within 1e-6
    // barycentric of an edge, bit1 another triangle was hit at exactly its
    // distance (three registers instead of keeping the winner's u and v)
    uint32_t fl = 0;

    __device__ __forceinline__ bool found() const { return btri != kNoHit; }
    __device__ __forceinline__ static uint32_t edge_flag(float u, float v)
    {
        return (u < 1e-6f || v < 1e-6f || u + v > 1.0f - 1e-6f) ? 1u : 0u;
    }
    // strict `t < best`; an exact-t tie with another triangle is flagged.
    // Offered only when `take`; written with selects, not branches.
    __device__ __forceinline__ void offer_if(bool take, float t, uint32_t tri, float u, float v)
    {
        const bool better = take & (!found() | (t < best));
        const bool same = take & !better & (t == best) & (tri != btri);
        best = better ? t : best;
        btri = better ? tri : btri;
        fl = better ? edge_flag(u, v) : (same ? (fl | 2u) : fl);
    }
    // flags: bit0 winning hit within 1e-6 barycentric of an edge, bit1 exact-t tie
    __device__ __forceinline__ bool result(float& t_out, uint32_t& tri_out, uint32_t& flags) const
    {
        if (!found()) return false;
        t_out = best;
        tri_out = btri;
        flags = fl;
        return true;
    }
};

// Mesh::raycast (mesh.rs:484-541): ordered traversal, near child popped
// first (ties to the left child), strict `t < best`, pop-time cull on
// `t_enter >= best`.  flags: bit0 winning hit within 1e-6 barycentric of
// an edge, bit1 exact-t tie with another original triangle.
template <bool COUNT>
__device__ __forceinline__ bool mesh_raycast(const GeomArgs& G, f3 o, f3 d, float& t_out, uint32_t& tri_out,
                             uint32_t& flags, TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t st_ref[kStackCap];
    float st_t[kStackCap];
    int sp = 1;
    st_ref[0] = G.root_ref;
    st_t[0] = 0.0f;
    BestHit H;
    while (sp > 0) {
        sp--;
        uint32_t ref = st_ref[sp];
        float te = st_t[sp];
        if (H.found() && te >= H.best) continue;
        if (ref & kLeafBit) {
            uint32_t start = ref & kLeafStartMask;
            uint32_t cnt_p = (ref >> kLeafCountShift) & 7u;
            for (uint32_t k = start; k < start + cnt_p; k++) {
                Prim p = load_prim(G.prims, k);
                if (COUNT) cnt.tri++;
                float t, u, v;
                H.offer_if(ray_tri(p, o, d, t, u, v), t, p.tri, u, v);
            }
        } else {
            const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes) + ref);
            float4 n0 = q[0], n1 = q[1], n2 = q[2], n3 = q[3];
            const float lmin[3] = {n0.x, n0.y, n0.z}, lmax[3] = {n1.x, n1.y, n1.z};
            const float rmin[3] = {n2.x, n2.y, n2.z}, rmax[3] = {n3.x, n3.y, n3.z};
            uint32_t lref = __float_as_uint(n0.w), rref = __float_as_uint(n1.w);
            float tmax = H.found() ? H.best : FLT_MAX;
            float a = 0.0f, b = 0.0f;
            if (COUNT) cnt.aabb += 2;
            bool hl = ray_aabb(lmin, lmax, o, inv, tmax, a);
            bool hr = ray_aabb(rmin, rmax, o, inv, tmax, b);
            if (hl && hr) {
                bool lnear = a <= b;
                st_ref[sp] = lnear ? rref : lref;
                st_t[sp] = lnear ? b : a;
                st_ref[sp + 1] = lnear ? lref : rref;
                st_t[sp + 1] = lnear ? a : b;
                sp += 2;
            } else if (hl) {
                st_ref[sp] = lref;
                st_t[sp] = a;
                sp += 1;
            } else if (hr) {
                st_ref[sp] = rref;
                st_t[sp] = b;
                sp += 1;
            }
        }
    }
    return H.result(t_out, tri_out, flags);
    return true;
}


// Pointers into LDS carry address space 3 explicitly.  With generic
// pointers the compiler may merge an LDS access and a scratch/global access
// of the two sides of a branch into one FLAT access through a selected
// pointer; a FLAT read of LDS is not ordered after an earlier DS write of
// the same wave, so a stack entry pushed with ds_write could be popped stale.
#define OMG_LDS __attribute__((address_space(3)))
typedef OMG_LDS float4 lds_float4;
typedef OMG_LDS uint32_t lds_u32;
typedef OMG_LDS float lds_f32;

__device__ __forceinline__ float4 lds_ld4(const lds_float4* p)
{
    const lds_f32* f = reinterpret_cast<const lds_f32*>(p);
    return make_float4(f[0], f[1], f[2], f[3]);
}

// LDS working set of one block for mesh_raycast_lds: a prefix of the node
// array (the top levels, nodes are stored breadth-first) and the top SK
// entries of every lane's traversal stack, laid out [depth][lane] so that
// lanes at the same depth hit distinct banks.
struct LdsTrav {
    const lds_float4* nodes;   // [nc][4]
    uint32_t nc;
    lds_u32* st_ref;           // [SK][bs]
    lds_f32* st_t;             // [SK][bs]
    uint32_t lane, bs;
};

__device__ __forceinline__ void load_node(const GeomArgs& G, const LdsTrav& L, uint32_t ref,
                                          float4& n0, float4& n1, float4& n2, float4& n3)
{
    if (ref < L.nc) {
        const lds_float4* q = L.nodes + 4 * ref;
        n0 = lds_ld4(q);
        n1 = lds_ld4(q + 1);
        n2 = lds_ld4(q + 2);
        n3 = lds_ld4(q + 3);
    } else {
        const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes) + ref);
        n0 = q[0]; n1 = q[1]; n2 = q[2]; n3 = q[3];
    }
}

// Mesh::raycast (mesh.rs:484-541) with the stack top in LDS.  The near child
// becomes the current node directly instead of being pushed and popped at
// once; its pop-time cull (`t_enter >= best`) is applied in the same place,
// so the visiting order, the culls and the result are the reference's.
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_lds(const GeomArgs& G, const LdsTrav& L, f3 o, f3 d,
                                                 float& t_out, uint32_t& tri_out, uint32_t& flags,
                                                 TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t ovf_ref[kStackCap - SK];
    float ovf_t[kStackCap - SK];
    int sp = 0;
    uint32_t cur = G.root_ref;
    float cur_t = 0.0f;
    BestHit H;
    while (true) {
        if (!(H.found() && cur_t >= H.best)) {
            if (cur & kLeafBit) {
                uint32_t start = cur & kLeafStartMask;
                uint32_t cnt_p = (cur >> kLeafCountShift) & 7u;
                for (uint32_t k = start; k < start + cnt_p; k++) {
                    Prim p = load_prim(G.prims, k);
                    if (COUNT) cnt.tri++;
                    float t, u, v;
                    H.offer_if(ray_tri(p, o, d, t, u, v), t, p.tri, u, v);
                }
            } else {
                float4 n0, n1, n2, n3;
                load_node(G, L, cur, n0, n1, n2, n3);
                const float lmin[3] = {n0.x, n0.y, n0.z}, lmax[3] = {n1.x, n1.y, n1.z};
                const float rmin[3] = {n2.x, n2.y, n2.z}, rmax[3] = {n3.x, n3.y, n3.z};
                uint32_t lref = __float_as_uint(n0.w), rref = __float_as_uint(n1.w);
                float tmax = H.found() ? H.best : FLT_MAX;
                float a = 0.0f, b = 0.0f;
                if (COUNT) cnt.aabb += 2;
                bool hl = ray_aabb(lmin, lmax, o, inv, tmax, a);
                bool hr = ray_aabb(rmin, rmax, o, inv, tmax, b);
                if (hl && hr) {
                    bool lnear = a <= b;
                    uint32_t far_ref = lnear ? rref : lref;
                    float far_t = lnear ? b : a;
                    if (sp < SK) {
                        L.st_ref[sp * L.bs + L.lane] = far_ref;
                        L.st_t[sp * L.bs + L.lane] = far_t;
                    } else {
                        ovf_ref[sp - SK] = far_ref;
                        ovf_t[sp - SK] = far_t;
                    }
                    sp++;
                    cur = lnear ? lref : rref;
                    cur_t = lnear ? a : b;
                    continue;
                }
                if (hl || hr) {
                    cur = hl ? lref : rref;
                    cur_t = hl ? a : b;
                    continue;
                }
            }
        }
        if (sp == 0) break;
        sp--;
        if (sp < SK) {
            cur = L.st_ref[sp * L.bs + L.lane];
            cur_t = L.st_t[sp * L.bs + L.lane];
        } else {
            cur = ovf_ref[sp - SK];
            cur_t = ovf_t[sp - SK];
        }
    }
    return H.result(t_out, tri_out, flags);
}

// ------------------------------------------------------------------ quad traversal
//
// Four lanes (a DPP quad) trace one ray together.  Every lane of the quad
// holds the same traversal state; the work of one step is split: at an
// internal node lane 0 tests the left child box and lane 1 the right one, at
// a leaf lane q tests primitive q.  Results are broadcast inside the quad
// with DPP quad_perm moves and folded in the reference's order (left/right,
// then primitives in leaf order), so the outcome is the one-lane traversal's.

// Broadcast of lane Q of each quad.  Every lane of a quad is active where
// it is used (the quad shares its control flow), so bound_ctrl may be set,
// which lets the compiler fold the move into the instruction using it.
template <int Q>
__device__ __forceinline__ uint32_t qb(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, Q | (Q << 2) | (Q << 4) | (Q << 6), 0xF, 0xF,
                                              true);
}
template <int Q>
__device__ __forceinline__ float qbf(float x) { return __uint_as_float(qb<Q>(__float_as_uint(x))); }

// LDS stack of one quad: written by lane 0, read by all four (same address,
// a broadcast read).  Laid out [depth][quad].
struct QuadTrav {
    const lds_float4* nodes;   // LDS node prefix [nc][4]
    uint32_t nc;
    lds_u32* st_ref;           // [SK][quads per block]
    lds_f32* st_t;
    uint32_t quad, nq, q;      // quad index in the block, quads per block, lane in quad
    uint32_t nq_log2;          // the block size is a power of two (256)
};

// Quad reductions by DPP (xor 1, then xor 2 inside the quad).  Every lane
// of the quad is active where they are used.
constexpr int kDppXor1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);   // quad_perm [2,3,0,1]

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t quad_min(uint32_t x)
{
    const uint32_t y = dpp_u<kDppXor1>(x);
    x = x < y ? x : y;
    const uint32_t z = dpp_u<kDppXor2>(x);
    return x < z ? x : z;
}
__device__ __forceinline__ uint32_t quad_or(uint32_t x)
{
    x |= dpp_u<kDppXor1>(x);
    return x | dpp_u<kDppXor2>(x);
}

// A leaf tested by a quad (lane q: primitive q of the leaf when q < cnt_p,
// hit_q / t_q / u_q / v_q its result) folded into H exactly as offering the
// primitives one by one in leaf order does (BestHit::offer_if): with m the
// least hit distance, the first primitive at m wins when m < best (and the
// tie flag is then whether another triangle is also at m), otherwise H
// stays and its tie flag also records a hit of another triangle at best.
// Hit distances exceed 1e-5, so positive floats order as their bits.
__device__ __forceinline__ void fold_leaf(uint32_t cnt_p, uint32_t hit_q, float t_q, uint32_t tri_q,
                                          float u_q, float v_q, BestHit& H)
{
    const uint32_t q = __lane_id() & 3u;
    const bool hit = (q < cnt_p) & (hit_q != 0u);
    const uint32_t key = hit ? __float_as_uint(t_q) : 0xFFFFFFFFu;
    const uint32_t m = quad_min(key);
    if (m == 0xFFFFFFFFu) return;
    const float tm = __uint_as_float(m);
    const bool is_m = key == m;
    const uint32_t w = quad_min(is_m ? q : 4u);
    const bool is_w = q == w;
    const uint32_t wtri = quad_or(is_w ? tri_q : 0u);
    const uint32_t wedge = quad_or(is_w ? BestHit::edge_flag(u_q, v_q) : 0u);
    const uint32_t tie_m = quad_or((is_m & !is_w & (tri_q != wtri)) ? 2u : 0u);
    const uint32_t tie_b = quad_or((hit & (t_q == H.best) & (tri_q != H.btri)) ? 2u : 0u);
    const bool better = !H.found() | (tm < H.best);
    H.best = better ? tm : H.best;
    H.btri = better ? wtri : H.btri;
    H.fl = better ? (wedge | tie_m) : (H.fl | tie_b);
}

template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast_quad(const GeomArgs& G, const QuadTrav& L, f3 o, f3 d,
                                                  float& t_out, uint32_t& tri_out, uint32_t& flags,
                                                  TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    uint32_t ovf_ref[kStackCap - SK];
    float ovf_t[kStackCap - SK];
    int sp = 0;
    uint32_t cur = G.root_ref;
    float cur_t = 0.0f;
    BestHit H;
    const uint32_t q = L.q;
    while (true) {
        if (!(H.found() && cur_t >= H.best)) {
            if (cur & kLeafBit) {
                uint32_t start = cur & kLeafStartMask;
                uint32_t cnt_p = (cur >> kLeafCountShift) & 7u;
                uint32_t hit_q = 0, tri_q = 0;
                float t_q = 0.0f, u_q = 0.0f, v_q = 0.0f;
                if (q < cnt_p) {
                    Prim p = load_prim(G.prims, start + q);
                    tri_q = p.tri;
                    hit_q = ray_tri(p, o, d, t_q, u_q, v_q) ? 1u : 0u;
                }
                if (COUNT && q == 0) cnt.tri += cnt_p;
                fold_leaf(cnt_p, hit_q, t_q, tri_q, u_q, v_q, H);
            } else {
                // lane 0: left child (q[0], q[1]); lane 1: right child (q[2], q[3])
                uint32_t half = q & 1u;
                float4 lo, hi;
                if (cur < L.nc) {
                    const lds_float4* src = L.nodes + 4 * cur;
                    lo = lds_ld4(src + 2 * half);
                    hi = lds_ld4(src + 2 * half + 1);
                } else {
                    const OMG_GLOBAL float4* src = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes) + cur);
                    lo = src[2 * half];
                    hi = src[2 * half + 1];
                }
                const float bmin[3] = {lo.x, lo.y, lo.z}, bmax[3] = {hi.x, hi.y, hi.z};
                float tmax = H.found() ? H.best : FLT_MAX;
                float te = 0.0f;
                uint32_t hit = ray_aabb(bmin, bmax, o, inv, tmax, te) ? 1u : 0u;
                uint32_t lref = qb<0>(__float_as_uint(lo.w)), rref = qb<0>(__float_as_uint(hi.w));
                uint32_t hl = qb<0>(hit), hr = qb<1>(hit);
                float a = qbf<0>(te), b = qbf<1>(te);
                if (COUNT && q == 0) cnt.aabb += 2;
                if (hl && hr) {
                    bool lnear = a <= b;
                    uint32_t far_ref = lnear ? rref : lref;
                    float far_t = lnear ? b : a;
                    if (sp < SK) {
                        if (q == 0) {
                            L.st_ref[((uint32_t)sp << L.nq_log2) + L.quad] = far_ref;
                            L.st_t[((uint32_t)sp << L.nq_log2) + L.quad] = far_t;
                        }
                    } else {
                        ovf_ref[sp - SK] = far_ref;
                        ovf_t[sp - SK] = far_t;
                    }
                    sp++;
                    cur = lnear ? lref : rref;
                    cur_t = lnear ? a : b;
                    continue;
                }
                if (hl || hr) {
                    cur = hl ? lref : rref;
                    cur_t = hl ? a : b;
                    continue;
                }
            }
        }
        if (sp == 0) break;
        sp--;
        if (sp < SK) {
            cur = L.st_ref[((uint32_t)sp << L.nq_log2) + L.quad];
            cur_t = L.st_t[((uint32_t)sp << L.nq_log2) + L.quad];
        } else {
            cur = ovf_ref[sp - SK];
            cur_t = ovf_t[sp - SK];
        }
    }
    return H.result(t_out, tri_out, flags);
}

// ------------------------------------------------------------------ 4-wide traversal
//
// The same tree collapsed two binary levels at a time into 128 B nodes
// (Node4): slots 0-1 are the binary node's left half (its left child's two
// children, or the left child itself when it is a leaf, slot 1 unused),
// slots 2-3 its right half.  At a node every slot's box is tested against
// the current best distance, and the hit slots are visited in the binary
// walk's order: the half with the nearer box first (the left one on equal
// entries), and inside a half the nearer slot first (again left on ties).
// The halves' boxes are not stored: a half's box is the union of its
// slots' boxes, whose slab entry follows from the slots' per-axis entries
// (half_entry, half_key).  Culling is the binary walk's (`t_enter >= best` at pop,
// the current best as tmax of the box test), so the 4-wide walk visits the
// binary walk's leaves in the binary walk's order with the same best
// distance at every step, and its nearest hit - triangle, distance and the
// winner among hits at equal distance - is the binary walk's, bit for bit.
// (A binary node whose children are visited next to each other has its
// children's boxes tested at the 4-wide node's time, with a best that is
// not smaller; a child that passes this test but not the binary walk's
// later one has an entry beyond the best at its pop and is culled there.)

// A slot in visiting order: entry distance, reference, hit.
struct Cand4 {
    float t;
    uint32_t ref;
    uint32_t hit;
};

__device__ __forceinline__ Cand4 csel(bool c, const Cand4& a, const Cand4& b)
{
    Cand4 r;
    r.t = c ? a.t : b.t;
    r.ref = c ? a.ref : b.ref;
    r.hit = c ? a.hit : b.hit;
    return r;
}

// Slab test of one slot box that also returns the per-axis entries (the
// lesser of the two plane distances, as the slab test takes them).  For a
// finite 1/d the plane distances are numbers or infinities, never NaN, so
// the swap of the reference's test is a min / max and the result is
// ray_aabb's.  (The quad and 16-lane walks hand other rays to the binary
// walk; the one-lane 4-wide walk takes the reference form, below, in a
// step where some lane at a node has an infinite 1/d component.)
__device__ __forceinline__ bool slab_entries(const float4& lo, const float4& hi, f3 o, f3 inv,
                                             float tmax, float e[3], float& t_enter)
{
    const float bmin[3] = {lo.x, lo.y, lo.z}, bmax[3] = {hi.x, hi.y, hi.z};
    float ex[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ta = (bmin[a] - get(o, a)) * get(inv, a), tb = (bmax[a] - get(o, a)) * get(inv, a);
        e[a] = fminf(ta, tb);
        ex[a] = fmaxf(ta, tb);
    }
    const float t0 = fmaxf(fmaxf(fmaxf(0.0f, e[0]), e[1]), e[2]);
    const float t1 = fminf(fminf(fminf(tmax, ex[0]), ex[1]), ex[2]);
    t_enter = t0;
    return !(t0 > t1);
}

// Slab test with per-axis entries for any 1/d, in the reference's form: a
// plane distance pair is swapped when the first is the greater, so for a
// zero direction component (1/d infinite) an entry is NaN exactly when the
// origin lies on that box plane, and the fmaxf of the entry distance then
// ignores it as ray_aabb does.  The one-lane 4-wide walk uses it for every
// ray.
__device__ __forceinline__ bool slab_entries_any(const float4& lo, const float4& hi, f3 o, f3 inv,
                                                 float tmax, float e[3], float& t_enter)
{
    const float bmin[3] = {lo.x, lo.y, lo.z}, bmax[3] = {hi.x, hi.y, hi.z};
    float t0 = 0.0f, t1 = tmax;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ta = (bmin[a] - get(o, a)) * get(inv, a), tb = (bmax[a] - get(o, a)) * get(inv, a);
        const bool sw = ta > tb;
        const float en = sw ? tb : ta, ex = sw ? ta : tb;
        e[a] = en;
        t0 = fmaxf(t0, en);
        t1 = fminf(t1, ex);
    }
    t_enter = t0;
    return !(t0 > t1);
}

// Least of two per-axis entries as an entry of the union box.  With a
// finite 1/d every entry is a number (an overflow gives an infinity) and
// rounding is monotone, so the union's entry is the least of its parts'.
// With an infinite 1/d an entry is NaN when the origin lies on that box
// plane; the slab test's fmaxf then ignores it, which is what -inf does,
// and the union's entry is NaN (ignored) exactly when one of its parts' is.
__device__ __forceinline__ float union_entry(float a, float b, bool infax)
{
    float m = fminf(a, b);
    if (infax) m = (a != a || b != b) ? -__builtin_inff() : m;
    return m;
}

// Entry distance of a half's box from its two slots' per-axis entries.
__device__ __forceinline__ float half_entry(const float ea[3], const float eb[3], bool infax)
{
    float t = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) t = fmaxf(t, union_entry(ea[a], eb[a], infax));
    return t;
}

// Entry keys, for the walks that reduce entries across lanes.  A slot's
// per-axis entry e enters as the key max(0, e) in the bits of the float
// (a NaN or -0 gives 0): non-negative floats and +inf order as their bit
// patterns, so the entry of a union box is the integer minimum of its
// parts' keys per axis and a box's entry distance the integer maximum over
// its axes.  This holds for a finite 1/d, where every entry is a number
// and rounding is monotone (max(0, .) commutes with the minimum).  A ray
// with an infinite 1/d component can meet NaN entries, for which neither
// holds; the quad and 16-lane walks hand such rays to the binary walk
// (mesh_raycast), which is the reference walk itself.
__device__ __forceinline__ uint32_t entry_key(float e) { return e > 0.0f ? __float_as_uint(e) : 0u; }

// Entry key of the union of two slots from their per-axis keys.
__device__ __forceinline__ uint32_t half_key(const uint32_t ka[3], const uint32_t kb[3])
{
    const uint32_t x = ka[0] < kb[0] ? ka[0] : kb[0];
    const uint32_t y = ka[1] < kb[1] ? ka[1] : kb[1];
    const uint32_t z = ka[2] < kb[2] ? ka[2] : kb[2];
    const uint32_t m = x > y ? x : y;
    return m > z ? m : z;
}

// The binary walk's visiting order of a node's four slots (p[0] first):
// tl / tr are the halves' entry distances, c[j] the slots'.  Misses keep
// their place; the walk skips them.
__device__ __forceinline__ void binary_order(const Cand4 c[4], float tl, float tr, Cand4 p[4])
{
    const bool rx = !(tl <= tr), r0 = !(c[0].t <= c[1].t), r1 = !(c[2].t <= c[3].t);
    const Cand4 ln = csel(r0, c[1], c[0]), lf = csel(r0, c[0], c[1]);
    const Cand4 rn = csel(r1, c[3], c[2]), rf = csel(r1, c[2], c[3]);
    p[0] = csel(rx, rn, ln);
    p[1] = csel(rx, rf, lf);
    p[2] = csel(rx, ln, rn);
    p[3] = csel(rx, lf, rf);
}

// Tests the four slots of a node in registers (slot j: box lo[j].xyz ..
// hi[j].xyz, ref lo[j].w) and puts them in the binary walk's order.  ANY:
// the reference form of the slab test and NaN-aware union entries, right
// for any 1/d (infax: some component of 1/d is infinite); otherwise the
// min / max form, right when every component of 1/d is finite (no plane
// distance is then NaN, and both forms give the same hits and entries).
template <bool ANY>
__device__ __forceinline__ void order_children(const float4 lo[4], const float4 hi[4], f3 o, f3 inv,
                                               bool infax, float tmax, Cand4 p[4], uint32_t& tested)
{
    Cand4 c[4];
    float e[4][3];
    tested = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t ref = __float_as_uint(lo[j].w);
        // an unused slot holds a zero box: its test runs but never counts
        const bool valid = ref != kNoChild;
        float te = 0.0f;
        const bool h = (ANY ? slab_entries_any(lo[j], hi[j], o, inv, tmax, e[j], te)
                            : slab_entries(lo[j], hi[j], o, inv, tmax, e[j], te)) && valid;
        tested += valid ? 1u : 0u;
        c[j].t = te;
        c[j].ref = ref;
        c[j].hit = h ? 1u : 0u;
    }
    // a half whose second slot is unused is its first slot (a leaf)
    const float ul = half_entry(e[0], e[1], ANY && infax), ur = half_entry(e[2], e[3], ANY && infax);
    const float tl = c[1].ref == kNoChild ? c[0].t : ul;
    const float tr = c[3].ref == kNoChild ? c[2].t : ur;
    binary_order(c, tl, tr, p);
}

// True when some component of 1/d is infinite.
__device__ __forceinline__ bool inf_axis(f3 inv)
{
    return isinf(inv.x) || isinf(inv.y) || isinf(inv.z);
}

struct Lds4Trav {
    const lds_float4* nodes;   // LDS node prefix [nc][8]
    uint32_t nc;
    lds_u32* st_ref;           // [SK][bs]
    lds_f32* st_t;             // [SK][bs]
    uint32_t lane, bs;
    uint32_t bs_log2;          // the block size is a power of two (256)
};

struct Quad4Trav {
    const lds_float4* nodes;   // LDS node prefix [nc][8]
    uint32_t nc;
    lds_u32* st_ref;           // [SK][quads per block]
    lds_f32* st_t;
    uint32_t quad, nq, q;
    uint32_t nq_log2;
};

// Resumable state of the 4-wide walk of one lane: the node in hand, its
// entry distance, the stack top, the leaf primitives still to test and the
// nearest hit so far.  sp is the LDS slot of the next push, depth x block
// size + lane (the stack is laid out [depth][lane]), so that one register
// holds both the depth and the lane's column.  pk packs the primitives as a
// leaf reference does - bits 0-27 the next one, bits 28-30 how many are
// left - and bit 31 says that some component of 1/d is infinite.
struct Walk4 {
    f3 inv;
    uint32_t cur;
    float cur_t;
    uint32_t sp;
    uint32_t pk;
    BestHit H;

    __device__ __forceinline__ bool infax() const { return (pk >> 31) != 0u; }
    __device__ __forceinline__ uint32_t left() const { return (pk >> kLeafCountShift) & 7u; }
};

// Starts a walk; false when there is nothing to walk (an empty mesh).
__device__ __forceinline__ bool walk4_begin(const GeomArgs& G, const Lds4Trav& L, Walk4& W, f3 d)
{
    W.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    W.pk = inf_axis(W.inv) ? 0x80000000u : 0u;
    W.cur = G.root4;
    W.cur_t = 0.0f;
    W.sp = L.lane;
    W.H = BestHit();
    return !G.empty;
}

// One step of the nearest-hit walk over the 4-wide nodes: one node tested,
// or one primitive of a leaf, then a pop when the lane has nothing in
// hand; false once the walk has ended.  A leaf's primitives take one step
// each (in order), so that the lanes of a wave that are in leaves and those
// in inner nodes share every step instead of the whole wave waiting for its
// largest leaf.  Nodes, primitives and their order are those of the walk
// that tests a leaf at once.  The stack's top SK entries live in LDS,
// deeper ones in ovf.
template <bool COUNT, int SK, bool ANY>
__device__ __forceinline__ bool walk4_step_t(const GeomArgs& G, const Lds4Trav& L, Walk4& W, f3 o,
                                             f3 d, uint32_t (&ovf_ref)[kStackCap - SK],
                                             float (&ovf_t)[kStackCap - SK], TravCount& cnt)
{
    if (W.left() == 0u && !(W.H.found() && W.cur_t >= W.H.best)) {
        if (W.cur & kLeafBit) {
            // a leaf reference is bit 31, the count in bits 28-30, the first primitive
            W.pk = (W.pk & 0x80000000u) | (W.cur & 0x7FFFFFFFu);
        } else {
            float4 lo[4], hi[4];
            if (W.cur < L.nc) {
                const lds_float4* q = L.nodes + 8 * W.cur;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    lo[j] = lds_ld4(q + 2 * j);
                    hi[j] = lds_ld4(q + 2 * j + 1);
                }
            } else {
                const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes4) + W.cur);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    lo[j] = q[2 * j];
                    hi[j] = q[2 * j + 1];
                }
            }
            Cand4 p[4];
            uint32_t tested;
            order_children<ANY>(lo, hi, o, W.inv, ANY && W.infax(), W.H.found() ? W.H.best : FLT_MAX,
                                p, tested);
            if (COUNT) cnt.aabb += tested;
            const bool h0 = p[0].hit != 0u, h1 = p[1].hit != 0u, h2 = p[2].hit != 0u,
                       h3 = p[3].hit != 0u;
            if (h0 | h1 | h2 | h3) {
                auto push = [&](uint32_t ref, float t) {
                    if (W.sp < SK * L.bs) {
                        L.st_ref[W.sp] = ref;
                        L.st_t[W.sp] = t;
                    } else {
                        const uint32_t k = (W.sp >> L.bs_log2) - SK;
                        ovf_ref[k] = ref;
                        ovf_t[k] = t;
                    }
                    W.sp += L.bs;
                };
                // every hit slot after the first is pushed, the last one deepest
                if (h3 & (h0 | h1 | h2)) push(p[3].ref, p[3].t);
                if (h2 & (h0 | h1)) push(p[2].ref, p[2].t);
                if (h1 & h0) push(p[1].ref, p[1].t);
                const Cand4 f = csel(h0, p[0], csel(h1, p[1], csel(h2, p[2], p[3])));
                W.cur = f.ref;
                W.cur_t = f.t;
                // a leaf visited next tests its first primitive in this step
                // (nothing changes before the next step's cull test)
                if (!(W.cur & kLeafBit) || (W.H.found() && W.cur_t >= W.H.best)) return true;
                W.pk = (W.pk & 0x80000000u) | (W.cur & 0x7FFFFFFFu);
            }
        }
    }
    if (W.left() != 0u) {
        Prim p = load_prim(G.prims, W.pk & kLeafStartMask);
        if (COUNT) cnt.tri++;
        float t, u, v;
        W.H.offer_if(ray_tri(p, o, d, t, u, v), t, p.tri, u, v);
        // the next primitive, one fewer left (none left: only the infinite-axis bit stays)
        W.pk = W.left() > 1u ? W.pk + 1u - (1u << kLeafCountShift) : (W.pk & 0x80000000u);
        if (W.left() != 0u) return true;
    }
    // pop; entries the current best culls are passed over here rather than
    // in a step each (a culled entry is skipped unopened either way)
    for (;;) {
        if (W.sp < L.bs) return false;
        W.sp -= L.bs;
        if (W.sp < SK * L.bs) {
            W.cur = L.st_ref[W.sp];
            W.cur_t = L.st_t[W.sp];
        } else {
            const uint32_t k = (W.sp >> L.bs_log2) - SK;
            W.cur = ovf_ref[k];
            W.cur_t = ovf_t[k];
        }
        if (!(W.H.found() && W.cur_t >= W.H.best)) return true;
    }
}

// The step in the min / max form unless some active lane has an infinite
// 1/d component (a wave-uniform choice: such rays are rare, and the whole
// step is duplicated so that neither form holds the other's registers).
template <bool COUNT, int SK>
__device__ __forceinline__ bool walk4_step(const GeomArgs& G, const Lds4Trav& L, Walk4& W, f3 o,
                                           f3 d, uint32_t (&ovf_ref)[kStackCap - SK],
                                           float (&ovf_t)[kStackCap - SK], TravCount& cnt)
{
    if (__ballot(W.infax()) != 0)
        return walk4_step_t<COUNT, SK, true>(G, L, W, o, d, ovf_ref, ovf_t, cnt);
    return walk4_step_t<COUNT, SK, false>(G, L, W, o, d, ovf_ref, ovf_t, cnt);
}

// Nearest hit over the 4-wide nodes, one lane per ray, stack top in LDS.
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast4_lds(const GeomArgs& G, const Lds4Trav& L, f3 o, f3 d,
                                                  float& t_out, uint32_t& tri_out, uint32_t& flags,
                                                  TravCount& cnt)
{
    uint32_t ovf_ref[kStackCap - SK];
    float ovf_t[kStackCap - SK];
    Walk4 W;
    if (!walk4_begin(G, L, W, d)) return false;
    while (walk4_step<COUNT, SK>(G, L, W, o, d, ovf_ref, ovf_t, cnt)) {
    }
    return W.H.result(t_out, tri_out, flags);
}

// quad_perm move inside a DPP quad: lane q reads lane q ^ 1 (its partner
// in the same half of a 4-wide node).  Every lane of the quad is active
// where it is used.
__device__ __forceinline__ uint32_t qswap(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 1 | (0 << 2) | (3 << 4) | (2 << 6), 0xF, 0xF,
                                              true);
}
__device__ __forceinline__ float qswapf(float x) { return __uint_as_float(qswap(__float_as_uint(x))); }

// Quad form of order_children's half entries: lane q holds slot q's
// per-axis entries e, its entry te and reference ref; every lane of the
// quad gets both halves' entry distances.
__device__ __forceinline__ void quad_half_entries(const float e[3], float te, uint32_t ref, float& tl,
                                                  float& tr)
{
    const uint32_t k[3] = {entry_key(e[0]), entry_key(e[1]), entry_key(e[2])};
    const uint32_t pk[3] = {qswap(k[0]), qswap(k[1]), qswap(k[2])};
    const uint32_t pref = qswap(ref);
    // on lanes 0 and 2: the half's entry (its second slot unused: the first slot's)
    const float th = pref == kNoChild ? te : __uint_as_float(half_key(k, pk));
    tl = qbf<0>(th);
    tr = qbf<2>(th);
}

// The same walk on a DPP quad: lane q loads and tests slot q (32 B) and
// primitive q of a leaf; the slots' results are broadcast inside the quad
// and every lane orders them identically.  Lane 0 writes the shared LDS
// stack.
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast4_quad(const GeomArgs& G, const Quad4Trav& L, f3 o, f3 d,
                                                   float& t_out, uint32_t& tri_out, uint32_t& flags,
                                                   TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (inf_axis(inv)) return mesh_raycast<COUNT>(G, o, d, t_out, tri_out, flags, cnt);
    uint32_t ovf_ref[kStackCap - SK];
    float ovf_t[kStackCap - SK];
    int sp = 0;
    uint32_t cur = G.root4;
    float cur_t = 0.0f;
    BestHit H;
    const uint32_t q = L.q;
    auto push = [&](uint32_t ref, float t) {
        if (sp < SK) {
            if (q == 0) {
                L.st_ref[((uint32_t)sp << L.nq_log2) + L.quad] = ref;
                L.st_t[((uint32_t)sp << L.nq_log2) + L.quad] = t;
            }
        } else {
            ovf_ref[sp - SK] = ref;
            ovf_t[sp - SK] = t;
        }
        sp++;
    };
    while (true) {
        if (!(H.found() && cur_t >= H.best)) {
            if (cur & kLeafBit) {
                uint32_t start = cur & kLeafStartMask;
                uint32_t cnt_p = (cur >> kLeafCountShift) & 7u;
                uint32_t hit_q = 0, tri_q = 0;
                float t_q = 0.0f, u_q = 0.0f, v_q = 0.0f;
                if (q < cnt_p) {
                    Prim p = load_prim(G.prims, start + q);
                    tri_q = p.tri;
                    hit_q = ray_tri(p, o, d, t_q, u_q, v_q) ? 1u : 0u;
                }
                if (COUNT && q == 0) cnt.tri += cnt_p;
                fold_leaf(cnt_p, hit_q, t_q, tri_q, u_q, v_q, H);
            } else {
                float4 lo, hi;
                if (cur < L.nc) {
                    const lds_float4* src = L.nodes + 8 * cur + 2 * q;
                    lo = lds_ld4(src);
                    hi = lds_ld4(src + 1);
                } else {
                    const OMG_GLOBAL float4* src = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes4) + cur) + 2 * q;
                    lo = src[0];
                    hi = src[1];
                }
                const uint32_t ref = __float_as_uint(lo.w);
                float e[3], te = 0.0f;
                const bool h = slab_entries(lo, hi, o, inv, H.found() ? H.best : FLT_MAX, e, te) &&
                               ref != kNoChild;
                float tl, tr;
                quad_half_entries(e, te, ref, tl, tr);
                const uint32_t hq = h ? 1u : 0u;
                Cand4 c[4];
                c[0] = Cand4{qbf<0>(te), qb<0>(ref), qb<0>(hq)};
                c[1] = Cand4{qbf<1>(te), qb<1>(ref), qb<1>(hq)};
                c[2] = Cand4{qbf<2>(te), qb<2>(ref), qb<2>(hq)};
                c[3] = Cand4{qbf<3>(te), qb<3>(ref), qb<3>(hq)};
                if (COUNT && q == 0)
                    cnt.aabb += (c[0].ref != kNoChild) + (c[1].ref != kNoChild) +
                                (c[2].ref != kNoChild) + (c[3].ref != kNoChild);
                Cand4 p[4];
                binary_order(c, tl, tr, p);
                const bool h0 = p[0].hit != 0u, h1 = p[1].hit != 0u, h2 = p[2].hit != 0u,
                           h3 = p[3].hit != 0u;
                if (h0 | h1 | h2 | h3) {
                    if (h3 & (h0 | h1 | h2)) push(p[3].ref, p[3].t);
                    if (h2 & (h0 | h1)) push(p[2].ref, p[2].t);
                    if (h1 & h0) push(p[1].ref, p[1].t);
                    const Cand4 f = csel(h0, p[0], csel(h1, p[1], csel(h2, p[2], p[3])));
                    cur = f.ref;
                    cur_t = f.t;
                    continue;
                }
            }
        }
        if (sp == 0) break;
        sp--;
        if (sp < SK) {
            cur = L.st_ref[((uint32_t)sp << L.nq_log2) + L.quad];
            cur_t = L.st_t[((uint32_t)sp << L.nq_log2) + L.quad];
        } else {
            cur = ovf_ref[sp - SK];
            cur_t = ovf_t[sp - SK];
        }
    }
    return H.result(t_out, tri_out, flags);
}

// The quad walk with a stack deep enough for the whole tree (SK at least the
// tree's worst-case stack, checked on the host), so nothing spills to
// scratch and the four lanes push in parallel: lane q finds its slot's
// place among the quad's hits in the binary walk's order, writes it at
// stack depth sp + nhit - 1 - rank unless it is the first, and the first
// slot's reference and distance reach every lane through an OR over the
// quad (only that lane contributes non-zero bits; node 0 is the root,
// never a child).
template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast4_quad_deep(const GeomArgs& G, const Quad4Trav& L, f3 o,
                                                        f3 d, float& t_out, uint32_t& tri_out,
                                                        uint32_t& flags, TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (inf_axis(inv)) return mesh_raycast<COUNT>(G, o, d, t_out, tri_out, flags, cnt);
    int sp = 0;
    uint32_t cur = G.root4;
    float cur_t = 0.0f;
    BestHit H;
    const uint32_t q = L.q;
    const uint32_t qshift = (__lane_id() & 63u) & ~3u;
    while (true) {
        bool at_leaf = false;
        if (!(H.found() && cur_t >= H.best)) {
            if (cur & kLeafBit) {
                at_leaf = true;
            } else {
                float4 lo, hi;
                if (cur < L.nc) {
                    const lds_float4* src = L.nodes + 8 * cur + 2 * q;
                    lo = lds_ld4(src);
                    hi = lds_ld4(src + 1);
                } else {
                    const OMG_GLOBAL float4* src = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes4) + cur) + 2 * q;
                    lo = src[0];
                    hi = src[1];
                }
                const uint32_t ref = __float_as_uint(lo.w);
                float e[3], te = 0.0f;
                const bool h = slab_entries(lo, hi, o, inv, H.found() ? H.best : FLT_MAX, e, te) &&
                               ref != kNoChild;
                if (COUNT && q == 0) {
                    uint32_t valid = (uint32_t)(__ballot(ref != kNoChild) >> qshift) & 0xFu;
                    cnt.aabb += __popc(valid);
                }
                float tl, tr;
                quad_half_entries(e, te, ref, tl, tr);
                const uint32_t hm = (uint32_t)(__ballot(h) >> qshift) & 0xFu;
                const int nhit = __popc(hm);
                if (nhit > 0) {
                    // this slot's place among the quad's hits in visiting order:
                    // the other half's hits come first when this half is the
                    // farther one, the partner's when it is the nearer slot
                    const float pte = qswapf(te);
                    const bool odd = (q & 1u) != 0u;
                    const bool half_far = ((q >> 1) != 0u) != !(tl <= tr);
                    const bool slot_far = odd != !((odd ? pte : te) <= (odd ? te : pte));
                    const uint32_t other = q < 2 ? 0xCu : 0x3u;
                    const int rank = (half_far ? __popc(hm & other) : 0) +
                                     ((slot_far && ((hm >> (q ^ 1u)) & 1u)) ? 1 : 0);
                    if (h && rank > 0) {
                        const int pos = sp + nhit - 1 - rank;
                        L.st_ref[((uint32_t)pos << L.nq_log2) + L.quad] = ref;
                        L.st_t[((uint32_t)pos << L.nq_log2) + L.quad] = te;
                    }
                    // the quad's one hit slot of rank 0 is the next node,
                    // read from its lane by the whole quad (all of it active)
                    const uint32_t fm = (uint32_t)(__ballot(h && rank == 0) >> qshift) & 0xFu;
                    const int src = (int)((qshift + (uint32_t)__builtin_ctz(fm)) << 2);
                    cur = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)ref);
                    cur_t = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(te)));
                    sp += nhit - 1;
                    // a leaf visited next is tested in this iteration
                    // (nothing changes before the next iteration's cull test)
                    if (!(cur & kLeafBit) || (H.found() && cur_t >= H.best)) continue;
                    at_leaf = true;
                }
            }
        }
        if (at_leaf) {
            uint32_t start = cur & kLeafStartMask;
            uint32_t cnt_p = (cur >> kLeafCountShift) & 7u;
            uint32_t hit_q = 0, tri_q = 0;
            float t_q = 0.0f, u_q = 0.0f, v_q = 0.0f;
            if (q < cnt_p) {
                Prim p = load_prim(G.prims, start + q);
                tri_q = p.tri;
                hit_q = ray_tri(p, o, d, t_q, u_q, v_q) ? 1u : 0u;
            }
            if (COUNT && q == 0) cnt.tri += cnt_p;
            fold_leaf(cnt_p, hit_q, t_q, tri_q, u_q, v_q, H);
        }
        // pop, passing over the entries the current best culls (a culled
        // entry is skipped unopened either way)
        bool done = false;
        for (;;) {
            if (sp == 0) {
                done = true;
                break;
            }
            sp--;
            cur = L.st_ref[((uint32_t)sp << L.nq_log2) + L.quad];
            cur_t = L.st_t[((uint32_t)sp << L.nq_log2) + L.quad];
            if (!(H.found() && cur_t >= H.best)) break;
        }
        if (done) break;
    }
    return H.result(t_out, tri_out, flags);
}

// ------------------------------------------------------------------ 16-wide traversal
//
// A 16-lane DPP row traces one ray over the 16-wide nodes (Node16): lane r
// loads and tests slot r (32 B; the row reads the node's four 128 B lines
// together).  The binary walk's order over the node's four binary levels
// comes from group entry distances: the union box of an aligned group of
// 2, 4 or 8 slots is entered where its slots' entry keys say (entry_key),
// reduced across the row with DPP (xor 1 and 2 by quad_perm, the other
// quad by row_half_mirror, the other 8-group by row_mirror).  At each of
// the four levels a slot lies in the farther of two sibling groups or in
// the nearer one (the left one on equal entries); the hit slots of every
// sibling group it lies behind come before it, which gives its place among
// the node's hits, and each hit slot pushes itself at that depth of the
// row's LDS stack (the first one becomes the current node, broadcast by an
// OR over the row).  Culling is the binary walk's, so the leaves, their
// order and the nearest hit are the binary walk's (see the 4-wide
// section).  Unused slots key +inf (they have no box); a leaf's primitives
// are tested by every quad of the row (lane r takes primitive r & 3), so
// the quads keep identical nearest hits without crossing quads.

// DPP moves inside a 16-lane row (dpp_u; every lane of the row is active
// where they are used, the row sharing its control flow): besides the quad
// moves, the other quad of an 8-group and the other 8-group.
constexpr int kDppHalfMirror = 0x141;                           // lane i <-> 7 - i in 8
constexpr int kDppMirror = 0x140;                               // lane i <-> 15 - i in 16

// LDS working set of a 16-lane row walk: the node prefix (top 16-wide
// nodes, breadth-first) and the stack [SK][rows per block].
struct Row16Trav {
    const lds_float4* nodes;   // [nc][32]
    uint32_t nc;
    lds_u32* st_ref;
    lds_f32* st_t;
    uint32_t row, nrow, r;     // row in the block, rows per block, lane in the row
    uint32_t nrow_log2;        // the block size is a power of two (256)
};

__device__ __forceinline__ uint32_t umin3(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax3(uint32_t x, uint32_t y, uint32_t z)
{
    const uint32_t m = x > y ? x : y;
    return m > z ? m : z;
}

// This slot's group is visited after its sibling group: right = it is the
// right one; keys of the two groups' entries.
__device__ __forceinline__ bool group_far(bool right, uint32_t mine, uint32_t sib)
{
    const uint32_t kl = right ? sib : mine, kr = right ? mine : sib;
    return right == (kl <= kr);
}

// Resumable state of a row's walk (the same in every lane of the row): the
// node in hand, its entry distance, the LDS stack depth and the nearest hit.
struct Walk16 {
    f3 inv;
    uint32_t cur;
    float cur_t;
    int sp;
    BestHit H;
};

// Starts a row's walk; false when it has already ended: an empty mesh, or a
// ray with an infinite 1/d component, which the binary walk (the reference
// walk itself) traces whole here.
template <bool COUNT>
__device__ __forceinline__ bool walk16_begin(const GeomArgs& G, Walk16& W, f3 o, f3 d, TravCount& cnt)
{
    W.H = BestHit();
    W.sp = 0;
    W.cur = G.root16;
    W.cur_t = 0.0f;
    if (G.empty) return false;
    W.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (inf_axis(W.inv)) {
        float t;
        uint32_t tri, fl;
        if (mesh_raycast<COUNT>(G, o, d, t, tri, fl, cnt)) {
            W.H.best = t;
            W.H.btri = tri;
            W.H.fl = fl;
        }
        return false;
    }
    return true;
}

// One step of a row's walk: the node in hand tested (a leaf chosen next is
// tested in the same step), or a leaf, then a pop that passes over the
// entries the current best culls; false once the walk has ended.
template <bool COUNT, int SK>
__device__ __forceinline__ bool walk16_step(const GeomArgs& G, const Row16Trav& L, Walk16& W, f3 o, f3 d,
                                            TravCount& cnt)
{
    const uint32_t r = L.r, q = r & 3u;
    const uint32_t rshift = (__lane_id() & 63u) & ~15u;
    bool at_leaf = false;
    if (!(W.H.found() && W.cur_t >= W.H.best)) {
        if (W.cur & kLeafBit) {
            at_leaf = true;
        } else {
            float4 lo, hi;
            if (W.cur < L.nc) {
                const lds_float4* src = L.nodes + 32 * W.cur + 2 * r;
                lo = lds_ld4(src);
                hi = lds_ld4(src + 1);
            } else {
                const OMG_GLOBAL float4* src = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes16) + W.cur) + 2 * r;
                lo = src[0];
                hi = src[1];
            }
            const uint32_t ref = __float_as_uint(lo.w);
            const bool valid = ref != kNoChild;
            float e[3], te = 0.0f;
            const bool h = slab_entries(lo, hi, o, W.inv, W.H.found() ? W.H.best : FLT_MAX, e, te) && valid;
            if (COUNT && r == 0) {
                uint32_t v = (uint32_t)(__ballot(valid) >> rshift) & 0xFFFFu;
                cnt.aabb += __popc(v);
            }
            const uint32_t hm = (uint32_t)(__ballot(h) >> rshift) & 0xFFFFu;
            const int nhit = __popc(hm);
            if (nhit > 0) {
                // per-axis keys of this slot's groups of 2, 4 and 8 (an
                // unused slot has no box: +inf)
                const uint32_t inf = 0x7F800000u;
                const uint32_t k0 = valid ? entry_key(e[0]) : inf;
                const uint32_t k1 = valid ? entry_key(e[1]) : inf;
                const uint32_t k2 = valid ? entry_key(e[2]) : inf;
                const uint32_t a0 = umin3(k0, dpp_u<kDppXor1>(k0));
                const uint32_t a1 = umin3(k1, dpp_u<kDppXor1>(k1));
                const uint32_t a2 = umin3(k2, dpp_u<kDppXor1>(k2));
                const uint32_t b0 = umin3(a0, dpp_u<kDppXor2>(a0));
                const uint32_t b1 = umin3(a1, dpp_u<kDppXor2>(a1));
                const uint32_t b2 = umin3(a2, dpp_u<kDppXor2>(a2));
                const uint32_t c0 = umin3(b0, dpp_u<kDppHalfMirror>(b0));
                const uint32_t c1 = umin3(b1, dpp_u<kDppHalfMirror>(b1));
                const uint32_t c2 = umin3(b2, dpp_u<kDppHalfMirror>(b2));
                const uint32_t g2 = umax3(a0, a1, a2), g4 = umax3(b0, b1, b2), g8 = umax3(c0, c1, c2);
                const uint32_t g1 = entry_key(te);
                // each level's sibling group
                const uint32_t s1 = dpp_u<kDppXor1>(g1), s2 = dpp_u<kDppXor2>(g2);
                const uint32_t s4 = dpp_u<kDppHalfMirror>(g4), s8 = dpp_u<kDppMirror>(g8);
                int rank = 0;
                rank += group_far(r & 1u, g1, s1) ? (int)((hm >> (r ^ 1u)) & 1u) : 0;
                rank += group_far(r & 2u, g2, s2) ? __popc(hm & (0x3u << ((r & ~1u) ^ 2u))) : 0;
                rank += group_far(r & 4u, g4, s4) ? __popc(hm & (0xFu << ((r & ~3u) ^ 4u))) : 0;
                rank += group_far(r & 8u, g8, s8) ? __popc(hm & (0xFFu << ((r & ~7u) ^ 8u))) : 0;
                if (h && rank > 0) {
                    const int pos = W.sp + nhit - 1 - rank;
                    L.st_ref[((uint32_t)pos << L.nrow_log2) + L.row] = ref;
                    L.st_t[((uint32_t)pos << L.nrow_log2) + L.row] = te;
                }
                // the row's one hit slot of rank 0 is the next node: every
                // lane of the row reads it from that lane (the whole row
                // is active here, so the source lane is)
                const uint32_t fm = (uint32_t)(__ballot(h && rank == 0) >> rshift) & 0xFFFFu;
                const int src = (int)((rshift + (uint32_t)__builtin_ctz(fm)) << 2);
                W.cur = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)ref);
                W.cur_t = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(te)));
                W.sp += nhit - 1;
                // a leaf visited next is tested in this step (nothing
                // changes before the next step's cull test)
                if (!(W.cur & kLeafBit) || (W.H.found() && W.cur_t >= W.H.best)) return true;
                at_leaf = true;
            }
        }
    }
    if (at_leaf) {
        uint32_t start = W.cur & kLeafStartMask;
        uint32_t cnt_p = (W.cur >> kLeafCountShift) & 7u;
        uint32_t hit_q = 0, tri_q = 0;
        float t_q = 0.0f, u_q = 0.0f, v_q = 0.0f;
        if (q < cnt_p) {
            Prim p = load_prim(G.prims, start + q);
            tri_q = p.tri;
            hit_q = ray_tri(p, o, d, t_q, u_q, v_q) ? 1u : 0u;
        }
        if (COUNT && r == 0) cnt.tri += cnt_p;
        fold_leaf(cnt_p, hit_q, t_q, tri_q, u_q, v_q, W.H);
    }
    // pop, passing over the entries the current best culls (a culled
    // entry is skipped unopened either way)
    for (;;) {
        if (W.sp == 0) return false;
        W.sp--;
        W.cur = L.st_ref[((uint32_t)W.sp << L.nrow_log2) + L.row];
        W.cur_t = L.st_t[((uint32_t)W.sp << L.nrow_log2) + L.row];
        if (!(W.H.found() && W.cur_t >= W.H.best)) return true;
    }
}

template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast16(const GeomArgs& G, const Row16Trav& L, f3 o, f3 d,
                                               float& t_out, uint32_t& tri_out, uint32_t& flags,
                                               TravCount& cnt)
{
    Walk16 W;
    if (walk16_begin<COUNT>(G, W, o, d, cnt))
        while (walk16_step<COUNT, SK>(G, L, W, o, d, cnt)) {
        }
    return W.H.result(t_out, tri_out, flags);
}

// The same walk on an 8-lane row: lane r loads and tests slots 2r and 2r+1
// (64 B), so a wave walks eight rays where the 16-lane row walks four.  The
// group levels are the 16-lane walk's with the first one inside the lane:
// slot s = 2r + j; pairs of slots are a lane, groups of 4 two lanes (xor
// 1), groups of 8 a quad (xor 2), and the two 8-groups the row's halves
// (row_half_mirror).  Ranks, pushes, culling and the leaf fold are the
// 16-lane walk's, so the nearest hit is the binary walk's bit for bit.
struct Row8Trav {
    const lds_float4* nodes;   // [nc][32]
    uint32_t nc;
    lds_u32* st_ref;
    lds_f32* st_t;
    uint32_t row, nrow, r;     // row in the block, rows per block, lane in the row
    uint32_t nrow_log2;        // the block size is a power of two (256)
};

// Bits 0-7 of x moved to the even bit positions 0, 2, ..., 14.
__device__ __forceinline__ uint32_t spread8(uint32_t x)
{
    x = (x | (x << 4)) & 0x0F0Fu;
    x = (x | (x << 2)) & 0x3333u;
    return (x | (x << 1)) & 0x5555u;
}

template <bool COUNT, int SK>
__device__ __forceinline__ bool mesh_raycast16x2(const GeomArgs& G, const Row8Trav& L, f3 o, f3 d,
                                                 float& t_out, uint32_t& tri_out, uint32_t& flags,
                                                 TravCount& cnt)
{
    if (G.empty) return false;
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (inf_axis(inv)) return mesh_raycast<COUNT>(G, o, d, t_out, tri_out, flags, cnt);
    int sp = 0;
    uint32_t cur = G.root16;
    float cur_t = 0.0f;
    BestHit H;
    const uint32_t r = L.r, q = r & 3u;
    const uint32_t rshift = (__lane_id() & 63u) & ~7u;
    while (true) {
        bool at_leaf = false;
        if (!(H.found() && cur_t >= H.best)) {
            if (cur & kLeafBit) {
                at_leaf = true;
            } else {
                float4 lo0, hi0, lo1, hi1;
                if (cur < L.nc) {
                    const lds_float4* src = L.nodes + 32 * cur + 4 * r;
                    lo0 = lds_ld4(src);
                    hi0 = lds_ld4(src + 1);
                    lo1 = lds_ld4(src + 2);
                    hi1 = lds_ld4(src + 3);
                } else {
                    const OMG_GLOBAL float4* src = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes16) + cur) + 4 * r;
                    lo0 = src[0];
                    hi0 = src[1];
                    lo1 = src[2];
                    hi1 = src[3];
                }
                const float tmax = H.found() ? H.best : FLT_MAX;
                const uint32_t ref0 = __float_as_uint(lo0.w), ref1 = __float_as_uint(lo1.w);
                const bool valid0 = ref0 != kNoChild, valid1 = ref1 != kNoChild;
                float e0[3], e1[3], te0 = 0.0f, te1 = 0.0f;
                const bool h0 = slab_entries(lo0, hi0, o, inv, tmax, e0, te0) && valid0;
                const bool h1 = slab_entries(lo1, hi1, o, inv, tmax, e1, te1) && valid1;
                if (COUNT && r == 0) {
                    const uint32_t v0 = (uint32_t)(__ballot(valid0) >> rshift) & 0xFFu;
                    const uint32_t v1 = (uint32_t)(__ballot(valid1) >> rshift) & 0xFFu;
                    cnt.aabb += __popc(v0) + __popc(v1);
                }
                // hit mask over the node's 16 slots, slot 2r + j at bit 2r + j
                const uint32_t hm = spread8((uint32_t)(__ballot(h0) >> rshift) & 0xFFu) |
                                    (spread8((uint32_t)(__ballot(h1) >> rshift) & 0xFFu) << 1);
                const int nhit = __popc(hm);
                if (nhit > 0) {
                    // per-axis keys of the lane's pair, then of its groups of 4 and 8
                    // (an unused slot has no box: +inf)
                    const uint32_t inf = 0x7F800000u;
                    uint32_t a[3], b[3], c[3];
#pragma unroll
                    for (int x = 0; x < 3; x++) {
                        const uint32_t k0 = valid0 ? entry_key(e0[x]) : inf;
                        const uint32_t k1 = valid1 ? entry_key(e1[x]) : inf;
                        a[x] = umin3(k0, k1);
                    }
#pragma unroll
                    for (int x = 0; x < 3; x++) b[x] = umin3(a[x], dpp_u<kDppXor1>(a[x]));
#pragma unroll
                    for (int x = 0; x < 3; x++) c[x] = umin3(b[x], dpp_u<kDppXor2>(b[x]));
                    const uint32_t g2 = umax3(a[0], a[1], a[2]), g4 = umax3(b[0], b[1], b[2]),
                                   g8 = umax3(c[0], c[1], c[2]);
                    const uint32_t g1a = entry_key(te0), g1b = entry_key(te1);
                    // each level's sibling group (the pair's other slot in the lane)
                    const uint32_t s2 = dpp_u<kDppXor1>(g2), s4 = dpp_u<kDppXor2>(g4);
                    const uint32_t s8 = dpp_u<kDppHalfMirror>(g8);
                    // levels shared by the lane's two slots (slot bits 1, 2, 3 = lane bits 0, 1, 2)
                    const uint32_t sl = 2u * r;
                    int rk = 0;
                    rk += group_far(r & 1u, g2, s2) ? __popc(hm & (0x3u << (sl ^ 2u))) : 0;
                    rk += group_far(r & 2u, g4, s4) ? __popc(hm & (0xFu << ((sl & ~3u) ^ 4u))) : 0;
                    rk += group_far(r & 4u, g8, s8) ? __popc(hm & (0xFFu << ((sl & ~7u) ^ 8u))) : 0;
                    // slot level: slot 1 of the pair is the right one
                    const int r0 = rk + (group_far(false, g1a, g1b) ? (int)((hm >> (sl + 1u)) & 1u) : 0);
                    const int r1 = rk + (group_far(true, g1b, g1a) ? (int)((hm >> sl) & 1u) : 0);
                    if (h0 && r0 > 0) {
                        const int pos = sp + nhit - 1 - r0;
                        L.st_ref[((uint32_t)pos << L.nrow_log2) + L.row] = ref0;
                        L.st_t[((uint32_t)pos << L.nrow_log2) + L.row] = te0;
                    }
                    if (h1 && r1 > 0) {
                        const int pos = sp + nhit - 1 - r1;
                        L.st_ref[((uint32_t)pos << L.nrow_log2) + L.row] = ref1;
                        L.st_t[((uint32_t)pos << L.nrow_log2) + L.row] = te1;
                    }
                    // the lane holding the row's one hit slot of rank 0 gives
                    // the next node to the whole row (all of it active here)
                    const bool f0 = h0 && r0 == 0, f1 = h1 && r1 == 0;
                    const uint32_t nr = f0 ? ref0 : ref1;
                    const uint32_t nt = f0 ? __float_as_uint(te0) : __float_as_uint(te1);
                    const uint32_t fm = (uint32_t)(__ballot(f0 || f1) >> rshift) & 0xFFu;
                    const int src = (int)((rshift + (uint32_t)__builtin_ctz(fm)) << 2);
                    cur = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)nr);
                    cur_t = __uint_as_float((uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)nt));
                    sp += nhit - 1;
                    // a leaf visited next is tested in this iteration
                    // (nothing changes before the next iteration's cull test)
                    if (!(cur & kLeafBit) || (H.found() && cur_t >= H.best)) continue;
                    at_leaf = true;
                }
            }
        }
        if (at_leaf) {
            uint32_t start = cur & kLeafStartMask;
            uint32_t cnt_p = (cur >> kLeafCountShift) & 7u;
            uint32_t hit_q = 0, tri_q = 0;
            float t_q = 0.0f, u_q = 0.0f, v_q = 0.0f;
            if (q < cnt_p) {
                Prim p = load_prim(G.prims, start + q);
                tri_q = p.tri;
                hit_q = ray_tri(p, o, d, t_q, u_q, v_q) ? 1u : 0u;
            }
            if (COUNT && r == 0) cnt.tri += cnt_p;
            fold_leaf(cnt_p, hit_q, t_q, tri_q, u_q, v_q, H);
        }
        // pop, passing over the entries the current best culls
        bool done = false;
        for (;;) {
            if (sp == 0) {
                done = true;
                break;
            }
            sp--;
            cur = L.st_ref[((uint32_t)sp << L.nrow_log2) + L.row];
            cur_t = L.st_t[((uint32_t)sp << L.nrow_log2) + L.row];
            if (!(H.found() && cur_t >= H.best)) break;
        }
        if (done) break;
    }
    return H.result(t_out, tri_out, flags);
}

// One crossing of a segment: position along it (0..1), original triangle,
// authored surface and material.
struct SegHit {
    float t;
    uint32_t tri;
    uint32_t surf;
    uint32_t mat;
};

// Mesh::segment_hits (mesh.rs:546-604) into a caller-provided list of
// capacity `cap`: all crossings strictly inside the open segment, stably
// sorted by t, then each same-t cluster reduced to its lowest surface id.
// The reference pops a node, tests its box and pushes both children (the
// second child is visited first); here a child's box is tested before it
// is pushed, which visits the same leaves in the same order.  Returns the
// count, or -1 when more than `cap` crossings were found.
template <bool COUNT>
__device__ __forceinline__ int segment_hits(const GeomArgs& G, f3 a, f3 b, SegHit* out, int cap,
                                            TravCount& cnt)
{
    if (G.empty) return 0;
    f3 d = sub(b, a);
    float len = length(d);
    if (len < 1e-6f) return 0;
    f3 dn = mul(d, 1.0f / len);
    f3 inv = mk(1.0f / dn.x, 1.0f / dn.y, 1.0f / dn.z);
    int n = 0;
    bool overflow = false;
    uint32_t stack[kStackCap];
    int sp = 0;
    float te;
    if (COUNT) cnt.aabb++;
    if (ray_aabb(G.root_min, G.root_max, a, inv, len, te)) stack[sp++] = G.root_ref;
    while (sp > 0) {
        uint32_t ref = stack[--sp];
        if (ref & kLeafBit) {
            uint32_t start = ref & kLeafStartMask;
            uint32_t cnt_p = (ref >> kLeafCountShift) & 7u;
            for (uint32_t k = start; k < start + cnt_p; k++) {
                Prim p = load_prim(G.prims, k);
                if (COUNT) cnt.tri++;
                float t, u, v;
                if (ray_tri(p, a, dn, t, u, v)) {
                    float tt = t / len;
                    if (tt > 1e-4f && tt < 1.0f - 1e-4f) {
                        if (n == cap) {
                            overflow = true;
                            continue;
                        }
                        // stable insertion: after every entry with t <= tt
                        int j = n;
                        while (j > 0 && out[j - 1].t > tt) {
                            out[j] = out[j - 1];
                            j--;
                        }
                        out[j] = SegHit{tt, p.tri, p.surf, p.mat};
                        n++;
                    }
                }
            }
        } else {
            const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes) + ref);
            float4 n0 = q[0], n1 = q[1], n2 = q[2], n3 = q[3];
            const float lmin[3] = {n0.x, n0.y, n0.z}, lmax[3] = {n1.x, n1.y, n1.z};
            const float rmin[3] = {n2.x, n2.y, n2.z}, rmax[3] = {n3.x, n3.y, n3.z};
            if (COUNT) cnt.aabb += 2;
            if (ray_aabb(lmin, lmax, a, inv, len, te)) stack[sp++] = __float_as_uint(n0.w);
            if (ray_aabb(rmin, rmax, a, inv, len, te)) stack[sp++] = __float_as_uint(n1.w);
        }
    }
    if (overflow) return -1;
    // same-t dedup, lowest surface id wins (mesh.rs:589-603)
    float tol = 1e-4f / fmaxf(len, 1.0f);
    int w = 0;
    for (int i = 0; i < n; i++) {
        if (w > 0 && fabsf(out[i].t - out[w - 1].t) < tol) {
            if (out[i].surf < out[w - 1].surf) {
                out[w - 1].tri = out[i].tri;
                out[w - 1].surf = out[i].surf;
                out[w - 1].mat = out[i].mat;
            }
        } else {
            out[w++] = out[i];
        }
    }
    return w;
}

}  // namespace omg
