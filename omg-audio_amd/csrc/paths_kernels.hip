// AutoPaths::find (crates/omg-core/src/paths.rs:117-332) on the device.
//
// Two launches per call.  The direct pass prices every query's straight
// segment, one lane per query (paths.rs:124-135), writes the direct path
// and appends the queries whose mid band is blocked to a work list.  The
// bent pass then gives each blocked query one 256-lane workgroup:
//   1. single bends: every candidate edge's golden-section bend point and
//      knife-edge bound, one lane per edge, ranked by an LDS bitonic sort
//      (paths.rs:164-189; more than 1024 edges are first cut to the 1024
//      nearest the corridor, :169-171);
//   2. double bends: the two corridor rankings (:238-239), then every
//      (near-source, near-listener) pair's alternating bend points and
//      bound, one lane per pair, ranked and cut to 64 (:240-279);
//   3. every surviving path's legs are priced by segment transmission, and
//      the 23 silhouette drops are cast, one lane per leg or drop
//      (:190-232, :280-323, :337-405);
//   4. bent paths are assembled, ranked by their mid-band gain and cut to
//      the budget (:325-331).
// Ranking keys are f32::total_cmp order on the value with the candidate's
// position as the low word, so the sort reproduces the reference's stable
// sort_by exactly.  All arithmetic keeps the reference's f32 expressions.
#include "paths_kernels.h"
#include "bvh_device.h"

#pragma clang fp contract(off)

namespace omg {

namespace {

constexpr uint64_t kNoKey = ~0ull;

// f32::total_cmp descending, ties by ascending position: an ascending u64 key
__device__ __forceinline__ uint64_t desc_key(float v, uint32_t ord)
{
    int32_t k = __float_as_int(v);
    k ^= (int32_t)((uint32_t)(k >> 31) >> 1);
    uint32_t u = (uint32_t)k ^ 0x80000000u;
    return ((uint64_t)(~u) << 32) | ord;
}

// Ascending bitonic sort of k[0, n), n a power of two, by the whole block.
__device__ __forceinline__ void block_sort(uint64_t* k, uint32_t n)
{
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t j = size >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                uint32_t p = i ^ j;
                if (p > i) {
                    uint64_t x = k[i], y = k[p];
                    bool up = (i & size) == 0;
                    if ((x > y) == up) {
                        k[i] = y;
                        k[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Number of keys other than kNoKey among the first `limit` of a sorted array.
__device__ __forceinline__ uint32_t count_valid(const uint64_t* k, uint32_t limit, uint32_t* s_cnt)
{
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < limit; i += blockDim.x) c += k[i] != kNoKey ? 1u : 0u;
    if (c) atomicAdd(s_cnt, c);
    __syncthreads();
    uint32_t r = *s_cnt;
    __syncthreads();
    return r;
}

__device__ __forceinline__ f3 load3(const float* p, uint32_t i)
{
    return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__device__ __forceinline__ void load_edge(const PathEdgeDev* e, uint32_t i, f3& a, f3& b, float& imp)
{
    const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(e + i));
    float4 x = q[0], y = q[1];
    a = mk(x.x, x.y, x.z);
    imp = x.w;
    b = mk(y.x, y.y, y.z);
}

// One crossing of a segment: its position along it (0..1) and the
// crossed triangle's surface << 16 | material.
struct Crossing {
    float t;
    uint32_t sm;
};

// Slab test of one box against the ray interval [tlo, thi] (ray_aabb,
// mesh.rs:409-426, with t0 starting at tlo): tlo = 0, thi = len is the
// reference's segment test.
__device__ __forceinline__ bool ray_range(const float bmin[3], const float bmax[3], f3 o, f3 inv,
                                          float tlo, float thi)
{
    float t0 = tlo, t1 = thi;
    ray_slab(bmin[0], bmax[0], o.x, inv.x, t0, t1);
    ray_slab(bmin[1], bmax[1], o.y, inv.y, t0, t1);
    ray_slab(bmin[2], bmax[2], o.z, inv.z, t0, t1);
    return !(t0 > t1);
}

// Mesh::segment_hits' traversal (mesh.rs:546-587) over the ray interval
// [tlo, thi], keeping the crossings with lo <= t/len < hi: stable insertion
// by t, so equal positions keep the visiting order.  The walk continues
// into the right child when its box is hit and pushes the left one only
// when both are (the reference pushes left then right and pops right
// first: the same leaves in the same order, with fewer stack operations).
// With the whole interval ([0, len], lo = -inf, hi = inf) it performs the
// reference's box and triangle tests exactly.  Returns the count, or -1
// when more than `cap` crossings were found.  PRIM_STEP: a leaf's
// primitives take one step each, so the lanes of a wave in leaves and
// those in inner nodes share every step (faster for the short, many-lane
// interval walks; the whole-leaf loop is faster for long one- and two-lane
// walks).  Either way the same tests in the same order.
template <bool COUNT, bool PRIM_STEP = false>
__device__ __forceinline__ int seg_walk(const GeomArgs& G, f3 a, f3 dn, f3 inv, float len, float tlo,
                                        float thi, float lo, float hi, Crossing* out, int cap,
                                        TravCount& tc)
{
    if (COUNT) tc.aabb++;
    if (!ray_range(G.root_min, G.root_max, a, inv, tlo, thi)) return 0;
    uint32_t stack[kStackCap];
    int sp = 0, n = 0;
    bool overflow = false;
    // cur: the inner node to test next; [k, kend): the primitives of the
    // leaf being tested (one per step with PRIM_STEP, else all in one)
    uint32_t cur = G.root_ref, k = 0, kend = 0;
    auto enter = [&](uint32_t ref) {
        if (ref & kLeafBit) {
            k = ref & kLeafStartMask;
            kend = k + ((ref >> kLeafCountShift) & 7u);
        } else {
            cur = ref;
        }
    };
    enter(cur);
    while (true) {
        if (k < kend) {
            do {
                Prim p = load_prim(G.prims, k);
                k++;
                if (COUNT) tc.tri++;
                float t, u, v;
                if (ray_tri(p, a, dn, t, u, v)) {
                    float tt = t / len;
                    if (tt > 1e-4f && tt < 1.0f - 1e-4f && tt >= lo && tt < hi) {
                        if (n == cap) {
                            overflow = true;
                        } else {
                            int j = n;
                            while (j > 0 && out[j - 1].t > tt) {
                                out[j] = out[j - 1];
                                j--;
                            }
                            out[j] = Crossing{tt, (p.surf << 16) | (p.mat & 0xFFFFu)};
                            n++;
                        }
                    }
                }
            } while (!PRIM_STEP && k < kend);
            if (k < kend) continue;   // the leaf's next primitive, next step
            if (sp == 0) break;
            enter(stack[--sp]);
            continue;
        }
        const OMG_GLOBAL float4* q = reinterpret_cast<const OMG_GLOBAL float4*>(glb(G.nodes) + cur);
        float4 n0 = q[0], n1 = q[1], n2 = q[2], n3 = q[3];
        const float lmin[3] = {n0.x, n0.y, n0.z}, lmax[3] = {n1.x, n1.y, n1.z};
        const float rmin[3] = {n2.x, n2.y, n2.z}, rmax[3] = {n3.x, n3.y, n3.z};
        if (COUNT) tc.aabb += 2;
        bool hl = ray_range(lmin, lmax, a, inv, tlo, thi);
        bool hr = ray_range(rmin, rmax, a, inv, tlo, thi);
        if (hr) {
            if (hl) stack[sp++] = __float_as_uint(n0.w);
            enter(__float_as_uint(n1.w));
        } else if (hl) {
            enter(__float_as_uint(n0.w));
        } else {
            if (sp == 0) break;
            enter(stack[--sp]);
        }
    }
    return overflow ? -1 : n;
}

// segment_hits' same-t dedup (mesh.rs:589-603): within 1e-4 / max(len, 1)
// of the cluster's first crossing the lowest surface id wins.
__device__ __forceinline__ int dedup(Crossing* c, int n, float len)
{
    float tol = 1e-4f / fmaxf(len, 1.0f);
    int w = 0;
    for (int i = 0; i < n; i++) {
        if (w > 0 && fabsf(c[i].t - c[w - 1].t) < tol) {
            if ((c[i].sm >> 16) < (c[w - 1].sm >> 16)) c[w - 1].sm = c[i].sm;
        } else {
            c[w++] = c[i];
        }
    }
    return w;
}

// AutoPaths::transmission's pairing (paths.rs:87-112): crossings paired
// entry/exit by material give the wall thickness of
// Material::transmission_at (material.rs:59-62).
__device__ __forceinline__ void pair_gains(const omg_material_dev* mats, const Crossing* c, int n,
                                           float len, float g[3])
{
    g[0] = g[1] = g[2] = 1.0f;
    int i = 0;
    while (i < n) {
        int enter = i;
        uint32_t m = c[enter].sm & 0xFFFFu;
        float th;
        if (i + 1 < n && (c[i + 1].sm & 0xFFFFu) == m) {
            i += 1;
            th = fmaxf((c[i].t - c[enter].t) * len, 0.01f);
        } else {
            th = 0.1f;
        }
        float k = fminf(0.15f / fmaxf(th, 0.02f), 2.0f);
        for (int band = 0; band < 3; band++) g[band] *= fminf(mats[m].transmission[band] * k, 0.9f);
        i += 1;
    }
}

// AutoPaths::transmission (paths.rs:84-113) of a -> b on one lane.
// Returns -1 when the crossing list overflowed.
template <bool COUNT>
__device__ __forceinline__ int transmission(const GeomArgs& G, const omg_material_dev* mats, f3 a, f3 b,
                                            float g[3], TravCount& tc)
{
    g[0] = g[1] = g[2] = 1.0f;
    if (G.empty) return 0;
    f3 d = sub(b, a);
    float len = length(d);
    if (len < 1e-6f) return 0;
    f3 dn = mul(d, 1.0f / len);
    f3 inv = mk(1.0f / dn.x, 1.0f / dn.y, 1.0f / dn.z);
    Crossing buf[kPathSegCap];
    int n = seg_walk<COUNT>(G, a, dn, inv, len, 0.0f, len, -INFINITY, INFINITY, buf, kPathSegCap, tc);
    if (n < 0) return -1;
    n = dedup(buf, n, len);
    pair_gains(mats, buf, n, len, g);
    return 0;
}

// rank_corridor (paths.rs:143-163) keys: edges whose midpoint projects into
// [lo_t, hi_t] of the segment, scored importance / (d^2 + 0.5).
__device__ __forceinline__ void corridor_keys(const PathsArgs& A, f3 src, f3 seg, float seg_len2,
                                              float lo_t, float hi_t, uint64_t* sk, uint32_t n_pad)
{
    for (uint32_t i = threadIdx.x; i < n_pad; i += blockDim.x) {
        uint64_t key = kNoKey;
        if (i < A.ne) {
            f3 a, b;
            float imp;
            load_edge(A.edges, i, a, b, imp);
            f3 c = mul(add(a, b), 0.5f);
            float t = dot(sub(c, src), seg) / seg_len2;
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
            if (!(t < lo_t || t > hi_t)) {
                f3 on = add(src, mul(seg, t));
                float d2 = dot(sub(c, on), sub(c, on));
                key = desc_key(imp / (d2 + 0.5f), i);
            }
        }
        sk[i] = key;
    }
    __syncthreads();
}

__device__ __forceinline__ void flush_counts(const PathsArgs& A, const TravCount& tc)
{
    if (tc.aabb) atomicAdd(reinterpret_cast<unsigned long long*>(A.trav), (unsigned long long)tc.aabb);
    if (tc.tri) atomicAdd(reinterpret_cast<unsigned long long*>(A.trav) + 1, (unsigned long long)tc.tri);
}

__device__ __forceinline__ void write_path(PathDev* o, uint32_t key, const f3* pts, uint32_t np,
                                           float length, const float g[3])
{
    o->key = key;
    o->n_points = np;
    o->length = length;
    for (int b = 0; b < 3; b++) o->gains[b] = g[b];
    for (uint32_t k = 0; k < (uint32_t)kPathMaxPoints; k++) {
        f3 p = k < np ? pts[k] : mk(0.0f, 0.0f, 0.0f);
        o->points[k][0] = p.x;
        o->points[k][1] = p.y;
        o->points[k][2] = p.z;
    }
}

// AutoPaths::transmission (paths.rs:84-113) of one segment per group of K
// lanes, called by every lane of a 64-lane block (QPB = 64 / K groups).
// Lane k of a group walks only the boxes the ray meets near the interval
// k/K <= t < (k+1)/K of the segment's parameter and keeps the crossings in
// it; the groups' lists, concatenated in interval order, are the whole
// walk's sorted list (equal positions fall in one interval and keep the
// visiting order), which the group's lane 0 dedups and pairs.  K = 1 is
// the reference's walk exactly (the counted pass).  Lane 0 gets the gains
// and returns -1 on overflow; other lanes return 0.
// Crossings one lane keeps, and a group's merged list: a lane whose
// interval holds more sends the segment to the one-lane walk, so a group
// never merges more than K x lane_cap (nor more than the walk's 64).
constexpr int lane_cap(int K) { return K == 1 ? kPathSegCap : (K >= 8 ? 16 : 8); }
constexpr int merge_cap(int K) { return K * lane_cap(K) < kPathSegCap ? K * lane_cap(K) : kPathSegCap; }

template <bool COUNT, int K>
__device__ __forceinline__ int split_transmission(const GeomArgs& G, const omg_material_dev* mats,
                                                  bool live, f3 a, f3 b, float g[3], TravCount& tc,
                                                  Crossing (*merged)[merge_cap(K)],
                                                  uint32_t (*lane_n)[K])
{
    constexpr int LANE_CAP = lane_cap(K);
    constexpr uint32_t kLaneOvf = 0xFFFFFFFFu;
    const uint32_t gq = threadIdx.x / K, k = threadIdx.x % K;
    Crossing buf[LANE_CAP];
    float len = 0.0f;
    int n = 0;
    if (live && !G.empty) {
        f3 d = sub(b, a);
        len = length(d);
        if (!(len < 1e-6f)) {
            f3 dn = mul(d, 1.0f / len);
            f3 inv = mk(1.0f / dn.x, 1.0f / dn.y, 1.0f / dn.z);
            float lo = k == 0 ? -INFINITY : (float)k / (float)K;
            float hi = k == K - 1 ? INFINITY : (float)(k + 1) / (float)K;
            float margin = 1e-4f * len + 1e-3f;
            float tlo = k == 0 ? 0.0f : lo * len - margin;
            float thi = k == K - 1 ? len : hi * len + margin;
            n = seg_walk<COUNT, (K >= 8)>(G, a, dn, inv, len, tlo, thi, lo, hi, buf, LANE_CAP, tc);
        }
    }
    lane_n[gq][k] = n < 0 ? kLaneOvf : (uint32_t)n;
    __syncthreads();
    uint32_t off = 0, total = 0;
    bool lane_ovf = false;
    for (int j = 0; j < K; j++) {
        uint32_t c = lane_n[gq][j];
        if (c == kLaneOvf) {
            lane_ovf = true;
        } else {
            total += c;
            if (j < (int)k) off += c;
        }
    }
    if (!lane_ovf && total <= (uint32_t)merge_cap(K))
        for (int i = 0; i < n; i++) merged[gq][off + i] = buf[i];
    __syncthreads();
    g[0] = g[1] = g[2] = 1.0f;
    if (!live || k != 0) return 0;
    if (lane_ovf)   // one interval held more crossings than a lane keeps: the whole walk here
        return transmission<false>(G, mats, a, b, g, tc);
    if (total > (uint32_t)merge_cap(K)) return -1;   // more than the walk's 64 crossings
    int m = dedup(merged[gq], (int)total, len);
    pair_gains(mats, merged[gq], m, len, g);
    return 0;
}

// The direct segment of every query (paths.rs:124-135), kDirectSplit lanes
// per segment (one in the counted pass); blocked queries go to the work list.
template <bool COUNT>
__global__ __launch_bounds__(64) void k_paths_direct(PathsArgs A)
{
    constexpr int K = COUNT ? 1 : kDirectSplit;
    constexpr int QPB = 64 / K;
    __shared__ Crossing merged[QPB][merge_cap(K)];
    __shared__ uint32_t lane_n[QPB][K];
    const uint32_t gq = threadIdx.x / K, k = threadIdx.x % K;
    const uint32_t q = blockIdx.x * QPB + gq;
    const bool live = q < A.n;
    TravCount tc;
    f3 s = mk(0.0f, 0.0f, 0.0f), l = s;
    if (live) {
        s = load3(A.src, q);
        l = load3(A.lis, q);
    }
    float g[3];
    int rc = split_transmission<COUNT, K>(A.g, A.mats, live, s, l, g, tc, merged, lane_n);
    if (live && k == 0) {
        if (rc < 0) atomicOr(&A.counters[1], 1u);
        f3 pts[2] = {s, l};
        PathDev* o = A.out + (size_t)q * (1 + A.max_paths);
        write_path(o, 0u, pts, 2, length(sub(l, s)), g);
        // the bent slots start cleared: the assembly pass fills the kept ones
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t i = 1; i <= A.max_paths; i++) write_path(o + i, 0u, pts, 0, 0.0f, zero);
        A.n_out[q] = 1;
        if (g[1] < 0.25f) {
            uint32_t slot = atomicAdd(&A.counters[0], 1u);
            A.work[slot] = q;
        }
    }
    if (COUNT) flush_counts(A, tc);
}

// Ranking pass, one blocked query per workgroup (paths.rs:136-279): the
// single-bend and pair rankings, the ranked candidates' bend points and
// the endpoints of every leg to price, written to the query's plan.  The
// bend-point chains are serial and VALU-bound: 8 waves per SIMD (64 VGPRs)
// keep more of them in flight.
__global__ __launch_bounds__(kPathBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_paths_rank(PathsArgs A)
{
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_cnt;
    if (blockIdx.x >= A.counters[0]) return;
    const RankLds lay = rank_lds(A.ne, A.edge_candidates, A.pair_edges);
    uint64_t* sk = reinterpret_cast<uint64_t*>(smem + lay.keys);           // sort keys
    uint32_t* pre = reinterpret_cast<uint32_t*>(smem + lay.pre);           // single-bend candidates
    uint32_t* near_s = reinterpret_cast<uint32_t*>(smem + lay.near_s);
    uint32_t* near_l = reinterpret_cast<uint32_t*>(smem + lay.near_l);
    uint32_t* sel1 = reinterpret_cast<uint32_t*>(smem + lay.sel);          // ranked single edges
    uint32_t (*pe)[3] = reinterpret_cast<uint32_t (*)[3]>(smem + lay.pe);  // ranked pairs (ei, ej, position)
    const PlanLayout pl = plan_layout(A.ne, A.edge_candidates, A.pair_edges);
    float* F = A.plan + (size_t)blockIdx.x * pl.words;
    uint32_t* P = reinterpret_cast<uint32_t*>(F);

    const uint32_t q = A.work[blockIdx.x];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const f3 src = load3(A.src, q), lis = load3(A.lis, q);
    const f3 seg = sub(lis, src);
    const float seg_len2 = fmaxf(dot(seg, seg), 1e-6f);
    const float direct_len = length(sub(lis, src));
    const uint32_t ne = A.ne;

    // ---- single bends (paths.rs:164-189)
    uint32_t npre;
    if (ne > kPathPrefilter) {
        uint32_t n_pad = pow2_ceil(ne);
        corridor_keys(A, src, seg, seg_len2, 0.0f, 1.0f, sk, n_pad);
        block_sort(sk, n_pad);
        npre = count_valid(sk, kPathPrefilter, &s_cnt);
        for (uint32_t i = tid; i < npre; i += bd) pre[i] = (uint32_t)sk[i];
    } else {
        npre = ne;
        for (uint32_t i = tid; i < npre; i += bd) pre[i] = i;
    }
    __syncthreads();
    uint32_t n_pad = pow2_ceil(npre);
    for (uint32_t i = tid; i < n_pad; i += bd) {
        uint64_t key = kNoKey;
        if (i < npre) {
            f3 a, b;
            float imp;
            load_edge(A.edges, pre[i], a, b, imp);
            f3 p = bend_point(a, b, src, lis);
            F[pl.csp + 3 * i] = p.x;
            F[pl.csp + 3 * i + 1] = p.y;
            F[pl.csp + 3 * i + 2] = p.z;
            float d1 = length(sub(p, src)), d2 = length(sub(lis, p));
            if (!(d1 < 0.2f || d2 < 0.2f)) {
                float detour = d1 + d2 - direct_len;
                key = desc_key(knife_edge_amp(fmaxf(detour, 1e-4f), band_hz(1)), i);
            }
        }
        sk[i] = key;
    }
    __syncthreads();
    block_sort(sk, n_pad);
    uint32_t n1 = count_valid(sk, n_pad < A.edge_candidates ? n_pad : A.edge_candidates, &s_cnt);
    // sel1: the ranked candidates' positions (edge = pre[position])
    for (uint32_t r = tid; r < n1; r += bd) sel1[r] = (uint32_t)sk[r];
    __syncthreads();

    // ---- double bends (paths.rs:233-279)
    uint32_t ns = 0, nl = 0;
    if (ne > 0 && A.pair_edges > 0) {
        uint32_t n_pad_e = pow2_ceil(ne);
        corridor_keys(A, src, seg, seg_len2, 0.0f, 0.65f, sk, n_pad_e);
        block_sort(sk, n_pad_e);
        ns = count_valid(sk, n_pad_e < A.pair_edges ? n_pad_e : A.pair_edges, &s_cnt);
        for (uint32_t r = tid; r < ns; r += bd) near_s[r] = (uint32_t)sk[r];
        __syncthreads();
        corridor_keys(A, src, seg, seg_len2, 0.35f, 1.0f, sk, n_pad_e);
        block_sort(sk, n_pad_e);
        nl = count_valid(sk, n_pad_e < A.pair_edges ? n_pad_e : A.pair_edges, &s_cnt);
        for (uint32_t r = tid; r < nl; r += bd) near_l[r] = (uint32_t)sk[r];
        __syncthreads();
    }
    const uint32_t npair = ns * nl;
    n_pad = pow2_ceil(npair);
    for (uint32_t i = tid; i < n_pad; i += bd) {
        uint64_t key = kNoKey;
        if (i < npair) {
            uint32_t ei = near_s[i / nl], ej = near_l[i % nl];
            if (ei != ej) {
                f3 e1a, e1b, e2a, e2b;
                float imp;
                load_edge(A.edges, ei, e1a, e1b, imp);
                load_edge(A.edges, ej, e2a, e2b, imp);
                f3 p2 = mul(add(e2a, e2b), 0.5f);
                f3 p1 = bend_point(e1a, e1b, src, p2);
                for (int it = 0; it < 3; it++) {
                    p2 = bend_point(e2a, e2b, p1, lis);
                    p1 = bend_point(e1a, e1b, src, p2);
                }
                float* cp = F + pl.cpp + 6 * i;
                cp[0] = p1.x; cp[1] = p1.y; cp[2] = p1.z;
                cp[3] = p2.x; cp[4] = p2.y; cp[5] = p2.z;
                float d1 = length(sub(p1, src)), dm = length(sub(p2, p1)), d2 = length(sub(lis, p2));
                if (!(d1 < 0.2f || dm < 0.2f || d2 < 0.2f)) {
                    float det1 = d1 + dm - length(sub(p2, src));
                    float det2 = dm + d2 - length(sub(lis, p1));
                    float bound = knife_edge_amp(fmaxf(det1, 1e-4f), band_hz(1)) *
                                  knife_edge_amp(fmaxf(det2, 1e-4f), band_hz(1));
                    key = desc_key(bound, i);
                }
            }
        }
        sk[i] = key;
    }
    __syncthreads();
    block_sort(sk, n_pad);
    uint32_t n2 = count_valid(sk, n_pad < kPathPairKeep ? n_pad : kPathPairKeep, &s_cnt);
    for (uint32_t r = tid; r < n2; r += bd) {
        uint32_t i = (uint32_t)sk[r];
        pe[r][0] = near_s[i / nl];
        pe[r][1] = near_l[i % nl];
        pe[r][2] = i;
    }
    __syncthreads();

    // ---- the plan: the ranked candidates' bend points and their legs
    const f3 flat = mk(lis.x - src.x, lis.y - src.y, 0.0f);
    const float run = length(flat);
    const uint32_t n_sil = run < 1.0f ? 0u : (uint32_t)(kSilSamples - 1);
    const uint32_t n_legs = 2 * n1 + 3 * n2 + n_sil;
    auto put = [&](uint32_t leg, f3 a, f3 b) {
        float* L = F + pl.legs + 6 * leg;
        L[0] = a.x; L[1] = a.y; L[2] = a.z;
        L[3] = b.x; L[4] = b.y; L[5] = b.z;
    };
    for (uint32_t r = tid; r < n1 + n2 + n_sil; r += bd) {
        if (r < n1) {
            const uint32_t i = sel1[r];
            f3 p = mk(F[pl.csp + 3 * i], F[pl.csp + 3 * i + 1], F[pl.csp + 3 * i + 2]);
            P[pl.sel + r] = pre[i];
            F[pl.sp + 3 * r] = p.x;
            F[pl.sp + 3 * r + 1] = p.y;
            F[pl.sp + 3 * r + 2] = p.z;
            f3 e1 = A.taps ? add(p, mul(normalize(sub(p, src)), 0.06f)) : p;
            f3 e2 = A.taps ? add(p, mul(normalize(sub(p, lis)), 0.06f)) : p;
            put(2 * r, src, e1);
            put(2 * r + 1, e2, lis);
        } else if (r < n1 + n2) {
            uint32_t k = r - n1;
            const float* cp = F + pl.cpp + 6 * pe[k][2];
            f3 p1 = mk(cp[0], cp[1], cp[2]), p2 = mk(cp[3], cp[4], cp[5]);
            P[pl.pe + 2 * k] = pe[k][0];
            P[pl.pe + 2 * k + 1] = pe[k][1];
            float* pp = F + pl.pp + 6 * k;
            pp[0] = p1.x; pp[1] = p1.y; pp[2] = p1.z;
            pp[3] = p2.x; pp[4] = p2.y; pp[5] = p2.z;
            f3 x1 = A.taps ? add(p1, mul(normalize(sub(p1, src)), 0.06f)) : p1;
            f3 m1 = A.taps ? add(p1, mul(normalize(sub(p1, p2)), 0.06f)) : p1;
            f3 m2 = A.taps ? add(p2, mul(normalize(sub(p2, p1)), 0.06f)) : p2;
            f3 x2 = A.taps ? add(p2, mul(normalize(sub(p2, lis)), 0.06f)) : p2;
            uint32_t base = 2 * n1 + 3 * k;
            put(base, src, x1);
            put(base + 1, m1, m2);
            put(base + 2, x2, lis);
        } else {
            // over_silhouette's drop i (paths.rs:344-350): its origin
            uint32_t i = r - n1 - n2 + 1;
            float t = (float)i / (float)kSilSamples;
            put(2 * n1 + 3 * n2 + i - 1, mk(src.x + flat.x * t, src.y + flat.y * t, 60.0f),
                mk(0.0f, 0.0f, -1.0f));
        }
    }
    if (tid == 0) {
        P[0] = n1;
        P[1] = n2;
        P[2] = n_legs;
        P[3] = n_sil;
    }
}

// Leg pass: every leg of every blocked query, K lanes per leg (one in the
// counted pass); a silhouette drop (paths.rs:350-358) is one downward
// raycast on the group's first lane.
template <bool COUNT, int K>
__global__ __launch_bounds__(64) void k_paths_legs(PathsArgs A)
{
    constexpr int QPB = 64 / K;
    __shared__ Crossing merged[QPB][merge_cap(K)];
    __shared__ uint32_t lane_n[QPB][K];
    const PlanLayout pl = plan_layout(A.ne, A.edge_candidates, A.pair_edges);
    const uint32_t gq = threadIdx.x / K, k = threadIdx.x % K;
    const uint64_t u = (uint64_t)blockIdx.x * QPB + gq;
    const uint32_t slot = (uint32_t)(u / pl.max_legs), leg = (uint32_t)(u % pl.max_legs);
    float* F = A.plan + (size_t)slot * pl.words;
    const uint32_t* P = reinterpret_cast<const uint32_t*>(F);
    bool live = false, drop = false;
    if (slot < A.counters[0]) {
        uint32_t n_legs = P[2];
        live = leg < n_legs;
        drop = live && leg >= n_legs - P[3];
    }
    TravCount tc;
    const float* L = F + pl.legs + 6 * leg;
    f3 a = mk(0.0f, 0.0f, 0.0f), b = a;
    if (live) {
        a = mk(L[0], L[1], L[2]);
        b = mk(L[3], L[4], L[5]);
    }
    float g[3];
    int rc = split_transmission<COUNT, K>(A.g, A.mats, live && !drop, a, b, g, tc, merged, lane_n);
    if (live && k == 0) {
        float* out = F + pl.lg + 3 * leg;
        if (drop) {
            float th, z = -INFINITY;
            uint32_t tri, flags;
            if (mesh_raycast<COUNT>(A.g, a, b, th, tri, flags, tc)) {
                // the drop's line height is the segment's at the same t (paths.rs:347-349)
                uint32_t i = leg - (P[2] - P[3]) + 1;
                float t = (float)i / (float)kSilSamples;
                const uint32_t q = A.work[slot];
                f3 src = load3(A.src, q), lis = load3(A.lis, q);
                float line_z = src.z + (lis.z - src.z) * t;
                float top = 60.0f - th;
                if (top > line_z + 0.05f) z = top + 0.03f;
            }
            out[0] = z;
        } else {
            if (rc < 0) atomicOr(&A.counters[1], 1u);
            out[0] = g[0];
            out[1] = g[1];
            out[2] = g[2];
        }
    }
    if (COUNT) flush_counts(A, tc);
}

// Assembly pass, one blocked query per workgroup (paths.rs:190-232,
// :280-331, :337-405): gains of the priced candidates, the silhouette hull,
// the final mid-band ranking and the output.
__global__ __launch_bounds__(kPathBlock) void k_paths_assemble(PathsArgs A)
{
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_cnt, s_hull;
    if (blockIdx.x >= A.counters[0]) return;
    const AsmLds lay = asm_lds(A.edge_candidates, A.pair_edges);
    uint64_t* sk = reinterpret_cast<uint64_t*>(smem + lay.keys);
    float (*bg)[4] = reinterpret_cast<float (*)[4]>(smem + lay.bg);        // gains, length
    uint32_t* bkey = reinterpret_cast<uint32_t*>(smem + lay.bkey);
    float (*hull)[3] = reinterpret_cast<float (*)[3]>(smem + lay.hull);
    const PlanLayout pl = plan_layout(A.ne, A.edge_candidates, A.pair_edges);
    const float* F = A.plan + (size_t)blockIdx.x * pl.words;
    const uint32_t* P = reinterpret_cast<const uint32_t*>(F);
    const uint32_t q = A.work[blockIdx.x];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const f3 src = load3(A.src, q), lis = load3(A.lis, q);
    const float direct_len = length(sub(lis, src));
    const uint32_t n1 = P[0], n2 = P[1], n_legs = P[2], n_sil = P[3];
    const float* lg = F + pl.lg;
    const uint32_t nb = n1 + n2 + 1;
    for (uint32_t r = tid; r < n1 + n2; r += bd) {
        bool ok;
        float g[3], len;
        uint32_t key;
        if (r < n1) {
            f3 p = mk(F[pl.sp + 3 * r], F[pl.sp + 3 * r + 1], F[pl.sp + 3 * r + 2]);
            const float* l1 = lg + 3 * (2 * r);
            const float* l2 = l1 + 3;
            float d1 = length(sub(p, src)), d2 = length(sub(lis, p));
            ok = !(l1[0] * l2[0] < 1e-4f);
            float detour = d1 + d2 - direct_len;
            for (int b = 0; b < 3; b++)
                g[b] = knife_edge_amp(fmaxf(detour, 1e-4f), band_hz(b)) * l1[b] * l2[b];
            len = d1 + d2;
            key = 1u + P[pl.sel + r];
        } else {
            uint32_t k = r - n1;
            const float* pp = F + pl.pp + 6 * k;
            f3 p1 = mk(pp[0], pp[1], pp[2]), p2 = mk(pp[3], pp[4], pp[5]);
            const float* l1 = lg + 3 * (2 * n1 + 3 * k);
            const float* lm = l1 + 3;
            const float* l2 = l1 + 6;
            float d1 = length(sub(p1, src)), dm = length(sub(p2, p1)), d2 = length(sub(lis, p2));
            ok = !(l1[0] * lm[0] * l2[0] < 1e-4f);
            float det1 = d1 + dm - length(sub(p2, src));
            float det2 = dm + d2 - length(sub(lis, p1));
            for (int b = 0; b < 3; b++)
                g[b] = knife_edge_amp(fmaxf(det1, 1e-4f), band_hz(b)) *
                       knife_edge_amp(fmaxf(det2, 1e-4f), band_hz(b)) * l1[b] * lm[b] * l2[b];
            len = d1 + dm + d2;
            key = 100000u + P[pl.pe + 2 * k] * 4096u + P[pl.pe + 2 * k + 1];
        }
        for (int b = 0; b < 3; b++) bg[r][b] = g[b];
        bg[r][3] = len;
        bkey[r] = key;
        sk[r] = ok ? desc_key(g[1], r) : kNoKey;
    }
    if (tid == bd - 1) {
        // over_silhouette (paths.rs:337-405): upper hull of the drops
        const uint32_t slot = n1 + n2;
        const f3 flat = mk(lis.x - src.x, lis.y - src.y, 0.0f);
        const float run = length(flat);
        uint64_t key = kNoKey;
        s_hull = 0;
        if (n_sil) {
            const float* drops = lg + 3 * (n_legs - n_sil);
            float ps[kSilSamples + 1], pz[kSilSamples + 1];
            int np = 0;
            ps[np] = 0.0f;
            pz[np] = src.z;
            np++;
            for (int i = 1; i < kSilSamples; i++) {
                float z = drops[3 * (i - 1)];
                if (z != -INFINITY) {
                    float t = (float)i / (float)kSilSamples;
                    ps[np] = run * t;
                    pz[np] = z;
                    np++;
                }
            }
            if (np > 1) {
                ps[np] = run;
                pz[np] = lis.z;
                np++;
                float hs[kSilSamples + 1], hz[kSilSamples + 1];
                int nh = 0;
                for (int k = 0; k < np; k++) {
                    while (nh >= 2) {
                        float ax = hs[nh - 2], az = hz[nh - 2], bx = hs[nh - 1], bz = hz[nh - 1];
                        if ((bx - ax) * (pz[k] - az) - (bz - az) * (ps[k] - ax) >= 0.0f)
                            nh--;
                        else
                            break;
                    }
                    hs[nh] = ps[k];
                    hz[nh] = pz[k];
                    nh++;
                }
                if (nh >= 3) {
                    auto d2f = [&](int i, int j) {
                        return sqrtf((hs[i] - hs[j]) * (hs[i] - hs[j]) + (hz[i] - hz[j]) * (hz[i] - hz[j]));
                    };
                    float g[3] = {1.0f, 1.0f, 1.0f};
                    float len = 0.0f;
                    for (int i = 0; i + 1 < nh; i++) len += d2f(i, i + 1);
                    for (int i = 1; i < nh - 1; i++) {
                        float detour = d2f(i - 1, i) + d2f(i, i + 1) - d2f(i - 1, i + 1);
                        for (int b = 0; b < 3; b++) g[b] *= knife_edge_amp(fmaxf(detour, 1e-4f), band_hz(b));
                    }
                    for (int i = 0; i < nh; i++) {
                        float t = hs[i] / run;
                        hull[i][0] = src.x + flat.x * t;
                        hull[i][1] = src.y + flat.y * t;
                        hull[i][2] = hz[i];
                    }
                    s_hull = (uint32_t)nh;
                    for (int b = 0; b < 3; b++) bg[slot][b] = g[b];
                    bg[slot][3] = len;
                    bkey[slot] = 90000u;
                    key = desc_key(g[1], slot);
                }
            }
        }
        sk[slot] = key;
    }
    __syncthreads();
    uint32_t n_pad = pow2_ceil(nb);
    for (uint32_t i = nb + tid; i < n_pad; i += bd) sk[i] = kNoKey;
    __syncthreads();
    block_sort(sk, n_pad);
    uint32_t keep = count_valid(sk, n_pad < A.max_paths ? n_pad : A.max_paths, &s_cnt);
    PathDev* o = A.out + (size_t)q * (1 + A.max_paths);
    for (uint32_t i = tid; i < keep; i += bd) {
        uint32_t r = (uint32_t)sk[i];
        f3 pts[kPathMaxPoints];
        uint32_t np;
        if (r < n1) {
            pts[0] = src;
            pts[1] = mk(F[pl.sp + 3 * r], F[pl.sp + 3 * r + 1], F[pl.sp + 3 * r + 2]);
            pts[2] = lis;
            np = 3;
        } else if (r < n1 + n2) {
            const float* pp = F + pl.pp + 6 * (r - n1);
            pts[0] = src;
            pts[1] = mk(pp[0], pp[1], pp[2]);
            pts[2] = mk(pp[3], pp[4], pp[5]);
            pts[3] = lis;
            np = 4;
        } else {
            np = s_hull;
            for (uint32_t k = 0; k < np; k++) pts[k] = mk(hull[k][0], hull[k][1], hull[k][2]);
        }
        float g[3] = {bg[r][0], bg[r][1], bg[r][2]};
        write_path(o + 1 + i, bkey[r], pts, np, bg[r][3], g);
    }
    if (tid == 0) A.n_out[q] = 1 + keep;
}

// Batched AutoPaths::transmission (paths.rs:84-113): each segment over
// kDirectSplit lanes, as the direct pass.
__global__ __launch_bounds__(64) void k_path_transmission(PathTransArgs A)
{
    constexpr int K = kDirectSplit;
    constexpr int QPB = 64 / K;
    __shared__ Crossing merged[QPB][merge_cap(K)];
    __shared__ uint32_t lane_n[QPB][K];
    const uint32_t gq = threadIdx.x / K, k = threadIdx.x % K;
    const uint32_t i = blockIdx.x * QPB + gq;
    const bool live = i < A.n;
    TravCount tc;
    f3 a = mk(0.0f, 0.0f, 0.0f), b = a;
    if (live) {
        a = load3(A.a, i);
        b = load3(A.b, i);
    }
    float g[3];
    int rc = split_transmission<false, K>(A.g, A.mats, live, a, b, g, tc, merged, lane_n);
    if (live && k == 0) {
        if (rc < 0) atomicOr(&A.counters[1], 1u);
        for (int band = 0; band < 3; band++) A.gains[3 * i + band] = g[band];
    }
}

template <bool COUNT>
hipError_t launch_paths_t(const PathsArgs& A, hipStream_t s)
{
    if (A.n == 0) return hipSuccess;
    const uint32_t qd = COUNT ? 64u : 64u / kDirectSplit;
    hipLaunchKernelGGL(k_paths_direct<COUNT>, dim3((A.n + qd - 1) / qd), dim3(64), 0, s, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const RankLds rl = rank_lds(A.ne, A.edge_candidates, A.pair_edges);
    hipLaunchKernelGGL(k_paths_rank, dim3(A.n), dim3(kPathBlock), rl.bytes, s, A);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // one group per (query slot, leg); slots past the blocked count exit.
    // Few legs are split over 8 lanes each (latency); when the legs give
    // at least 2M lanes at 2 per leg, over 2 (less repeated descent through
    // the top of the tree).
    const PlanLayout pl = plan_layout(A.ne, A.edge_candidates, A.pair_edges);
    const uint64_t units = (uint64_t)A.n * pl.max_legs;
    const uint32_t K = COUNT ? 1u : (units * kLegSplitMany >= (2ull << 20) ? kLegSplitMany : kLegSplitFew);
    const uint64_t ql = 64u / K;
    const uint64_t blocks = (units + ql - 1) / ql;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    if (COUNT)
        hipLaunchKernelGGL((k_paths_legs<COUNT, 1>), dim3((uint32_t)blocks), dim3(64), 0, s, A);
    else if (K == kLegSplitMany)
        hipLaunchKernelGGL((k_paths_legs<COUNT, kLegSplitMany>), dim3((uint32_t)blocks), dim3(64), 0, s, A);
    else
        hipLaunchKernelGGL((k_paths_legs<COUNT, kLegSplitFew>), dim3((uint32_t)blocks), dim3(64), 0, s, A);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const AsmLds al = asm_lds(A.edge_candidates, A.pair_edges);
    hipLaunchKernelGGL(k_paths_assemble, dim3(A.n), dim3(kPathBlock), al.bytes, s, A);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_paths(const PathsArgs& A, bool count, hipStream_t s)
{
    return count ? launch_paths_t<true>(A, s) : launch_paths_t<false>(A, s);
}

hipError_t launch_path_transmission(const PathTransArgs& A, hipStream_t s)
{
    if (A.n == 0) return hipSuccess;
    const uint32_t qpb = 64u / kDirectSplit;
    hipLaunchKernelGGL(k_path_transmission, dim3((A.n + qpb - 1) / qpb), dim3(64), 0, s, A);
    return hipGetLastError();
}

}  // namespace omg
