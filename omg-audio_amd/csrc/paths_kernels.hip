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

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n)
{
    uint32_t p = 2;
    while (p < n) p <<= 1;
    return p;
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

// AutoPaths::transmission (paths.rs:84-113): crossings paired entry/exit by
// material give the wall thickness of Material::transmission_at
// (material.rs:59-62).  Returns -1 when the crossing list overflowed.
template <bool COUNT>
__device__ __forceinline__ int transmission(const GeomArgs& G, const omg_material_dev* mats, f3 a, f3 b,
                                            float g[3], TravCount& tc)
{
    SegHit buf[kPathSegCap];
    g[0] = g[1] = g[2] = 1.0f;
    int n = segment_hits<COUNT>(G, a, b, buf, kPathSegCap, tc);
    if (n < 0) return -1;
    float len = length(sub(b, a));
    int i = 0;
    while (i < n) {
        int enter = i;
        uint32_t m = buf[enter].mat;
        float th;
        if (i + 1 < n && buf[i + 1].mat == m) {
            i += 1;
            th = fmaxf((buf[i].t - buf[enter].t) * len, 0.01f);
        } else {
            th = 0.1f;
        }
        float k = fminf(0.15f / fmaxf(th, 0.02f), 2.0f);
        for (int band = 0; band < 3; band++) g[band] *= fminf(mats[m].transmission[band] * k, 0.9f);
        i += 1;
    }
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

// The direct segment of every query (paths.rs:124-135).
template <bool COUNT>
__global__ __launch_bounds__(64) void k_paths_direct(PathsArgs A)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    TravCount tc;
    if (q < A.n) {
        f3 s = load3(A.src, q), l = load3(A.lis, q);
        float g[3];
        if (transmission<COUNT>(A.g, A.mats, s, l, g, tc) < 0) atomicOr(&A.counters[1], 1u);
        f3 pts[2] = {s, l};
        write_path(A.out + (size_t)q * (1 + A.max_paths), 0u, pts, 2, length(sub(l, s)), g);
        A.n_out[q] = 1;
        if (g[1] < 0.25f) {
            uint32_t slot = atomicAdd(&A.counters[0], 1u);
            A.work[slot] = q;
        }
    }
    if (COUNT) flush_counts(A, tc);
}

// One blocked query per workgroup (paths.rs:136-331).
template <bool COUNT>
__global__ __launch_bounds__(kPathBlock) void k_paths_bent(PathsArgs A)
{
    extern __shared__ uint64_t sk[];                                   // sort keys
    __shared__ uint32_t pre[kPathPrefilter];                           // single-bend candidates
    __shared__ uint32_t near_s[kPathMaxPairEdges], near_l[kPathMaxPairEdges];
    __shared__ uint32_t sel1[kPathMaxCandidates];                      // ranked single edges
    __shared__ float sp[kPathMaxCandidates][3];                        // their bend points
    __shared__ float sl[kPathMaxCandidates][6];                        // leg gains l1, l2
    __shared__ uint32_t pe[kPathPairKeep][2];                          // ranked pairs (ei, ej)
    __shared__ float pp[kPathPairKeep][6];                             // p1, p2
    __shared__ float pl[kPathPairKeep][9];                             // leg gains l1, lm, l2
    __shared__ float sil_z[kSilSamples];                               // roof height + 0.03, or -inf
    __shared__ float bg[kPathMaxCandidates + kPathPairKeep + 1][4];    // gains, length
    __shared__ uint32_t bkey[kPathMaxCandidates + kPathPairKeep + 1];
    __shared__ float hull[kPathMaxPoints][3];
    __shared__ uint32_t s_cnt, s_hull;

    if (blockIdx.x >= A.counters[0]) return;
    const uint32_t q = A.work[blockIdx.x];
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const f3 src = load3(A.src, q), lis = load3(A.lis, q);
    const f3 seg = sub(lis, src);
    const float seg_len2 = fmaxf(dot(seg, seg), 1e-6f);
    const float direct_len = length(sub(lis, src));
    const uint32_t ne = A.ne;
    TravCount tc;

    // ---- 1. single bends (paths.rs:164-189)
    uint32_t npre;
    if (ne > kPathPrefilter) {
        uint32_t n_pad = pow2_at_least(ne);
        corridor_keys(A, src, seg, seg_len2, 0.0f, 1.0f, sk, n_pad);
        block_sort(sk, n_pad);
        npre = count_valid(sk, kPathPrefilter, &s_cnt);
        for (uint32_t i = tid; i < npre; i += bd) pre[i] = (uint32_t)sk[i];
    } else {
        npre = ne;
        for (uint32_t i = tid; i < npre; i += bd) pre[i] = i;
    }
    __syncthreads();
    uint32_t n_pad = pow2_at_least(npre);
    for (uint32_t i = tid; i < n_pad; i += bd) {
        uint64_t key = kNoKey;
        if (i < npre) {
            f3 a, b;
            float imp;
            load_edge(A.edges, pre[i], a, b, imp);
            f3 p = bend_point(a, b, src, lis);
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
    for (uint32_t r = tid; r < n1; r += bd) sel1[r] = pre[(uint32_t)sk[r]];
    __syncthreads();

    // ---- 2. double bends (paths.rs:233-279)
    uint32_t ns = 0, nl = 0;
    if (ne > 0 && A.pair_edges > 0) {
        uint32_t n_pad_e = pow2_at_least(ne);
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
    n_pad = pow2_at_least(npair);
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
    }
    __syncthreads();

    // bend points of the ranked candidates (recomputed: the same f32 result)
    for (uint32_t r = tid; r < n1 + n2; r += bd) {
        if (r < n1) {
            f3 a, b;
            float imp;
            load_edge(A.edges, sel1[r], a, b, imp);
            f3 p = bend_point(a, b, src, lis);
            sp[r][0] = p.x;
            sp[r][1] = p.y;
            sp[r][2] = p.z;
        } else {
            uint32_t k = r - n1;
            f3 e1a, e1b, e2a, e2b;
            float imp;
            load_edge(A.edges, pe[k][0], e1a, e1b, imp);
            load_edge(A.edges, pe[k][1], e2a, e2b, imp);
            f3 p2 = mul(add(e2a, e2b), 0.5f);
            f3 p1 = bend_point(e1a, e1b, src, p2);
            for (int it = 0; it < 3; it++) {
                p2 = bend_point(e2a, e2b, p1, lis);
                p1 = bend_point(e1a, e1b, src, p2);
            }
            pp[k][0] = p1.x;
            pp[k][1] = p1.y;
            pp[k][2] = p1.z;
            pp[k][3] = p2.x;
            pp[k][4] = p2.y;
            pp[k][5] = p2.z;
        }
    }
    __syncthreads();

    // ---- 3. legs and silhouette drops, one lane each
    const f3 flat = mk(lis.x - src.x, lis.y - src.y, 0.0f);
    const float run = length(flat);
    const uint32_t n_sil = run < 1.0f ? 0u : (uint32_t)(kSilSamples - 1);
    const uint32_t n_tasks = 2 * n1 + 3 * n2 + n_sil;
    for (uint32_t task = tid; task < n_tasks; task += bd) {
        if (task < 2 * n1) {
            uint32_t r = task >> 1, leg = task & 1;
            f3 p = mk(sp[r][0], sp[r][1], sp[r][2]);
            f3 e = p;
            if (A.taps) e = add(p, mul(normalize(sub(p, leg == 0 ? src : lis)), 0.06f));
            float g[3];
            int rc = leg == 0 ? transmission<COUNT>(A.g, A.mats, src, e, g, tc)
                              : transmission<COUNT>(A.g, A.mats, e, lis, g, tc);
            if (rc < 0) atomicOr(&A.counters[1], 1u);
            for (int b = 0; b < 3; b++) sl[r][3 * leg + b] = g[b];
        } else if (task < 2 * n1 + 3 * n2) {
            uint32_t t2 = task - 2 * n1, r = t2 / 3, leg = t2 % 3;
            f3 p1 = mk(pp[r][0], pp[r][1], pp[r][2]), p2 = mk(pp[r][3], pp[r][4], pp[r][5]);
            f3 a, b;
            if (leg == 0) {
                a = src;
                b = A.taps ? add(p1, mul(normalize(sub(p1, src)), 0.06f)) : p1;
            } else if (leg == 1) {
                a = A.taps ? add(p1, mul(normalize(sub(p1, p2)), 0.06f)) : p1;
                b = A.taps ? add(p2, mul(normalize(sub(p2, p1)), 0.06f)) : p2;
            } else {
                a = A.taps ? add(p2, mul(normalize(sub(p2, lis)), 0.06f)) : p2;
                b = lis;
            }
            float g[3];
            if (transmission<COUNT>(A.g, A.mats, a, b, g, tc) < 0) atomicOr(&A.counters[1], 1u);
            for (int k = 0; k < 3; k++) pl[r][3 * leg + k] = g[k];
        } else {
            // over_silhouette's drop i (paths.rs:344-360)
            uint32_t i = task - 2 * n1 - 3 * n2 + 1;
            float t = (float)i / (float)kSilSamples;
            float x = src.x + flat.x * t;
            float y = src.y + flat.y * t;
            float line_z = src.z + (lis.z - src.z) * t;
            float th;
            uint32_t tri, flags;
            float z = -INFINITY;
            if (mesh_raycast<COUNT>(A.g, mk(x, y, 60.0f), mk(0.0f, 0.0f, -1.0f), th, tri, flags, tc)) {
                float top = 60.0f - th;
                if (top > line_z + 0.05f) z = top + 0.03f;
            }
            sil_z[i] = z;
        }
    }
    __syncthreads();

    // ---- 4. assemble the bent paths (paths.rs:190-232, :280-331)
    const uint32_t nb = n1 + n2 + 1;
    for (uint32_t r = tid; r < n1 + n2; r += bd) {
        bool ok;
        float g[3], len;
        uint32_t key;
        if (r < n1) {
            f3 p = mk(sp[r][0], sp[r][1], sp[r][2]);
            float d1 = length(sub(p, src)), d2 = length(sub(lis, p));
            ok = !(sl[r][0] * sl[r][3] < 1e-4f);
            float detour = d1 + d2 - direct_len;
            for (int b = 0; b < 3; b++)
                g[b] = knife_edge_amp(fmaxf(detour, 1e-4f), band_hz(b)) * sl[r][b] * sl[r][3 + b];
            len = d1 + d2;
            key = 1u + sel1[r];
        } else {
            uint32_t k = r - n1;
            f3 p1 = mk(pp[k][0], pp[k][1], pp[k][2]), p2 = mk(pp[k][3], pp[k][4], pp[k][5]);
            float d1 = length(sub(p1, src)), dm = length(sub(p2, p1)), d2 = length(sub(lis, p2));
            ok = !(pl[k][0] * pl[k][3] * pl[k][6] < 1e-4f);
            float det1 = d1 + dm - length(sub(p2, src));
            float det2 = dm + d2 - length(sub(lis, p1));
            for (int b = 0; b < 3; b++)
                g[b] = knife_edge_amp(fmaxf(det1, 1e-4f), band_hz(b)) *
                       knife_edge_amp(fmaxf(det2, 1e-4f), band_hz(b)) * pl[k][b] * pl[k][3 + b] *
                       pl[k][6 + b];
            len = d1 + dm + d2;
            key = 100000u + pe[k][0] * 4096u + pe[k][1];
        }
        for (int b = 0; b < 3; b++) bg[r][b] = g[b];
        bg[r][3] = len;
        bkey[r] = key;
        sk[r] = ok ? desc_key(g[1], r) : kNoKey;
    }
    if (tid == bd - 1) {
        // over_silhouette (paths.rs:337-405): upper hull of the drops
        uint32_t slot = n1 + n2;
        uint64_t key = kNoKey;
        s_hull = 0;
        if (n_sil) {
            float ps[kSilSamples + 1], pz[kSilSamples + 1];
            int np = 0;
            ps[np] = 0.0f;
            pz[np] = src.z;
            np++;
            for (int i = 1; i < kSilSamples; i++) {
                if (sil_z[i] != -INFINITY) {
                    float t = (float)i / (float)kSilSamples;
                    ps[np] = run * t;
                    pz[np] = sil_z[i];
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
    n_pad = pow2_at_least(nb);
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
            pts[1] = mk(sp[r][0], sp[r][1], sp[r][2]);
            pts[2] = lis;
            np = 3;
        } else if (r < n1 + n2) {
            uint32_t k = r - n1;
            pts[0] = src;
            pts[1] = mk(pp[k][0], pp[k][1], pp[k][2]);
            pts[2] = mk(pp[k][3], pp[k][4], pp[k][5]);
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
    if (COUNT) flush_counts(A, tc);
}

// Batched AutoPaths::transmission (paths.rs:84-113), one lane per segment.
__global__ __launch_bounds__(64) void k_path_transmission(PathTransArgs A)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    TravCount tc;
    float g[3];
    if (transmission<false>(A.g, A.mats, load3(A.a, i), load3(A.b, i), g, tc) < 0)
        atomicOr(&A.counters[1], 1u);
    for (int b = 0; b < 3; b++) A.gains[3 * i + b] = g[b];
}

template <bool COUNT>
hipError_t launch_paths_t(const PathsArgs& A, hipStream_t s)
{
    if (A.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_paths_direct<COUNT>, dim3((A.n + 63) / 64), dim3(64), 0, s, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the sort buffer holds the largest of: all edges, the prefiltered
    // candidates, the pairs, the bent paths
    uint32_t n_keys = A.ne > kPathPrefilter ? A.ne : kPathPrefilter;
    uint32_t pairs = A.pair_edges * A.pair_edges;
    if (pairs > n_keys) n_keys = pairs;
    uint32_t bent = A.edge_candidates + kPathPairKeep + 1;
    if (bent > n_keys) n_keys = bent;
    uint32_t n_pad = 2;
    while (n_pad < n_keys) n_pad <<= 1;
    hipLaunchKernelGGL(k_paths_bent<COUNT>, dim3(A.n), dim3(kPathBlock), n_pad * sizeof(uint64_t), s, A);
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
    hipLaunchKernelGGL(k_path_transmission, dim3((A.n + 63) / 64), dim3(64), 0, s, A);
    return hipGetLastError();
}

}  // namespace omg
