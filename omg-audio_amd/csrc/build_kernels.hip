// Device-side scene build (see build_kernels.h).
//
// The host build (scene_build.cpp, after crates/omg-core/src/mesh.rs:
// 241-405) is recursive and sequential.  Here the tessellation runs in
// breadth-first rounds over the triangles that split (each round splits
// every open patch once, so a large triangle spreads over many threads),
// and every level of the tree is a handful of launches over all
// primitives:
//   box     segment boxes and centroid boxes (order-preserving u32 keys,
//           atomics pre-reduced per wave when a wave lies in one segment)
//   seg     the level's nodes: parent links in breadth-first numbering,
//           split axis and centroid width
//   bin     16-bin SAH histograms of the splitting segments
//   split   one thread per segment: the host's cost sweep and cut
//   flag / scan / ptable / permute
//           the host's two-pointer partition (swap from the back) in
//           closed form: with nL primitives on the left, the c right-side
//           primitives in the first nL positions and the c left-side ones
//           behind them trade places pairwise (the k-th of each from the
//           front and back respectively), the k-th front right-side one
//           landing just before the (k-1)-th back left-side one, and every
//           other right-side primitive moving down by one; left-side
//           primitives in front stay where they are.
// Min/max are exact in any order, and the cost sweep runs in the host's
// order on one thread, so the tree is the host's bit for bit (boxes equal
// as values: a zero may carry the other sign).
#include "build_kernels.h"

#include <rocprim/device/device_scan.hpp>

#include "omg_numerics.h"

#pragma clang fp contract(off)

namespace omg {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t okey(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float ofloat(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

struct TriV {
    f3 a, b, c;
};

// One step of the host's LIFO tessellation (mesh.rs:251-279): false when
// the longest edge is within the limit, else the two halves of the
// longest-edge midpoint split in push order.
__device__ bool tess_split(const TriV& w, float lim, TriV& first, TriV& second)
{
    float lab = length(sub(w.b, w.a)), lbc = length(sub(w.c, w.b)), lca = length(sub(w.a, w.c));
    float longest = fmaxf(fmaxf(lab, lbc), lca);
    if (!(longest > lim)) return false;
    if (lab >= lbc && lab >= lca) {
        f3 m = mul(add(w.a, w.b), 0.5f);
        first = TriV{w.a, m, w.c};
        second = TriV{m, w.b, w.c};
    } else if (lbc >= lca) {
        f3 m = mul(add(w.b, w.c), 0.5f);
        first = TriV{w.a, w.b, m};
        second = TriV{w.a, m, w.c};
    } else {
        f3 m = mul(add(w.c, w.a), 0.5f);
        first = TriV{w.a, w.b, m};
        second = TriV{m, w.b, w.c};
    }
    return true;
}

__device__ __forceinline__ bool finite3(f3 v)
{
    return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}

__device__ __forceinline__ f3 vtx(const float* pos, uint32_t i)
{
    return mk(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]);
}

__device__ __forceinline__ float tri_limit(const float* tri_max_edge, uint32_t t)
{
    return tri_max_edge ? tri_max_edge[t] : 4.0f;
}

// Triangle checks, normals, and whether each triangle splits at all; one
// that does not is its own single patch.  Slots are in reversed triangle
// order: the host pops the last triangle first.
__global__ __launch_bounds__(kBlock) void k_tess_count(TessArgs A)
{
    uint32_t t = gid();
    if (t == 0) {
        A.cnt_rev[A.nt] = 0;
        A.big[A.nt] = 0;
    }
    if (t >= A.nt) return;
    uint32_t slot = A.nt - 1 - t;
    uint32_t i0 = A.idx[3 * (size_t)t], i1 = A.idx[3 * (size_t)t + 1], i2 = A.idx[3 * (size_t)t + 2];
    uint32_t m = A.tri_mat[t];
    uint32_t err = 0, cnt = 0, big = 0;
    if (i0 >= A.nv || i1 >= A.nv || i2 >= A.nv) err |= kBuildErrIndex;
    if (m >= A.nm) err |= kBuildErrMaterial;
    if (!(err & kBuildErrIndex)) {
        f3 a = vtx(A.pos, i0), b = vtx(A.pos, i1), c = vtx(A.pos, i2);
        if (!finite3(a) || !finite3(b) || !finite3(c)) {
            err |= kBuildErrFinite;
        } else {
            f3 n = normalize(cross(sub(b, a), sub(c, a)));
            A.tris[t] = TriInfo{{n.x, n.y, n.z}, m};
            TriV f, s;
            big = tess_split(TriV{a, b, c}, tri_limit(A.tri_max_edge, t), f, s) ? 1u : 0u;
            cnt = 1u - big;
        }
    }
    if (A.tri_surf32) A.tri_surf32[t] = A.tri_surf ? (uint32_t)A.tri_surf[t] : 0u;
    if (err) {
        atomicOr(&A.meta->err, err);
        cnt = big = 0;
    }
    A.cnt_rev[slot] = cnt;
    A.big[slot] = big;
}

// The work list: every splitting triangle in slot order.
__global__ void k_tess_seed(TessArgs A, TessItem* items)
{
    uint32_t t = gid();
    if (t >= A.nt) return;
    uint32_t slot = A.nt - 1 - t;
    if (!A.big[slot]) return;
    f3 a = vtx(A.pos, A.idx[3 * (size_t)t]), b = vtx(A.pos, A.idx[3 * (size_t)t + 1]),
       c = vtx(A.pos, A.idx[3 * (size_t)t + 2]);
    TessItem it;
    it.a[0] = a.x; it.a[1] = a.y; it.a[2] = a.z;
    it.b[0] = b.x; it.b[1] = b.y; it.b[2] = b.z;
    it.c[0] = c.x; it.c[1] = c.y; it.c[2] = c.z;
    it.slot = slot;
    it.done = 0;
    it.pad = 0;
    items[A.big_off[slot]] = it;
}

__device__ __forceinline__ TriV item_tri(const TessItem& it)
{
    return TriV{mk(it.a[0], it.a[1], it.a[2]), mk(it.b[0], it.b[1], it.b[2]),
                mk(it.c[0], it.c[1], it.c[2])};
}

__device__ __forceinline__ TessItem tri_item(const TriV& w, uint32_t slot, uint32_t done)
{
    TessItem it;
    it.a[0] = w.a.x; it.a[1] = w.a.y; it.a[2] = w.a.z;
    it.b[0] = w.b.x; it.b[1] = w.b.y; it.b[2] = w.b.z;
    it.c[0] = w.c.x; it.c[1] = w.c.y; it.c[2] = w.c.z;
    it.slot = slot;
    it.done = done;
    it.pad = 0;
    return it;
}

// Expanding every open item by one split keeps the list in the host's pop
// order: items stay grouped by slot, and within a triangle the halves are
// placed second first, which is the order the LIFO list pops them.
__global__ __launch_bounds__(kBlock) void k_tess_round_count(TessRound R)
{
    uint32_t i = gid();
    if (i == 0) R.oc[R.m] = 0;
    if (i >= R.m) return;
    TessItem it = R.in[i];
    uint32_t oc = 1;
    if (!it.done) {
        TriV f, s;
        if (tess_split(item_tri(it), tri_limit(R.tri_max_edge, R.nt - 1 - it.slot), f, s)) oc = 2;
    }
    R.oc[i] = oc;
}

__global__ __launch_bounds__(kBlock) void k_tess_round_write(TessRound R)
{
    uint32_t i = gid();
    if (i >= R.m) return;
    TessItem it = R.in[i];
    uint32_t at = R.pos[i];
    if (it.done) {
        R.out[at] = it;
        return;
    }
    TriV f, s;
    if (tess_split(item_tri(it), tri_limit(R.tri_max_edge, R.nt - 1 - it.slot), f, s)) {
        R.out[at] = tri_item(s, it.slot, 0);
        R.out[at + 1] = tri_item(f, it.slot, 0);
    } else {
        it.done = 1;
        R.out[at] = it;
    }
}

// Patch counts and first list entry of the splitting triangles.  A
// triangle's items are contiguous, so each wave adds the length of every
// run of equal slots once, from the run's last lane.
__global__ __launch_bounds__(kBlock) void k_tess_leaves(TessArgs A, const TessItem* items, uint32_t m)
{
    const uint32_t i = gid();
    const uint32_t lane = __lane_id();
    const bool act = i < m;
    const uint32_t slot = act ? items[i].slot : 0xFFFFFFFFu;
    const uint32_t prev = __shfl_up(slot, 1);
    const uint32_t next = __shfl_down(slot, 1);
    const bool head = lane == 0 || prev != slot;
    const bool tail = lane == 63 || next != slot;
    const uint64_t heads = __ballot(head);
    if (!act) return;
    if (i == 0 || items[i - 1].slot != slot) A.first_of[slot] = i;
    if (tail) {
        uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
        uint32_t start = 63u - (uint32_t)__clzll((long long)upto);
        atomicAdd(&A.cnt_rev[slot], lane - start + 1);
    }
}

__global__ void k_tess_check(const uint32_t* cnt, const uint32_t* expect, uint32_t nt, BuildMeta* meta)
{
    uint32_t i = gid();
    if (i < nt && cnt[i] != expect[i]) atomicOr(&meta->err, kBuildErrRefit);
}

// A packed primitive (mesh.rs:272-278) and its build record; the box and
// centroid follow the host builder's expressions from a, e1, e2.
__device__ void emit_patch(const TessArgs& A, uint32_t t, const TriV& w, uint32_t at)
{
    f3 e1 = sub(w.b, w.a), e2 = sub(w.c, w.a);
    Prim pr;
    pr.a[0] = w.a.x; pr.a[1] = w.a.y; pr.a[2] = w.a.z;
    pr.tri = t;
    pr.e1[0] = e1.x; pr.e1[1] = e1.y; pr.e1[2] = e1.z;
    pr.mat = A.tri_mat[t];
    pr.e2[0] = e2.x; pr.e2[1] = e2.y; pr.e2[2] = e2.z;
    pr.surf = A.tri_surf ? (uint32_t)A.tri_surf[t] : 0u;
    A.prims[A.leaf_of ? A.leaf_of[at] : at] = pr;
    if (A.recs) {
        f3 pb = add(w.a, e1), pc = add(w.a, e2);
        f3 lo = vmin(vmin(w.a, pb), pc), hi = vmax(vmax(w.a, pb), pc);
        f3 cen = add(w.a, mul(add(e1, e2), 1.0f / 3.0f));
        BRec r;
        r.lo[0] = lo.x; r.lo[1] = lo.y; r.lo[2] = lo.z;
        r.cx = cen.x;
        r.hi[0] = hi.x; r.hi[1] = hi.y; r.hi[2] = hi.z;
        r.cy = cen.y;
        r.cz = cen.z;
        r.idx = at;
        r.seg = 0;
        r.bin = 0;
        A.recs[at] = r;
    }
}

// Threads [0, nt) emit the triangles that do not split, threads [nt, nt+m)
// the final items of those that do.
__global__ __launch_bounds__(kBlock) void k_tess_emit(TessArgs A, const TessItem* items, uint32_t m)
{
    uint32_t g = gid();
    if (g < A.nt) {
        uint32_t t = g, slot = A.nt - 1 - t;
        if (A.big[slot]) return;
        f3 a = vtx(A.pos, A.idx[3 * (size_t)t]), b = vtx(A.pos, A.idx[3 * (size_t)t + 1]),
           c = vtx(A.pos, A.idx[3 * (size_t)t + 2]);
        emit_patch(A, t, TriV{a, b, c}, A.off_rev[slot]);
        return;
    }
    uint32_t i = g - A.nt;
    if (i >= m) return;
    const TessItem& it = items[i];
    uint32_t slot = it.slot;
    emit_patch(A, A.nt - 1 - slot, item_tri(it), A.off_rev[slot] + (i - A.first_of[slot]));
}

__global__ __launch_bounds__(kBlock) void k_box_init(LevelArgs A)
{
    uint32_t s = gid();
    if (s >= A.n_seg) return;
    uint32_t* k = A.seg_box + 12 * (size_t)s;
#pragma unroll
    for (int j = 0; j < 12; j++) k[j] = (j < 3 || (j >= 6 && j < 9)) ? 0xFFFFFFFFu : 0u;
}

// word j of a segment box: 0..2 box min, 3..5 box max, 6..8 centroid min,
// 9..11 centroid max
__device__ __forceinline__ bool is_min_word(int j) { return j < 3 || (j >= 6 && j < 9); }

// Each wave walks kChunksPerWave consecutive 64-primitive chunks and keeps
// the reduction of a run of chunks that lie in one segment in registers, so
// the global atomics per level fall from one set per wave of primitives to
// one per run (the top levels are single large segments).
constexpr uint32_t kChunksPerWave = 16;
// segments up to this size are built by one thread each (k_small)
constexpr uint32_t kSmallSeg = 64;

inline dim3 grid_chunks(uint32_t n)
{
    size_t chunks = ((size_t)n + 63) / 64;
    size_t waves = (chunks + kChunksPerWave - 1) / kChunksPerWave;
    return grid_for(waves * 64);
}

__device__ __forceinline__ void box_atomics(uint32_t* seg_box, uint32_t seg, const uint32_t* k)
{
    uint32_t* dst = seg_box + 12 * (size_t)seg;
#pragma unroll
    for (int j = 0; j < 12; j++) {
        if (is_min_word(j)) atomicMin(dst + j, k[j]);
        else atomicMax(dst + j, k[j]);
    }
}

__global__ __launch_bounds__(kBlock) void k_box(LevelArgs A)
{
    const uint32_t lane = __lane_id();
    const uint32_t wave = gid() >> 6;
    const uint32_t n_chunks = (A.n + 63) / 64;
    const uint32_t c0 = wave * kChunksPerWave;
    const uint32_t c1 = min(c0 + kChunksPerWave, n_chunks);
    uint32_t acc[12];
#pragma unroll
    for (int j = 0; j < 12; j++) acc[j] = is_min_word(j) ? 0xFFFFFFFFu : 0u;
    uint32_t acc_seg = kSegInactive;
    for (uint32_t c = c0; c < c1; c++) {
        uint32_t p = c * 64 + lane;
        bool act = p < A.n;
        BRec r{};
        if (act) {
            r = A.cur[p];
            act = r.seg != kSegInactive && A.seg_count[r.seg] > kSmallSeg;
        }
        uint64_t rem = __ballot(act);
        // the (few) large segments of this chunk, in order
        while (rem) {
            uint32_t l = (uint32_t)__ffsll((unsigned long long)rem) - 1;
            uint32_t s0 = __shfl(r.seg, (int)l);
            bool mine = act && r.seg == s0;
            uint64_t mm = __ballot(mine);
            uint32_t k[12];
            k[0] = okey(r.lo[0]); k[1] = okey(r.lo[1]); k[2] = okey(r.lo[2]);
            k[3] = okey(r.hi[0]); k[4] = okey(r.hi[1]); k[5] = okey(r.hi[2]);
            k[6] = okey(r.cx); k[7] = okey(r.cy); k[8] = okey(r.cz);
            k[9] = k[6]; k[10] = k[7]; k[11] = k[8];
            if (!mine) {
#pragma unroll
                for (int j = 0; j < 12; j++) k[j] = is_min_word(j) ? 0xFFFFFFFFu : 0u;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
                for (int j = 0; j < 12; j++) {
                    uint32_t o = __shfl_xor(k[j], off);
                    k[j] = is_min_word(j) ? min(k[j], o) : max(k[j], o);
                }
            }
            if (s0 != acc_seg) {
                if (acc_seg != kSegInactive && lane == 0) box_atomics(A.seg_box, acc_seg, acc);
#pragma unroll
                for (int j = 0; j < 12; j++) acc[j] = k[j];
                acc_seg = s0;
            } else {
#pragma unroll
                for (int j = 0; j < 12; j++) acc[j] = is_min_word(j) ? min(acc[j], k[j]) : max(acc[j], k[j]);
            }
            rem &= ~mm;
        }
    }
    if (acc_seg != kSegInactive && lane == 0) box_atomics(A.seg_box, acc_seg, acc);
}

__device__ __forceinline__ float sah_area(f3 lo, f3 hi)
{
    f3 e = vmax(sub(hi, lo), mk(0.0f, 0.0f, 0.0f));
    return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
}

__device__ __forceinline__ uint32_t bin_of(const BRec& r, uint32_t axis, float cmin_a, float cwidth)
{
    float cv = axis == 0 ? r.cx : (axis == 1 ? r.cy : r.cz);
    float f = (cv - cmin_a) / cwidth * 16.0f;
    return sat_u32(fminf(f, 15.0f));
}

// The axis rule of the host builder: the widest centroid extent, x before
// y before z on ties.
__device__ __forceinline__ uint32_t split_axis(f3 cext)
{
    return (cext.x >= cext.y && cext.x >= cext.z) ? 0u : (cext.y >= cext.z ? 1u : 2u);
}

// Segments of at most kSmallSeg primitives: one thread does what the host
// builder does for the node -- box and centroid box, the 16 bins (in LDS,
// with the host's initial values), the cost sweep and the two-pointer
// partition in place -- and records the split as already placed.
constexpr int kSmallBlock = 64;

__global__ __launch_bounds__(kSmallBlock) void k_small(LevelArgs A)
{
    __shared__ float s_min[16][3][kSmallBlock];
    __shared__ float s_max[16][3][kSmallBlock];
    __shared__ uint32_t s_n[16][kSmallBlock];
    const uint32_t s = gid();
    const uint32_t tl = threadIdx.x;
    if (s >= A.n_seg) return;
    const uint32_t count = A.seg_count[s];
    if (count > kSmallSeg) return;
    const uint32_t start = A.seg_start[s], end = start + count;
    BRec* R = A.cur;
    f3 bmin = mk(FLT_MAX, FLT_MAX, FLT_MAX), bmax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    f3 cmin = bmin, cmax = bmax;
    for (uint32_t p = start; p < end; p++) {
        const BRec r = R[p];
        bmin = vmin(bmin, mk(r.lo[0], r.lo[1], r.lo[2]));
        bmax = vmax(bmax, mk(r.hi[0], r.hi[1], r.hi[2]));
        f3 c = mk(r.cx, r.cy, r.cz);
        cmin = vmin(cmin, c);
        cmax = vmax(cmax, c);
    }
    uint32_t* k = A.seg_box + 12 * (size_t)s;
    k[0] = okey(bmin.x); k[1] = okey(bmin.y); k[2] = okey(bmin.z);
    k[3] = okey(bmax.x); k[4] = okey(bmax.y); k[5] = okey(bmax.z);
    k[6] = okey(cmin.x); k[7] = okey(cmin.y); k[8] = okey(cmin.z);
    k[9] = okey(cmax.x); k[10] = okey(cmax.y); k[11] = okey(cmax.z);
    if (count <= 4) return;
    const uint32_t i = A.seg_irank[s];
    f3 cext = sub(cmax, cmin);
    uint32_t axis = split_axis(cext);
    float cwidth = get(cext, (int)axis);
    float cmin_a = get(cmin, (int)axis);
    uint32_t mid = start + count / 2;
    if (!(cwidth < 1e-6f)) {
        for (int b = 0; b < 16; b++) {
            for (int d = 0; d < 3; d++) {
                s_min[b][d][tl] = FLT_MAX;
                s_max[b][d][tl] = -FLT_MAX;
            }
            s_n[b][tl] = 0;
        }
        for (uint32_t p = start; p < end; p++) {
            const BRec r = R[p];
            uint32_t b = bin_of(r, axis, cmin_a, cwidth);
            for (int d = 0; d < 3; d++) {
                s_min[b][d][tl] = fminf(s_min[b][d][tl], r.lo[d]);
                s_max[b][d][tl] = fmaxf(s_max[b][d][tl], r.hi[d]);
            }
            s_n[b][tl] += 1;
        }
        float right_cost[16];
#pragma unroll
        for (int b = 0; b < 16; b++) right_cost[b] = 0.0f;
        f3 amin = mk(FLT_MAX, FLT_MAX, FLT_MAX), amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        uint32_t an = 0;
#pragma unroll
        for (int b = 15; b >= 1; b--) {
            amin = vmin(amin, mk(s_min[b][0][tl], s_min[b][1][tl], s_min[b][2][tl]));
            amax = vmax(amax, mk(s_max[b][0][tl], s_max[b][1][tl], s_max[b][2][tl]));
            an += s_n[b][tl];
            right_cost[b - 1] = an > 0 ? sah_area(amin, amax) * (float)an : 0.0f;
        }
        int best_cut = -1;
        float best_cost = FLT_MAX;
        amin = mk(FLT_MAX, FLT_MAX, FLT_MAX);
        amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        an = 0;
#pragma unroll
        for (int b = 0; b < 15; b++) {
            amin = vmin(amin, mk(s_min[b][0][tl], s_min[b][1][tl], s_min[b][2][tl]));
            amax = vmax(amax, mk(s_max[b][0][tl], s_max[b][1][tl], s_max[b][2][tl]));
            an += s_n[b][tl];
            if (an == 0 || an == count) continue;
            float cost = sah_area(amin, amax) * (float)an + right_cost[b];
            if (cost < best_cost) {
                best_cost = cost;
                best_cut = b;
            }
        }
        if (best_cut >= 0) {
            // two-pointer partition, swapping from the back (mesh.rs:384-395)
            uint32_t a = start, e = end;
            while (a < e) {
                if (bin_of(R[a], axis, cmin_a, cwidth) <= (uint32_t)best_cut) {
                    a++;
                } else {
                    e--;
                    BRec t = R[a];
                    R[a] = R[e];
                    R[e] = t;
                }
            }
            mid = a;
        }
    }
    mid = min(max(mid, start + 1), end - 1);
    ISplit q;
    q.start = start;
    q.count = count;
    q.axis = axis;
    q.mode = 1;
    q.cmin_a = cmin_a;
    q.cwidth = cwidth;
    q.cut = kNoCut;   // already in place: children by position
    q.mid = mid;
    A.split[i] = q;
    A.nseg_start[2 * (size_t)i] = start;
    A.nseg_count[2 * (size_t)i] = mid - start;
    A.nseg_start[2 * (size_t)i + 1] = mid;
    A.nseg_count[2 * (size_t)i + 1] = end - mid;
    A.nseg_int[2 * (size_t)i] = (mid - start) > 4 ? 1u : 0u;
    A.nseg_int[2 * (size_t)i + 1] = (end - mid) > 4 ? 1u : 0u;
}

// The level's nodes: each segment's box goes into its parent's paired node
// (breadth-first numbering: level base + rank among the level's splitting
// segments), and each splitting segment gets its axis and bins.
__global__ __launch_bounds__(kBlock) void k_seg(LevelArgs A)
{
    uint32_t s = gid();
    if (s >= A.n_seg) return;
    uint32_t start = A.seg_start[s], count = A.seg_count[s];
    const uint32_t* k = A.seg_box + 12 * (size_t)s;
    f3 lo = mk(ofloat(k[0]), ofloat(k[1]), ofloat(k[2]));
    f3 hi = mk(ofloat(k[3]), ofloat(k[4]), ofloat(k[5]));
    f3 cmin = mk(ofloat(k[6]), ofloat(k[7]), ofloat(k[8]));
    f3 cmax = mk(ofloat(k[9]), ofloat(k[10]), ofloat(k[11]));
    uint32_t ref = count > 4 ? A.base + A.seg_irank[s] : leaf_ref(start, count);
    if (A.level == 0) {
        A.meta->root_ref = ref;
        A.meta->root_min[0] = lo.x; A.meta->root_min[1] = lo.y; A.meta->root_min[2] = lo.z;
        A.meta->root_max[0] = hi.x; A.meta->root_max[1] = hi.y; A.meta->root_max[2] = hi.z;
    } else {
        Node2* o = A.nodes + A.base_prev + (s >> 1);
        if ((s & 1u) == 0) {
            o->lmin[0] = lo.x; o->lmin[1] = lo.y; o->lmin[2] = lo.z;
            o->lref = ref;
            o->lmax[0] = hi.x; o->lmax[1] = hi.y; o->lmax[2] = hi.z;
            o->pad0 = 0;
            o->pad1 = 0;
        } else {
            o->rmin[0] = lo.x; o->rmin[1] = lo.y; o->rmin[2] = lo.z;
            o->rref = ref;
            o->rmax[0] = hi.x; o->rmax[1] = hi.y; o->rmax[2] = hi.z;
        }
    }
    if (count <= kSmallSeg) return;
    uint32_t i = A.seg_irank[s];
    f3 cext = sub(cmax, cmin);
    uint32_t axis = split_axis(cext);
    float cwidth = get(cext, (int)axis);
    ISplit q;
    q.start = start;
    q.count = count;
    q.axis = axis;
    q.mode = cwidth < 1e-6f ? 1u : 0u;
    q.cmin_a = get(cmin, (int)axis);
    q.cwidth = cwidth;
    q.cut = kNoCut;
    q.mid = 0;
    A.split[i] = q;
    if (q.mode == 0) {
        uint32_t* b = A.bins + 128 * (size_t)i;
        for (int j = 0; j < 16; j++) {
            b[8 * j + 0] = b[8 * j + 1] = b[8 * j + 2] = 0xFFFFFFFFu;
            b[8 * j + 3] = b[8 * j + 4] = b[8 * j + 5] = 0u;
            b[8 * j + 6] = 0u;
            b[8 * j + 7] = 0u;
        }
    }
}

// field j of a bin: 0..2 box min keys, 3..5 box max keys, 6 count
__device__ __forceinline__ void bin_atomic(uint32_t* dst, uint32_t j, uint32_t v)
{
    if (j < 3) atomicMin(dst, v);
    else if (j < 6) atomicMax(dst, v);
    else atomicAdd(dst, v);
}

// The wave's 16-bin table of the current run of chunks lives in LDS, in
// the global layout (8 words per bin); the lanes fold their primitives in
// with LDS atomics, and the table goes to the segment's global bins with
// one atomic per touched word when the run ends.
__device__ __forceinline__ uint32_t tab_identity(uint32_t w) { return (w & 7u) < 3 ? 0xFFFFFFFFu : 0u; }

__device__ __forceinline__ void tab_reset(uint32_t* tab, uint32_t lane)
{
    tab[lane] = tab_identity(lane);
    tab[lane + 64] = tab_identity(lane + 64);
}

__device__ __forceinline__ void tab_flush(const uint32_t* tab, uint32_t* bins, uint32_t i, uint32_t lane)
{
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        uint32_t w = lane + 64 * h, j = w & 7u;
        uint32_t v = tab[w];
        if (j < 7 && v != tab_identity(w)) bin_atomic(bins + 128 * (size_t)i + w, j, v);
    }
}

__global__ __launch_bounds__(kBlock) void k_bin(LevelArgs A)
{
    __shared__ uint32_t s_tab[kBlock / 64][128];
    const uint32_t lane = __lane_id();
    uint32_t* tab = s_tab[threadIdx.x >> 6];
    const uint32_t wave = gid() >> 6;
    const uint32_t n_chunks = (A.n + 63) / 64;
    const uint32_t c0 = wave * kChunksPerWave;
    const uint32_t c1 = min(c0 + kChunksPerWave, n_chunks);
    tab_reset(tab, lane);
    uint32_t acc_i = kSegInactive;
    for (uint32_t c = c0; c < c1; c++) {
        uint32_t p = c * 64 + lane;
        bool act = p < A.n;
        BRec r{};
        uint32_t i = 0, k = 0;
        if (act) {
            r = A.cur[p];
            act = r.seg != kSegInactive && A.seg_count[r.seg] > kSmallSeg;
            if (act) {
                i = A.seg_irank[r.seg];
                ISplit q = A.split[i];
                act = q.mode == 0;
                if (act) {
                    k = bin_of(r, q.axis, q.cmin_a, q.cwidth);
                    A.cur[p].bin = k;
                }
            }
        }
        uint64_t segs = __ballot(act);
        // the (few) large segments of this chunk, in order
        while (segs) {
            uint32_t ls = (uint32_t)__ffsll((unsigned long long)segs) - 1;
            uint32_t i0 = __shfl(i, (int)ls);
            bool in_seg = act && i == i0;
            segs &= ~__ballot(in_seg);
            if (i0 != acc_i) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if (acc_i != kSegInactive) tab_flush(tab, A.bins, acc_i, lane);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                tab_reset(tab, lane);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                acc_i = i0;
            }
            if (in_seg) {
                uint32_t* t = tab + 8 * k;
                atomicMin(t + 0, okey(r.lo[0]));
                atomicMin(t + 1, okey(r.lo[1]));
                atomicMin(t + 2, okey(r.lo[2]));
                atomicMax(t + 3, okey(r.hi[0]));
                atomicMax(t + 4, okey(r.hi[1]));
                atomicMax(t + 5, okey(r.hi[2]));
                atomicAdd(t + 6, 1u);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (acc_i != kSegInactive) tab_flush(tab, A.bins, acc_i, lane);
}

// The host's SAH sweep over the 16 bins (mesh.rs:353-380) and the split
// point, clamped to leave both sides non-empty.
__global__ __launch_bounds__(kBlock) void k_split(LevelArgs A)
{
    uint32_t i = gid();
    if (i == 0) A.nseg_int[2 * (size_t)A.n_int] = 0;
    if (i >= A.n_int) return;
    ISplit q = A.split[i];
    if (q.count <= kSmallSeg) return;
    uint32_t start = q.start, count = q.count, end = start + count;
    uint32_t mid = start + count / 2;
    uint32_t cut = kNoCut;
    if (q.mode == 0) {
        const uint32_t* b = A.bins + 128 * (size_t)i;
        f3 bnmin[16], bnmax[16];
        uint32_t bnn[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            bnn[k] = b[8 * k + 6];
            if (bnn[k] == 0) {
                bnmin[k] = mk(FLT_MAX, FLT_MAX, FLT_MAX);
                bnmax[k] = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            } else {
                bnmin[k] = mk(ofloat(b[8 * k]), ofloat(b[8 * k + 1]), ofloat(b[8 * k + 2]));
                bnmax[k] = mk(ofloat(b[8 * k + 3]), ofloat(b[8 * k + 4]), ofloat(b[8 * k + 5]));
            }
        }
        float right_cost[16];
#pragma unroll
        for (int k = 0; k < 16; k++) right_cost[k] = 0.0f;
        f3 amin = mk(FLT_MAX, FLT_MAX, FLT_MAX), amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        uint32_t an = 0;
#pragma unroll
        for (int k = 15; k >= 1; k--) {
            amin = vmin(amin, bnmin[k]);
            amax = vmax(amax, bnmax[k]);
            an += bnn[k];
            right_cost[k - 1] = an > 0 ? sah_area(amin, amax) * (float)an : 0.0f;
        }
        int best_cut = -1;
        float best_cost = FLT_MAX;
        uint32_t best_n = 0;
        amin = mk(FLT_MAX, FLT_MAX, FLT_MAX);
        amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        an = 0;
#pragma unroll
        for (int k = 0; k < 15; k++) {
            amin = vmin(amin, bnmin[k]);
            amax = vmax(amax, bnmax[k]);
            an += bnn[k];
            if (an == 0 || an == count) continue;
            float cost = sah_area(amin, amax) * (float)an + right_cost[k];
            if (cost < best_cost) {
                best_cost = cost;
                best_cut = k;
                best_n = an;
            }
        }
        if (best_cut >= 0) {
            mid = start + best_n;
            cut = (uint32_t)best_cut;
        }
    }
    mid = min(max(mid, start + 1), end - 1);
    A.split[i].cut = cut;
    A.split[i].mid = mid;
    A.nseg_start[2 * (size_t)i] = start;
    A.nseg_count[2 * (size_t)i] = mid - start;
    A.nseg_start[2 * (size_t)i + 1] = mid;
    A.nseg_count[2 * (size_t)i + 1] = end - mid;
    A.nseg_int[2 * (size_t)i] = (mid - start) > 4 ? 1u : 0u;
    A.nseg_int[2 * (size_t)i + 1] = (end - mid) > 4 ? 1u : 0u;
}

// the split of a primitive's segment when it is a SAH cut, else null
__device__ __forceinline__ bool sah_split(const LevelArgs& A, const BRec& r, ISplit& q, uint32_t& i)
{
    if (r.seg == kSegInactive || A.seg_count[r.seg] <= 4) return false;
    i = A.seg_irank[r.seg];
    q = A.split[i];
    return q.cut != kNoCut;
}

__global__ __launch_bounds__(kBlock) void k_flag(LevelArgs A)
{
    uint32_t p = gid();
    if (p == 0) A.flags[A.n] = 0;
    if (p >= A.n) return;
    BRec r = A.cur[p];
    ISplit q;
    uint32_t i;
    uint32_t f = 0;
    if (sah_split(A, r, q, i)) f = r.bin <= q.cut ? 1u : 0u;
    A.flags[p] = f;
}

__global__ __launch_bounds__(kBlock) void k_ptable(LevelArgs A)
{
    uint32_t p = gid();
    if (p >= A.n) return;
    BRec r = A.cur[p];
    ISplit q;
    uint32_t i;
    if (!sah_split(A, r, q, i)) return;
    uint32_t s0 = q.start, nl = q.mid - s0;
    uint32_t lb = A.lscan[p] - A.lscan[s0];
    if (A.flags[p]) {
        if (p >= s0 + nl) A.bL[s0 + (nl - 1 - lb)] = p;
    } else if (p < s0 + nl) {
        A.aR[s0 + ((p - s0) - lb)] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_permute(LevelArgs A)
{
    uint32_t p = gid();
    if (p >= A.n) return;
    BRec r = A.cur[p];
    BRec* dst = A.out;
    if (r.seg == kSegInactive) {
        dst[p] = r;
        return;
    }
    if (A.seg_count[r.seg] <= 4) {
        r.seg = kSegInactive;
        dst[p] = r;
        return;
    }
    uint32_t i = A.seg_irank[r.seg];
    ISplit q = A.split[i];
    uint32_t d = p;
    if (q.cut != kNoCut) {
        uint32_t s0 = q.start, e = s0 + q.count, nl = q.mid - s0;
        uint32_t lb = A.lscan[p] - A.lscan[s0];
        if (A.flags[p]) {
            d = p < s0 + nl ? p : A.aR[s0 + (nl - 1 - lb)];
        } else {
            uint32_t rb = (p - s0) - lb;
            uint32_t c = nl - (A.lscan[s0 + nl] - A.lscan[s0]);
            uint32_t bprev = rb == 0 ? e : A.bL[s0 + rb - 1];
            if (rb < c) {
                d = bprev - 1;
            } else {
                uint32_t bc = c == 0 ? e : A.bL[s0 + c - 1];
                d = (rb == c && p < bc) ? bc - 1 : p - 1;
            }
        }
    }
    r.seg = 2 * i + (d >= q.mid ? 1u : 0u);
    dst[d] = r;
}

__global__ void k_gather(const Prim* tess, const BRec* recs, uint32_t n, Prim* out, uint32_t* leaf_of)
{
    uint32_t p = gid();
    if (p >= n) return;
    uint32_t t = recs[p].idx;
    out[p] = tess[t];
    if (leaf_of) leaf_of[t] = p;
}

struct Kid {
    float mn[3], mx[3];
    uint32_t ref, src;
};

__device__ __forceinline__ Kid kid_of(const Node2& n, uint32_t node, uint32_t side)
{
    Kid k;
    if (side == 0) {
        k.mn[0] = n.lmin[0]; k.mn[1] = n.lmin[1]; k.mn[2] = n.lmin[2];
        k.mx[0] = n.lmax[0]; k.mx[1] = n.lmax[1]; k.mx[2] = n.lmax[2];
        k.ref = n.lref;
    } else {
        k.mn[0] = n.rmin[0]; k.mn[1] = n.rmin[1]; k.mn[2] = n.rmin[2];
        k.mx[0] = n.rmax[0]; k.mx[1] = n.rmax[1]; k.mx[2] = n.rmax[2];
        k.ref = n.rref;
    }
    k.src = node << 1 | side;
    return k;
}

// One 4-wide node from its paired node, two binary levels (the host
// collapse): slots 0-1 hold the left child's two children, or the leaf
// itself in slot 0 and slot 1 unused; slots 2-3 the right child likewise.
// Internal children keep their paired-node reference until k_collapse_link.
__global__ void k_collapse_expand(Collapse4Args A)
{
    uint32_t j = gid();
    if (j == A.q) A.n_int[A.q] = 0;
    if (j >= A.q) return;
    uint32_t nref = A.queue[j];
    const Node2 n = A.nodes[nref];
    Kid k[4];
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        const Kid half = kid_of(n, nref, h);
        if (half.ref & kLeafBit) {
            k[2 * h] = half;
            k[2 * h + 1] = half;
            k[2 * h + 1].ref = kNoChild;
        } else {
            const Node2 m = A.nodes[half.ref];
            k[2 * h] = kid_of(m, half.ref, 0);
            k[2 * h + 1] = kid_of(m, half.ref, 1);
        }
    }
    Node4 out;
    uint32_t ni = 0;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        Child4& c = out.c[t];
        c.pad = 0;
        if (k[t].ref == kNoChild) {
            c.mn[0] = c.mn[1] = c.mn[2] = 0.0f;
            c.mx[0] = c.mx[1] = c.mx[2] = 0.0f;
            c.ref = kNoChild;
            A.src4[4 * (size_t)(A.base4 + j) + t] = kNoChild;
            continue;
        }
        c.mn[0] = k[t].mn[0]; c.mn[1] = k[t].mn[1]; c.mn[2] = k[t].mn[2];
        c.mx[0] = k[t].mx[0]; c.mx[1] = k[t].mx[1]; c.mx[2] = k[t].mx[2];
        c.ref = k[t].ref;
        A.src4[4 * (size_t)(A.base4 + j) + t] = k[t].src;
        if (!(k[t].ref & kLeafBit)) ni++;
    }
    A.nodes4[A.base4 + j] = out;
    A.n_int[j] = ni;
}

// Breadth-first numbering of the next level's 4-wide nodes and the
// worst-case stack: the largest sum of (children - 1) along a path.
__global__ void k_collapse_link(Collapse4Args A)
{
    uint32_t j = gid();
    if (j >= A.q) return;
    Node4* o = A.nodes4 + A.base4 + j;
    uint32_t off = A.int_scan[j];
    uint32_t nk = 0;
    for (int t = 0; t < 4; t++)
        if (o->c[t].ref != kNoChild) nk++;
    uint32_t here = A.path[j] + nk - 1;
    for (int t = 0; t < 4; t++) {
        uint32_t ref = o->c[t].ref;
        if (ref == kNoChild || (ref & kLeafBit)) continue;
        A.queue_next[off] = ref;
        A.path_next[off] = here;
        o->c[t].ref = A.base4_next + off;
        off++;
    }
    atomicMax(&A.meta->stack4, here);
}

// One 16-wide node from its 4-wide node: slot group 4k..4k+3 is the
// 4-wide node's slot k expanded once more (the four slots of the 4-wide
// node it refers to; a leaf in slot 4k; or four unused slots).  Internal
// children keep their 4-wide reference until k_derive16_link.
__global__ void k_derive16_expand(Derive16Args A)
{
    uint32_t j = gid();
    if (j == A.q) A.n_int[A.q] = 0;
    if (j >= A.q) return;
    const Node4 a = A.nodes4[A.queue[j]];
    Node16 o;
    uint32_t ni = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const Child4 c = a.c[k];
        Child4 none;
        none.mn[0] = none.mn[1] = none.mn[2] = 0.0f;
        none.mx[0] = none.mx[1] = none.mx[2] = 0.0f;
        none.ref = kNoChild;
        none.pad = 0;
        if (c.ref == kNoChild || (c.ref & kLeafBit)) {
            o.c[4 * k] = c;
            o.c[4 * k].pad = 0;
            o.c[4 * k + 1] = none;
            o.c[4 * k + 2] = none;
            o.c[4 * k + 3] = none;
        } else {
            const Node4 m = A.nodes4[c.ref];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                o.c[4 * k + i] = m.c[i];
                o.c[4 * k + i].pad = 0;
                const uint32_t r = m.c[i].ref;
                if (r != kNoChild && !(r & kLeafBit)) ni++;
            }
        }
    }
    A.nodes16[A.base16 + j] = o;
    A.n_int[j] = ni;
}

// Breadth-first numbering of the next level's 16-wide nodes and the
// worst-case stack: the largest sum of (children - 1) along a path.
__global__ void k_derive16_link(Derive16Args A)
{
    uint32_t j = gid();
    if (j >= A.q) return;
    Node16* o = A.nodes16 + A.base16 + j;
    uint32_t off = A.int_scan[j];
    uint32_t nk = 0;
    for (int t = 0; t < 16; t++)
        if (o->c[t].ref != kNoChild) nk++;
    uint32_t here = A.path[j] + nk - 1;
    for (int t = 0; t < 16; t++) {
        uint32_t ref = o->c[t].ref;
        if (ref == kNoChild || (ref & kLeafBit)) continue;
        A.queue_next[off] = ref;
        A.path_next[off] = here;
        o->c[t].ref = A.base16_next + off;
        off++;
    }
    atomicMax(A.stack16, here);
}

__device__ __forceinline__ void leaf_box(const Prim* prims, uint32_t ref, f3& lo, f3& hi)
{
    uint32_t start = ref & kLeafStartMask, cnt = (ref >> kLeafCountShift) & 7u;
    lo = mk(FLT_MAX, FLT_MAX, FLT_MAX);
    hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (uint32_t i = 0; i < cnt; i++) {
        const Prim& p = prims[start + i];
        f3 a = mk(p.a[0], p.a[1], p.a[2]);
        f3 b = add(a, mk(p.e1[0], p.e1[1], p.e1[2]));
        f3 c = add(a, mk(p.e2[0], p.e2[1], p.e2[2]));
        lo = vmin(lo, vmin(vmin(a, b), c));
        hi = vmax(hi, vmax(vmax(a, b), c));
    }
}

__device__ __forceinline__ void child_box(const Node2* nodes, const Prim* prims, uint32_t ref,
                                          f3& lo, f3& hi)
{
    if (ref & kLeafBit) {
        leaf_box(prims, ref, lo, hi);
        return;
    }
    const Node2& c = nodes[ref];
    lo = vmin(mk(c.lmin[0], c.lmin[1], c.lmin[2]), mk(c.rmin[0], c.rmin[1], c.rmin[2]));
    hi = vmax(mk(c.lmax[0], c.lmax[1], c.lmax[2]), mk(c.rmax[0], c.rmax[1], c.rmax[2]));
}

__global__ void k_refit_level(Node2* nodes, const Prim* prims, uint32_t first, uint32_t count)
{
    uint32_t j = gid();
    if (j >= count) return;
    Node2& o = nodes[first + j];
    f3 lo, hi;
    child_box(nodes, prims, o.lref, lo, hi);
    o.lmin[0] = lo.x; o.lmin[1] = lo.y; o.lmin[2] = lo.z;
    o.lmax[0] = hi.x; o.lmax[1] = hi.y; o.lmax[2] = hi.z;
    child_box(nodes, prims, o.rref, lo, hi);
    o.rmin[0] = lo.x; o.rmin[1] = lo.y; o.rmin[2] = lo.z;
    o.rmax[0] = hi.x; o.rmax[1] = hi.y; o.rmax[2] = hi.z;
}

__global__ void k_refit_root(const Node2* nodes, const Prim* prims, uint32_t root_ref, BuildMeta* meta)
{
    f3 lo, hi;
    child_box(nodes, prims, root_ref, lo, hi);
    meta->root_min[0] = lo.x; meta->root_min[1] = lo.y; meta->root_min[2] = lo.z;
    meta->root_max[0] = hi.x; meta->root_max[1] = hi.y; meta->root_max[2] = hi.z;
}

__global__ void k_refit4(const Node2* nodes, const uint32_t* src4, Node4* nodes4, uint32_t n4)
{
    uint32_t g = gid();
    if (g >= 4 * n4) return;
    uint32_t src = src4[g];
    if (src == kNoChild) return;
    const Node2& n = nodes[src >> 1];
    Child4& c = nodes4[g >> 2].c[g & 3];
    if (src & 1u) {
        c.mn[0] = n.rmin[0]; c.mn[1] = n.rmin[1]; c.mn[2] = n.rmin[2];
        c.mx[0] = n.rmax[0]; c.mx[1] = n.rmax[1]; c.mx[2] = n.rmax[2];
    } else {
        c.mn[0] = n.lmin[0]; c.mn[1] = n.lmin[1]; c.mn[2] = n.lmin[2];
        c.mx[0] = n.lmax[0]; c.mx[1] = n.lmax[1]; c.mx[2] = n.lmax[2];
    }
}

}  // namespace

hipError_t launch_derive16_expand(const Derive16Args& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_derive16_expand, grid_for((size_t)A.q + 1), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_derive16_link(const Derive16Args& A, hipStream_t s)
{
    if (!A.q) return hipSuccess;
    hipLaunchKernelGGL(k_derive16_link, grid_for(A.q), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_tess_count(const TessArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_tess_count, grid_for((size_t)A.nt + 1), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_tess_seed(const TessArgs& A, TessItem* items, hipStream_t s)
{
    if (!A.nt) return hipSuccess;
    hipLaunchKernelGGL(k_tess_seed, grid_for(A.nt), dim3(kBlock), 0, s, A, items);
    return hipGetLastError();
}

hipError_t launch_tess_round_count(const TessRound& R, hipStream_t s)
{
    hipLaunchKernelGGL(k_tess_round_count, grid_for((size_t)R.m + 1), dim3(kBlock), 0, s, R);
    return hipGetLastError();
}

hipError_t launch_tess_round_write(const TessRound& R, hipStream_t s)
{
    if (!R.m) return hipSuccess;
    hipLaunchKernelGGL(k_tess_round_write, grid_for(R.m), dim3(kBlock), 0, s, R);
    return hipGetLastError();
}

hipError_t launch_tess_leaves(const TessArgs& A, const TessItem* items, uint32_t m, hipStream_t s)
{
    if (!m) return hipSuccess;
    hipLaunchKernelGGL(k_tess_leaves, grid_for(m), dim3(kBlock), 0, s, A, items, m);
    return hipGetLastError();
}

hipError_t launch_tess_check(const uint32_t* cnt, const uint32_t* expect, uint32_t nt,
                             BuildMeta* meta, hipStream_t s)
{
    if (!nt) return hipSuccess;
    hipLaunchKernelGGL(k_tess_check, grid_for(nt), dim3(kBlock), 0, s, cnt, expect, nt, meta);
    return hipGetLastError();
}

hipError_t launch_tess_emit(const TessArgs& A, const TessItem* items, uint32_t m, hipStream_t s)
{
    if (!A.nt) return hipSuccess;
    hipLaunchKernelGGL(k_tess_emit, grid_for((size_t)A.nt + m), dim3(kBlock), 0, s, A, items, m);
    return hipGetLastError();
}

hipError_t launch_box_init(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_box_init, grid_for(A.n_seg), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_box(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_box, grid_chunks(A.n), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_small(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_small, dim3((A.n_seg + kSmallBlock - 1) / kSmallBlock), dim3(kSmallBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_seg(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_seg, grid_for(A.n_seg), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_bin(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_bin, grid_chunks(A.n), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_split(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_split, grid_for((size_t)A.n_int + 1), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_flag(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_flag, grid_for((size_t)A.n + 1), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_ptable(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_ptable, grid_for(A.n), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_permute(const LevelArgs& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_permute, grid_for(A.n), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_gather_prims(const Prim* tess, const BRec* recs, uint32_t n, Prim* out,
                               uint32_t* leaf_of, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_gather, grid_for(n), dim3(kBlock), 0, s, tess, recs, n, out, leaf_of);
    return hipGetLastError();
}

hipError_t launch_collapse_expand(const Collapse4Args& A, hipStream_t s)
{
    hipLaunchKernelGGL(k_collapse_expand, grid_for((size_t)A.q + 1), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_collapse_link(const Collapse4Args& A, hipStream_t s)
{
    if (!A.q) return hipSuccess;
    hipLaunchKernelGGL(k_collapse_link, grid_for(A.q), dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_refit_level(Node2* nodes, const Prim* prims, uint32_t first, uint32_t count,
                              hipStream_t s)
{
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_refit_level, grid_for(count), dim3(kBlock), 0, s, nodes, prims, first, count);
    return hipGetLastError();
}

hipError_t launch_refit_root(const Node2* nodes, const Prim* prims, uint32_t root_ref,
                             BuildMeta* meta, hipStream_t s)
{
    hipLaunchKernelGGL(k_refit_root, dim3(1), dim3(1), 0, s, nodes, prims, root_ref, meta);
    return hipGetLastError();
}

hipError_t launch_refit4(const Node2* nodes, const uint32_t* src4, Node4* nodes4, uint32_t n4,
                         hipStream_t s)
{
    if (!n4) return hipSuccess;
    hipLaunchKernelGGL(k_refit4, grid_for(4 * (size_t)n4), dim3(kBlock), 0, s, nodes, src4, nodes4, n4);
    return hipGetLastError();
}

hipError_t scan_u32(void*& tmp, size_t& tmp_bytes, const uint32_t* in, uint32_t* out, size_t n,
                    hipStream_t s)
{
    size_t need = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, need, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
    if (e != hipSuccess) return e;
    if (need > tmp_bytes) {
        if (tmp) (void)hipFree(tmp);
        tmp = nullptr;
        tmp_bytes = 0;
        size_t want = need + need / 2 + 256;
        e = hipMalloc(&tmp, want);
        if (e != hipSuccess) return e;
        tmp_bytes = want;
    }
    return rocprim::exclusive_scan(tmp, need, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

}  // namespace omg
