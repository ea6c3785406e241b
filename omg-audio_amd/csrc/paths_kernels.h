// Knife-edge path search launch layout (see paths_kernels.hip).
#pragma once
#include "dome_kernels.h"

namespace omg {

constexpr int kPathSegCap = 64;              // crossings kept per transmission segment
constexpr int kPathMaxPoints = 25;           // the silhouette hull's worst case (24 samples + 1)
constexpr uint32_t kPathMaxEdges = 4096;     // edge graph cap: one LDS sort of all edges
constexpr uint32_t kPathMaxCandidates = 256; // PathBudget::edge_candidates cap
constexpr uint32_t kPathMaxPairEdges = 64;   // PathBudget::pair_edges cap (4096 pairs)
constexpr uint32_t kPathPrefilter = 1024;    // single-bend prefilter size (paths.rs:186-193)
constexpr uint32_t kPathPairKeep = 64;       // double-bend pairs priced (paths.rs:276)
constexpr int kSilSamples = 24;              // silhouette samples along the run (paths.rs:341)
constexpr int kPathBlock = 256;
constexpr int kDirectSplit = 16;             // lanes per direct segment (timed pass)
constexpr int kLegSplitFew = 8;              // lanes per leg of a bent path (timed pass, few legs)
constexpr int kLegSplitMany = 2;             // ... when the legs alone fill the chip several times

// DiffractionEdge (mesh.rs:626-634), padded to two 16 B loads.
struct PathEdgeDev {
    float a[3];
    float importance;
    float b[3];
    uint32_t pad;
};

// FoundPath (paths.rs:48-60) with a fixed point array.
struct PathDev {
    uint32_t key, n_points;
    float length;
    float gains[3];
    float points[kPathMaxPoints][3];
};

struct PathsArgs {
    GeomArgs g;
    const omg_material_dev* mats;
    const PathEdgeDev* edges;
    uint32_t ne;
    const float* src;          // [n][3]
    const float* lis;          // [n][3]
    uint32_t n;
    uint32_t edge_candidates, pair_edges, max_paths, taps;
    PathDev* out;              // [n][1 + max_paths]
    uint32_t* n_out;           // [n]
    uint32_t* work;            // [n] blocked queries (filled by the direct pass)
    float* plan;               // [n][plan_layout().words] per blocked query
    uint32_t* counters;        // [0] blocked count, [1] overflow flag (sticky)
    unsigned long long* trav;  // optional [2]: AABB and triangle tests of all traversals
};

struct PathTransArgs {
    GeomArgs g;
    const omg_material_dev* mats;
    const float* a;            // [n][3]
    const float* b;            // [n][3]
    uint32_t n;
    float* gains;              // [n][3]
    uint32_t* counters;        // [1] overflow flag
};

// bend_point (paths.rs:411-436): golden section on t in [0, 1].  Each step
// keeps one probe and evaluates one new one; the side is chosen with
// selects, not a branch, so lanes that keep different sides evaluate their
// one cost together.  Every probe and bound is the reference's expression
// (hi - phi * (hi - lo) or lo + phi * (hi - lo)), so the result is the same.
OMG_HD f3 bend_point(f3 a, f3 b, f3 src, f3 lis)
{
    f3 ab = sub(b, a);
    auto cost = [&](float t) {
        return length(sub(add(a, mul(ab, t)), src)) + length(sub(lis, add(a, mul(ab, t))));
    };
    const float phi = 0.618034f;
    float lo = 0.0f, hi = 1.0f;
    float m1 = hi - phi * (hi - lo), m2 = lo + phi * (hi - lo);
    float c1 = cost(m1), c2 = cost(m2);
    for (int k = 0; k < 24; k++) {
        const bool left = c1 <= c2;
        const float nhi = left ? m2 : hi;
        const float nlo = left ? lo : m1;
        const float keep_m = left ? m1 : m2;
        const float keep_c = left ? c1 : c2;
        const float t = left ? nhi - phi * (nhi - nlo) : nlo + phi * (nhi - nlo);
        const float c = cost(t);
        hi = nhi;
        lo = nlo;
        m1 = left ? t : keep_m;
        c1 = left ? c : keep_c;
        m2 = left ? keep_m : t;
        c2 = left ? keep_c : c;
    }
    return add(a, mul(ab, 0.5f * (lo + hi)));
}

OMG_HD uint32_t pow2_ceil(uint32_t n)
{
    uint32_t p = 2;
    while (p < n) p <<= 1;
    return p;
}

OMG_HD uint32_t pair_keep(uint32_t pair_edges)
{
    uint32_t pairs = pair_edges * pair_edges;
    return pairs < kPathPairKeep ? pairs : kPathPairKeep;
}

// Per blocked query, the ranking pass leaves a plan in HBM for the leg and
// assembly passes (32-bit words): n1, n2, n_legs, n_drops; the ranked
// single-bend edges and their bend points; the ranked pairs and their two
// bend points; every leg's endpoints (a silhouette drop's origin in a) and
// its three band gains (a drop's roof height + 0.03, or -inf, in the first).
// Legs: 2 per single bend, 3 per pair, then the drops.
// The ranking pass also keeps the bend point of every single-bend
// candidate (csp, by candidate position) and both bend points of every
// pair it ranked (cpp, by pair position), so the ranked ones are not
// recomputed.
struct PlanLayout {
    uint32_t sel, sp, pe, pp, legs, lg, csp, cpp, words, max_legs;
};

OMG_HD PlanLayout plan_layout(uint32_t ne, uint32_t ec, uint32_t pair_edges)
{
    uint32_t kp = pair_keep(pair_edges);
    uint32_t pre = ne < kPathPrefilter ? ne : kPathPrefilter;
    PlanLayout L;
    L.max_legs = 2 * ec + 3 * kp + (kSilSamples - 1);
    uint32_t o = 4;
    L.csp = o;  o += 3 * pre;
    L.cpp = o;  o += 6 * pair_edges * pair_edges;
    L.sel = o;  o += ec;
    L.sp = o;   o += 3 * ec;
    L.pe = o;   o += 2 * kp;
    L.pp = o;   o += 6 * kp;
    L.legs = o; o += 6 * L.max_legs;
    L.lg = o;   o += 3 * L.max_legs;
    L.words = (o + 3u) & ~3u;
    return L;
}

// LDS of the ranking pass (bytes): the sort keys (8 B x the largest sort:
// all edges, the prefiltered candidates, the pairs), the single-bend
// candidates, the two corridors, the ranked singles and pairs.
struct RankLds {
    uint32_t keys, pre, near_s, near_l, sel, pe, bytes;
};

OMG_HD RankLds rank_lds(uint32_t ne, uint32_t ec, uint32_t pair_edges)
{
    uint32_t pairs = pair_edges * pair_edges;
    uint32_t n_keys = ne > pairs ? ne : pairs;
    uint32_t pre = ne < kPathPrefilter ? ne : kPathPrefilter;
    auto al = [](uint32_t x) { return (x + 15u) & ~15u; };
    auto one = [](uint32_t x) { return x ? x : 1u; };
    RankLds L;
    uint32_t o = 0;
    L.keys = o;   o = al(o + 8 * pow2_ceil(n_keys));
    L.pre = o;    o = al(o + 4 * one(pre));
    L.near_s = o; o = al(o + 4 * one(pair_edges));
    L.near_l = o; o = al(o + 4 * one(pair_edges));
    L.sel = o;    o = al(o + 4 * one(ec));
    L.pe = o;     o = al(o + 12 * one(pair_keep(pair_edges)));
    L.bytes = o;
    return L;
}

// LDS of the assembly pass: the final sort keys, gains and lengths, keys
// of the bent paths, the silhouette hull.
struct AsmLds {
    uint32_t keys, bg, bkey, hull, bytes;
};

OMG_HD AsmLds asm_lds(uint32_t ec, uint32_t pair_edges)
{
    uint32_t nb = ec + pair_keep(pair_edges) + 1;
    auto al = [](uint32_t x) { return (x + 15u) & ~15u; };
    AsmLds L;
    uint32_t o = 0;
    L.keys = o; o = al(o + 8 * pow2_ceil(nb));
    L.bg = o;   o = al(o + 16 * nb);
    L.bkey = o; o = al(o + 4 * nb);
    L.hull = o; o = al(o + 12 * kPathMaxPoints);
    L.bytes = o;
    return L;
}

// Direct, ranking, leg and assembly passes; counters must be zero on entry.
hipError_t launch_paths(const PathsArgs& A, bool count, hipStream_t s);
hipError_t launch_path_transmission(const PathTransArgs& A, hipStream_t s);

}  // namespace omg
