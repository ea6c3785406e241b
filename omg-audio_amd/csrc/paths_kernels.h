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

// bend_point (paths.rs:411-436): golden section over t in [0, 1]
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
        if (c1 <= c2) {
            hi = m2;
            m2 = m1;
            c2 = c1;
            m1 = hi - phi * (hi - lo);
            c1 = cost(m1);
        } else {
            lo = m1;
            m1 = m2;
            c1 = c2;
            m2 = lo + phi * (hi - lo);
            c2 = cost(m2);
        }
    }
    return add(a, mul(ab, 0.5f * (lo + hi)));
}

// Direct pass then the bent-path pass; counters must be zero on entry.
hipError_t launch_paths(const PathsArgs& A, bool count, hipStream_t s);
hipError_t launch_path_transmission(const PathTransArgs& A, hipStream_t s);

}  // namespace omg
