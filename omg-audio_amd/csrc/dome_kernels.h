// Ambient dome launch layout (see dome_kernels.hip).
#pragma once
#include "trace_kernels.h"

namespace omg {

constexpr uint32_t kDomeBins = 9;     // 8 azimuth sectors + zenith (dome.rs:30)
constexpr int kDomeSegCap = 96;       // crossings kept per straight-through ray

// Raw per-band material (material.rs:7-17) as the dome reads it.
struct omg_material_dev {
    float absorption[kMaxBands];
    float scattering;
    float transmission[kMaxBands];
};

struct DomePanelDev {
    uint32_t axis;
    float plane;
    float lat0, lat1, z0, z1;
    uint32_t glass;
    float trans[kMaxBands];
};

struct DomeRayDev {
    uint32_t escaped, events, used_through, bin;
    float tp[kMaxBands];
    float dir[3];
};

struct DomeArgs {
    GeomArgs g;
    const omg_material_dev* mats;
    const DomePanelDev* panels;
    uint32_t n_panels;
    const float* eyes;        // [n_eyes][3]
    uint32_t n_eyes;
    const float* fan;         // [n_fan][3] golden-spiral directions
    const uint32_t* fan_bin;  // [n_fan] sector of each direction
    uint32_t n_fan, n_rays, max_events, nbands;
    float horizon, through_len;
    DomeRayDev* rays;         // [n_eyes][n_rays]
    float* energy;            // [n_eyes][9][nbands]
    float* dsum;              // [n_eyes][9][3]
    uint32_t* status;         // bit0: a crossing list overflowed
    unsigned long long* counters;  // [2] AABB and triangle tests, or null
};

hipError_t launch_dome(const DomeArgs& A, bool count, hipStream_t s);
// One tiny launch that makes the device load this file's kernels (a code
// object loads on the first launch of any of its kernels): writes 0 to p[0].
hipError_t launch_dome_prime(float* p, hipStream_t s);

}  // namespace omg
