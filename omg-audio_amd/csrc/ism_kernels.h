// Image-source discovery and solve launch layout (see ism_kernels.hip).
#pragma once
#include "dome_kernels.h"

namespace omg {

constexpr int kSolveSegCap = 64;   // crossings kept per leg

struct SurfaceDev {              // pt_mesh.rs:29-41
    float n[3];
    float d;
    float refl[kMaxBands];
    float trans[kMaxBands];
    uint32_t has_rect;
    float rmin[3], rmax[3];
};

struct BlockerDev {              // pt.rs:31-35
    float min[3], max[3];
    float trans[kMaxBands];
};

struct MeshRecordDev {           // pt_mesh.rs:111-120
    uint32_t valid;
    float delay;
    float dir[3];
    float gains[kMaxBands];
};

struct DiscoverArgs {
    GeomArgs g;
    const uint32_t* tri_surface;
    float listener[3];
    float ga;            // golden angle pi*(3 - sqrt 5) in f32
    float rot_phase;     // rot * 0.61803398875 * TAU in f32
    const float* jitter; // [n_rays] xorshift jitter in [0, 1)
    uint32_t n_rays;
    const float* boxes;  // [n_boxes][6] overlay boxes (min, max) that reflect too
    uint32_t n_boxes;
    uint32_t base;       // surface id of the first box's first face
    uint32_t* out;       // [n_rays][2]: chain0 | chain1 << 16, chain2 | levels << 16
};

struct SolveArgs {
    GeomArgs g;
    const uint32_t* tri_surface;
    const omg_material_dev* mats;
    const SurfaceDev* surfaces;
    uint32_t n_surf;
    const float* sources;
    uint32_t n_src;
    const uint16_t* chains;   // [n_chains][3]
    const uint8_t* orders;
    uint32_t n_chains;
    float listener[3];
    const BlockerDev* extras;
    uint32_t n_extras;
    float air[kMaxBands];
    uint32_t nbands;
    MeshRecordDev* records;   // [n_src][n_chains]
    uint32_t* status;
};

hipError_t launch_discover(const DiscoverArgs& A, hipStream_t s);
hipError_t launch_solve(const SolveArgs& A, hipStream_t s);

}  // namespace omg
