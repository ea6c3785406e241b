// Host runtime internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/omg_hip.h"
#include "omg_layout.h"
#include "scene_build.h"
#include "trace_kernels.h"

namespace omg {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define OMG_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::omg::fail(OMG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    // grows to exactly `bytes` (no slack): for arenas whose size is budgeted
    hipError_t ensure_exact(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&ptr, bytes);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const { return reinterpret_cast<T*>(ptr); }
};

}  // namespace omg

struct omg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    // per-call staging (one user thread at a time per context)
    omg::DevBuf job_mats, job_ovl, job_src, job_lis, job_bins, job_dirs, job_rec, job_cnt;
    omg::DevBuf aux0, aux1, aux2, aux3, aux4;
    // dome fan of the last size used, and the sticky overflow word of async launches
    omg::DevBuf fan_dirs, fan_bins, dome_status, dome_panels;
    uint32_t fan_n = 0;
    // scenes and batches alive on this context; omg_shutdown defers the
    // teardown until the last is destroyed
    uint32_t live_scenes = 0;
    // autotuned batch launch shapes by batch class
    std::map<uint64_t, omg::BatchShape> tuned;
    // tracer launch configuration (omg_set_launch_config)
    uint32_t node_cache = 512;
    uint32_t stack_lds = 16;
    uint32_t lanes = 0;                 // 0: pick per launch (pick_lanes)
    uint32_t width = 4;                 // BVH width of the plain launch (4 where built, else 2)
    uint64_t quad_below = 1ull << 20;   // launches with fewer rays trace a quad per ray
    uint32_t n_cu = 0;                  // compute units (persistent launches size their grid to it)
    // plain launches of at least route_rays rays run on the persistent
    // regenerating kernel (0: never); its job view, items and cursor.  The
    // default is quad_below: a launch too large for a quad per ray runs
    // persistent, never one lane per ray waiting for its block's longest ray
    uint64_t route_rays = 1ull << 20;
    omg::DevBuf route_job, route_items, route_cursor;
    // quad-per-ray plain launches hand their rays still alive after
    // handoff_bounce bounces to the 16-lane walk (0: never); the saved
    // states and the count / cursor words, last used on handoff_stream
    uint32_t handoff_bounce = 4;
    uint32_t handoff_nc = 32;           // 16-wide nodes in LDS per block of the resume kernel
    uint32_t handoff_wps = 6;           // its resident waves per SIMD
    uint32_t handoff_lanes = 0;         // its lanes per ray: 16, 8 (two node slots per lane), 0 by size
    omg::DevBuf handoff, handoff_ctl;
    hipStream_t handoff_stream = nullptr;
    bool closing = false;
    // the *_async entry points leave kernels in flight that read the staged
    // job tables (job_mats, job_ovl) and dome fan / panels; staging the same
    // bytes again is skipped, and anything that overwrites a buffer such a
    // kernel may read first waits for the device (omg::quiesce)
    bool async_inflight = false;
    bool tables_valid = false;
    std::vector<uint8_t> tables_key;
};

struct omg_scene {
    omg_ctx* ctx = nullptr;
    uint32_t nbands = 3;
    uint32_t n_tris = 0, n_prims = 0, n_nodes = 0, max_depth = 0, root_ref = 0;
    float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
    bool empty = false;
    std::vector<omg_material> materials;       // host copy (overlay tables append to it)
    std::vector<omg::MatParams> mat_params;    // tracer constants of the scene materials
    std::vector<uint16_t> tri_surface;         // host copy for chain ids
    std::vector<float> positions;              // host copy for the surface table
    std::vector<uint32_t> indices;
    std::vector<uint16_t> tri_material;
    std::vector<omg::TriInfo> tris_host;
    omg::DevBuf nodes, prims, tris, mats, mats_raw, tri_surf, nodes4, nodes16;
    uint32_t n_nodes4 = 0, root4 = 0, stack4 = 0;
    bool has4 = false;
    // 16-wide nodes derived from the 4-wide ones (omg::derive16); has16
    // when the worst-case stack fits the 16-lane walk's LDS stack
    uint32_t n_nodes16 = 0, root16 = 0, stack16 = 0;
    bool has16 = false;
    uint64_t device_bytes = 0;
    // scenes built on the device keep what a refit needs: the inputs, the
    // per-triangle patch counts and offsets, each patch's leaf position,
    // the first node of every tree level and each 4-wide child's source
    bool device_built = false;
    uint32_t nv = 0;
    omg::DevBuf b_pos, b_idx, b_mat, b_surf, b_maxe, b_cnt, b_off, leaf_of, src4;
    std::vector<uint32_t> level_base;
};

namespace omg {
hipError_t launch_trace(const TraceArgs& A, bool mesh, bool count, hipStream_t s);
hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s);
hipError_t launch_fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s);
// Copies `bytes` (a multiple of 16, both ends 16-byte aligned) with a kernel on s.
hipError_t launch_copy16(void* dst, const void* src, size_t bytes, hipStream_t s);
int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s);
uint32_t pick_lanes(const omg_ctx* c, uint64_t n_rays);
// lanes per ray, BVH width, LDS node prefix and stack of a plain launch
void plain_shape(const omg_ctx* c, const omg_scene* s, const omg_trace_params* p, uint32_t n_src,
                 bool counting, TraceArgs& A);
// the hand-off buffers of a plain launch that plain_shape gave a hand-off
// bounce: sized for its rays, counters cleared on `st`
int setup_handoff(omg_ctx* c, TraceArgs& A, hipStream_t st);
void fill_geom(const omg_scene* s, GeomArgs& g);
int validate_params(const omg_trace_params* p, uint32_t nbands, uint32_t n_src, uint32_t n_lis,
                    uint32_t n_ovl, uint32_t* nbins);
void job_tables(const std::vector<MatParams>& base_mats, const omg_overlay* ovl, uint32_t n_ovl,
                uint32_t nbands, std::vector<MatParams>& mats, std::vector<Overlay>& ov);
void job_scalars(const omg_trace_params* p, uint32_t nbands, uint32_t n_ovl, JobView& A);
// drops one scene/batch/server reference; tears the context down after omg_shutdown
void ctx_release(omg_ctx* c);
void make_fan(uint32_t n_fan, std::vector<float>& fan, std::vector<uint32_t>& bins);
// waits for in-flight *_async work before shared staging is overwritten
int quiesce(omg_ctx* c);
// Resolves `device`'s copy of the allocation-free launch's scratch array
// (omg_trace_world_async_noalloc); called by omg_init with the device current.
int noalloc_setup(int device);
// derives the scene's 16-wide nodes from its resident 4-wide nodes on
// the device (stream-ordered; waits for the stream at the end)
int derive16(omg_scene* s, hipStream_t st);
// host copies of the mesh and its triangle normals (scenes built on the
// device download and compute them on first use)
int ensure_host_mesh(omg_scene* s);
}  // namespace omg
