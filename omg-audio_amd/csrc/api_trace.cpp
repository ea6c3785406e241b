// C ABI: lifecycle, scenes, the late-field tracer and its post-processing.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "runtime.h"

#pragma clang fp contract(off)

namespace omg {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace omg

using namespace omg;

extern "C" const char* omg_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ lifecycle

extern "C" int omg_init(int device, uint32_t layout_version, omg_ctx** out)
{
    if (!out) return fail(OMG_ERR_INVALID, "out is null");
    *out = nullptr;
    if (layout_version != OMG_LAYOUT_VERSION)
        return fail(OMG_ERR_VERSION, "layout version mismatch");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return fail(OMG_ERR_NO_DEVICE, "no usable HIP device");
    OMG_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    OMG_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    omg_ctx* c = new omg_ctx();
    c->device = device;
    c->name = prop.gcnArchName;
    c->n_cu = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 1u;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(OMG_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (int rc = noalloc_setup(device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return OMG_OK;
}

static void ctx_teardown(omg_ctx* c)
{
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    // async launches on the caller's streams may still read the context's buffers
    if (c->async_inflight) (void)hipDeviceSynchronize();
    for (DevBuf* b : {&c->job_mats, &c->job_ovl, &c->job_src, &c->job_lis, &c->job_bins,
                      &c->job_dirs, &c->job_rec, &c->job_cnt, &c->aux0, &c->aux1, &c->aux2,
                      &c->aux3, &c->aux4, &c->fan_dirs, &c->fan_bins, &c->dome_status,
                      &c->dome_panels, &c->route_job, &c->route_items, &c->route_cursor,
                      &c->handoff, &c->handoff_ctl})
        b->release();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

// The context outlives its scenes whatever order the caller releases them in:
// with scenes alive, the teardown waits for the last omg_scene_destroy.
extern "C" void omg_shutdown(omg_ctx* c)
{
    if (!c) return;
    c->closing = true;
    if (c->live_scenes == 0) ctx_teardown(c);
}

extern "C" int omg_set_launch_config(omg_ctx* c, uint32_t node_cache, uint32_t stack_lds,
                                     uint32_t lanes, uint32_t width)
{
    if (!c) return fail(OMG_ERR_INVALID, "bad argument");
    if (width != 2 && width != 4 && width != 16) return fail(OMG_ERR_INVALID, "width must be 2, 4 or 16");
    if (stack_lds != 16 && stack_lds != 32) return fail(OMG_ERR_INVALID, "stack_lds must be 16 or 32");
    if (lanes != 0 && lanes != 1 && lanes != 4 && lanes != 16)
        return fail(OMG_ERR_INVALID, "lanes must be 0, 1, 4 or 16");
    size_t lds = (size_t)node_cache * sizeof(Node2) + (size_t)stack_lds * 256 * 8;
    if (lds > 160 * 1024) return fail(OMG_ERR_CAP, "LDS working set above 160 KiB");
    c->node_cache = node_cache;
    c->stack_lds = stack_lds;
    c->lanes = lanes;
    c->width = width;
    return OMG_OK;
}

extern "C" int omg_set_route_rays(omg_ctx* c, uint64_t min_rays)
{
    if (!c) return fail(OMG_ERR_INVALID, "bad argument");
    c->route_rays = min_rays;
    return OMG_OK;
}

extern "C" int omg_set_handoff_lanes(omg_ctx* c, uint32_t lanes)
{
    if (!c) return fail(OMG_ERR_INVALID, "bad argument");
    if (lanes != 0 && lanes != 8 && lanes != 16) return fail(OMG_ERR_INVALID, "lanes must be 0, 8 or 16");
    c->handoff_lanes = lanes;
    return OMG_OK;
}

extern "C" int omg_set_handoff(omg_ctx* c, uint32_t bounce, uint32_t node_cache16,
                               uint32_t waves_per_simd)
{
    if (!c) return fail(OMG_ERR_INVALID, "bad argument");
    if (waves_per_simd < 1 || waves_per_simd > 8) return fail(OMG_ERR_INVALID, "waves_per_simd must be 1..8");
    if ((size_t)node_cache16 * sizeof(Node16) + (size_t)kStack16 * 16 * 8 > 64 * 1024)
        return fail(OMG_ERR_CAP, "LDS working set above 64 KiB");
    c->handoff_bounce = bounce;
    c->handoff_nc = node_cache16;
    c->handoff_wps = waves_per_simd;
    return OMG_OK;
}

// Lanes per ray for a launch of n rays: a quad per ray while one lane per
// ray would leave the 1024 SIMDs with few waves to switch between.
uint32_t omg::pick_lanes(const omg_ctx* c, uint64_t n_rays)
{
    if (c->lanes) return c->lanes;
    return n_rays < c->quad_below ? 4u : 1u;
}

void omg::plain_shape(const omg_ctx* c, const omg_scene* s, const omg_trace_params* p,
                      uint32_t n_src, bool counting, TraceArgs& A)
{
    const uint64_t rays = (uint64_t)p->n_rays * n_src;
    A.node_cache = s->n_nodes < c->node_cache ? s->n_nodes : c->node_cache;
    A.stack_lds = c->stack_lds;
    A.lanes = pick_lanes(c, rays);
    A.width = 2;
    if (counting) {
        // traversal counters are the reference walk's: counted on the binary tree
        if (A.lanes == 16) A.lanes = 4;
        return;
    }
    if (A.lanes == 16 || c->width == 16) {
        if (s->has16 && s->n_nodes16) {
            // a 16-lane row per ray over the 16-wide nodes, the whole stack in LDS
            A.lanes = 16;
            A.width = 16;
            A.stack_lds = kStack16;
            A.node_cache = std::min(s->n_nodes16, c->node_cache / 8);
            return;
        }
        A.lanes = 4;
    }
    if (c->width >= 4 && s->has4) {
        A.width = 4;
        A.node_cache = std::min(s->n_nodes4, c->node_cache / 2);
        if (A.lanes == 4 && s->stack4 <= kDeepStack && c->stack_lds == 16) {
            // the quad walk keeps its whole stack in LDS: 20 KB per block, so
            // the node prefix shrinks to keep a block within 40 KB
            A.stack_lds = kDeepStack;
            A.node_cache = std::min(A.node_cache, 160u);
            if (c->handoff_bounce && s->has16 && s->n_nodes16 && p->max_bounces > c->handoff_bounce &&
                !(c->route_rays && rays >= c->route_rays)) {
                // the rays still alive after the first bounces continue on a
                // 16-lane row each (see k_trace_resume16)
                A.handoff_bounce = c->handoff_bounce;
                A.resume_blocks = c->n_cu * c->handoff_wps;
                A.node_cache16 = std::min(s->n_nodes16, c->handoff_nc);
                // 8-lane rows (two slots per lane) from 256 k rays: with that
                // many rays handed off the kernel is bound by instruction
                // issue, below it by each ray's chain of steps (a 53k-
                // triangle city block, 16 sources: 64 k rays 16 lanes
                // 1.43 ms, 8 lanes 1.47 ms; 256 k rays 3.86 / 3.60 ms)
                A.resume_lanes = c->handoff_lanes ? c->handoff_lanes : rays >= (1u << 18) ? 8u : 16u;
                A.resume_stack = std::max(1u, s->stack16);
                A.resume_simds = c->n_cu * 4u;
            }
        }
    }
}

int omg::setup_handoff(omg_ctx* c, TraceArgs& A, hipStream_t st)
{
    if (!A.handoff_bounce) return OMG_OK;
    const size_t bytes = (size_t)A.ray_count * kHandoffF4 * sizeof(float4);
    // the buffers are the context's: a launch on another stream, or one that
    // regrows them, waits for the launches that may still use them
    if ((bytes > c->handoff.cap || st != c->handoff_stream) && c->handoff_stream) {
        int rc = quiesce(c);
        if (rc) return rc;
        OMG_HIP_CHECK(hipStreamSynchronize(c->handoff_stream));
    }
    OMG_HIP_CHECK(c->handoff.ensure(bytes));
    OMG_HIP_CHECK(c->handoff_ctl.ensure(kHandoffCtlWords * sizeof(uint32_t)));
    c->handoff_stream = st;
    A.handoff = c->handoff.as<float4>();
    A.handoff_ctl = c->handoff_ctl.as<uint32_t>();
    OMG_HIP_CHECK(launch_fill_zero(reinterpret_cast<float*>(A.handoff_ctl), kHandoffCtlWords, st));
    return OMG_OK;
}

extern "C" int omg_device_name(omg_ctx* c, char* buf, uint32_t len)
{
    if (!c || !buf || len == 0) return fail(OMG_ERR_INVALID, "bad argument");
    std::strncpy(buf, c->name.c_str(), len - 1);
    buf[len - 1] = 0;
    return OMG_OK;
}

// ------------------------------------------------------------------ scenes

int omg::upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s)
{
    OMG_HIP_CHECK(b.ensure(bytes ? bytes : 16));
    if (bytes) OMG_HIP_CHECK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
    return OMG_OK;
}

extern "C" int omg_scene_create(omg_ctx* c, const float* positions, uint32_t nv,
                                const uint32_t* indices, uint32_t nt, const uint16_t* tri_material,
                                const uint16_t* tri_surface, const float* tri_max_edge,
                                const omg_material* materials, uint32_t nm, uint32_t nbands,
                                omg_scene** out)
{
    if (!c || !out || (nt && (!positions || !indices || !tri_material)) || (nm && !materials))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (nbands < 1 || nbands > OMG_MAX_BANDS) return fail(OMG_ERR_INVALID, "nbands out of range");
    for (uint32_t t = 0; t < nt; t++)
        if (tri_material[t] >= nm) return fail(OMG_ERR_INVALID, "tri_material out of range");
    *out = nullptr;
    HostScene H;
    std::string err;
    if (build_scene(positions, nv, indices, nt, tri_material, tri_surface, tri_max_edge, H, err))
        return fail(OMG_ERR_INVALID, err);
    if (H.max_depth + 2 > (uint32_t)kStackCap)
        return fail(OMG_ERR_CAP, "BVH deeper than the traversal stack");
    OMG_HIP_CHECK(hipSetDevice(c->device));
    omg_scene* s = new omg_scene();
    s->ctx = c;
    s->nbands = nbands;
    s->n_tris = nt;
    s->n_prims = (uint32_t)H.prims.size();
    s->n_nodes = (uint32_t)H.nodes.size();
    s->max_depth = H.max_depth;
    s->root_ref = H.root_ref;
    for (int k = 0; k < 3; k++) {
        s->root_min[k] = H.root_min[k];
        s->root_max[k] = H.root_max[k];
    }
    s->empty = H.empty;
    s->has4 = H.has4;
    s->root4 = H.root4;
    s->stack4 = H.stack4;
    s->n_nodes4 = (uint32_t)H.nodes4.size();
    s->materials.assign(materials, materials + nm);
    s->mat_params.resize(nm);
    for (uint32_t m = 0; m < nm; m++)
        mat_params(materials[m].absorption, materials[m].scattering, materials[m].transmission,
                   nbands, s->mat_params[m]);
    s->tri_surface.assign(nt, 0);
    if (tri_surface) s->tri_surface.assign(tri_surface, tri_surface + nt);
    s->positions.assign(positions, positions + (size_t)nv * 3);
    s->indices.assign(indices, indices + (size_t)nt * 3);
    s->tri_material.assign(tri_material, tri_material + nt);
    s->tris_host = H.tris;
    int rc = upload(s->nodes, H.nodes.data(), H.nodes.size() * sizeof(Node2), c->stream);
    if (!rc) rc = upload(s->prims, H.prims.data(), H.prims.size() * sizeof(Prim), c->stream);
    if (!rc) rc = upload(s->tris, H.tris.data(), H.tris.size() * sizeof(TriInfo), c->stream);
    if (!rc) rc = upload(s->mats, s->mat_params.data(), s->mat_params.size() * sizeof(MatParams), c->stream);
    if (!rc) rc = upload(s->mats_raw, s->materials.data(), s->materials.size() * sizeof(omg_material), c->stream);
    std::vector<uint32_t> ts32(s->tri_surface.begin(), s->tri_surface.end());
    if (!rc) rc = upload(s->tri_surf, ts32.data(), ts32.size() * sizeof(uint32_t), c->stream);
    if (!rc && s->has4) rc = upload(s->nodes4, H.nodes4.data(), H.nodes4.size() * sizeof(Node4), c->stream);
    if (!rc) rc = derive16(s, c->stream);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "scene upload");
    if (rc) {
        s->nodes.release();
        s->prims.release();
        s->tris.release();
        s->mats.release();
        s->mats_raw.release();
        s->tri_surf.release();
        s->nodes4.release();
        s->nodes16.release();
        delete s;
        return rc;
    }
    s->device_bytes = s->nodes.cap + s->prims.cap + s->tris.cap + s->mats.cap + s->mats_raw.cap +
                      s->tri_surf.cap + s->nodes4.cap + s->nodes16.cap;
    c->live_scenes++;
    *out = s;
    return OMG_OK;
}

extern "C" int omg_scene_create_soup(omg_ctx* c, const float* tri_verts, uint32_t nt,
                                     const uint16_t* tri_material, const uint16_t* tri_surface,
                                     const float* tri_max_edge, const omg_material* materials,
                                     uint32_t nm, uint32_t nbands, omg_scene** out)
{
    if (nt && !tri_verts) return fail(OMG_ERR_INVALID, "tri_verts is null");
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    uint32_t nv = dedup_soup(tri_verts, nt, pos, idx);
    return omg_scene_create(c, pos.data(), nv, idx.data(), nt, tri_material, tri_surface,
                            tri_max_edge, materials, nm, nbands, out);
}

extern "C" int omg_scene_info_get(omg_scene* s, omg_scene_info* info)
{
    if (!s || !info) return fail(OMG_ERR_INVALID, "bad argument");
    info->n_tris = s->n_tris;
    info->n_prims = s->n_prims;
    info->n_nodes = s->n_nodes;
    info->max_depth = s->max_depth;
    info->device_bytes = s->device_bytes;
    return OMG_OK;
}

extern "C" int omg_dedup_soup(const float* tri_verts, uint32_t nt, float* positions_out,
                              uint32_t* indices_out, uint32_t* nv_out)
{
    if ((nt && !tri_verts) || !nv_out) return fail(OMG_ERR_INVALID, "bad argument");
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    uint32_t nv = dedup_soup(tri_verts, nt, pos, idx);
    if (positions_out) std::memcpy(positions_out, pos.data(), pos.size() * sizeof(float));
    if (indices_out) std::memcpy(indices_out, idx.data(), idx.size() * sizeof(uint32_t));
    *nv_out = nv;
    return OMG_OK;
}

extern "C" int omg_bvh_build_host(const float* positions, uint32_t nv, const uint32_t* indices,
                                  uint32_t nt, const uint16_t* tri_material,
                                  const uint16_t* tri_surface, const float* tri_max_edge,
                                  float* prims_out, uint32_t prim_cap, uint32_t* n_prims,
                                  float* nodes_out, uint32_t node_cap, uint32_t* n_nodes,
                                  uint32_t* root_ref, uint32_t* max_depth)
{
    if ((nt && (!positions || !indices || !tri_material)) || !n_prims || !n_nodes || !root_ref ||
        !max_depth)
        return fail(OMG_ERR_INVALID, "bad argument");
    HostScene H;
    std::string err;
    if (build_scene(positions, nv, indices, nt, tri_material, tri_surface, tri_max_edge, H, err))
        return fail(OMG_ERR_INVALID, err);
    *n_prims = (uint32_t)H.prims.size();
    *n_nodes = (uint32_t)H.nodes.size();
    *root_ref = H.root_ref;
    *max_depth = H.max_depth;
    if (prims_out) {
        if (prim_cap < H.prims.size()) return fail(OMG_ERR_CAP, "prim buffer too small");
        std::memcpy(prims_out, H.prims.data(), H.prims.size() * sizeof(Prim));
    }
    if (nodes_out) {
        if (node_cap < H.nodes.size()) return fail(OMG_ERR_CAP, "node buffer too small");
        std::memcpy(nodes_out, H.nodes.data(), H.nodes.size() * sizeof(Node2));
    }
    return OMG_OK;
}

extern "C" int omg_bvh4_build_host(const float* positions, uint32_t nv, const uint32_t* indices,
                                   uint32_t nt, const uint16_t* tri_material,
                                   const uint16_t* tri_surface, const float* tri_max_edge,
                                   float* nodes4_out, uint32_t node_cap, uint32_t* n_nodes4,
                                   uint32_t* root4, uint32_t* stack4, uint32_t* usable)
{
    if ((nt && (!positions || !indices || !tri_material)) || !n_nodes4 || !root4 || !stack4 ||
        !usable)
        return fail(OMG_ERR_INVALID, "bad argument");
    HostScene H;
    std::string err;
    if (build_scene(positions, nv, indices, nt, tri_material, tri_surface, tri_max_edge, H, err))
        return fail(OMG_ERR_INVALID, err);
    *n_nodes4 = (uint32_t)H.nodes4.size();
    *root4 = H.root4;
    *stack4 = H.stack4;
    *usable = H.has4 ? 1u : 0u;
    if (nodes4_out) {
        if (node_cap < H.nodes4.size()) return fail(OMG_ERR_CAP, "node buffer too small");
        std::memcpy(nodes4_out, H.nodes4.data(), H.nodes4.size() * sizeof(Node4));
    }
    return OMG_OK;
}

extern "C" void omg_scene_destroy(omg_scene* s)
{
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    s->nodes.release();
    s->prims.release();
    s->tris.release();
    s->mats.release();
    s->mats_raw.release();
    s->tri_surf.release();
    s->nodes4.release();
    s->nodes16.release();
    for (DevBuf* b : {&s->b_pos, &s->b_idx, &s->b_mat, &s->b_surf, &s->b_maxe, &s->b_cnt,
                      &s->b_off, &s->leaf_of, &s->src4})
        b->release();
    omg_ctx* c = s->ctx;
    delete s;
    ctx_release(c);
}

int omg::quiesce(omg_ctx* c)
{
    if (!c->async_inflight) return OMG_OK;
    OMG_HIP_CHECK(hipDeviceSynchronize());
    c->async_inflight = false;
    return OMG_OK;
}

void omg::ctx_release(omg_ctx* c)
{
    c->live_scenes--;
    if (c->closing && c->live_scenes == 0) ctx_teardown(c);
}

// ------------------------------------------------------------------ tracer

namespace {
uint32_t nbins_of(float max_time, float bin_dt) { return sat_u32(max_time / bin_dt); }
}  // namespace

int omg::validate_params(const omg_trace_params* p, uint32_t nbands, uint32_t n_src,
                         uint32_t n_lis, uint32_t n_ovl, uint32_t* nbins)
{
    if (!p) return fail(OMG_ERR_INVALID, "params is null");
    if (p->nbands != nbands) return fail(OMG_ERR_INVALID, "params.nbands differs from the scene");
    if (p->dir_band >= nbands) return fail(OMG_ERR_INVALID, "dir_band out of range");
    if (p->n_rays == 0 || n_src == 0) return fail(OMG_ERR_INVALID, "no rays");
    if (n_lis == 0) return fail(OMG_ERR_INVALID, "no listeners");
    if (n_lis > OMG_MAX_LISTENERS) return fail(OMG_ERR_CAP, "too many listeners");
    if (n_ovl > OMG_MAX_OVERLAYS) return fail(OMG_ERR_CAP, "too many overlays");
    if ((uint64_t)p->n_rays * n_src >= (1ull << 32)) return fail(OMG_ERR_CAP, "too many rays");
    if (p->max_bounces == 0 || p->max_bounces > 4096) return fail(OMG_ERR_INVALID, "max_bounces");
    if (!(p->bin_dt > 0.0f) || !(p->max_time > 0.0f)) return fail(OMG_ERR_INVALID, "bin_dt/max_time");
    *nbins = nbins_of(p->max_time, p->bin_dt);
    if (*nbins == 0 || *nbins > (1u << 22)) return fail(OMG_ERR_INVALID, "echogram length");
    return OMG_OK;
}

// The job's material table (scene materials, then one per overlay) and
// its overlay boxes.
void omg::job_tables(const std::vector<MatParams>& base_mats, const omg_overlay* ovl,
                     uint32_t n_ovl, uint32_t nbands, std::vector<MatParams>& mats,
                     std::vector<Overlay>& ov)
{
    mats = base_mats;
    ov.assign(n_ovl, Overlay{});
    for (uint32_t i = 0; i < n_ovl; i++) {
        MatParams mp;
        mat_params(ovl[i].mat.absorption, ovl[i].mat.scattering, ovl[i].mat.transmission, nbands, mp);
        mats.push_back(mp);
        for (int k = 0; k < 3; k++) {
            ov[i].mn[k] = ovl[i].min[k];
            ov[i].mx[k] = ovl[i].max[k];
        }
        ov[i].mat = (uint32_t)(base_mats.size() + i);
        ov[i].pad = 0;
    }
}

// Launch-invariant scalars of a job (tracer.rs:92-103 with the constants
// of tracer.rs:12-16 as parameters).
void omg::job_scalars(const omg_trace_params* p, uint32_t nbands, uint32_t n_ovl, JobView& A)
{
    A.n_ovl = n_ovl;
    A.n_rays = p->n_rays;
    A.nbands = nbands;
    A.max_bounces = p->max_bounces;
    A.dir_band = p->dir_band;
    A.key0 = (uint32_t)p->seed;
    A.key1 = (uint32_t)(p->seed >> 32);
    A.bin_dt = p->bin_dt;
    A.max_time = p->max_time;
    A.radius = p->radius;
    float per_ray = 1.0f / (float)p->n_rays;
    A.floor_e = p->floor_rel * per_ray;
    for (int b = 0; b < kMaxBands; b++) A.energy0[b] = (uint32_t)b < nbands ? per_ray * p->energy[b] : 0.0f;
}

namespace {

// Uploads the job's material table and overlays and fills the job scalars.
int prepare_job(omg_ctx* c, const omg_trace_params* p, const std::vector<MatParams>& base_mats,
                const omg_overlay* ovl, uint32_t n_ovl, uint32_t nbands, TraceArgs& A)
{
    std::vector<MatParams> mats;
    std::vector<Overlay> ov;
    job_tables(base_mats, ovl, n_ovl, nbands, mats, ov);
    std::vector<uint8_t> key(mats.size() * sizeof(MatParams) + ov.size() * sizeof(Overlay) + 8);
    uint32_t sizes[2] = {(uint32_t)mats.size(), (uint32_t)ov.size()};
    std::memcpy(key.data(), sizes, 8);
    if (!mats.empty()) std::memcpy(key.data() + 8, mats.data(), mats.size() * sizeof(MatParams));
    if (!ov.empty())
        std::memcpy(key.data() + 8 + mats.size() * sizeof(MatParams), ov.data(), ov.size() * sizeof(Overlay));
    if (!(c->tables_valid && key == c->tables_key)) {
        int rc = quiesce(c);
        if (!rc) rc = upload(c->job_mats, mats.data(), mats.size() * sizeof(MatParams), c->stream);
        if (!rc) rc = upload(c->job_ovl, ov.data(), ov.size() * sizeof(Overlay), c->stream);
        // the staging vectors die with this frame: wait for the copies to land
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(OMG_ERR_DEVICE, "job upload");
        c->tables_valid = rc == OMG_OK;
        if (rc) return rc;
        c->tables_key.swap(key);
    }
    A.mats = c->job_mats.as<MatParams>();
    A.overlays = c->job_ovl.as<Overlay>();
    job_scalars(p, nbands, n_ovl, A);
    return OMG_OK;
}

// Plain mesh jobs of at least route_rays rays run on the persistent
// regenerating kernel: one launch over the whole GPU whose lanes take new
// rays as theirs end, where the single-job kernel would wait for the
// longest ray of every block.  Same rays, same per-ray records.
bool routed(const omg_ctx* c, bool mesh, uint64_t rays, bool hits)
{
    return mesh && !hits && c->route_rays != 0 && rays >= c->route_rays;
}

int launch_routed(omg_ctx* c, const TraceArgs& A, uint32_t nbands, bool count, hipStream_t st)
{
    // the job view, items and cursor are the context's: wait for async work
    // that may still read them before they are rewritten or regrown
    int rc = quiesce(c);
    if (rc) return rc;
    const uint64_t items = (uint64_t)((A.n_rays + kBatchItemRays - 1) / kBatchItemRays) * A.n_src;
    OMG_HIP_CHECK(c->route_job.ensure(sizeof(JobView)));
    OMG_HIP_CHECK(c->route_items.ensure(items * sizeof(BatchItem)));
    OMG_HIP_CHECK(c->route_cursor.ensure(sizeof(uint32_t)));
    OMG_HIP_CHECK(launch_trace_routed(A, nbands, A.width, c->n_cu, c->route_job.as<JobView>(),
                                      c->route_items.as<BatchItem>(), c->route_cursor.as<uint32_t>(),
                                      count, st));
    return OMG_OK;
}

int run_trace_host(omg_ctx* c, TraceArgs& A, bool mesh, uint32_t nbins, uint32_t nbands,
                   const float* sources, uint32_t n_src, const float* listeners, uint32_t n_lis,
                   float* bins, float* dirs, omg_ray_record* records, omg_counters* counters,
                   uint32_t* hit_ids = nullptr, float* hit_t = nullptr)
{
    size_t total = (size_t)A.n_rays * n_src;
    size_t nb_f = (size_t)n_src * n_lis * nbins * nbands, nd_f = (size_t)n_src * n_lis * nbins * 3;
    size_t nh = total * A.max_bounces;
    int rc = upload(c->job_src, sources, (size_t)n_src * 3 * sizeof(float), c->stream);
    if (!rc) rc = upload(c->job_lis, listeners, (size_t)n_lis * 3 * sizeof(float), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(c->job_bins.ensure(nb_f * sizeof(float)));
    OMG_HIP_CHECK(c->job_dirs.ensure(nd_f * sizeof(float)));
    if (records) OMG_HIP_CHECK(c->job_rec.ensure(total * sizeof(omg_ray_record)));
    if (counters) OMG_HIP_CHECK(c->job_cnt.ensure(3 * sizeof(unsigned long long)));
    if (hit_ids) {
        OMG_HIP_CHECK(c->aux0.ensure(nh * sizeof(uint32_t)));
        OMG_HIP_CHECK(c->aux1.ensure(nh * sizeof(float)));
        A.hit_ids = c->aux0.as<uint32_t>();
        A.hit_t = c->aux1.as<float>();
        // rays that end early leave their later slots as "no hit"
        OMG_HIP_CHECK(hipMemsetD32Async((hipDeviceptr_t)A.hit_ids, 0xFFFFFFFFu, nh, c->stream));
        OMG_HIP_CHECK(launch_fill_zero(A.hit_t, nh, c->stream));
    }
    A.sources = c->job_src.as<float>();
    A.listeners = c->job_lis.as<float>();
    A.n_src = n_src;
    A.n_lis = n_lis;
    A.nbins = nbins;
    A.bins = c->job_bins.as<float>();
    A.dirs = c->job_dirs.as<float>();
    A.records = records ? c->job_rec.as<uint32_t>() : nullptr;
    A.counters = counters ? c->job_cnt.as<unsigned long long>() : nullptr;
    A.ray_base = 0;
    A.ray_count = (uint32_t)total;
    OMG_HIP_CHECK(launch_fill_zero(A.bins, nb_f, c->stream));
    OMG_HIP_CHECK(launch_fill_zero(A.dirs, nd_f, c->stream));
    if (counters) OMG_HIP_CHECK(launch_fill_zero_u64(A.counters, 3, c->stream));
    if (routed(c, mesh, total, hit_ids != nullptr)) {
        rc = launch_routed(c, A, nbands, counters != nullptr, c->stream);
        if (rc) return rc;
    } else {
        rc = setup_handoff(c, A, c->stream);
        if (rc) return rc;
        OMG_HIP_CHECK(launch_trace(A, mesh, counters != nullptr, c->stream));
    }
    OMG_HIP_CHECK(hipMemcpyAsync(bins, A.bins, nb_f * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(dirs, A.dirs, nd_f * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (records)
        OMG_HIP_CHECK(hipMemcpyAsync(records, A.records, total * sizeof(omg_ray_record),
                                     hipMemcpyDeviceToHost, c->stream));
    unsigned long long cnt[3] = {0, 0, 0};
    if (counters)
        OMG_HIP_CHECK(hipMemcpyAsync(cnt, A.counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    if (hit_ids) {
        OMG_HIP_CHECK(hipMemcpyAsync(hit_ids, A.hit_ids, nh * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        OMG_HIP_CHECK(hipMemcpyAsync(hit_t, A.hit_t, nh * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (counters) {
        counters->aabb_tests = cnt[0];
        counters->tri_tests = cnt[1];
        counters->bounce_rays = cnt[2];
    }
    return OMG_OK;
}

}  // namespace

void omg::fill_geom(const omg_scene* s, GeomArgs& g)
{
    g.nodes = s->nodes.as<Node2>();
    g.prims = s->prims.as<Prim>();
    g.tris = s->tris.as<TriInfo>();
    g.nodes4 = s->has4 ? s->nodes4.as<Node4>() : nullptr;
    g.nodes16 = s->has16 ? s->nodes16.as<Node16>() : nullptr;
    g.root_ref = s->root_ref;
    g.root4 = s->root4;
    g.root16 = s->root16;
    g.empty = s->empty ? 1u : 0u;
    g.box[0] = g.box[1] = g.box[2] = 0.0f;
    for (int k = 0; k < 3; k++) {
        g.root_min[k] = s->root_min[k];
        g.root_max[k] = s->root_max[k];
    }
}

extern "C" int omg_trace_world(omg_scene* s, const omg_trace_params* p, const float* sources,
                               uint32_t n_src, const float* listeners, uint32_t n_lis,
                               const omg_overlay* overlays, uint32_t n_ovl, float* bins,
                               float* dirs, omg_ray_record* records, omg_counters* counters)
{
    if (!s || !sources || !listeners || !bins || !dirs || (n_ovl && !overlays))
        return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t nbins;
    int rc = validate_params(p, s->nbands, n_src, n_lis, n_ovl, &nbins);
    if (rc) return rc;
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    TraceArgs A{};
    fill_geom(s, A.g);
    plain_shape(c, s, p, n_src, counters != nullptr, A);
    rc = prepare_job(c, p, s->mat_params, overlays, n_ovl, s->nbands, A);
    if (rc) return rc;
    return run_trace_host(c, A, true, nbins, s->nbands, sources, n_src, listeners, n_lis, bins,
                          dirs, records, counters);
}

extern "C" int omg_trace_world_hits(omg_scene* s, const omg_trace_params* p, const float* sources,
                                    uint32_t n_src, const float* listeners, uint32_t n_lis,
                                    const omg_overlay* overlays, uint32_t n_ovl, float* bins,
                                    float* dirs, omg_ray_record* records, uint32_t* hit_ids,
                                    float* hit_t)
{
    if (!s || !sources || !listeners || !bins || !dirs || !records || !hit_ids || !hit_t ||
        (n_ovl && !overlays))
        return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t nbins;
    int rc = validate_params(p, s->nbands, n_src, n_lis, n_ovl, &nbins);
    if (rc) return rc;
    if ((uint64_t)p->n_rays * n_src * p->max_bounces > (1ull << 28))
        return fail(OMG_ERR_CAP, "per-bounce hit record too large");
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    TraceArgs A{};
    fill_geom(s, A.g);
    plain_shape(c, s, p, n_src, false, A);
    rc = prepare_job(c, p, s->mat_params, overlays, n_ovl, s->nbands, A);
    if (rc) return rc;
    return run_trace_host(c, A, true, nbins, s->nbands, sources, n_src, listeners, n_lis, bins,
                          dirs, records, nullptr, hit_ids, hit_t);
}

extern "C" int omg_trace_box(omg_ctx* c, const float size[3], const omg_material walls[6],
                             uint32_t nbands, const omg_trace_params* p, const float* sources,
                             uint32_t n_src, const float* listeners, uint32_t n_lis, float* bins,
                             float* dirs, omg_ray_record* records, omg_counters* counters)
{
    if (!c || !size || !walls || !sources || !listeners || !bins || !dirs)
        return fail(OMG_ERR_INVALID, "bad argument");
    if (nbands < 1 || nbands > OMG_MAX_BANDS) return fail(OMG_ERR_INVALID, "nbands out of range");
    uint32_t nbins;
    int rc = validate_params(p, nbands, n_src, n_lis, 0, &nbins);
    if (rc) return rc;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    std::vector<MatParams> wm(6);
    for (int w = 0; w < 6; w++)
        mat_params(walls[w].absorption, walls[w].scattering, walls[w].transmission, nbands, wm[w]);
    TraceArgs A{};
    A.g.nodes = nullptr;
    A.g.prims = nullptr;
    A.g.tris = nullptr;
    A.g.root_ref = 0;
    A.g.empty = 0;
    for (int k = 0; k < 3; k++) A.g.box[k] = size[k];
    rc = prepare_job(c, p, wm, nullptr, 0, nbands, A);
    if (rc) return rc;
    return run_trace_host(c, A, false, nbins, nbands, sources, n_src, listeners, n_lis, bins, dirs,
                          records, counters);
}

extern "C" int omg_trace_world_async(omg_scene* s, const omg_trace_params* p,
                                     const float* d_sources, uint32_t n_src,
                                     const float* d_listeners, uint32_t n_lis,
                                     const omg_overlay* overlays, uint32_t n_ovl, float* d_bins,
                                     float* d_dirs, omg_ray_record* d_records, void* stream)
{
    if (!s || !d_sources || !d_listeners || !d_bins || !d_dirs || (n_ovl && !overlays))
        return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t nbins;
    int rc = validate_params(p, s->nbands, n_src, n_lis, n_ovl, &nbins);
    if (rc) return rc;
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = stream ? (hipStream_t)stream : c->stream;
    TraceArgs A{};
    fill_geom(s, A.g);
    plain_shape(c, s, p, n_src, false, A);
    // the job tables are staged on the context stream, then ordered before the launch
    rc = prepare_job(c, p, s->mat_params, overlays, n_ovl, s->nbands, A);
    if (rc) return rc;
    if (st != c->stream) OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    size_t nb_f = (size_t)n_src * n_lis * nbins * s->nbands, nd_f = (size_t)n_src * n_lis * nbins * 3;
    A.sources = d_sources;
    A.listeners = d_listeners;
    A.n_src = n_src;
    A.n_lis = n_lis;
    A.nbins = nbins;
    A.bins = d_bins;
    A.dirs = d_dirs;
    A.records = reinterpret_cast<uint32_t*>(d_records);
    A.counters = nullptr;
    A.ray_base = 0;
    A.ray_count = p->n_rays * n_src;
    if (routed(c, true, (uint64_t)p->n_rays * n_src, false)) {
        rc = quiesce(c);
        if (rc) return rc;
        OMG_HIP_CHECK(launch_fill_zero(d_bins, nb_f, st));
        OMG_HIP_CHECK(launch_fill_zero(d_dirs, nd_f, st));
        rc = launch_routed(c, A, s->nbands, false, st);
        c->async_inflight = true;
        return rc;
    }
    rc = setup_handoff(c, A, st);
    if (rc) return rc;
    OMG_HIP_CHECK(launch_fill_zero(d_bins, nb_f, st));
    OMG_HIP_CHECK(launch_fill_zero(d_dirs, nd_f, st));
    c->async_inflight = true;
    OMG_HIP_CHECK(launch_trace(A, true, false, st));
    return OMG_OK;
}

// ------------------------------------------------------------------ post-processing (host)

extern "C" int omg_estimate_reverb(const float* bins, uint32_t nbins, uint32_t nbands,
                                   float bin_dt, float* rt60, float* level)
{
    if (!bins || !rt60 || !level || nbands == 0 || nbands > OMG_MAX_BANDS || !(bin_dt > 0.0f))
        return fail(OMG_ERR_INVALID, "bad argument");
    std::vector<double> edc(nbins ? nbins : 1);
    auto db = [&](size_t i, double total) {
        double r = edc[i] / total;
        return 10.0 * std::log10(r > 1e-12 ? r : 1e-12);
    };
    for (uint32_t b = 0; b < nbands; b++) {
        rt60[b] = 0.5f;
        double total = 0.0;
        for (int64_t i = (int64_t)nbins - 1; i >= 0; i--) {
            total += (double)bins[(size_t)i * nbands + b];
            edc[i] = total;
        }
        if (total <= 0.0) continue;
        int64_t t5 = -1, t25 = -1;
        for (uint32_t i = 0; i < nbins; i++) {
            double d = db(i, total);
            if (t5 < 0 && d <= -5.0) t5 = i;
            if (t25 < 0 && d <= -25.0) {
                t25 = i;
                break;
            }
        }
        if (t5 >= 0 && t25 > t5) {
            float dt = (float)(t25 - t5) * bin_dt;
            float ddb = (float)(db(t25, total) - db(t5, total));
            rt60[b] = fminf(fmaxf(-60.0f * dt / ddb, 0.1f), 5.0f);
        }
    }
    uint32_t late_start = sat_u32(0.080f / bin_dt);
    for (uint32_t b = 0; b < nbands; b++) {
        float late_e = 0.0f;
        for (uint32_t i = late_start; i < nbins; i++) late_e += bins[(size_t)i * nbands + b];
        level[b] = fminf(sqrtf(late_e), 0.7f);
    }
    return OMG_OK;
}

extern "C" int omg_late_direction(const float* bins, const float* dirs, uint32_t nbins,
                                  uint32_t nbands, float bin_dt, float from_s, float dir_out[3],
                                  float* aniso)
{
    if (!bins || !dirs || !dir_out || !aniso || nbands < 2 || !(bin_dt > 0.0f))
        return fail(OMG_ERR_INVALID, "bad argument");
    uint32_t start = sat_u32(from_s / bin_dt);
    if (start > nbins) start = nbins;
    float v[3] = {0.0f, 0.0f, 0.0f}, e = 0.0f;
    for (uint32_t i = start; i < nbins; i++) {
        for (int k = 0; k < 3; k++) v[k] += dirs[(size_t)i * 3 + k];
        e += bins[(size_t)i * nbands + 1];
    }
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len < 1e-12f || e < 1e-12f) {
        dir_out[0] = 1.0f;
        dir_out[1] = 0.0f;
        dir_out[2] = 0.0f;
        *aniso = 0.0f;
        return OMG_OK;
    }
    dir_out[0] = v[0] / len;
    dir_out[1] = v[1] / len;
    dir_out[2] = v[2] / len;
    *aniso = fminf(len / e, 1.0f);
    return OMG_OK;
}

// ism::image_source_taps (ism.rs:15-84), host-side closed form.
extern "C" int omg_ism_box(const float size[3], const omg_material walls[6], uint32_t nb,
                           const float src[3], const float lis[3], uint32_t max_order,
                           float* taps, uint32_t cap, uint32_t* n_out)
{
    static const float kAir[3] = {1.0e-4f, 5.0e-4f, 3.0e-3f};   // material.rs:75
    if (!size || !walls || !src || !lis || !n_out || nb == 0 || nb > OMG_MAX_BANDS || (cap && !taps))
        return fail(OMG_ERR_INVALID, "bad argument");
    int n_max = ((int)max_order + 1) / 2 + 1;
    float refl[6][OMG_MAX_BANDS];
    for (int w = 0; w < 6; w++)
        for (uint32_t b = 0; b < nb; b++) refl[w][b] = sqrtf(fmaxf(1.0f - walls[w].absorption[b], 0.0f));
    f3 S = mk(src[0], src[1], src[2]), L = mk(lis[0], lis[1], lis[2]);
    uint32_t n = 0;
    const uint32_t row = 5 + nb;
    for (int px = 0; px < 2; px++)
        for (int py = 0; py < 2; py++)
            for (int pz = 0; pz < 2; pz++)
                for (int nx = -n_max; nx <= n_max; nx++)
                    for (int ny = -n_max; ny <= n_max; ny++)
                        for (int nz = -n_max; nz <= n_max; nz++) {
                            uint32_t hits[6] = {(uint32_t)std::abs(nx - px), (uint32_t)std::abs(nx),
                                                (uint32_t)std::abs(ny - py), (uint32_t)std::abs(ny),
                                                (uint32_t)std::abs(nz - pz), (uint32_t)std::abs(nz)};
                            uint32_t order = hits[0] + hits[1] + hits[2] + hits[3] + hits[4] + hits[5];
                            if (order > max_order) continue;
                            f3 img = mk((float)(1 - 2 * px) * S.x + 2.0f * (float)nx * size[0],
                                        (float)(1 - 2 * py) * S.y + 2.0f * (float)ny * size[1],
                                        (float)(1 - 2 * pz) * S.z + 2.0f * (float)nz * size[2]);
                            f3 to_img = sub(img, L);
                            float dist = fmaxf(length(to_img), 0.3f);
                            f3 dir = normalize(to_img);
                            if (n < cap) {
                                float* o = taps + (size_t)n * row;
                                o[0] = (float)n;
                                o[1] = dist / kSpeedOfSound;
                                o[2] = dir.x;
                                o[3] = dir.y;
                                o[4] = dir.z;
                                for (uint32_t b = 0; b < nb; b++) {
                                    float g = expf(-kAir[b < 3 ? b : 2] * dist) / dist;
                                    for (int w = 0; w < 6; w++)
                                        for (uint32_t k = 0; k < hits[w]; k++) g *= refl[w][b];
                                    o[5 + b] = g;
                                }
                            }
                            n++;
                        }
    *n_out = n < cap ? n : cap;
    if (n > cap) return fail(OMG_ERR_CAP, "tap buffer too small");
    return OMG_OK;
}
