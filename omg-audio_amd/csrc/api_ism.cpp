// C ABI: image sources on the BVH -- the surface table, chain discovery
// (WorldDiscovery) and the batched solve (WorldSolve).
#include <cmath>
#include <cstring>
#include <unordered_set>

#include "ism_kernels.h"
#include "runtime.h"

#pragma clang fp contract(off)

using namespace omg;

static_assert(sizeof(omg_surface) == sizeof(SurfaceDev), "surface layout");
static_assert(sizeof(omg_blocker) == sizeof(BlockerDev), "blocker layout");
static_assert(sizeof(omg_mesh_record) == sizeof(MeshRecordDev), "record layout");

// SurfaceTable::build (pt_mesh.rs:51-80): the first triangle of each surface
// id defines its plane and material; ids never used keep the placeholder.
extern "C" int omg_surface_table(omg_scene* s, omg_surface* out, uint32_t cap, uint32_t* n_out)
{
    if (!s || !n_out || (cap && !out)) return fail(OMG_ERR_INVALID, "bad argument");
    if (int rc = ensure_host_mesh(s)) return rc;
    std::vector<omg_surface> tab;
    for (uint32_t t = 0; t < s->n_tris; t++) {
        uint32_t sid = s->tri_surface[t];
        while (tab.size() <= sid) {
            omg_surface z;
            std::memset(&z, 0, sizeof z);
            z.n[2] = 1.0f;
            z.d = FLT_MAX;
            tab.push_back(z);
        }
        omg_surface& su = tab[sid];
        if (su.d == FLT_MAX) {
            const TriInfo& ti = s->tris_host[t];
            f3 n = mk(ti.n[0], ti.n[1], ti.n[2]);
            const float* p = &s->positions[3 * (size_t)s->indices[3 * (size_t)t]];
            const omg_material& m = s->materials[s->tri_material[t]];
            su.n[0] = n.x;
            su.n[1] = n.y;
            su.n[2] = n.z;
            su.d = dot(n, mk(p[0], p[1], p[2]));
            for (uint32_t b = 0; b < s->nbands; b++) {
                su.refl[b] = sqrtf(fmaxf(1.0f - m.absorption[b], 0.0f));
                su.trans[b] = m.transmission[b];
            }
            su.has_rect = 0;
        }
    }
    *n_out = (uint32_t)tab.size();
    if (tab.size() > cap) return fail(OMG_ERR_CAP, "surface buffer too small");
    if (!tab.empty()) std::memcpy(out, tab.data(), tab.size() * sizeof(omg_surface));
    return OMG_OK;
}

// SurfaceTable::append_box (pt_mesh.rs:87-106): the six faces of an overlay
// box as rect-bounded reflectors with the box's material, in the order
// x min, x max, y min, y max, z min, z max (the discovery's face ids).
extern "C" int omg_surface_append_box(const float mn[3], const float mx[3], const omg_material* m,
                                      uint32_t nbands, omg_surface* out6)
{
    if (!mn || !mx || !m || !out6 || nbands < 1 || nbands > OMG_MAX_BANDS)
        return fail(OMG_ERR_INVALID, "bad argument");
    int k = 0;
    for (int axis = 0; axis < 3; axis++) {
        for (int side = 0; side < 2; side++) {
            omg_surface& su = out6[k++];
            std::memset(&su, 0, sizeof su);
            float n[3] = {0.0f, 0.0f, 0.0f};
            n[axis] = side == 0 ? -1.0f : 1.0f;
            const float* corner = side == 0 ? mn : mx;
            for (int a = 0; a < 3; a++) {
                su.n[a] = n[a];
                su.rmin[a] = a == axis ? corner[axis] : mn[a];
                su.rmax[a] = a == axis ? corner[axis] : mx[a];
            }
            su.d = dot(mk(n[0], n[1], n[2]), mk(corner[0], corner[1], corner[2]));
            for (uint32_t b = 0; b < nbands; b++) {
                su.refl[b] = sqrtf(fmaxf(1.0f - m->absorption[b], 0.0f));
                su.trans[b] = m->transmission[b];
            }
            su.has_rect = 1;
        }
    }
    return OMG_OK;
}

extern "C" int omg_discover(omg_scene* s, const float* boxes, uint32_t n_boxes, uint16_t base,
                            const float listener[3], uint32_t n_rays, uint32_t rot,
                            uint16_t* chains, uint8_t* orders, uint32_t cap, uint32_t* n_out)
{
    if (!s || !listener || !n_out || (cap && (!chains || !orders)) || (n_boxes && !boxes))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (n_boxes > OMG_MAX_OVERLAYS) return fail(OMG_ERR_CAP, "more than 64 overlay boxes");
    *n_out = 0;
    if (n_rays == 0) return OMG_OK;
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    int qrc = quiesce(c);   // the context's staging may still be read by *_async work
    if (qrc) return qrc;
    // the rotating fan's jitter is one sequential xorshift64* stream (pt_mesh.rs:386-392)
    std::vector<float> jitter(n_rays);
    uint64_t x = (0xC6Bull ^ (uint64_t)rot) | 1ull;
    for (uint32_t i = 0; i < n_rays; i++) {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        uint64_t v = x * 0x2545F4914F6CDD1DULL;
        jitter[i] = (float)(v >> 40) / (float)(1ull << 24);
    }
    int rc = upload(c->aux0, jitter.data(), n_rays * sizeof(float), c->stream);
    if (!rc && n_boxes) rc = upload(c->aux2, boxes, (size_t)n_boxes * 6 * sizeof(float), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(c->aux1.ensure((size_t)n_rays * 8));
    DiscoverArgs A{};
    fill_geom(s, A.g);
    A.tri_surface = s->tri_surf.as<uint32_t>();
    for (int k = 0; k < 3; k++) A.listener[k] = listener[k];
    A.ga = kPi * (3.0f - sqrtf(5.0f));
    A.rot_phase = (float)rot * 0.61803398875f * kTau;
    A.jitter = c->aux0.as<float>();
    A.n_rays = n_rays;
    A.boxes = n_boxes ? c->aux2.as<float>() : nullptr;
    A.n_boxes = n_boxes;
    A.base = base;
    A.out = c->aux1.as<uint32_t>();
    OMG_HIP_CHECK(launch_discover(A, c->stream));
    std::vector<uint32_t> raw((size_t)n_rays * 2);
    OMG_HIP_CHECK(hipMemcpyAsync(raw.data(), A.out, raw.size() * 4, hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    // first-seen dedup of chain prefixes in ray order (the reference's HashSet)
    std::unordered_set<uint64_t> seen;
    uint32_t n = 0;
    for (uint32_t i = 0; i < n_rays; i++) {
        uint32_t ch[3] = {raw[2 * i] & 0xFFFFu, raw[2 * i] >> 16, raw[2 * i + 1] & 0xFFFFu};
        uint32_t levels = raw[2 * i + 1] >> 16;
        for (uint32_t k = 0; k < levels; k++) {
            uint64_t key = 1;
            for (uint32_t q = 0; q <= k; q++) key = key * 65536u + ch[q] + 1u;
            if (!seen.insert(key).second) continue;
            if (n < cap) {
                for (uint32_t q = 0; q < 3; q++) chains[3 * n + q] = (uint16_t)(q <= k ? ch[q] : 0xFFFFu);
                orders[n] = (uint8_t)(k + 1);
            }
            n++;
        }
    }
    *n_out = n < cap ? n : cap;
    if (n > cap) return fail(OMG_ERR_CAP, "chain buffer too small");
    return OMG_OK;
}

extern "C" int omg_solve_batch(omg_scene* s, const omg_surface* surfaces, uint32_t n_surf,
                               const float* sources, uint32_t n_src, const uint16_t* chains,
                               const uint8_t* orders, uint32_t n_chains, const float listener[3],
                               const omg_blocker* extras, uint32_t n_extras,
                               const float* air_per_m, omg_mesh_record* records)
{
    if (!s || !surfaces || !listener || !records || !air_per_m || (n_src && !sources) ||
        (n_chains && (!chains || !orders)) || (n_extras && !extras))
        return fail(OMG_ERR_INVALID, "bad argument");
    if (n_extras > 256) return fail(OMG_ERR_CAP, "too many extras");   // layout.rs:167
    uint64_t total = (uint64_t)n_src * n_chains;
    if (total == 0) return OMG_OK;
    if (total >= (1ull << 31)) return fail(OMG_ERR_CAP, "too many (source, chain) pairs");
    omg_ctx* c = s->ctx;
    OMG_HIP_CHECK(hipSetDevice(c->device));
    // the extras go to job_ovl, which an in-flight trace may read as overlays
    int rc = quiesce(c);
    if (rc) return rc;
    c->tables_valid = false;
    rc = upload(c->aux0, surfaces, n_surf * sizeof(omg_surface), c->stream);
    if (!rc) rc = upload(c->aux1, sources, (size_t)n_src * 3 * sizeof(float), c->stream);
    if (!rc) rc = upload(c->aux2, chains, (size_t)n_chains * 3 * sizeof(uint16_t), c->stream);
    if (!rc) rc = upload(c->aux3, orders, n_chains, c->stream);
    if (!rc) rc = upload(c->job_ovl, extras, n_extras * sizeof(omg_blocker), c->stream);
    if (rc) return rc;
    OMG_HIP_CHECK(c->aux4.ensure(total * sizeof(MeshRecordDev)));
    OMG_HIP_CHECK(c->job_cnt.ensure(4 * sizeof(unsigned long long)));
    SolveArgs A{};
    fill_geom(s, A.g);
    A.tri_surface = s->tri_surf.as<uint32_t>();
    A.mats = s->mats_raw.as<omg_material_dev>();
    A.surfaces = c->aux0.as<SurfaceDev>();
    A.n_surf = n_surf;
    A.sources = c->aux1.as<float>();
    A.n_src = n_src;
    A.chains = c->aux2.as<uint16_t>();
    A.orders = c->aux3.as<uint8_t>();
    A.n_chains = n_chains;
    for (int k = 0; k < 3; k++) A.listener[k] = listener[k];
    A.extras = c->job_ovl.as<BlockerDev>();
    A.n_extras = n_extras;
    for (uint32_t b = 0; b < kMaxBands; b++) A.air[b] = b < s->nbands ? air_per_m[b] : 0.0f;
    A.nbands = s->nbands;
    A.records = c->aux4.as<MeshRecordDev>();
    unsigned long long* cnt = c->job_cnt.as<unsigned long long>();
    A.status = reinterpret_cast<uint32_t*>(cnt);
    OMG_HIP_CHECK(launch_fill_zero_u64(cnt, 1, c->stream));
    OMG_HIP_CHECK(launch_solve(A, c->stream));
    OMG_HIP_CHECK(hipMemcpyAsync(records, A.records, total * sizeof(MeshRecordDev),
                                 hipMemcpyDeviceToHost, c->stream));
    unsigned long long st = 0;
    OMG_HIP_CHECK(hipMemcpyAsync(&st, cnt, sizeof st, hipMemcpyDeviceToHost, c->stream));
    OMG_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (st & 1ull) return fail(OMG_ERR_OVERFLOW, "a leg crossed too many surfaces");
    return OMG_OK;
}
