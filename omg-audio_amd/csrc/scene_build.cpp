// Host-side scene construction: MeshBuilder vertex dedup, tessellation of
// oversized triangles, 16-bin SAH BVH, and the flatten into the paired
// 64 B node layout resident in HBM.
//
// The primitive set and the tree must be the reference's exactly (same
// patches, same partition, same child order), so the f32 expressions and
// loop orders below follow crates/omg-core/src/mesh.rs:108-405; the
// container choices and the flatten are this library's own.
#include "scene_build.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>

#pragma clang fp contract(off)

namespace omg {

namespace {

struct QKey {
    int64_t x, y, z;
    bool operator==(const QKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct QHash {
    size_t operator()(const QKey& k) const
    {
        uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k.y + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.z + 0x8CB92BA72F3D8DD7ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// (v * 1e4).round() as i64, saturating (mesh.rs:109)
int64_t quant(float v)
{
    float r = roundf(v * 1e4f);
    if (r != r) return 0;
    if (r >= 9223372036854775808.0f) return INT64_MAX;
    if (r <= -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)r;
}

struct BNode {
    f3 bmin, bmax;
    int32_t left = -1, right = -1;
    uint32_t start = 0, count = 0;
};

struct Builder {
    std::vector<Prim>* prims;
    std::vector<BNode> nodes;
    uint32_t max_depth = 0;

    static void bounds(const Prim& p, f3& lo, f3& hi)
    {
        f3 a = mk(p.a[0], p.a[1], p.a[2]);
        f3 b = add(a, mk(p.e1[0], p.e1[1], p.e1[2]));
        f3 c = add(a, mk(p.e2[0], p.e2[1], p.e2[2]));
        lo = vmin(vmin(a, b), c);
        hi = vmax(vmax(a, b), c);
    }
    static f3 centroid(const Prim& p)
    {
        f3 a = mk(p.a[0], p.a[1], p.a[2]);
        f3 e = add(mk(p.e1[0], p.e1[1], p.e1[2]), mk(p.e2[0], p.e2[1], p.e2[2]));
        return add(a, mul(e, 1.0f / 3.0f));
    }
    static float area(f3 lo, f3 hi)
    {
        f3 e = vmax(sub(hi, lo), mk(0.0f, 0.0f, 0.0f));
        return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
    static uint32_t bin_of(const Prim& p, int axis, f3 cmin, float cwidth)
    {
        float f = (get(centroid(p), axis) - get(cmin, axis)) / cwidth * 16.0f;
        return sat_u32(fminf(f, 15.0f));
    }

    int32_t build(size_t start, size_t end, uint32_t depth)
    {
        std::vector<Prim>& P = *prims;
        f3 bmin = mk(FLT_MAX, FLT_MAX, FLT_MAX), bmax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        f3 cmin = bmin, cmax = bmax;
        for (size_t i = start; i < end; i++) {
            f3 lo, hi;
            bounds(P[i], lo, hi);
            bmin = vmin(bmin, lo);
            bmax = vmax(bmax, hi);
            f3 c = centroid(P[i]);
            cmin = vmin(cmin, c);
            cmax = vmax(cmax, c);
        }
        int32_t idx = (int32_t)nodes.size();
        nodes.push_back(BNode{});
        nodes[idx].bmin = bmin;
        nodes[idx].bmax = bmax;
        max_depth = std::max(max_depth, depth);
        size_t count = end - start;
        if (count <= 4) {
            nodes[idx].start = (uint32_t)start;
            nodes[idx].count = (uint32_t)count;
            return idx;
        }
        f3 cext = sub(cmax, cmin);
        int axis = (cext.x >= cext.y && cext.x >= cext.z) ? 0 : (cext.y >= cext.z ? 1 : 2);
        float cwidth = get(cext, axis);
        size_t mid;
        if (cwidth < 1e-6f) {
            mid = start + count / 2;
        } else {
            f3 bnmin[16], bnmax[16];
            uint32_t bnn[16];
            for (int k = 0; k < 16; k++) {
                bnmin[k] = mk(FLT_MAX, FLT_MAX, FLT_MAX);
                bnmax[k] = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
                bnn[k] = 0;
            }
            for (size_t i = start; i < end; i++) {
                uint32_t k = bin_of(P[i], axis, cmin, cwidth);
                f3 lo, hi;
                bounds(P[i], lo, hi);
                bnmin[k] = vmin(bnmin[k], lo);
                bnmax[k] = vmax(bnmax[k], hi);
                bnn[k]++;
            }
            float right_cost[16] = {0};
            f3 amin = mk(FLT_MAX, FLT_MAX, FLT_MAX), amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            uint32_t an = 0;
            for (int k = 15; k >= 1; k--) {
                amin = vmin(amin, bnmin[k]);
                amax = vmax(amax, bnmax[k]);
                an += bnn[k];
                right_cost[k - 1] = an > 0 ? area(amin, amax) * (float)an : 0.0f;
            }
            int best_cut = -1;
            float best_cost = FLT_MAX;
            amin = mk(FLT_MAX, FLT_MAX, FLT_MAX);
            amax = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
            an = 0;
            for (int k = 0; k < 15; k++) {
                amin = vmin(amin, bnmin[k]);
                amax = vmax(amax, bnmax[k]);
                an += bnn[k];
                if (an == 0 || an == (uint32_t)count) continue;
                float cost = area(amin, amax) * (float)an + right_cost[k];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_cut = k;
                }
            }
            if (best_cut < 0) {
                mid = start + count / 2;
            } else {
                // two-pointer partition, swapping from the back (mesh.rs:384-395)
                size_t i = start, j = end;
                while (i < j) {
                    if (bin_of(P[i], axis, cmin, cwidth) <= (uint32_t)best_cut) {
                        i++;
                    } else {
                        j--;
                        std::swap(P[i], P[j]);
                    }
                }
                mid = i;
            }
        }
        mid = std::min(std::max(mid, start + 1), end - 1);
        int32_t l = build(start, mid, depth + 1);
        int32_t r = build(mid, end, depth + 1);
        nodes[idx].left = l;
        nodes[idx].right = r;
        return idx;
    }
};

}  // namespace

uint32_t dedup_soup(const float* tv, uint32_t nt, std::vector<float>& pos,
                    std::vector<uint32_t>& idx)
{
    std::unordered_map<QKey, uint32_t, QHash> lookup;
    lookup.reserve((size_t)nt * 2);
    pos.clear();
    idx.resize((size_t)nt * 3);
    for (size_t i = 0; i < (size_t)nt * 3; i++) {
        const float* p = tv + 3 * i;
        QKey k{quant(p[0]), quant(p[1]), quant(p[2])};
        auto it = lookup.find(k);
        if (it == lookup.end()) {
            uint32_t id = (uint32_t)(pos.size() / 3);
            lookup.emplace(k, id);
            pos.insert(pos.end(), p, p + 3);
            idx[i] = id;
        } else {
            idx[i] = it->second;
        }
    }
    return (uint32_t)(pos.size() / 3);
}

void mat_params(const float* absorption, float scattering, const float* transmission,
                uint32_t nbands, MatParams& out)
{
    std::memset(&out, 0, sizeof out);
    float p_t = 0.0f;
    for (uint32_t b = 0; b < nbands; b++) p_t = fmaxf(p_t, transmission[b] * transmission[b]);
    float p = p_t > 1e-5f ? fminf(fmaxf(p_t, 0.02f), 0.5f) : 0.0f;
    out.p = p;
    out.scattering = scattering;
    for (uint32_t b = 0; b < nbands; b++) {
        float tt = transmission[b] * transmission[b];
        out.tw[b] = p > 0.0f ? tt / p : 0.0f;
        float keep = fmaxf(1.0f - absorption[b] - tt, 0.0f);
        out.rw[b] = keep / (1.0f - p);
    }
}

void collapse4(HostScene& S);

int build_scene(const float* positions, uint32_t nv, const uint32_t* indices, uint32_t nt,
                const uint16_t* tri_material, const uint16_t* tri_surface,
                const float* tri_max_edge, HostScene& S, std::string& err)
{
    S = HostScene{};
    for (size_t i = 0; i < (size_t)nt * 3; i++) {
        if (indices[i] >= nv) {
            err = "triangle index out of range";
            return -1;
        }
    }
    auto P = [&](uint32_t v) { return mk(positions[3 * v], positions[3 * v + 1], positions[3 * v + 2]); };
    S.tris.resize(nt);
    for (uint32_t t = 0; t < nt; t++) {
        f3 a = P(indices[3 * t]), b = P(indices[3 * t + 1]), c = P(indices[3 * t + 2]);
        f3 n = normalize(cross(sub(b, a), sub(c, a)));
        S.tris[t] = TriInfo{{n.x, n.y, n.z}, tri_material[t]};
    }
    // tessellation: LIFO work list, longest-edge midpoint split (mesh.rs:244-279)
    struct Work {
        f3 a, b, c;
        uint32_t t;
    };
    std::vector<Work> work;
    work.reserve((size_t)nt + 64);
    for (uint32_t t = 0; t < nt; t++)
        work.push_back(Work{P(indices[3 * t]), P(indices[3 * t + 1]), P(indices[3 * t + 2]), t});
    std::vector<Prim>& prims = S.prims;
    prims.reserve((size_t)nt * 2);
    while (!work.empty()) {
        Work w = work.back();
        work.pop_back();
        float lab = length(sub(w.b, w.a)), lbc = length(sub(w.c, w.b)), lca = length(sub(w.a, w.c));
        float longest = fmaxf(fmaxf(lab, lbc), lca);
        float lim = tri_max_edge ? tri_max_edge[w.t] : 4.0f;
        if (longest > lim) {
            if (lab >= lbc && lab >= lca) {
                f3 m = mul(add(w.a, w.b), 0.5f);
                work.push_back(Work{w.a, m, w.c, w.t});
                work.push_back(Work{m, w.b, w.c, w.t});
            } else if (lbc >= lca) {
                f3 m = mul(add(w.b, w.c), 0.5f);
                work.push_back(Work{w.a, w.b, m, w.t});
                work.push_back(Work{w.a, m, w.c, w.t});
            } else {
                f3 m = mul(add(w.c, w.a), 0.5f);
                work.push_back(Work{w.a, w.b, m, w.t});
                work.push_back(Work{m, w.b, w.c, w.t});
            }
            continue;
        }
        f3 e1 = sub(w.b, w.a), e2 = sub(w.c, w.a);
        Prim pr;
        pr.a[0] = w.a.x; pr.a[1] = w.a.y; pr.a[2] = w.a.z;
        pr.e1[0] = e1.x; pr.e1[1] = e1.y; pr.e1[2] = e1.z;
        pr.e2[0] = e2.x; pr.e2[1] = e2.y; pr.e2[2] = e2.z;
        pr.tri = w.t;
        pr.mat = tri_material[w.t];
        pr.surf = tri_surface ? tri_surface[w.t] : 0u;
        prims.push_back(pr);
    }
    if (prims.size() >= (size_t)kLeafStartMask) {
        err = "too many primitives for the leaf encoding";
        return -1;
    }
    S.root_ref = 0;
    S.max_depth = 0;
    if (prims.empty()) {
        S.empty = true;
        S.has4 = true;   // nothing to walk at either width
        return 0;
    }
    Builder B;
    B.prims = &prims;
    B.nodes.reserve(prims.size() / 2 + 16);
    B.build(0, prims.size(), 1);
    S.max_depth = B.max_depth;
    S.n_bin_nodes = (uint32_t)B.nodes.size();
    // flatten: internal nodes in breadth-first order become Node2 entries, so
    // the first entries are the top levels every traversal crosses (the
    // kernels keep a prefix of them in LDS)
    std::vector<int32_t> remap(B.nodes.size(), -1);
    uint32_t n_int = 0;
    {
        std::vector<int32_t> queue;
        queue.reserve(B.nodes.size());
        if (B.nodes[0].left >= 0) queue.push_back(0);
        for (size_t q = 0; q < queue.size(); q++) {
            int32_t bi = queue[q];
            remap[bi] = (int32_t)n_int++;
            for (int32_t c : {B.nodes[bi].left, B.nodes[bi].right})
                if (B.nodes[c].left >= 0) queue.push_back(c);
        }
    }
    auto ref_of = [&](int32_t bi) -> uint32_t {
        const BNode& n = B.nodes[bi];
        if (n.left < 0) return leaf_ref(n.start, n.count);
        return (uint32_t)remap[bi];
    };
    S.nodes.resize(n_int);
    for (size_t i = 0; i < B.nodes.size(); i++) {
        const BNode& n = B.nodes[i];
        if (n.left < 0) continue;
        Node2& o = S.nodes[remap[i]];
        std::memset(&o, 0, sizeof o);
        const BNode& l = B.nodes[n.left];
        const BNode& r = B.nodes[n.right];
        o.lmin[0] = l.bmin.x; o.lmin[1] = l.bmin.y; o.lmin[2] = l.bmin.z;
        o.lmax[0] = l.bmax.x; o.lmax[1] = l.bmax.y; o.lmax[2] = l.bmax.z;
        o.rmin[0] = r.bmin.x; o.rmin[1] = r.bmin.y; o.rmin[2] = r.bmin.z;
        o.rmax[0] = r.bmax.x; o.rmax[1] = r.bmax.y; o.rmax[2] = r.bmax.z;
        o.lref = ref_of(n.left);
        o.rref = ref_of(n.right);
    }
    S.root_ref = ref_of(0);
    S.root_min[0] = B.nodes[0].bmin.x; S.root_min[1] = B.nodes[0].bmin.y; S.root_min[2] = B.nodes[0].bmin.z;
    S.root_max[0] = B.nodes[0].bmax.x; S.root_max[1] = B.nodes[0].bmax.y; S.root_max[2] = B.nodes[0].bmax.z;
    collapse4(S);
    return 0;
}

// Collapses the paired binary nodes into 4-wide nodes, breadth-first, two
// binary levels per node: slots 0-1 hold the binary node's left child (its
// two children when it is internal, else the leaf in slot 0 and slot 1
// unused), slots 2-3 its right child likewise.  The fixed halves let the
// walk rebuild the binary walk's visiting order from the slots alone.
// stack4 is the worst-case traversal stack: the largest sum of
// (children - 1) along a root-to-leaf path.
void collapse4(HostScene& S)
{
    S.nodes4.clear();
    S.root4 = S.root_ref;
    S.stack4 = 0;
    S.has4 = true;
    if (S.empty || (S.root_ref & kLeafBit)) return;
    struct Kid {
        float mn[3], mx[3];
        uint32_t ref;
    };
    auto left = [](const Node2& n) {
        return Kid{{n.lmin[0], n.lmin[1], n.lmin[2]}, {n.lmax[0], n.lmax[1], n.lmax[2]}, n.lref};
    };
    auto right = [](const Node2& n) {
        return Kid{{n.rmin[0], n.rmin[1], n.rmin[2]}, {n.rmax[0], n.rmax[1], n.rmax[2]}, n.rref};
    };
    const Kid none{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, kNoChild};
    std::vector<uint32_t> queue{S.root_ref};  // binary node of each 4-wide node
    S.nodes4.reserve(S.nodes.size() / 2 + 1);
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const Node2& n = S.nodes[queue[qi]];
        Kid k[4];
        const Kid half[2] = {left(n), right(n)};
        for (int h = 0; h < 2; h++) {
            if (half[h].ref & kLeafBit) {
                k[2 * h] = half[h];
                k[2 * h + 1] = none;
            } else {
                const Node2& m = S.nodes[half[h].ref];
                k[2 * h] = left(m);
                k[2 * h + 1] = right(m);
            }
        }
        Node4 o;
        std::memset(&o, 0, sizeof o);
        for (int j = 0; j < 4; j++) {
            Child4& c = o.c[j];
            for (int a = 0; a < 3; a++) {
                c.mn[a] = k[j].mn[a];
                c.mx[a] = k[j].mx[a];
            }
            if (k[j].ref == kNoChild || (k[j].ref & kLeafBit)) {
                c.ref = k[j].ref;
            } else {
                c.ref = (uint32_t)queue.size();
                queue.push_back(k[j].ref);
            }
        }
        S.nodes4.push_back(o);
    }
    S.root4 = 0;
    std::vector<uint32_t> need(S.nodes4.size(), 0);
    for (size_t i = S.nodes4.size(); i-- > 0;) {
        uint32_t nk = 0, deeper = 0;
        for (const Child4& c : S.nodes4[i].c) {
            if (c.ref == kNoChild) continue;
            nk++;
            if (!(c.ref & kLeafBit)) deeper = std::max(deeper, need[c.ref]);
        }
        need[i] = (nk - 1) + deeper;
    }
    S.stack4 = need[0];
    S.has4 = S.stack4 + 1 <= (uint32_t)kStackCap;
}

}  // namespace omg
