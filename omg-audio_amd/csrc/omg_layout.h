// Device-resident data layout of a scene and of one tracer launch.
//
// BVH: the reference's binary SAH tree (mesh.rs:292-405) flattened into
// 64-byte "paired" nodes: an internal node stores BOTH children's boxes and
// references, so one cache-line-aligned 64 B fetch serves the two AABB
// tests the reference performs when it pops an internal node
// (mesh.rs:510-537).  Leaves are not stored as nodes: a child reference
// with bit 31 set encodes the leaf's primitive range directly.
#pragma once
#include <stdint.h>

namespace omg {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafCountShift = 28;          // count (1..4) in bits 28..30
constexpr uint32_t kLeafStartMask = 0x0FFFFFFFu;  // start in bits 0..27
constexpr int kMaxBands = 8;
constexpr int kStackCap = 64;                     // the reference's stack (mesh.rs:492)

inline __host__ __device__ uint32_t leaf_ref(uint32_t start, uint32_t count)
{
    return kLeafBit | (count << kLeafCountShift) | start;
}

struct alignas(64) Node2 {
    float lmin[3];
    uint32_t lref;
    float lmax[3];
    uint32_t rref;
    float rmin[3];
    uint32_t pad0;
    float rmax[3];
    uint32_t pad1;
};
static_assert(sizeof(Node2) == 64, "Node2 is one 64 B line");

// 4-wide node: the same tree collapsed two binary levels at a time, so
// that one 128 B fetch serves four child boxes.  Slots 0-1 hold the binary
// node's left half (its left child's two children, or that child itself
// in slot 0 when it is a leaf, slot 1 unused), slots 2-3 the right half.
// An unused slot has ref == kNoChild.  The fixed halves let the walk visit
// the slots in the binary walk's order (see bvh_device.h).
constexpr uint32_t kNoChild = 0xFFFFFFFFu;

struct alignas(16) Child4 {
    float mn[3];
    uint32_t ref;
    float mx[3];
    uint32_t pad;
};

struct alignas(128) Node4 {
    Child4 c[4];
};
static_assert(sizeof(Node4) == 128, "Node4 is one 128 B line");

// 16-wide node: four binary levels, one slot per lane of a 16-lane row.
// Slot group 4k..4k+3 is slot k of the 4-wide node expanded once more:
// the four slots of the 4-wide node that slot k refers to, or the leaf in
// slot 4k (4k+1..4k+3 unused), or four unused slots.  Hence slots [0, 8)
// are the binary node's left child's subtree, [0, 4) its left child's,
// and so on down to single slots.  512 B = four 128 B lines.
struct alignas(128) Node16 {
    Child4 c[16];
};
static_assert(sizeof(Node16) == 512, "Node16 is four 128 B lines");

// Packed primitive (mesh.rs:43-50): vertex a, edges e1 = b - a, e2 = c - a,
// the ORIGINAL triangle id and its material.  48 B = three 16 B loads.
struct alignas(16) Prim {
    float a[3];
    uint32_t tri;
    float e1[3];
    uint32_t mat;
    float e2[3];
    uint32_t surf;
};
static_assert(sizeof(Prim) == 48, "Prim is 48 B");

// Per original triangle: unit normal of normalize(cross(b-a, c-a))
// (mesh.rs:458-461) and the material index.
struct alignas(16) TriInfo {
    float n[3];
    uint32_t mat;
};

// Per-material tracer constants precomputed on the host with the
// reference's f32 expressions (tracer.rs:148-185): the transmission
// probability p, scattering, the per-band transmitted weight (t*t)/p and
// the reflected weight keep/(1-p).  The kernel multiplies by these, which
// is bit-identical to evaluating the reference's ratios per bounce.
struct MatParams {
    float p;
    float scattering;
    float tw[kMaxBands];
    float rw[kMaxBands];
};

struct Overlay {
    float mn[3];
    float mx[3];
    uint32_t mat;   // index into the launch's material table
    uint32_t pad;
};

}  // namespace omg
