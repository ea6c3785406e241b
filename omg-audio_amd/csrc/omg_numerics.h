// Host/device numerics shared by the HIP kernels and the host runtime.
//
// Everything here must produce the same f32 bits on the CPU and on gfx950:
// no FMA contraction (the whole library is built with -ffp-contract=off),
// IEEE division and square root, fmaxf/fminf with IEEE maxNum semantics
// (Rust's f32::max/min), and the evaluation order of the reference's Vec3
// (crates/omg-core/src/vec3.rs:13-36).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <math.h>

#pragma clang fp contract(off)

#define OMG_HD __host__ __device__ __forceinline__

namespace omg {

struct f3 {
    float x, y, z;
};

OMG_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
OMG_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
OMG_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
OMG_HD f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
OMG_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OMG_HD float length(f3 a) { return sqrtf(dot(a, a)); }
OMG_HD f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize multiplies by the reciprocal of the length (vec3.rs:21-28)
OMG_HD f3 normalize(f3 a)
{
    float l = length(a);
    if (l > 1e-12f) return mul(a, 1.0f / l);
    return mk(1.0f, 0.0f, 0.0f);
}
OMG_HD f3 vmin(f3 a, f3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
OMG_HD f3 vmax(f3 a, f3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
OMG_HD float get(f3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }

constexpr float kTau = 6.28318548202514648438f;   // core::f32::consts::TAU
constexpr float kPi = 3.14159274101257324219f;    // core::f32::consts::PI
constexpr float kSpeedOfSound = 343.0f;           // crates/omg-core/src/lib.rs:18

// Rust `f as usize` (truncate, NaN and negatives to 0), saturated at 2^32-1:
// callers only compare against bin counts far below that.
OMG_HD uint32_t sat_u32(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
OMG_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform in [0, 1) from the top 24 bits of a word (the reference's
// next_f32 keeps the top 24 bits of its u64, rng.rs:19-21).
OMG_HD float u24(uint32_t w) { return (float)(w >> 8) / 16777216.0f; }

// Project sin/cos: Cody-Waite reduction by pi/2 (three-part constant) and
// Taylor polynomials of degree 11 (sin) and 12 (cos) on [-pi/4, pi/4],
// Horner in f32 without FMA.  Bit-identical on host and device.
OMG_HD void sincos_cw(float x, float* s, float* c)
{
    const float two_over_pi = 0x1.45f306p-1f;
    const float p1 = 0x1.92p+0f;
    const float p2 = 0x1.fb4p-12f;
    const float p3 = 0x1.4442d2p-24f;
    float kf = floorf(x * two_over_pi + 0.5f);
    float r = x - kf * p1;
    r = r - kf * p2;
    r = r - kf * p3;
    float r2 = r * r;
    float ps = -0x1.ae6456p-26f;
    ps = ps * r2 + 0x1.71de3ap-19f;
    ps = ps * r2 + -0x1.a01a02p-13f;
    ps = ps * r2 + 0x1.111112p-7f;
    ps = ps * r2 + -0x1.555556p-3f;
    float sr = r + r * (r2 * ps);
    float pc = 0x1.1eed8ep-29f;
    pc = pc * r2 + -0x1.27e4fcp-22f;
    pc = pc * r2 + 0x1.a01a02p-16f;
    pc = pc * r2 + -0x1.6c16c2p-10f;
    pc = pc * r2 + 0x1.555556p-5f;
    float cr = 1.0f - 0.5f * r2 + (r2 * r2) * pc;
    int q = ((int)kf) & 3;
    float so = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    float co = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
    *s = so;
    *c = co;
}

// Rng::unit_sphere (rng.rs:24-29) from two uniforms.
OMG_HD f3 unit_sphere(float u1, float u2)
{
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    float phi = kTau * u2;
    float s, c;
    sincos_cw(phi, &s, &c);
    return mk(r * c, r * s, z);
}

// Kurze-Anderson knife-edge amplitude (crates/omg-core/src/diffraction.rs:27-42)
// for a path detour in metres at one frequency.  tanh, tan, log10 and pow
// are evaluated in f64 and rounded once to f32, on the host and on the
// device alike (correctly rounded but for double-rounding cases), instead
// of the platform's f32 functions.
// band centre frequencies of the knife-edge kernel (diffraction.rs:15)
OMG_HD float band_hz(int b) { return b == 0 ? 125.0f : (b == 1 ? 790.0f : 5600.0f); }

OMG_HD float knife_edge_amp(float detour_m, float freq_hz)
{
    float n = 2.0f * detour_m * freq_hz / kSpeedOfSound;
    if (n <= -0.1916f) return 1.0f;
    float x = sqrtf(2.0f * kPi * fabsf(n));
    float ratio;
    if (x < 1e-3f)
        ratio = 1.0f;
    else if (n >= 0.0f)
        ratio = x / (float)tanh((double)x);
    else
        ratio = x / (float)tan((double)x);
    float a_db = 5.0f + 20.0f * (float)log10((double)ratio);
    return (float)pow(10.0, (double)(-a_db / 20.0f));
}

OMG_HD uint32_t fnv1a_u32(uint32_t h, uint32_t v)
{
    h = (h ^ (v & 0xFFu)) * 16777619u;
    h = (h ^ ((v >> 8) & 0xFFu)) * 16777619u;
    h = (h ^ ((v >> 16) & 0xFFu)) * 16777619u;
    h = (h ^ (v >> 24)) * 16777619u;
    return h;
}

}  // namespace omg
