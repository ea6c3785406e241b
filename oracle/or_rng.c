/*
 * Oracle numerics: the reference RNG, the counter-based RNG and the
 * project's sin/cos.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>

/* Rng::next_u64 -- xorshift64* (Vigna 2014), crates/omg-core/src/rng.rs:10-16 */
uint64_t or_xorshift_next(uint64_t *state)
{
    uint64_t x = *state;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *state = x;
    return x * 0x2545F4914F6CDD1DULL;
}

/* Rng::next_f32, rng.rs:19-21: top 24 bits / 2^24 */
float or_xorshift_f32(uint64_t *state)
{
    return (float)(or_xorshift_next(state) >> 40) / (float)(1ULL << 24);
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11).  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl
 * key increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* The project's sin/cos, shared bit-for-bit with the GPU kernels.
 * Method: Cody-Waite argument reduction by pi/2 with a three-part
 * constant (11 + 11 + 24 significant bits), k = floor(x*2/pi + 1/2), and
 * Taylor polynomials of sin (degree 11) and cos (degree 12) on
 * [-pi/4, pi/4] evaluated by Horner in f32 without FMA.  Truncation
 * error is < 1e-11 there, so the result is within a few ulp of the
 * correctly rounded value for |x| < 2^13 (tested against libm). */
void or_sincosf(float x, float *s, float *c)
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
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}
