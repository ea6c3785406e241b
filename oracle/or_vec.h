/*
 * Vec3 with the reference's evaluation order (crates/omg-core/src/vec3.rs:1-82).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef OR_VEC_H
#define OR_VEC_H
#include <math.h>
#include <float.h>
#include <stdint.h>

typedef struct {
    float x, y, z;
} v3;

static inline v3 V(float x, float y, float z)
{
    v3 r = {x, y, z};
    return r;
}
static inline v3 vload(const float *p) { return V(p[0], p[1], p[2]); }
static inline void vstore(float *p, v3 a)
{
    p[0] = a.x;
    p[1] = a.y;
    p[2] = a.z;
}
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
/* vec3.rs:13-15 */
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
/* vec3.rs:21-28: multiply by the reciprocal, never divide */
static inline v3 vnorm(v3 a)
{
    float l = vlen(a);
    if (l > 1e-12f)
        return vmul(a, 1.0f / l);
    return V(1.0f, 0.0f, 0.0f);
}
/* vec3.rs:30-36 */
static inline v3 vcross(v3 a, v3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 vmin(v3 a, v3 b) { return V(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return V(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float vget(v3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
static inline void vset(v3 *a, int axis, float v)
{
    if (axis == 0)
        a->x = v;
    else if (axis == 1)
        a->y = v;
    else
        a->z = v;
}

/* Rust `f as usize` for f32: truncation, NaN -> 0, saturating. */
static inline uint64_t sat_usize(float f)
{
    if (!(f > 0.0f))
        return 0;
    if (f >= 18446744073709551616.0f)
        return UINT64_MAX;
    return (uint64_t)f;
}

#define OR_TAU 6.28318548202514648438f  /* core::f32::consts::TAU */
#define OR_PI 3.14159274101257324219f   /* core::f32::consts::PI */
#define OR_SPEED_OF_SOUND 343.0f        /* lib.rs:18 */

#endif
