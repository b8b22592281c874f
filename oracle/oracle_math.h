/*
 * ORACLE -- test infrastructure only. Nothing in the product may include,
 * link or call this code; tests, the smoke check and the benchmark's CPU
 * baseline use it as the checker.
 *
 * Scalar/vector helpers for the C restatement of the reference's Slang path.
 * REAL is float (FP32 "mirror" build, same operation order as the shaders,
 * compiled with -ffp-contract=off) or double (FP64 converged reference).
 * Slang intrinsics are restated with their HLSL definitions:
 * rsqrt(x) = 1/sqrt(x), normalize(v) = v*rsqrt(dot(v,v)), frac(x) = x-floor(x),
 * lerp(a,b,t) = a+(b-a)*t, saturate = clamp(x,0,1),
 * refract(i,n,eta): k = 1-eta^2(1-dot(n,i)^2); k<0 ? 0 : eta*i-(eta*dot(n,i)+sqrt(k))*n.
 */
#ifndef ORACLE_MATH_H
#define ORACLE_MATH_H

#include <math.h>
#include <stdint.h>

#ifndef REAL
#define REAL float
#endif

typedef REAL real;

#if defined(ORACLE_F64)
#define RSQRT_(x) (1.0 / sqrt(x))
#define SQRT_(x) sqrt(x)
#define COS_(x) cos(x)
#define SIN_(x) sin(x)
#define EXP_(x) exp(x)
#define EXP2_(x) exp2(x)
#define LOG_(x) log(x)
#define POW_(a, b) pow(a, b)
#define ACOS_(x) acos(x)
#define ATAN2_(a, b) atan2(a, b)
#define FLOOR_(x) floor(x)
#define FABS_(x) fabs(x)
#else
#define RSQRT_(x) (1.0f / sqrtf(x))
#define SQRT_(x) sqrtf(x)
#define COS_(x) cosf(x)
#define SIN_(x) sinf(x)
#define EXP_(x) expf(x)
#define EXP2_(x) exp2f(x)
#define LOG_(x) logf(x)
#define POW_(a, b) powf(a, b)
#define ACOS_(x) acosf(x)
#define ATAN2_(a, b) atan2f(a, b)
#define FLOOR_(x) floorf(x)
#define FABS_(x) fabsf(x)
#endif

#define R(x) ((real)(x))

typedef struct {
    real x, y;
} v2;
typedef struct {
    real x, y, z;
} v3;
typedef struct {
    real x, y, z, w;
} v4;

static inline real rmin(real a, real b) { return a < b ? a : b; }
static inline real rmax(real a, real b) { return a > b ? a : b; }
static inline real clampr(real x, real lo, real hi) { return rmin(rmax(x, lo), hi); }
static inline real sat(real x) { return clampr(x, R(0), R(1)); }
static inline real fracr(real x) { return x - FLOOR_(x); }
static inline real lerpr(real a, real b, real t) { return a + (b - a) * t; }

static inline v2 mk2(real x, real y) { v2 r = {x, y}; return r; }
static inline v3 mk3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v4 mk4(real x, real y, real z, real w) { v4 r = {x, y, z, w}; return r; }
static inline v3 s3(real s) { return mk3(s, s, s); }
static inline v4 s4(real s) { return mk4(s, s, s, s); }

static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 div3(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl3(v3 a, real s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 divs3(v3 a, real s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline real dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real length3(v3 a) { return SQRT_(dot3(a, a)); }
static inline v3 normalize3(v3 a) { return scl3(a, RSQRT_(dot3(a, a))); }
static inline v3 sat3(v3 a) { return mk3(sat(a.x), sat(a.y), sat(a.z)); }
static inline v3 max3s(v3 a, real s) { return mk3(rmax(a.x, s), rmax(a.y, s), rmax(a.z, s)); }
static inline v3 lerp3(v3 a, v3 b, real t) { return mk3(lerpr(a.x, b.x, t), lerpr(a.y, b.y, t), lerpr(a.z, b.z, t)); }
static inline real maxc3(v3 a) { return rmax(a.x, rmax(a.y, a.z)); }
static inline int any3gt0(v3 a) { return a.x > R(0) || a.y > R(0) || a.z > R(0); }

static inline v4 add4(v4 a, v4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline v4 mul4(v4 a, v4 b) { return mk4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
static inline v4 scl4(v4 a, real s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
static inline v4 divs4(v4 a, real s) { return mk4(a.x / s, a.y / s, a.z / s, a.w / s); }
static inline real dot4one(v4 a) { return a.x + a.y + a.z + a.w; }
static inline real maxc4(v4 a) { return rmax(rmax(a.x, a.y), rmax(a.z, a.w)); }
static inline int any4gt0(v4 a) { return a.x > R(0) || a.y > R(0) || a.z > R(0) || a.w > R(0); }
static inline v4 sat4(v4 a) { return mk4(sat(a.x), sat(a.y), sat(a.z), sat(a.w)); }

#endif
