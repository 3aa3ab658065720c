/*
 * vtmath.h - elementary functions and the per-sample random stream of the
 * voxel path tracer.
 *
 * The same text is compiled by hipcc for the gfx950 kernels and by a plain C
 * compiler for the CPU checker, so that every transcendental the integrator
 * evaluates returns the same bits on both sides (contraction off, IEEE
 * single precision, correctly rounded division and square root).  The
 * includer defines VT_FN first: `__device__ __forceinline__` for HIP code,
 * `static inline` for C.
 *
 * The reference evaluates these through GLSL built-ins (sin, cos, acos,
 * atan, pow in src/shaders/shared/coordinates.h:85-116,
 * src/shaders/shared/sampling.h:2-36, src/shaders/bsdf/microfacet.h:24-68,
 * src/shaders/integrator/pathTracer.fs:294) whose accuracy is left to the
 * driver.  Here each one is a short polynomial written from first
 * principles: Cody-Waite reduction by pi/2 then odd/even polynomials for
 * sin and cos; octant reduction plus an odd polynomial for atan2; the
 * half-angle identity acos(x) = 2 asin(sqrt((1-x)/2)) plus an odd
 * polynomial for asin; the atanh series in t=(m-1)/(m+1) for log2; a
 * degree-6 polynomial on [-1/2,1/2] for exp2; pow = exp2(y log2 x).
 * Coefficients are weighted least-squares fits on Chebyshev nodes to the
 * exact Taylor series of each function.  Accuracy against libm in double is
 * checked by the test suite (a few ulp on the argument ranges used).
 *
 * The random stream replaces the reference's noise-texture walk
 * (src/shaders/shared/random.h:13-27): PCG32 (O'Neill, "PCG: A Family of
 * Simple Fast Space-Efficient Statistically Good Algorithms for Random
 * Number Generation", 2014; pcg32_srandom_r / pcg32_random_r), one stream per
 * (seed, pixel, frame).  Each reference rand() call returns 4 consecutive
 * outputs mapped to (x >> 8) * 2^-24 in [0,1).
 */
#ifndef VT_FN
#error "define VT_FN before including vtmath.h"
#endif

#ifndef VTMATH_H_
#define VTMATH_H_

#include <stdint.h>

#define VT_PI        3.14159265358979323846f
#define VT_PI_F      3.14159265359f      /* constants.h:1 literal, same float */
#define VT_TWO_PI_F  6.28318530718f      /* constants.h:2 */
#define VT_INV_TWOPI_F 0.15915494309f    /* constants.h:3 */

VT_FN float vt_bits2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
VT_FN uint32_t vt_f2bits(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
VT_FN int vt_isnan(float x) { return x != x; }
VT_FN float vt_absf(float x) { return vt_bits2f(vt_f2bits(x) & 0x7fffffffu); }
VT_FN float vt_nanf(void) { return vt_bits2f(0x7fc00000u); }
VT_FN float vt_inff(void) { return vt_bits2f(0x7f800000u); }

/* GLSL min/max/step/sign as the specification defines them (no NaN
 * canonicalisation): min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x,
 * step(edge,x) = x < edge ? 0 : 1. */
VT_FN float vt_minf(float x, float y) { return y < x ? y : x; }
VT_FN float vt_maxf(float x, float y) { return x < y ? y : x; }
VT_FN float vt_stepf(float edge, float x) { return x < edge ? 0.0f : 1.0f; }
VT_FN float vt_signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* ---- sin / cos ---------------------------------------------------------- */
#define VT_PIO2_1 1.5703125f                  /* 11 significant bits */
#define VT_PIO2_2 4.837512969970703125e-4f    /* 11 significant bits */
#define VT_PIO2_3 7.549790126404332e-08f
#define VT_SIN_P0 -1.666666418e-01f
#define VT_SIN_P1  8.332744241e-03f
#define VT_SIN_P2 -1.958730136e-04f
#define VT_COS_P0  4.166666418e-02f
#define VT_COS_P1 -1.388830016e-03f
#define VT_COS_P2  2.454755850e-05f

VT_FN void vt_sincosf(float x, float* s_out, float* c_out)
{
    if (vt_isnan(x) || vt_absf(x) == vt_inff()) {
        *s_out = vt_nanf(); *c_out = vt_nanf(); return;
    }
    const float q = floorf(x * 0.636619747f + 0.5f);
    const int k = (int)q;
    const float r = ((x - q * VT_PIO2_1) - q * VT_PIO2_2) - q * VT_PIO2_3;
    const float z = r * r;
    const float sp = r + (r * z) * (VT_SIN_P0 + z * (VT_SIN_P1 + z * VT_SIN_P2));
    const float cp = (1.0f - 0.5f * z) + (z * z) * (VT_COS_P0 + z * (VT_COS_P1 + z * VT_COS_P2));
    switch (k & 3) {
        case 0:  *s_out = sp;  *c_out = cp;  break;
        case 1:  *s_out = cp;  *c_out = -sp; break;
        case 2:  *s_out = -sp; *c_out = -cp; break;
        default: *s_out = -cp; *c_out = sp;  break;
    }
}
VT_FN float vt_sinf(float x) { float s, c; vt_sincosf(x, &s, &c); return s; }
VT_FN float vt_cosf(float x) { float s, c; vt_sincosf(x, &s, &c); return c; }

/* ---- atan2 -------------------------------------------------------------- */
#define VT_ATAN_P0 -3.333328664e-01f
#define VT_ATAN_P1  1.999116987e-01f
#define VT_ATAN_P2 -1.402309537e-01f
#define VT_ATAN_P3  8.516459167e-02f

VT_FN float vt_atan2f(float y, float x)
{
    if (vt_isnan(x) || vt_isnan(y)) return x + y;
    const float ax = vt_absf(x), ay = vt_absf(y);
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    float r;
    if (mx == 0.0f) {
        r = 0.0f;
    } else if (mx == vt_inff()) {
        r = (mn == vt_inff()) ? 0.785398163f : 0.0f;
    } else {
        float a = mn / mx;
        float off = 0.0f;
        if (a > 0.414213562f) { a = (a - 1.0f) / (a + 1.0f); off = 0.785398163f; }
        const float z = a * a;
        r = off + (a + (a * z) * (VT_ATAN_P0 + z * (VT_ATAN_P1 + z * (VT_ATAN_P2 + z * VT_ATAN_P3))));
    }
    if (ay > ax) r = 1.57079633f - r;
    if (x < 0.0f) r = VT_PI - r;
    return y < 0.0f ? -r : r;
}

/* ---- acos --------------------------------------------------------------- */
#define VT_ASIN_P0 1.666667312e-01f
#define VT_ASIN_P1 7.498858869e-02f
#define VT_ASIN_P2 4.500060156e-02f
#define VT_ASIN_P3 2.655955590e-02f
#define VT_ASIN_P4 3.807495162e-02f

VT_FN float vt_asin_poly(float a, float z)
{
    return a + (a * z) * (VT_ASIN_P0 + z * (VT_ASIN_P1 + z * (VT_ASIN_P2 + z * (VT_ASIN_P3 + z * VT_ASIN_P4))));
}

VT_FN float vt_acosf(float x)
{
    if (vt_isnan(x) || vt_absf(x) > 1.0f) return vt_nanf();
    const float ax = vt_absf(x);
    if (ax <= 0.5f) {
        return 1.57079633f - vt_asin_poly(x, x * x);
    }
    const float z = (1.0f - ax) * 0.5f;
    const float s = sqrtf(z);
    const float r = 2.0f * vt_asin_poly(s, z);
    return x < 0.0f ? VT_PI - r : r;
}

/* ---- log2 / exp2 / pow --------------------------------------------------- */
#define VT_LOG_P0 6.666666865e-01f
#define VT_LOG_P1 4.000012279e-01f
#define VT_LOG_P2 2.855083644e-01f
#define VT_LOG_P3 2.333009392e-01f
#define VT_EXP2_P0 1.000000000e+00f
#define VT_EXP2_P1 6.931471825e-01f
#define VT_EXP2_P2 2.402264625e-01f
#define VT_EXP2_P3 5.550329015e-02f
#define VT_EXP2_P4 9.618519805e-03f
#define VT_EXP2_P5 1.339985989e-03f
#define VT_EXP2_P6 1.533757750e-04f

VT_FN float vt_log2f(float x)
{
    if (vt_isnan(x) || x < 0.0f) return vt_nanf();
    if (x == 0.0f) return -vt_inff();
    if (x == vt_inff()) return x;
    int e = 0;
    uint32_t b = vt_f2bits(x);
    if (b < 0x00800000u) { x = x * 8388608.0f; b = vt_f2bits(x); e = -23; }
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = vt_bits2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float z = t * t;
    const float ln = 2.0f * t + (t * z) * (VT_LOG_P0 + z * (VT_LOG_P1 + z * (VT_LOG_P2 + z * VT_LOG_P3)));
    return (float)e + ln * 1.44269502f;
}

VT_FN float vt_exp2f(float x)
{
    if (vt_isnan(x)) return x;
    if (x >= 128.0f) return vt_inff();
    if (x < -150.0f) return 0.0f;
    const float n = floorf(x + 0.5f);
    const float f = x - n;
    const float p = VT_EXP2_P0 + f * (VT_EXP2_P1 + f * (VT_EXP2_P2 + f * (VT_EXP2_P3 +
                    f * (VT_EXP2_P4 + f * (VT_EXP2_P5 + f * VT_EXP2_P6)))));
    const int ni = (int)n;
    const int n1 = ni / 2;
    const int n2 = ni - n1;
    return (p * vt_bits2f((uint32_t)(n1 + 127) << 23)) * vt_bits2f((uint32_t)(n2 + 127) << 23);
}

/* pow as GLSL leaves it: NaN for x < 0 (the reference relies on this for
 * its negative-cosine paths), exp2(y log2 x) otherwise. */
VT_FN float vt_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (vt_isnan(x) || vt_isnan(y)) return vt_nanf();
    if (x < 0.0f) return vt_nanf();
    if (x == 0.0f) return y > 0.0f ? 0.0f : vt_inff();
    return vt_exp2f(y * vt_log2f(x));
}

/* ---- PCG32 -------------------------------------------------------------- */
typedef struct vt_pcg32 { uint64_t state; uint64_t inc; } vt_pcg32;

VT_FN uint32_t vt_pcg32_next(vt_pcg32* r)
{
    const uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

VT_FN void vt_pcg32_seed(vt_pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    vt_pcg32_next(r);
    r->state += initstate;
    vt_pcg32_next(r);
}

VT_FN uint64_t vt_splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* One stream per (seed, pixel, frame): pixel = y*W + x with y counted from
 * the bottom row (gl_FragCoord convention), frame = the reference's
 * sampleCount uniform (src/renderer/renderer.cpp:594). */
VT_FN void vt_rng_init(vt_pcg32* r, uint64_t seed, uint32_t pixel, uint32_t frame)
{
    vt_pcg32_seed(r, vt_splitmix64(seed ^ ((uint64_t)frame << 32) ^ (uint64_t)pixel), (uint64_t)pixel);
}

VT_FN float vt_u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-08f; }

/* The reference rand(): four consecutive draws. */
VT_FN void vt_rand4(vt_pcg32* r, float* u)
{
    u[0] = vt_u01(vt_pcg32_next(r));
    u[1] = vt_u01(vt_pcg32_next(r));
    u[2] = vt_u01(vt_pcg32_next(r));
    u[3] = vt_u01(vt_pcg32_next(r));
}

#endif /* VTMATH_H_ */
