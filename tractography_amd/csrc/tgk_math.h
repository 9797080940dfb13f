/*
 * tgk_math.h -- the FP32 elementary-function contract of the tracking engine.
 *
 * The tracking kernels and their host restatement must round identically, so
 * neither may call a platform libm for sin/cos/atan2/acos/exp/log: the GPU
 * device library and glibc differ in the last bit. Every function here is
 * written with IEEE single-precision +, -, *, /, sqrt, fma, floor and rint only,
 * each of which is correctly rounded on both x86-64 (SSE) and gfx950, so the
 * same source yields bit-identical results on the host and on the device.
 *
 * Method: Cody & Waite argument reduction (Software Manual for the Elementary
 * Functions, 1980) to a short interval, then a polynomial whose float32
 * coefficients were fitted by fit_math_poly.py (Lawson-iterated weighted least
 * squares in the relative-error norm, coefficients rounded lowest order first).
 * Accuracy against float64 (asserted by the test suite): sin and cos within
 * 1.5 ulp on |x| <= 1e4, atan2 within 2.1 ulp, acos, exp and log within
 * 1.1 ulp on their full domains; sqrt, divide and the IEEE special values
 * (inf, nan, signed zeros) are exact. Beyond |x| = 2^20 sin/cos stay finite
 * and deterministic but lose accuracy.
 *
 * Compile every translation unit that includes this header with
 * -ffp-contract=off: the functions spell out each fma they use and must not
 * gain others from the compiler.
 *
 * The header is valid C11 (host restatement) and HIP C++ (host and device).
 */
#ifndef TGK_MATH_H
#define TGK_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define TGK_FN static inline __host__ __device__
#else
#define TGK_FN static inline
#endif

#define TGK_INFINITY_BITS 0x7f800000u

/* Split constants: hi is the float nearest the value, lo the float nearest
 * the remainder. Produced by fit_math_poly.split_constant. */
#define TGK_TWO_OVER_PI 6.366197467e-01f
#define TGK_PIO2_1 1.570796371e+00f
#define TGK_PIO2_2 -4.371138829e-08f
#define TGK_PIO2_3 -1.715124510e-15f
#define TGK_PIO4_HI 7.853981853e-01f
#define TGK_PIO4_LO -2.185569414e-08f
#define TGK_PIO2_HI 1.570796371e+00f
#define TGK_PIO2_LO -4.371138829e-08f
#define TGK_3PIO4_HI 2.356194496e+00f
#define TGK_3PIO4_LO -5.962440319e-09f
#define TGK_PI_HI 3.141592741e+00f
#define TGK_PI_LO -8.742277657e-08f
#define TGK_LOG2E 1.442695022e+00f
#define TGK_LN2_HI 6.931471825e-01f
#define TGK_LN2_LO -1.904654212e-09f
#define TGK_TWO_PI_D 6.283185307179586

TGK_FN uint32_t tgk_as_u32(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, sizeof u);
    return u;
}

TGK_FN float tgk_as_f32(uint32_t u)
{
    float f;
    __builtin_memcpy(&f, &u, sizeof f);
    return f;
}

TGK_FN float tgk_inf(void) { return tgk_as_f32(TGK_INFINITY_BITS); }

TGK_FN float tgk_copysignf(float magnitude, float sign)
{
    return tgk_as_f32((tgk_as_u32(magnitude) & 0x7fffffffu) | (tgk_as_u32(sign) & 0x80000000u));
}

TGK_FN int tgk_signbitf(float x) { return (tgk_as_u32(x) >> 31) != 0; }

TGK_FN float tgk_sqrtf(float x) { return sqrtf(x); }

/* sin and cos of x. Reduction r = x - k*pi/2 with k = rint(2x/pi) and a
 * three-piece pi/2, |r| <= pi/4; polynomials sin(r) = r + r^3 P(r^2),
 * cos(r) = 1 + r^2 Q(r^2); the quadrant k mod 4 permutes and negates them.
 * Arguments beyond 2^20 are first folded by an exact double fmod(x, 2pi). */
TGK_FN void tgk_sincosf(float x, float *s, float *c)
{
    if (!(fabsf(x) < tgk_inf())) {
        *s = x - x;
        *c = x - x;
        return;
    }
    if (fabsf(x) > 1048576.0f)
        x = (float)fmod((double)x, TGK_TWO_PI_D);
    const float k = rintf(x * TGK_TWO_OVER_PI);
    float r = fmaf(-k, TGK_PIO2_1, x);
    r = fmaf(-k, TGK_PIO2_2, r);
    r = fmaf(-k, TGK_PIO2_3, r);
    const float w = r * r;

    float ps = fmaf(w, 2.832745849e-06f, -1.985307608e-04f);
    ps = fmaf(w, ps, 8.333379403e-03f);
    ps = fmaf(w, ps, -1.666666716e-01f);
    const float sr = fmaf(r * w, ps, r);

    float pc = fmaf(w, -2.888124584e-07f, 2.482202217e-05f);
    pc = fmaf(w, pc, -1.388898003e-03f);
    pc = fmaf(w, pc, 4.166666791e-02f);
    pc = fmaf(w, pc, -5.000000000e-01f);
    const float cr = fmaf(w, pc, 1.0f);

    const int q = (int)(k - 4.0f * floorf(k * 0.25f));
    if (q == 0) {
        *s = sr;
        *c = cr;
    } else if (q == 1) {
        *s = cr;
        *c = -sr;
    } else if (q == 2) {
        *s = -sr;
        *c = -cr;
    } else {
        *s = -cr;
        *c = sr;
    }
}

TGK_FN float tgk_sinf(float x)
{
    float s, c;
    tgk_sincosf(x, &s, &c);
    return s;
}

TGK_FN float tgk_cosf(float x)
{
    float s, c;
    tgk_sincosf(x, &s, &c);
    return c;
}

/* atan(t) for |t| <= tan(pi/8): t + t^3 P(t^2). */
TGK_FN float tgk_atan_kernel(float t)
{
    const float w = t * t;
    float p = fmaf(w, -8.545340598e-02f, 9.817933291e-02f);
    p = fmaf(w, p, -9.654480964e-02f);
    p = fmaf(w, p, 1.117930934e-01f);
    p = fmaf(w, p, -1.428978592e-01f);
    p = fmaf(w, p, 2.000011057e-01f);
    p = fmaf(w, p, -3.333333433e-01f);
    return fmaf(t * w, p, t);
}

/* atan2(y, x) with the IEEE-754 / C99 Annex F special cases. The ratio
 * a = min(|x|,|y|)/max(|x|,|y|) is reduced to |t| <= tan(pi/8) with
 * atan(a) = pi/4 + atan(t), t = (min - max)/(min + max); the octant is folded back as
 * K + sgn*atan(t) with K a multiple of pi/4 held as hi + lo. */
TGK_FN float tgk_atan2f(float y, float x)
{
    if (x != x || y != y)
        return x + y;
    const float ax = fabsf(x);
    const float ay = fabsf(y);
    const float inf = tgk_inf();
    if (ax == inf || ay == inf) {
        float r;
        if (ax == inf && ay == inf)
            r = tgk_signbitf(x) ? TGK_3PIO4_HI : TGK_PIO4_HI;
        else if (ay == inf)
            r = TGK_PIO2_HI;
        else
            r = tgk_signbitf(x) ? TGK_PI_HI : 0.0f;
        return tgk_copysignf(r, y);
    }
    if (ay == 0.0f)
        return tgk_signbitf(x) ? tgk_copysignf(TGK_PI_HI, y) : y;
    if (ax == 0.0f)
        return tgk_copysignf(TGK_PIO2_HI, y);

    const int swap = ay > ax;
    const float num = swap ? ax : ay;
    const float den = swap ? ay : ax;
    const float a = num / den;
    const int reduced = a > 0.41421356f;
    const float t = reduced ? (num - den) / (num + den) : a;
    const float at = tgk_atan_kernel(t);

    /* angle = K + sign * atan(t), K in multiples of pi/4 */
    int octant = reduced ? 1 : 0;  /* first octant result: octant*pi/4 + at */
    float sign = 1.0f;
    if (swap) {                    /* pi/2 - (octant*pi/4 + at) */
        octant = 2 - octant;
        sign = -sign;
    }
    if (tgk_signbitf(x)) {         /* pi - angle */
        octant = 4 - octant;
        sign = -sign;
    }
    float hi, lo;
    if (octant == 0) {
        hi = 0.0f;
        lo = 0.0f;
    } else if (octant == 1) {
        hi = TGK_PIO4_HI;
        lo = TGK_PIO4_LO;
    } else if (octant == 2) {
        hi = TGK_PIO2_HI;
        lo = TGK_PIO2_LO;
    } else if (octant == 3) {
        hi = TGK_3PIO4_HI;
        lo = TGK_3PIO4_LO;
    } else {
        hi = TGK_PI_HI;
        lo = TGK_PI_LO;
    }
    const float r = hi + (sign * at + lo);
    return tgk_copysignf(r, y);
}

/* asin(x) for |x| <= 1/2: x + x^3 P(x^2). */
TGK_FN float tgk_asin_kernel(float x, float w)
{
    float p = fmaf(w, 3.770212084e-02f, 1.431217324e-02f);
    p = fmaf(w, p, 3.183952346e-02f);
    p = fmaf(w, p, 4.451335222e-02f);
    p = fmaf(w, p, 7.500521839e-02f);
    p = fmaf(w, p, 1.666665971e-01f);
    return fmaf(x * w, p, x);
}

/* acos(x): pi/2 - asin(x) for |x| <= 1/2, 2 asin(sqrt((1-|x|)/2)) (and pi
 * minus it for x < 0) otherwise. NaN outside [-1, 1]. */
TGK_FN float tgk_acosf(float x)
{
    if (x != x)
        return x;
    const float ax = fabsf(x);
    if (ax > 1.0f)
        return (x - x) / (x - x);
    if (ax <= 0.5f) {
        const float w = x * x;
        float p = fmaf(w, 3.770212084e-02f, 1.431217324e-02f);
        p = fmaf(w, p, 3.183952346e-02f);
        p = fmaf(w, p, 4.451335222e-02f);
        p = fmaf(w, p, 7.500521839e-02f);
        p = fmaf(w, p, 1.666665971e-01f);
        return TGK_PIO2_HI - (x + fmaf(x * w, p, -TGK_PIO2_LO));
    }
    const float z = (1.0f - ax) * 0.5f;
    const float s = sqrtf(z);
    const float as = tgk_asin_kernel(s, z);
    if (x > 0.0f)
        return 2.0f * as;
    return TGK_PI_HI - (2.0f * as - TGK_PI_LO);
}

/* exp(x) = 2^k exp(r), k = rint(x log2 e), r = x - k ln2 (two-piece ln2),
 * exp(r) = 1 + r + r^2 P(r); the 2^k scaling is split so that results in the
 * subnormal range are rounded once. */
TGK_FN float tgk_expf(float x)
{
    if (x != x)
        return x;
    if (x > 89.0f)
        return tgk_inf();
    if (x < -104.0f)
        return 0.0f;
    const float k = rintf(x * TGK_LOG2E);
    float r = fmaf(-k, TGK_LN2_HI, x);
    r = fmaf(-k, TGK_LN2_LO, r);
    float p = fmaf(r, 1.974024926e-04f, 1.393507817e-03f);
    p = fmaf(r, p, 8.333570324e-03f);
    p = fmaf(r, p, 4.166646302e-02f);
    p = fmaf(r, p, 1.666666567e-01f);
    p = fmaf(r, p, 5.000000000e-01f);
    const float e = fmaf(r * r, p, r) + 1.0f;
    const int ki = (int)k;
    if (ki > 127)
        return (e * 2.0f) * tgk_as_f32((uint32_t)(127 + 127) << 23);
    if (ki < -126)
        return (e * tgk_as_f32((uint32_t)(ki + 126 + 127) << 23)) * tgk_as_f32(0x00800000u);
    return e * tgk_as_f32((uint32_t)(ki + 127) << 23);
}

/* log(x) = e ln2 + log(m), x = m 2^e with m in [sqrt(1/2), sqrt(2)).
 * With f = m - 1 (exact) and s = f/(2+f): log(m) = 2 atanh(s)
 * = f - s (f - s^2 P(s^2)), so the rounding of s only enters the small
 * correction term. */
TGK_FN float tgk_logf(float x)
{
    if (x != x)
        return x;
    if (x < 0.0f)
        return (x - x) / (x - x);
    if (x == 0.0f)
        return -tgk_inf();
    if (x == tgk_inf())
        return x;
    uint32_t ix = tgk_as_u32(x);
    int e = 0;
    if (ix < 0x00800000u) {
        ix = tgk_as_u32(x * 8388608.0f);
        e = -23;
    }
    e += (int)(ix >> 23) - 127;
    uint32_t mb = (ix & 0x007fffffu) | 0x3f800000u;
    if (mb > 0x3fb504f3u) {
        mb -= 0x00800000u;
        e += 1;
    }
    const float f = tgk_as_f32(mb) - 1.0f;
    const float s = f / (2.0f + f);
    const float w = s * s;
    float p = fmaf(w, 2.428623587e-01f, 2.849848568e-01f);
    p = fmaf(w, p, 4.000097513e-01f);
    p = fmaf(w, p, 6.666666269e-01f);
    const float correction = s * fmaf(-w, p, f);
    const float ef = (float)e;
    return fmaf(ef, TGK_LN2_HI, f - fmaf(-ef, TGK_LN2_LO, correction));
}

#endif /* TGK_MATH_H */
