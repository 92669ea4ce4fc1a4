/* rmath.h - the project's FP32 elementary-function library.
 *
 * One source text, compiled by gcc for the CPU oracle and host code and by
 * hipcc for gfx950 kernels, so that transcendental results are bit-identical
 * on both sides.  The OpenCL builtins that the reference kernels call
 * (sin, cos, exp, log, pow, tanh, atan2, acos, asin, fmod, round, sign,
 * degrees, and the vector helpers in rvec.h) are implementation-defined in
 * OpenCL; this library fixes one implementation.
 *
 * Methods (all written here from the mathematics):
 *   - sin/cos: Cody & Waite additive range reduction by pi/2 (three-part
 *     constant, fused multiply-adds), minimax kernels on [-pi/4, pi/4].
 *   - exp: reduction x = k ln2 + r (two-part ln2), minimax kernel, exact
 *     scaling by 2^k split to stay exact into the subnormal range.
 *   - log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), log(1+f) through
 *     s = f/(2+f) and the atanh series (minimax remainder).
 *   - pow: exp2(y log2 x) with y*e formed exactly by an FMA two-product.
 *   - tanh: odd minimax kernel below 0.625, 1 - 2/(e^{2|x|}+1) above.
 *   - atan/atan2: reduction to [0, tan(pi/8)] via atan(u) = pi/4 +
 *     atan((u-1)/(u+1)), odd minimax kernel; C99 Annex F special cases.
 *   - asin/acos: odd minimax kernel on [0, 1/2], half-angle identity
 *     asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above.
 *   - fmod: exact integer long division of the significands.
 * The kernel coefficients are produced by gen_rmath_coeffs.py (Lawson
 * minimax fits); measured errors are checked by tests/test_rmath.py.
 */
#ifndef RAMP_RMATH_H
#define RAMP_RMATH_H

#include "rdefs.h"

/* ---- kernel coefficients (gen_rmath_coeffs.py) ------------------------- */
#define RM_SIN0 -0x1.555556p-3f
#define RM_SIN1 0x1.111106p-7f
#define RM_SIN2 -0x1.a00dccp-13f
#define RM_SIN3 0x1.6c6d7cp-19f
#define RM_COS0 0x1.555556p-5f
#define RM_COS1 -0x1.6c16b6p-10f
#define RM_COS2 0x1.a00f6p-16f
#define RM_COS3 -0x1.23d2aep-22f
#define RM_EXP0 0x1p-1f
#define RM_EXP1 0x1.555554p-3f
#define RM_EXP2 0x1.55546ep-5f
#define RM_EXP3 0x1.1112acp-7f
#define RM_EXP4 0x1.6daa08p-10f
#define RM_EXP5 0x1.9ef19p-13f
#define RM_LOG0 0x1.555554p-1f
#define RM_LOG1 0x1.999a8ap-2f
#define RM_LOG2 0x1.243486p-2f
#define RM_LOG3 0x1.e3a376p-3f
#define RM_TANH0 -0x1.555554p-2f
#define RM_TANH1 0x1.111066p-3f
#define RM_TANH2 -0x1.b9ee9p-5f
#define RM_TANH3 0x1.638b58p-6f
#define RM_TANH4 -0x1.0bf4fep-7f
#define RM_TANH5 0x1.18e1ep-9f
#define RM_ATAN0 -0x1.555554p-2f
#define RM_ATAN1 0x1.9998fap-3f
#define RM_ATAN2 -0x1.247b48p-3f
#define RM_ATAN3 0x1.c40944p-4f
#define RM_ATAN4 -0x1.5929f4p-4f
#define RM_ATAN5 0x1.7c1a18p-5f
#define RM_ASIN0 0x1.55554cp-3f
#define RM_ASIN1 0x1.33387cp-4f
#define RM_ASIN2 0x1.6caed2p-5f
#define RM_ASIN3 0x1.0491ap-5f
#define RM_ASIN4 0x1.d91fa8p-7f
#define RM_ASIN5 0x1.3355cp-5f

/* ---- constants --------------------------------------------------------- */
#define RM_TWO_OVER_PI 0x1.45f306p-1f
#define RM_PIO2_1 0x1.921fb6p+0f
#define RM_PIO2_2 -0x1.777a5cp-25f
#define RM_PIO2_3 -0x1.ee59dap-50f
#define RM_PI_HI 0x1.921fb6p+1f
#define RM_PI_LO -0x1.777a5cp-24f
#define RM_PIO2_HI 0x1.921fb6p+0f
#define RM_PIO2_LO -0x1.777a5cp-25f
#define RM_PIO4_HI 0x1.921fb6p-1f
#define RM_PIO4_LO -0x1.777a5cp-26f
#define RM_LN2_HI 0x1.62e43p-1f
#define RM_LN2_LO -0x1.05c61p-29f
#define RM_LN2_HI16 0x1.62e4p-1f
#define RM_LN2_LO16 0x1.7f7d1cp-20f
#define RM_LOG2E 0x1.715476p+0f
#define RM_LN2 0x1.62e43p-1f
#define RM_RAD2DEG 0x1.ca5dc2p+5f
#define RM_EXP_MAX 0x1.62e42ep+6f
#define RM_EXP_MIN -0x1.9fe368p+6f
#define RM_INF (__builtin_inff())
#define RM_NAN (__builtin_nanf(""))

/* ---- exact primitives -------------------------------------------------- */
RHD float rm_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RHD float rm_sqrt(float x) { return __builtin_sqrtf(x); }
RHD float rm_fabs(float x) { return __builtin_fabsf(x); }
RHD float rm_floor(float x) { return __builtin_floorf(x); }
RHD float rm_ceil(float x) { return __builtin_ceilf(x); }
RHD float rm_trunc(float x) { return __builtin_truncf(x); }
RHD float rm_rint(float x) { return __builtin_rintf(x); }
RHD float rm_copysign(float x, float y) { return __builtin_copysignf(x, y); }
RHD int rm_signbit(float x) { return (int)(rm_f2u(x) >> 31); }
RHD int rm_isnan(float x) { return x != x; }
RHD int rm_isinf(float x) { return (rm_f2u(x) & 0x7fffffffu) == 0x7f800000u; }
RHD float rm_fmin(float a, float b) { return __builtin_fminf(a, b); }
RHD float rm_fmax(float a, float b) { return __builtin_fmaxf(a, b); }

/* round half away from zero (C roundf / OpenCL round) */
RHD float rm_round(float x)
{
    float t = rm_trunc(x);
    if (rm_fabs(x - t) >= 0.5f) t += rm_copysign(1.0f, x);
    return t;
}

/* OpenCL sign(): 1, -1, signed zero kept, NaN -> 0 */
RHD float rm_sign(float x)
{
    if (x > 0.0f) return 1.0f;
    if (x < 0.0f) return -1.0f;
    if (x != x) return 0.0f;
    return x;
}

RHD float rm_degrees(float x) { return x * RM_RAD2DEG; }

/* 2^k for -126 <= k <= 127 */
RHD float rm_pow2i(int k) { return rm_u2f((uint32_t)(k + 127) << 23); }

RHD float rm_ldexp(float x, int k)
{
    if (k > 127) {
        x *= 0x1p127f; k -= 127;
        if (k > 127) { x *= 0x1p127f; k -= 127; if (k > 127) k = 127; }
    } else if (k < -126) {
        x *= 0x1p-126f; k += 126;
        if (k < -126) { x *= 0x1p-126f; k += 126; if (k < -126) k = -126; }
    }
    return x * rm_pow2i(k);
}

/* ---- sin / cos --------------------------------------------------------- */
RHD float rm_sin_kernel(float r)
{
    float z = r * r;
    float p = rm_fma(z, RM_SIN3, RM_SIN2);
    p = rm_fma(z, p, RM_SIN1);
    p = rm_fma(z, p, RM_SIN0);
    return rm_fma(r * z, p, r);
}

RHD float rm_cos_kernel(float r)
{
    float z = r * r;
    float p = rm_fma(z, RM_COS3, RM_COS2);
    p = rm_fma(z, p, RM_COS1);
    p = rm_fma(z, p, RM_COS0);
    return rm_fma(z, rm_fma(z, p, -0.5f), 1.0f);
}

typedef struct { float r; int q; } rm_red;

/* x = k*pi/2 + r, q = k mod 4.  Accurate for |x| up to about 1e5. */
RHD rm_red rm_reduce_pio2(float x)
{
    rm_red o;
    float k = rm_rint(x * RM_TWO_OVER_PI);
    float r = rm_fma(-k, RM_PIO2_1, x);
    r = rm_fma(-k, RM_PIO2_2, r);
    r = rm_fma(-k, RM_PIO2_3, r);
    o.r = r;
    o.q = (int)(k - 4.0f * rm_floor(k * 0.25f));
    return o;
}

RHD float rm_sin(float x)
{
    if (!(rm_fabs(x) < 1.0e30f)) return x - x; /* NaN for inf/nan */
    rm_red d = rm_reduce_pio2(x);
    float s = rm_sin_kernel(d.r), c = rm_cos_kernel(d.r);
    float v = (d.q & 1) ? c : s;
    return (d.q & 2) ? -v : v;
}

RHD float rm_cos(float x)
{
    if (!(rm_fabs(x) < 1.0e30f)) return x - x;
    rm_red d = rm_reduce_pio2(x);
    float s = rm_sin_kernel(d.r), c = rm_cos_kernel(d.r);
    float v = (d.q & 1) ? s : c;
    return ((d.q + 1) & 2) ? -v : v;
}

/* ---- exp ---------------------------------------------------------------- */
RHD float rm_exp_kernel(float r) /* e^r for |r| <= 0.36 */
{
    float p = rm_fma(r, RM_EXP5, RM_EXP4);
    p = rm_fma(r, p, RM_EXP3);
    p = rm_fma(r, p, RM_EXP2);
    p = rm_fma(r, p, RM_EXP1);
    p = rm_fma(r, p, RM_EXP0);
    return rm_fma(r, rm_fma(r, p, 1.0f), 1.0f);
}

RHD float rm_exp(float x)
{
    if (x != x) return x;
    if (x > RM_EXP_MAX) return RM_INF;
    if (x < RM_EXP_MIN) return 0.0f;
    float k = rm_rint(x * RM_LOG2E);
    float r = rm_fma(-k, RM_LN2_HI, x);
    r = rm_fma(-k, RM_LN2_LO, r);
    return rm_ldexp(rm_exp_kernel(r), (int)k);
}

/* ---- log ---------------------------------------------------------------- */
typedef struct { float f; int e; } rm_logred;

/* x = 2^e (1+f) with 1+f in [sqrt(1/2), sqrt(2)); x finite and > 0 */
RHD rm_logred rm_log_reduce(float x)
{
    rm_logred o;
    uint32_t ix = rm_f2u(x);
    int e = 0;
    if (ix < 0x00800000u) { x *= 0x1p23f; ix = rm_f2u(x); e = -23; }
    e += (int)(ix >> 23) - 127;
    uint32_t mb = ix & 0x007fffffu;
    float m;
    if (mb > 0x3504f3u) { m = rm_u2f(mb | 0x3f000000u); e += 1; }
    else m = rm_u2f(mb | 0x3f800000u);
    o.f = m - 1.0f;
    o.e = e;
    return o;
}

/* log(1+f) for 1+f in [sqrt(1/2), sqrt(2)) */
RHD float rm_log1p_kernel(float f)
{
    float s = f / (2.0f + f);
    float z = s * s;
    float p = rm_fma(z, RM_LOG3, RM_LOG2);
    p = rm_fma(z, p, RM_LOG1);
    p = rm_fma(z, p, RM_LOG0);
    float R = z * p;
    float hfsq = 0.5f * f * f;
    return f - (hfsq - s * (hfsq + R));
}

RHD float rm_log(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return RM_NAN;
    if (x == 0.0f) return -RM_INF;
    if (rm_isinf(x)) return x;
    rm_logred d = rm_log_reduce(x);
    float s = d.f / (2.0f + d.f);
    float z = s * s;
    float p = rm_fma(z, RM_LOG3, RM_LOG2);
    p = rm_fma(z, p, RM_LOG1);
    p = rm_fma(z, p, RM_LOG0);
    float R = z * p;
    float hfsq = 0.5f * d.f * d.f;
    float ef = (float)d.e;
    return ef * RM_LN2_HI16 - ((hfsq - (s * (hfsq + R) + ef * RM_LN2_LO16)) - d.f);
}

/* ---- pow ---------------------------------------------------------------- */
RHD int rm_is_int(float y) { return rm_trunc(y) == y; }
RHD int rm_is_odd_int(float y)
{
    if (!rm_is_int(y) || rm_fabs(y) >= 0x1p24f) return 0;
    float h = y * 0.5f;
    return rm_trunc(h) != h;
}

RHD float rm_pow(float x, float y)
{
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    if (y == 1.0f) return x;
    if (y == 2.0f) return x * x;
    float ax = rm_fabs(x);
    int odd = rm_is_odd_int(y);
    if (rm_isinf(y)) {
        if (ax == 1.0f) return 1.0f;
        if ((ax > 1.0f) == (y > 0.0f)) return RM_INF;
        return 0.0f;
    }
    if (ax == 0.0f || rm_isinf(ax)) {
        float r = ((ax == 0.0f) == (y < 0.0f)) ? RM_INF : 0.0f;
        return (x < 0.0f && odd) ? -r : r;
    }
    if (x < 0.0f && !rm_is_int(y)) return RM_NAN;
    /* |x| finite, nonzero: t = y*log2|x| = y*e + y*l */
    rm_logred d = rm_log_reduce(ax);
    float l = rm_log1p_kernel(d.f) * RM_LOG2E;
    float ef = (float)d.e;
    float p1 = y * ef, e1 = rm_fma(y, ef, -p1);
    float p2 = y * l, e2 = rm_fma(y, l, -p2);
    float t = p1 + p2;
    float res;
    if (t > 128.5f) res = RM_INF;
    else if (t < -151.0f) res = 0.0f;
    else {
        float k = rm_rint(t);
        float fr = ((p1 - k) + p2) + (e1 + e2);
        res = rm_ldexp(rm_exp_kernel(fr * RM_LN2), (int)k);
    }
    return (x < 0.0f && odd) ? -res : res;
}

/* ---- tanh --------------------------------------------------------------- */
RHD float rm_tanh(float x)
{
    float ax = rm_fabs(x);
    if (ax != ax) return x;
    if (ax < 0.625f) {
        float z = x * x;
        float p = rm_fma(z, RM_TANH5, RM_TANH4);
        p = rm_fma(z, p, RM_TANH3);
        p = rm_fma(z, p, RM_TANH2);
        p = rm_fma(z, p, RM_TANH1);
        p = rm_fma(z, p, RM_TANH0);
        return rm_fma(x * z, p, x);
    }
    float r;
    if (ax > 9.1f) r = 1.0f;
    else {
        float e = rm_exp(2.0f * ax);
        r = 1.0f - 2.0f / (e + 1.0f);
    }
    return rm_copysign(r, x);
}

/* ---- atan / atan2 ------------------------------------------------------- */
RHD float rm_atan_kernel(float v) /* |v| <= tan(pi/8) + 0.01 */
{
    float z = v * v;
    float p = rm_fma(z, RM_ATAN5, RM_ATAN4);
    p = rm_fma(z, p, RM_ATAN3);
    p = rm_fma(z, p, RM_ATAN2);
    p = rm_fma(z, p, RM_ATAN1);
    p = rm_fma(z, p, RM_ATAN0);
    return rm_fma(v * z, p, v);
}

RHD float rm_atan01(float u) /* u in [0, 1] */
{
    if (u > 0.41421356f) {
        float v = (u - 1.0f) / (u + 1.0f);
        return RM_PIO4_HI + (rm_atan_kernel(v) + RM_PIO4_LO);
    }
    return rm_atan_kernel(u);
}

RHD float rm_atan(float x)
{
    float ax = rm_fabs(x);
    if (ax != ax) return x;
    float r;
    if (ax > 1.0f) r = RM_PIO2_HI + (RM_PIO2_LO - rm_atan01(1.0f / ax));
    else r = rm_atan01(ax);
    return rm_copysign(r, x);
}

RHD float rm_atan2(float y, float x)
{
    if (x != x || y != y) return x + y;
    float ax = rm_fabs(x), ay = rm_fabs(y);
    float r;
    if (ay == 0.0f) {
        r = rm_signbit(x) ? RM_PI_HI : 0.0f;
        return rm_copysign(r, y);
    }
    if (ax == 0.0f) return rm_copysign(RM_PIO2_HI, y);
    if (rm_isinf(ax)) {
        if (rm_isinf(ay)) r = rm_signbit(x) ? 0x1.2d97c8p+1f : RM_PIO4_HI;
        else r = rm_signbit(x) ? RM_PI_HI : 0.0f;
        return rm_copysign(r, y);
    }
    if (rm_isinf(ay)) return rm_copysign(RM_PIO2_HI, y);
    if (ay > ax) {
        r = rm_atan01(ax / ay);
        r = RM_PIO2_HI + (RM_PIO2_LO - r);
    } else {
        r = rm_atan01(ay / ax);
    }
    if (rm_signbit(x)) r = RM_PI_HI + (RM_PI_LO - r);
    return rm_copysign(r, y);
}

/* ---- asin / acos -------------------------------------------------------- */
RHD float rm_asin_kernel(float x) /* |x| <= 0.5 */
{
    float z = x * x;
    float p = rm_fma(z, RM_ASIN5, RM_ASIN4);
    p = rm_fma(z, p, RM_ASIN3);
    p = rm_fma(z, p, RM_ASIN2);
    p = rm_fma(z, p, RM_ASIN1);
    p = rm_fma(z, p, RM_ASIN0);
    return rm_fma(x * z, p, x);
}

RHD float rm_asin(float x)
{
    float ax = rm_fabs(x);
    if (!(ax <= 1.0f)) return RM_NAN;
    if (ax <= 0.5f) return rm_asin_kernel(x);
    float z = (1.0f - ax) * 0.5f;
    float s = rm_sqrt(z);
    float r = rm_asin_kernel(s);
    r = RM_PIO2_HI + (RM_PIO2_LO - 2.0f * r);
    return rm_copysign(r, x);
}

RHD float rm_acos(float x)
{
    float ax = rm_fabs(x);
    if (!(ax <= 1.0f)) return RM_NAN;
    if (ax <= 0.5f) return RM_PIO2_HI + (RM_PIO2_LO - rm_asin_kernel(x));
    float z = (1.0f - ax) * 0.5f;
    float s = rm_sqrt(z);
    float r = 2.0f * rm_asin_kernel(s);
    if (x > 0.0f) return r;
    return RM_PI_HI + (RM_PI_LO - r);
}

/* ---- fmod (exact) ------------------------------------------------------- */
RHD float rm_fmod(float x, float y)
{
    uint32_t ux = rm_f2u(x), uy = rm_f2u(y);
    uint32_t sx = ux & 0x80000000u;
    ux &= 0x7fffffffu;
    uy &= 0x7fffffffu;
    if (uy == 0u || ux >= 0x7f800000u || uy > 0x7f800000u) return RM_NAN;
    if (ux < uy) return x;
    if (ux == uy) return rm_u2f(sx);
    int ex = (int)(ux >> 23), ey = (int)(uy >> 23);
    uint32_t mx, my;
    if (ex == 0) { mx = ux; ex = 1; while (!(mx & 0x00800000u)) { mx <<= 1; ex--; } }
    else mx = (ux & 0x007fffffu) | 0x00800000u;
    if (ey == 0) { my = uy; ey = 1; while (!(my & 0x00800000u)) { my <<= 1; ey--; } }
    else my = (uy & 0x007fffffu) | 0x00800000u;
    for (; ex > ey; ex--) {
        if (mx >= my) { mx -= my; if (mx == 0u) return rm_u2f(sx); }
        mx <<= 1;
    }
    if (mx >= my) { mx -= my; if (mx == 0u) return rm_u2f(sx); }
    while (!(mx & 0x00800000u)) { mx <<= 1; ex--; }
    uint32_t out;
    if (ex > 0) out = ((uint32_t)ex << 23) | (mx & 0x007fffffu);
    else out = mx >> (1 - ex);
    return rm_u2f(out | sx);
}

/* ---- FP64 exp and cos ----------------------------------------------------
 * For the reference's double-precision sites that call a transcendental
 * (the monochromator's Box-Muller cos(2.0f*M_PI*u), the isotropic sample's
 * detailed-balance exp(-w/(kB*T))): one implementation of basic IEEE double
 * operations, so the oracle and the kernels agree bit for bit.
 *   exp: x = k ln2 + r, |r| <= ln2/2, with ln2 split into a 42-bit head (k
 *        times it is exact for |k| < 2^11) and a tail; e^r by its Taylor
 *        series to degree 13 (next term < 4e-18); exact scaling by 2^k.
 *   cos: x = k pi/2 + r, |r| <= pi/4, three-part pi/2 (33 + 33 bits + tail,
 *        exact products for |k| < 2^20, i.e. |x| below about 1.6e6); the
 *        Taylor series of cos r (degree 20) or sin r (degree 21) chosen by
 *        k mod 4 (next terms < 5e-24).
 * The series coefficients are 1/n! rounded to double. */
#define RD_INV_LN2 0x1.71547652b82fep+0
#define RD_LN2_HI 0x1.62e42fefa38p-1
#define RD_LN2_LO 0x1.ef35793c7673p-45
#define RD_EXP_MAX 0x1.62e42fefa39efp+9
#define RD_2_OVER_PI 0x1.45f306dc9c883p-1
#define RD_PIO2_1 0x1.921fb544p+0
#define RD_PIO2_2 0x1.0b4611a6p-34
#define RD_PIO2_3 0x1.3198a2e037073p-69
#define RD_F2 0x1.0000000000000p-1
#define RD_F3 0x1.5555555555555p-3
#define RD_F4 0x1.5555555555555p-5
#define RD_F5 0x1.1111111111111p-7
#define RD_F6 0x1.6c16c16c16c17p-10
#define RD_F7 0x1.a01a01a01a01ap-13
#define RD_F8 0x1.a01a01a01a01ap-16
#define RD_F9 0x1.71de3a556c734p-19
#define RD_F10 0x1.27e4fb7789f5cp-22
#define RD_F11 0x1.ae64567f544e4p-26
#define RD_F12 0x1.1eed8eff8d898p-29
#define RD_F13 0x1.6124613a86d09p-33
#define RD_F14 0x1.93974a8c07c9dp-37
#define RD_F15 0x1.ae7f3e733b81fp-41
#define RD_F16 0x1.ae7f3e733b81fp-45
#define RD_F17 0x1.952c77030ad4ap-49
#define RD_F18 0x1.6827863b97d97p-53
#define RD_F19 0x1.2f49b46814157p-57
#define RD_F20 0x1.e542ba4020225p-62
#define RD_F21 0x1.71b8ef6dcf572p-66

RHD double rd_exp(double x)
{
    if (!(x == x)) return x;
    if (x > RD_EXP_MAX) return 1.0 / 0.0;
    if (x < -745.2) return 0.0;
    double kd = __builtin_rint(x * RD_INV_LN2);
    double r = (x - kd * RD_LN2_HI) - kd * RD_LN2_LO;
    double p = RD_F13;
    p = p * r + RD_F12;
    p = p * r + RD_F11;
    p = p * r + RD_F10;
    p = p * r + RD_F9;
    p = p * r + RD_F8;
    p = p * r + RD_F7;
    p = p * r + RD_F6;
    p = p * r + RD_F5;
    p = p * r + RD_F4;
    p = p * r + RD_F3;
    p = p * r + RD_F2;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return __builtin_ldexp(p, (int)kd);
}

RHD double rd_cos_kernel(double r)
{
    double z = r * r;
    double p = RD_F20;
    p = p * z - RD_F18;
    p = p * z + RD_F16;
    p = p * z - RD_F14;
    p = p * z + RD_F12;
    p = p * z - RD_F10;
    p = p * z + RD_F8;
    p = p * z - RD_F6;
    p = p * z + RD_F4;
    p = p * z - RD_F2;
    return p * z + 1.0;
}

RHD double rd_sin_kernel(double r)
{
    double z = r * r;
    double p = RD_F21;
    p = p * z - RD_F19;
    p = p * z + RD_F17;
    p = p * z - RD_F15;
    p = p * z + RD_F13;
    p = p * z - RD_F11;
    p = p * z + RD_F9;
    p = p * z - RD_F7;
    p = p * z + RD_F5;
    p = p * z - RD_F3;
    return r + (r * z) * p;
}

/* a / c rounded to nearest double, for a divisor c whose reciprocal
 * rc = 1/c (rounded to nearest) is computed once on the host: q = a*rc is
 * within an ulp of a/c, the residual a - c*q is exact in one FMA, and
 * q + r*rc rounds to the double nearest a/c (Markstein's theorem).  Valid
 * for nonzero finite c and finite a with magnitudes in the float range (the
 * residual cannot underflow); a zero a gives +0. */
RHD double rd_div_rc(double a, double c, double rc)
{
    double q = a * rc;
    double r = __builtin_fma(-q, c, a);
    return __builtin_fma(r, rc, q);
}

RHD double rd_cos(double x)
{
    if (!(x - x == 0.0)) return x - x;   /* NaN for NaN and infinities */
    double kd = __builtin_rint(x * RD_2_OVER_PI);
    double r = ((x - kd * RD_PIO2_1) - kd * RD_PIO2_2) - kd * RD_PIO2_3;
    switch ((int)((long long)kd & 3)) {
    case 0: return rd_cos_kernel(r);
    case 1: return -rd_sin_kernel(r);
    case 2: return -rd_cos_kernel(r);
    default: return rd_sin_kernel(r);
    }
}

#endif
