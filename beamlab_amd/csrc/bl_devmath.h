// Host/device arithmetic of the soft-body step, written so that every float operation is the one the reference's
// CPU code performs, in the same order (compile with -ffp-contract=off: no fused multiply-adds are formed).
//   - vec3 operators follow src/core/math.h:46-83 (operator/(float) multiplies by the reciprocal, :57);
//   - smin/smax are std::min/std::max (b < a ? b : a, a < b ? b : a), including what they do with NaN;
//   - bl_expf is glibc's expf algorithm (Szabolcs Nagy's table method from ARM's optimized-routines, as shipped in
//     glibc 2.27+: exp(x) = 2^(k/32) * 2^(r/32), a 32-entry table of 2^(i/32) and a cubic in double precision),
//     evaluated with fused multiply-adds as glibc's FMA build does; it equals the host libm's expf on every float
//     input (checked exhaustively by tests/test_devmath.py).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BL_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define BL_HD static inline
#endif

#pragma clang fp contract(off)

namespace blg {

struct f3 {
    float x, y, z;
};

BL_HD f3 mk(float x, float y, float z) { f3 r; r.x = x, r.y = y, r.z = z; return r; }
BL_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
BL_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
BL_HD f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
BL_HD f3 divs(f3 a, float s) { float i = 1.0f / s; return mk(a.x * i, a.y * i, a.z * i); }
BL_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BL_HD float smin(float a, float b) { return b < a ? b : a; }
BL_HD float smax(float a, float b) { return a < b ? b : a; }
BL_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
BL_HD float signf(float v) { return v < 0.0f ? -1.0f : 1.0f; }
BL_HD f3 vmin(f3 a, f3 b) { return mk(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
BL_HD f3 vmax(f3 a, f3 b) { return mk(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }

#if defined(__HIPCC__)
BL_HD float fsqrt(float x) { return sqrtf(x); }
BL_HD float ffabs(float x) { return fabsf(x); }
BL_HD double dfma(double a, double b, double c) { return fma(a, b, c); }
BL_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
BL_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
BL_HD uint64_t d2u(double f) { return __builtin_bit_cast(uint64_t, f); }
BL_HD double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }
#else
BL_HD float fsqrt(float x) { return sqrtf(x); }
BL_HD float ffabs(float x) { return fabsf(x); }
BL_HD double dfma(double a, double b, double c) { return fma(a, b, c); }
BL_HD uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
BL_HD float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
BL_HD uint64_t d2u(double f) { uint64_t u; memcpy(&u, &f, 8); return u; }
BL_HD double u2d(uint64_t u) { double f; memcpy(&f, &u, 8); return f; }
#endif

BL_HD float length(f3 a) { return fsqrt(dot(a, a)); }

// sqrt of a double, rounded to nearest as the host's: the hardware result, then of it and its two neighbours the one
// whose square is nearest x (exact residuals by fma)
BL_HD double sqrt_cr(double x) {
    const double s = sqrt(x);
    if (!(s > 0) || !(s < 1.7976931348623157e308)) return s;
    const double dn = u2d(d2u(s) - 1), up = u2d(d2u(s) + 1);
    const double rs = fabs(dfma(-s, s, x)), rd = fabs(dfma(-dn, dn, x)), ru = fabs(dfma(-up, up, x));
    return rd < rs ? dn : (ru < rs ? up : s);
}

// 2^(i/32) rounded to double, less i << 47 (the exponent part the integer k adds back): the table of the method.
// Generated by hand from the definition (exp2l in extended precision, rounded to double); tests/test_devmath.py
// recomputes it.
#define BL_EXP2_TAB                                                                                                   \
    {0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,                   \
     0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,                   \
     0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,                   \
     0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,                   \
     0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,                   \
     0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,                   \
     0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,                   \
     0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull}

#if defined(__HIPCC__)
__constant__ const uint64_t k_exp2_tab[32] = BL_EXP2_TAB;
#define BL_EXP2_TABLE k_exp2_tab
#else
static const uint64_t k_exp2_tab_host[32] = BL_EXP2_TAB;
#define BL_EXP2_TABLE k_exp2_tab_host
#endif

// expf(x), bit-identical to glibc's (see the header comment).
BL_HD float bl_expf(float x) {
    const double inv_ln2_n = 0x1.71547652b82fep+0 * 32; // 32 / ln 2
    const double shift = 0x1.8p+52;                      // rounds to an integer in the low mantissa bits
    const double c0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, c1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
                 c2 = 0x1.62e42ff0c52d6p-1 / 32;         // the cubic for 2^(r/32), r in [-1/2, 1/2]
    const uint32_t ux = f2u(x);
    const uint32_t abstop = (ux >> 20) & 0x7ff;
    if (abstop >= (0x42b00000u >> 20)) { // |x| >= 88 or NaN
        if (ux == 0xff800000u) return 0.0f;     // -inf
        if (abstop >= (0x7f800000u >> 20)) return x + x;
        if (x > 0x1.62e42ep6f) return u2f(0x7f800000u);  // overflow
        if (x < -0x1.9fe368p6f) return 0.0f;              // underflow
    }
    const double xd = (double)x;
    double kd = dfma(inv_ln2_n, xd, shift);
    const uint64_t ki = d2u(kd);
    kd -= shift;
    const double r = dfma(inv_ln2_n, xd, -kd);
    uint64_t t = BL_EXP2_TABLE[ki % 32];
    t += ki << 47;
    const double s = u2d(t);
    double z = dfma(c0, r, c1);
    const double r2 = r * r;
    double y = dfma(c2, r, 1.0);
    y = dfma(z, r2, y);
    y = y * s;
    return (float)y;
}

// ---------------------------------------------------------------------------------------------- sinf and cosf
// glibc's sinf / cosf method (Szabolcs Nagy's, from ARM's optimized-routines, in glibc 2.28+): the argument reduced to
// [-pi/4, pi/4] in double (one multiply-add below 120, a 96-bit product with the bits of 2/pi above), then odd and even
// polynomials in double, rounded once to float; glibc's x86-64 FMA build forms the multiply-adds fused, as here. The
// coefficients are the method's published ones. Equal to the host libm's sinf / cosf on every float input (checked
// exhaustively by tests/test_devmath.py).
#define BL_SC_S1 (-0x1.555545995a603p-3)
#define BL_SC_S2 (0x1.1107605230bc4p-7)
#define BL_SC_S3 (-0x1.994eb3774cf24p-13)
#define BL_SC_C1 (-0x1.ffffffd0c621cp-2)
#define BL_SC_C2 (0x1.55553e1068f19p-5)
#define BL_SC_C3 (-0x1.6c087e89a359dp-10)
#define BL_SC_C4 (0x1.99343027bf8c3p-16)
// The bits of 2/pi (0.a2f9836e 4e441529 ...), a 32-bit window every 8 bits, three leading partial words.
#define BL_INV_PIO4                                                                                                   \
    {0xa2u,        0xa2f9u,      0xa2f983u,    0xa2f9836eu,  0xf9836e4eu,  0x836e4e44u,  0x6e4e4415u,  0x4e441529u,   \
     0x441529fcu,  0x1529fc27u,  0x29fc2757u,  0xfc2757d1u,  0x2757d1f5u,  0x57d1f534u,  0xd1f534ddu,  0xf534ddc0u,   \
     0x34ddc0dbu,  0xddc0db62u,  0xc0db6295u,  0xdb629599u,  0x6295993cu,  0x95993c43u,  0x993c4390u,  0x3c439041u}
#if defined(__HIPCC__)
__constant__ const uint32_t k_inv_pio4[24] = BL_INV_PIO4;
#define BL_INV_PIO4_TABLE k_inv_pio4
#else
static const uint32_t k_inv_pio4_host[24] = BL_INV_PIO4;
#define BL_INV_PIO4_TABLE k_inv_pio4_host
#endif

// The polynomials on the reduced argument: odd (sine) when n is even, even (cosine, times cs = +-1) when n is odd.
BL_HD float sc_poly(double x, double x2, int n, double cs) {
    if ((n & 1) == 0) {
        const double x3 = x * x2;
        const double s1 = dfma(x2, BL_SC_S3, BL_SC_S2);
        const double x7 = x3 * x2;
        const double s = dfma(x3, BL_SC_S1, x);
        return (float)dfma(x7, s1, s);
    }
    const double x4 = x2 * x2;
    const double c2 = dfma(x2, cs * BL_SC_C4, cs * BL_SC_C3);
    const double c1 = dfma(x2, cs * BL_SC_C1, cs);
    const double x6 = x4 * x2;
    const double c = dfma(x4, cs * BL_SC_C2, c1);
    return (float)dfma(x6, c2, c);
}
// |y| < 120: y = x + n pi/2, x in [-pi/4, pi/4] (2/pi scaled by 2^24, so the quadrant is bits 24..31 of the product)
BL_HD double sc_reduce_fast(double x, int* np) {
    const double r = x * 0x1.45f306dc9c883p+23;
    const int n = ((int32_t)r + 0x800000) >> 24;
    *np = n;
    return dfma(-(double)n, 0x1.921fb54442d18p0, x);
}
// |y| >= 120: the mantissa times 96 bits of 2/pi (from the window the exponent selects)
BL_HD double sc_reduce_large(uint32_t xi, int* np) {
    const uint32_t* arr = &BL_INV_PIO4_TABLE[(xi >> 26) & 15];
    const int shift = (xi >> 23) & 7;
    xi = (xi & 0xffffffu) | 0x800000u;
    xi <<= shift;
    uint64_t res0 = (uint64_t)(uint32_t)(xi * arr[0]);
    const uint64_t res1 = (uint64_t)xi * arr[4];
    const uint64_t res2 = (uint64_t)xi * arr[8];
    res0 = (res2 >> 32) | (res0 << 32);
    res0 += res1;
    const uint64_t n = (res0 + (1ull << 61)) >> 62;
    res0 -= n << 62;
    *np = (int)n;
    return (double)(int64_t)res0 * 0x1.921fb54442d18p-62;
}
BL_HD float bl_sinf(float y) {
    const uint32_t top = (f2u(y) >> 20) & 0x7ff;
    double x = (double)y;
    int n;
    if (top < (0x3f490fdbu >> 20)) { // |y| < 0.75 (the top 12 bits of pi/4's)
        if (top < (0x39800000u >> 20)) return y;  // |y| < 2^-12
        return sc_poly(x, x * x, 0, 1.0);
    }
    if (top < (0x42f00000u >> 20)) { // |y| < 120
        x = sc_reduce_fast(x, &n);
        const double s = (n & 3) == 1 || (n & 3) == 2 ? -1.0 : 1.0;
        return sc_poly(x * s, x * x, n, (n & 2) ? -1.0 : 1.0);
    }
    if (top < (0x7f800000u >> 20)) {
        const uint32_t xi = f2u(y);
        const int sign = (int)(xi >> 31);
        x = sc_reduce_large(xi, &n);
        const int m = (n + sign) & 3;
        const double s = m == 1 || m == 2 ? -1.0 : 1.0;
        return sc_poly(x * s, x * x, n, (m & 2) ? -1.0 : 1.0);
    }
    return (y - y) / (y - y); // inf or NaN: NaN
}
BL_HD float bl_cosf(float y) {
    const uint32_t top = (f2u(y) >> 20) & 0x7ff;
    double x = (double)y;
    int n;
    if (top < (0x3f490fdbu >> 20)) {
        if (top < (0x39800000u >> 20)) return 1.0f;
        return sc_poly(x, x * x, 1, 1.0);
    }
    if (top < (0x42f00000u >> 20)) {
        x = sc_reduce_fast(x, &n);
        const double s = (n & 3) == 1 || (n & 3) == 2 ? -1.0 : 1.0;
        return sc_poly(x * s, x * x, n ^ 1, (n & 2) ? -1.0 : 1.0);
    }
    if (top < (0x7f800000u >> 20)) {
        const uint32_t xi = f2u(y);
        const int sign = (int)(xi >> 31);
        x = sc_reduce_large(xi, &n);
        const int m = (n + sign) & 3;
        const double s = m == 1 || m == 2 ? -1.0 : 1.0;
        return sc_poly(x * s, x * x, n ^ 1, (m & 2) ? -1.0 : 1.0);
    }
    return (y - y) / (y - y);
}

// ---------------------------------------------------------------------------------------------- atanf and atan2f
// glibc's float atanf / atan2f (sysdeps/ieee754/flt-32 s_atanf.c, e_atan2f.c: fdlibm's, Sun Microsystems 1993,
// float version by Cygnus): the argument reduced about 0, 1/2, 1, 3/2 or infinity, an 11-term polynomial in float
// split into odd and even parts, atan(c) in two parts added back. Every operation in float, unfused, as glibc's
// generic build performs it. Equal to the host libm's on every float input (atanf) and on the pairs
// tests/test_devmath.py draws (atan2f).
BL_HD float bl_atanf(float x) {
    const float atanhi[4] = {4.6364760399e-01f, 7.8539812565e-01f, 9.8279368877e-01f, 1.5707962513e+00f};
    const float atanlo[4] = {5.0121582440e-09f, 3.7748947079e-08f, 3.4473217170e-08f, 7.5497894159e-08f};
    const float aT[11] = {3.3333334327e-01f, -2.0000000298e-01f, 1.4285714924e-01f, -1.1111110449e-01f,
                          9.0908870101e-02f, -7.6918758452e-02f, 6.6610731184e-02f, -5.8335702866e-02f,
                          4.9768779427e-02f, -3.6531571299e-02f, 1.6285819933e-02f};
    const uint32_t hx = f2u(x), ix = hx & 0x7fffffffu;
    int id;
    if (ix >= 0x4c000000u) { // |x| >= 2^25
        if (ix > 0x7f800000u) return x + x;
        return (hx >> 31) ? -atanhi[3] - atanlo[3] : atanhi[3] + atanlo[3];
    }
    if (ix < 0x3ee00000u) { // |x| < 0.4375
        if (ix < 0x31000000u) return x; // |x| < 2^-29
        id = -1;
    } else {
        x = ffabs(x);
        if (ix < 0x3f980000u) {      // |x| < 1.1875
            if (ix < 0x3f300000u) {  // 7/16 <= |x| < 11/16
                id = 0, x = (2.0f * x - 1.0f) / (2.0f + x);
            } else {                 // 11/16 <= |x| < 19/16
                id = 1, x = (x - 1.0f) / (x + 1.0f);
            }
        } else if (ix < 0x401c0000u) { // |x| < 2.4375
            id = 2, x = (x - 1.5f) / (1.0f + 1.5f * x);
        } else {                       // 2.4375 <= |x| < 2^25
            id = 3, x = -1.0f / x;
        }
    }
    const float z = x * x, w = z * z;
    const float s1 = z * (aT[0] + w * (aT[2] + w * (aT[4] + w * (aT[6] + w * (aT[8] + w * aT[10])))));
    const float s2 = w * (aT[1] + w * (aT[3] + w * (aT[5] + w * (aT[7] + w * aT[9]))));
    if (id < 0) return x - x * (s1 + s2);
    const float r = atanhi[id] - ((x * (s1 + s2) - atanlo[id]) - x);
    return (hx >> 31) ? -r : r;
}
BL_HD float bl_atan2f(float y, float x) {
    const float pi_o_2 = 1.5707963705e+00f, pi = 3.1415927410e+00f, pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
    const uint32_t hx = f2u(x), hy = f2u(y), ix = hx & 0x7fffffffu, iy = hy & 0x7fffffffu;
    if (ix > 0x7f800000u || iy > 0x7f800000u) return x + y; // NaN
    if (hx == 0x3f800000u) return bl_atanf(y);               // x = 1
    const int m = (int)((hy >> 31) & 1) | (int)((hx >> 30) & 2);
    if (iy == 0) {
        if (m <= 1) return y;
        return m == 2 ? pi + tiny : -pi - tiny;
    }
    if (ix == 0) return (hy >> 31) ? -pi_o_2 - tiny : pi_o_2 + tiny;
    if (ix == 0x7f800000u) {
        const float pi_o_4 = 7.8539818525e-01f;
        if (iy == 0x7f800000u) {
            switch (m) {
            case 0: return pi_o_4 + tiny;
            case 1: return -pi_o_4 - tiny;
            case 2: return 3.0f * pi_o_4 + tiny;
            default: return -3.0f * pi_o_4 - tiny;
            }
        }
        switch (m) {
        case 0: return 0.0f;
        case 1: return -0.0f;
        case 2: return pi + tiny;
        default: return -pi - tiny;
        }
    }
    if (iy == 0x7f800000u) return (hy >> 31) ? -pi_o_2 - tiny : pi_o_2 + tiny;
    const int k = ((int)iy - (int)ix) >> 23;
    float z;
    if (k > 60) z = pi_o_2 + 0.5f * pi_lo;          // |y/x| > 2^60
    else if ((hx >> 31) && k < -60) z = 0.0f;       // |y|/x < -2^60
    else z = bl_atanf(ffabs(y / x));
    switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
    }
}

// ---------------------------------------------------------------------------------------------- acosf
// glibc's float acosf (sysdeps/ieee754/flt-32 e_acosf.c: fdlibm's, float version by Cygnus): acos(x) = pi/2 - x - x r(x^2)
// below |x| 0.5, 2 asin(sqrt((1 - |x|) / 2)) above, r a rational approximation (degree 6 over 4) in float; every
// operation in float, unfused. Equal to the host libm's on every float input (tests/test_devmath.py).
BL_HD float bl_acosf(float x) {
    const float one = 1.0f, pi = 3.1415925026e+00f, pio2_hi = 1.5707962513e+00f, pio2_lo = 7.5497894159e-08f;
    const float pS0 = 1.6666667163e-01f, pS1 = -3.2556581497e-01f, pS2 = 2.0121252537e-01f, pS3 = -4.0055535734e-02f,
                pS4 = 7.9153501429e-04f, pS5 = 3.4793309169e-05f, qS1 = -2.4033949375e+00f, qS2 = 2.0209457874e+00f,
                qS3 = -6.8828397989e-01f, qS4 = 7.7038154006e-02f;
    const uint32_t hx = f2u(x), ix = hx & 0x7fffffffu;
    if (ix == 0x3f800000u) return (hx >> 31) ? pi + 2.0f * pio2_lo : 0.0f; // |x| = 1
    if (ix > 0x3f800000u) return (x - x) / (x - x);                         // |x| > 1: NaN
    if (ix < 0x3f000000u) {                                                 // |x| < 0.5
        if (ix <= 0x23000000u) return pio2_hi + pio2_lo;                   // |x| < 2^-57
        const float z = x * x;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float r = p / q;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if (hx >> 31) { // x < -0.5
        const float z = (one + x) * 0.5f;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float s = fsqrt(z);
        const float r = p / q;
        const float w = r * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    // x > 0.5
    const float z = (one - x) * 0.5f;
    const float s = fsqrt(z);
    const float df = u2f(f2u(s) & 0xfffff000u);
    const float c = (z - df * df) / (s + df);
    const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const float r = p / q;
    const float w = r * s + c;
    return 2.0f * (df + w);
}

// The ground models (src/phys/static_world.cpp:85-115: RoR's ground_models.cfg values; alpha 2, strength 1 for all).
struct ground_model {
    float va, ms, mc, t2, vs, alpha, strength, soft_depth, fluid_drag, fluid_density;
};
#define BL_SURF_N 11
#define BL_GROUND_MODELS                                                                                              \
    {{3, 0.80f, 0.55f, 0.005f, 7, 2, 1, 0, 0, 0},        {3, 0.80f, 0.58f, 0.006f, 4, 2, 1, 0, 0, 0},              \
     {3, 1.20f, 0.75f, 0.010f, 6, 2, 1, 0, 0, 0},        {3, 1.20f, 0.75f, 0.010f, 6, 2, 1, 0, 0, 0},              \
     {3, 0.85f, 0.60f, 0.006f, 3, 2, 1, 0, 0, 0},        {2, 0.50f, 0.35f, 0.004f, 3, 2, 1, 0.12f, 60, 1500},      \
     {3, 0.70f, 0.55f, 0.0001f, 9, 2, 1, 0.03f, 20, 1600}, {3, 0.95f, 0.70f, 0.007f, 8, 2, 1, 0, 0, 0},            \
     {3, 0.90f, 0.60f, 0.005f, 5, 2, 1, 0, 0, 0},        {3, 0.70f, 0.40f, 0.001f, 3, 2, 1, 0, 0, 0},              \
     {1, 0.30f, 0.20f, 0.0001f, 3, 2, 1, 0, 0, 0}}

// (sv / vs)^alpha for the models' alpha = 2: x * x, which equals libm's powf(x, 2) for every finite x (checked
// exhaustively by tests/test_devmath.py); the models have no other exponent.
BL_HD float pow_alpha(float x, float alpha) {
    (void)alpha;
    return x * x;
}

// RoR's primitiveCollision (src/phys/static_world.cpp:8-51).
BL_HD f3 primitive_collision(f3 F, f3 vel, float mass, f3 normal, float dt, const ground_model& gm, float pen,
                             float friction_coef, float push_max, float slop, float restitution) {
    f3 out = mk(0, 0, 0);
    const float Vn = dot(vel, normal);
    const float Fn = dot(F, normal);
    const float sgl = gm.soft_depth;
    if (sgl > 0 && pen >= 0) {
        out = add(out, mul(vel, -gm.fluid_drag * smin(1.0f, pen / sgl + 0.2f)));
        float Fb = gm.fluid_density * smin(pen, sgl) * 9.807f * 0.02f;
        if (Vn >= 0 && Fn < 0 && Fb > -Fn) Fb = -Fn;
        out = add(out, mul(normal, Fb));
    }
    if (pen >= sgl) {
        float R = -Fn;
        const float take = restitution >= 0 ? 1.0f + smin(restitution, 0.95f) : 0.8f;
        if (Vn < 0) R -= (take * Vn - smin(0.2f * smax(0.0f, pen - sgl - slop) / dt, push_max)) * mass / dt;
        if (R > 0) {
            const f3 slipF = sub(F, mul(normal, Fn));
            f3 slip = sub(vel, mul(normal, Vn));
            const float sv = length(slip);
            if (sv > 1e-9f) slip = divs(slip, sv);
            else slip = mk(0, 0, 0);
            const float G = R * gm.strength * friction_coef;
            const float stop = mass * sv / dt;
            if (sv < gm.va && G > 0 && length(slipF) <= gm.ms * G) {
                out = add(out, sub(sub(mul(normal, R), mul(slip, smin(gm.ms * G * (1.0f - bl_expf(-sv / gm.va)), stop))), slipF));
            } else {
                const float g = gm.mc + (gm.ms - gm.mc) * bl_expf(-pow_alpha(sv / gm.vs, gm.alpha));
                out = add(out, sub(mul(normal, R), mul(slip, smin((g + smin(gm.t2 * sv, 5.0f)) * G, smax(0.0f, stop + dot(slipF, slip))))));
            }
        }
    }
    return out;
}

// tyre_contact (src/phys/static_world.cpp:53-80): a tyre node's normal force, the tangential force that would hold it
// (need) up to its grip (cap), and the force it slides with
struct tyre_hit {
    f3 normal_force, need, slide;
    float cap;
    bool touching;
};
BL_HD tyre_hit tyre_contact(f3 F, f3 vel, float mass, f3 normal, float dt, const ground_model& gm, float pen,
                            float friction_coef) {
    tyre_hit t{mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0), 0.0f, false};
    const float Vn = dot(vel, normal);
    const float Fn = dot(F, normal);
    const float sgl = gm.soft_depth;
    if (sgl > 0 && pen >= 0) {
        t.normal_force = add(t.normal_force, mul(vel, -gm.fluid_drag * smin(1.0f, pen / sgl + 0.2f)));
        float Fb = gm.fluid_density * smin(pen, sgl) * 9.807f * 0.02f;
        if (Vn >= 0 && Fn < 0 && Fb > -Fn) Fb = -Fn;
        t.normal_force = add(t.normal_force, mul(normal, Fb));
    }
    if (pen < sgl) return t;
    float R = -Fn;
    if (Vn < 0) R -= (0.8f * Vn - smin(0.2f * (pen - sgl) / dt, 2.0f)) * mass / dt;
    if (R <= 0) return t;
    t.touching = true;
    t.normal_force = add(t.normal_force, mul(normal, R));
    const f3 slipF = sub(F, mul(normal, Fn));
    const f3 slipV = sub(vel, mul(normal, Vn));
    const float G = R * gm.strength * friction_coef;
    t.need = sub(mk(-slipF.x, -slipF.y, -slipF.z), mul(slipV, 0.5f * mass / dt));
    t.cap = gm.ms * G;
    const float sv = length(slipV);
    const float mu_slide = smax(gm.mc, 0.85f * gm.ms);
    const float g = mu_slide + (gm.ms - mu_slide) * bl_expf(-pow_alpha(sv / gm.vs, gm.alpha)) + smin(gm.t2 * sv, 5.0f);
    const float nl = length(t.need);
    const f3 dir = sv > 0.05f ? mul(slipV, -1.0f / sv) : (nl > 1e-6f ? mul(t.need, 1.0f / nl) : mk(0, 0, 0));
    t.slide = mul(dir, g * G);
    return t;
}

} // namespace blg
