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

} // namespace blg
