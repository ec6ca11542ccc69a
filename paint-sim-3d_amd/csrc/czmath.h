/*
 * czmath.h - FP32 elementary functions and f16 conversions of the film step,
 * written for this project from published methods so that the host build
 * and the gfx950 build of the same call give the same bits.
 *
 * The reference evaluates exp/log/pow/atan/atan2 through the WGSL built-ins
 * (the shaders, lowered by naga 30.0.1 and the driver); WGSL bounds their
 * accuracy only loosely (exp 3+2|x| ULP, pow inherited from
 * exp2(y*log2(x)), atan 4096 ULP).  These restatements are far tighter:
 *   exp   Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-7 Taylor
 *         polynomial in Horner form, scaling by 2^k split in two factors so
 *         that subnormal results round once.                (<= 2 ULP)
 *   log   x = 2^e m, m in [sqrt(1/2), sqrt(2)], log(m) = 2 atanh(s),
 *         s = (m-1)/(m+1), odd series to s^11.               (<= 2 ULP)
 *   pow   exp(y log x), the WGSL definition; x^2.5 as x^2 sqrt(x).
 *   atan  reduction to |t| <= tan(pi/8) by 1/x and (x-1)/(x+1), odd Taylor
 *         series to t^19.                                    (<= 2 ULP)
 *   f16   IEEE round-to-nearest-even with subnormals (pack2x16float), and
 *         the exact inverse (unpack2x16float); the device build uses the
 *         hardware conversions, which round the same way.
 * Every operation is an IEEE-754 single operation or an explicit fmaf;
 * translation units that include this header are built without FP
 * contraction, so host and device round identically.
 *
 * sqrt and division are the correctly rounded IEEE operations on both sides.
 */
#ifndef CZMATH_H
#define CZMATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define CZ_HD __host__ __device__
#else
#define CZ_HD
#endif

#define CZ_INLINE static inline CZ_HD

CZ_INLINE float cz_bits2f(uint32_t u) {
    union { uint32_t u; float f; } v;
    v.u = u;
    return v.f;
}

CZ_INLINE uint32_t cz_f2bits(float f) {
    union { uint32_t u; float f; } v;
    v.f = f;
    return v.u;
}

/* WGSL max/min: max(a, b) = b if a < b else a; min(a, b) = b if b < a else a. */
CZ_INLINE float cz_max(float a, float b) { return (a < b) ? b : a; }
CZ_INLINE float cz_min(float a, float b) { return (b < a) ? b : a; }
CZ_INLINE float cz_clamp(float x, float lo, float hi) { return cz_min(cz_max(x, lo), hi); }
CZ_INLINE float cz_sign(float x) { return (x > 0.0f) ? 1.0f : ((x < 0.0f) ? -1.0f : 0.0f); }
CZ_INLINE float cz_abs(float x) { return fabsf(x); }
/* WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3. */
CZ_INLINE float cz_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
CZ_INLINE float cz_len2(float x, float y) { return sqrtf(x * x + y * y); }
CZ_INLINE int cz_max_i(int a, int b) { return a < b ? b : a; }
CZ_INLINE int cz_min_i(int a, int b) { return b < a ? b : a; }
CZ_INLINE int cz_clamp_i(int x, int lo, int hi) { return cz_min_i(cz_max_i(x, lo), hi); }

CZ_INLINE float cz_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return cz_bits2f(0x7f800000u);
    if (x < -103.972084045410156f) return 0.0f;
    float k = rintf(x * 1.44269502162933350f);
    float r = fmaf(-k, 0.693145751953125f, x);
    r = fmaf(-k, 1.42860676533018704e-06f, r);
    float p = 1.98412698e-4f;
    p = fmaf(p, r, 1.38888889e-3f);
    p = fmaf(p, r, 8.33333333e-3f);
    p = fmaf(p, r, 4.16666667e-2f);
    p = fmaf(p, r, 1.66666667e-1f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int ki = (int)k;
    int k1 = ki / 2;
    int k2 = ki - k1;
    float s1 = cz_bits2f((uint32_t)(k1 + 127) << 23);
    float s2 = cz_bits2f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

CZ_INLINE float cz_logf(float x) {
    if (!(x > 0.0f)) {
        if (x == 0.0f) return cz_bits2f(0xff800000u);
        return cz_bits2f(0x7fc00000u);
    }
    uint32_t u = cz_f2bits(x);
    if (u >= 0x7f800000u) return x;
    int e = 0;
    if (u < 0x00800000u) {
        x = x * 8388608.0f;
        u = cz_f2bits(x);
        e = -23;
    }
    e += (int)(u >> 23) - 127;
    float m = cz_bits2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 9.09090909e-2f;
    p = fmaf(p, z, 1.11111111e-1f);
    p = fmaf(p, z, 1.42857143e-1f);
    p = fmaf(p, z, 2.0e-1f);
    p = fmaf(p, z, 3.33333333e-1f);
    float two_s = 2.0f * s;
    float lg = fmaf(two_s * z, p, two_s);
    float fe = (float)e;
    return fmaf(fe, 0.693145751953125f, fmaf(fe, 1.42860676533018704e-06f, lg));
}

CZ_INLINE float cz_powf(float x, float y) { return cz_expf(y * cz_logf(x)); }

/* x^y for x >= 0 and the capillary cap's exponent y: x^2.5 (its default) as
 * x^2 sqrt(x), equal in real arithmetic and within 2 ULP of the exact
 * power (three roundings); any other exponent through cz_powf. */
CZ_INLINE float cz_pow_fcap(float x, float y) { return (y == 2.5f) ? (x * x) * sqrtf(x) : cz_powf(x, y); }

/* max(k, 1)^0.7 for the feeder count k = 0..4 of a face: the values cz_powf
 * gives (1 and the bit patterns of cz_powf(2, 0.7), cz_powf(3, 0.7),
 * cz_powf(4, 0.7)), so a lookup returns the same bits as the call. */
CZ_INLINE float cz_feed_pow(uint32_t k) {
    return (k <= 1u) ? 1.0f
         : (k == 2u) ? cz_bits2f(0x3fcfefc6u)
         : (k == 3u) ? cz_bits2f(0x400a1741u)
                     : cz_bits2f(0x4028e5a3u);
}

CZ_INLINE float cz_atanf(float x) {
    if (x != x) return x;
    float ax = fabsf(x);
    int inv = ax > 1.0f;
    if (inv) ax = 1.0f / ax;
    int red = ax > 0.414213562f;
    if (red) ax = (ax - 1.0f) / (ax + 1.0f);
    float z = ax * ax;
    float p = -5.26315789e-2f;       /* -1/19 */
    p = fmaf(p, z, 5.88235294e-2f);  /*  1/17 */
    p = fmaf(p, z, -6.66666667e-2f); /* -1/15 */
    p = fmaf(p, z, 7.69230769e-2f);  /*  1/13 */
    p = fmaf(p, z, -9.09090909e-2f); /* -1/11 */
    p = fmaf(p, z, 1.11111111e-1f);  /*  1/9  */
    p = fmaf(p, z, -1.42857143e-1f); /* -1/7  */
    p = fmaf(p, z, 2.0e-1f);         /*  1/5  */
    p = fmaf(p, z, -3.33333333e-1f); /* -1/3  */
    float r = fmaf(ax * z, p, ax);
    if (red) r = r + 0.785398163f;
    if (inv) r = 1.57079633f - r;
    return (x < 0.0f) ? -r : r;
}

CZ_INLINE float cz_atan2f(float y, float x) {
    if (x > 0.0f) return cz_atanf(y / x);
    if (x < 0.0f) {
        float a = cz_atanf(y / x);
        return (y >= 0.0f) ? a + 3.14159265f : a - 3.14159265f;
    }
    if (y > 0.0f) return 1.57079633f;
    if (y < 0.0f) return -1.57079633f;
    return 0.0f;
}

/* f32 -> binary16 bits, round to nearest even, with subnormals. */
CZ_INLINE uint32_t cz_f2h_sw(float f) {
    uint32_t x = cz_f2bits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u);
    if (ax >= 0x477ff000u) return sign | 0x7c00u;
    if (ax >= 0x38800000u) {
        uint32_t mant = ax & 0x7fffffu;
        uint32_t e = (ax >> 23) - 112u;
        uint32_t h = (e << 10) | (mant >> 13);
        uint32_t rem = mant & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
        return sign | h;
    }
    if (ax < 0x33000000u) return sign;
    uint32_t e = ax >> 23;
    uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126u - e;
    uint32_t mm = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (mm & 1u))) mm += 1u;
    return sign | mm;
}

/* binary16 bits -> f32 (exact). */
CZ_INLINE float cz_h2f_sw(uint32_t h) {
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0u) {
        float v = (float)m * 5.9604644775390625e-08f;
        return sign ? -v : v;
    }
    if (e == 31u) return cz_bits2f(sign | 0x7f800000u | (m << 13));
    return cz_bits2f(sign | ((e + 112u) << 23) | (m << 13));
}

#if defined(__HIP_DEVICE_COMPILE__)
/* On the device the same two conversions are single instructions:
 * v_cvt_f16_f32 rounds to nearest even and keeps f16 subnormals (the default
 * MODE of HIP kernels), v_cvt_f32_f16 is exact.  They give the software
 * forms' bits for every non-NaN input (cz_f16_selftest checks all 2^32 f32
 * and 2^16 f16 patterns on the device); only NaN payloads may differ, and
 * the film state holds no NaN. */
CZ_INLINE uint32_t cz_f2h(float f) { return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)f); }
CZ_INLINE float cz_h2f(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (unsigned short)h); }
#else
CZ_INLINE uint32_t cz_f2h(float f) { return cz_f2h_sw(f); }
CZ_INLINE float cz_h2f(uint32_t h) { return cz_h2f_sw(h); }
#endif

/* WGSL unpack2x16float / pack2x16float. */
CZ_INLINE float cz_unpack_lo(uint32_t v) { return cz_h2f(v & 0xffffu); }
CZ_INLINE float cz_unpack_hi(uint32_t v) { return cz_h2f(v >> 16); }
CZ_INLINE uint32_t cz_pack2(float a, float b) { return cz_f2h(a) | (cz_f2h(b) << 16); }

#endif /* CZMATH_H */
