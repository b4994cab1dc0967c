// sppm_device.h -- device-side building blocks of the SPPM pass core (gfx950).
//
// Float semantics follow the reference exactly (FriendsRender R, SPPM/SPPM/):
//  * PTUtility::Vec3 op order (Vec.h:179-243): dot = (x*x' + y*y') + z*z',
//    normalized() divides by (float)sqrt((double)norm2), which is the
//    correctly rounded float sqrt (sqrt_narrowed below).
//  * Unqualified sqrt/cos/sin/exp of a float in the reference bind to the
//    double functions; they are evaluated in FP64 here and narrowed.
//  * std::pow(float, 2.0f) is x*x (g++ folds it), std::pow(float, int) is
//    a double x*x, std::pow(float, -2.0f) is powf.
//  * No FMA contraction: everything is built with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_layout.h"

namespace sppm {

#define SPPM_PI 3.14159f          // PTUtility::PI, Vec.h:10
#define SPPM_RAY_OFFSET 1e-4f     // Ray::RAY_OFFSET, Primitive.h:31
#define SPPM_SMIN 1e-16f          // Sampling::sMin, Sampling.h:8
#define SPPM_DELTA_PDF 1.0f       // Sampling::DeltaPDF, Sampling.h:7

// --------------------------------------------------------------------------
// Vec3 with the reference's operation order
// --------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
// Vec3 * float and float * Vec3 both compute s*x per component (Vec.h:303-308)
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float norm2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// The reference's (float)sqrt((double)x) of a float x equals the correctly
// rounded float sqrt of x: rounding the exact root to double (53 bits) and
// then to float (24) never differs from rounding it to float once, since
// 53 >= 2*24 + 2 (Figueroa, "When is double rounding innocuous?", 1995).
// gfx950's v_sqrt_f32 is refined to the correctly rounded result by the
// compiler, so the float form is bit-identical and avoids an FP64 sequence.
__device__ __forceinline__ float sqrt_narrowed(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float norm(f3 a) { return sqrt_narrowed(norm2(a)); }
__device__ __forceinline__ f3 normalized(f3 a) {
    float d = norm(a);
    return mk(a.x / d, a.y / d, a.z / d);
}
// The reference compares floats with double constants ((double)x > c).  For
// a float x that equals one float comparison with c's float neighbour on the
// right side: x > c  <=>  x > lo(c), the largest float <= c, and x < c  <=>
// x < hi(c), the smallest float >= c (NaN is false both ways).  No FP64
// conversion or comparison is then issued.
constexpr float LO_0P999999 = 0x1.ffffdep-1f;  // lo(0.999999)
constexpr float LO_0P99 = 0x1.fae146p-1f;      // lo(0.99)
constexpr float LO_0P9 = 0x1.ccccccp-1f;       // lo(0.9)
constexpr float LO_1EM4 = 0x1.a36e2ep-14f;     // lo(1e-4)
constexpr float LO_1EM5 = 0x1.4f8b58p-17f;     // lo(1e-5)
constexpr float HI_1EM5 = 0x1.4f8b5ap-17f;     // hi(1e-5)
constexpr float LO_1EM6 = 0x1.0c6f7ap-20f;     // lo(1e-6)
constexpr float HI_1EM6 = 0x1.0c6f7cp-20f;     // hi(1e-6)
constexpr float LO_1EM8 = 0x1.5798eep-27f;     // lo(1e-8)
constexpr float HI_1EM8 = 0x1.5798f0p-27f;     // hi(1e-8)

__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ f3 ld3(const float4 v) { return mk(v.x, v.y, v.z); }
// std::max(a,b) = (a<b)?b:a, std::min(a,b) = (b<a)?b:a
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fabs_(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ float sqrtf_(float a) { return __builtin_sqrtf(a); }

// --------------------------------------------------------------------------
// Per-path std::mt19937 + generate_canonical<double,53> (MT.h:7-44).
//
// Each path seeds a fresh engine.  Output k (< 227) of the first twist only
// needs initial-state words k, k+1 and k+397, so two forward cursors over
// the seeding recurrence replace the 624-word state: 5 registers per path.
// --------------------------------------------------------------------------
struct MTCursor {
    uint32_t a0, a1, b;
    uint32_t ia, ib;
    uint32_t k;
};
__device__ __forceinline__ uint32_t mt_init_step(uint32_t prev, uint32_t i) {
    return 1812433253u * (prev ^ (prev >> 30)) + i;
}
__device__ __forceinline__ void mt_seed(MTCursor& m, uint32_t seed) {
    m.a0 = seed;
    m.a1 = mt_init_step(seed, 1u);
    m.ia = 1u;
    uint32_t t = m.a1;
#pragma unroll 8
    for (uint32_t i = 2; i <= 397u; ++i) t = mt_init_step(t, i);
    m.b = t;
    m.ib = 397u;
    m.k = 0;
}
// Returns the next tempered 32-bit output; sets *ovf when the path would need
// more than 227 outputs (never happens within the reference's loop bounds).
__device__ __forceinline__ uint32_t mt_next(MTCursor& m, uint32_t* ovf) {
    const uint32_t UP = 0x80000000u, LO = 0x7fffffffu, A = 0x9908b0dfu;
    if (m.k >= 227u) *ovf = 1u;
    uint32_t y = (m.a0 & UP) | (m.a1 & LO);
    uint32_t z = m.b ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
    m.a0 = m.a1;
    m.ia += 1u;
    m.a1 = mt_init_step(m.a1, m.ia);
    m.ib += 1u;
    m.b = mt_init_step(m.b, m.ib);
    m.k += 1u;
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}
__device__ __forceinline__ double canonical(uint32_t w0, uint32_t w1) {
    double sum = (double)w0;
    sum = sum + (double)w1 * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;  // nextafter(1.0, 0.0)
    return r;
}
// genrand64_real1; genrand64_real2 is identical (1 - DBL_MIN rounds to 1)
__device__ __forceinline__ double real1(MTCursor& m, uint32_t* ovf) {
    uint32_t w0 = mt_next(m, ovf);
    uint32_t w1 = mt_next(m, ovf);
    return canonical(w0, w1);
}

// --------------------------------------------------------------------------
// Sampling.cpp / RayUtility.cpp
// --------------------------------------------------------------------------
__device__ __forceinline__ f3 invalid_dir() { return mk(1e-3f, 1e-3f, 1e-3f); }  // RUtil::INVALID_DIR

__device__ __forceinline__ f3 reflect_sampling(f3 n, f3 in) {  // Sampling.cpp:293-295
    return normalized(in - n * (2.0f * dot(n, in)));
}
__device__ inline float fresnel_reflectance(f3 n, f3 in, float ii, float io) {  // :297-313
    float n1n2 = ii / io;
    float cosA = fabs_(dot(n, in));
    float sinA = sqrtf_(1.0000001f - cosA * cosA);
    float sinB = n1n2 * sinA;
    if (sinB > LO_0P999999) return 1.0f;  // (double)sinB > 0.999999
    float cosB = sqrtf_(1.0000001f - sinB * sinB);
    float rp = (io * cosA - ii * cosB) / (io * cosA + ii * cosB);
    float rs = (ii * cosA - io * cosB) / (ii * cosA + io * cosB);
    return (0.5f * rs * rs) + (0.5f * rp * rp);
}
__device__ inline f3 refract_sampling(f3 n, f3 in, float ii, float io) {  // Sampling.cpp:329-337
    float eta = ii / io;
    f3 NN = (dot(in, n) > 0) ? -n : n;
    float LN = -dot(in, NN);
    float sq = smax(0.00000001f, 1.0f - eta * eta * (1.0f - LN * LN));
    return normalized((-NN) * sqrtf_(sq) - ((-in) - NN * LN) * eta);
}
__device__ inline f3 rutil_refract(f3 in, f3 n, float from, float into) {  // RayUtility.cpp:10-17
    f3 NN = (dot(in, n) < 0) ? n : -n;
    float eta = from / into;
    float LN = -dot(in, NN);
    float sq = smax(0.0f, 1.0f - eta * eta * (1.0f - LN * LN));
    return normalized((-NN) * sqrtf_(sq) - ((-in) - NN * LN) * eta);
}
__device__ __forceinline__ bool rutil_is_reflect(f3 in, f3 out, f3 n) {  // RayUtility.cpp:34-36
    return dot(n, in) * dot(n, out) < 0.0f;
}
__device__ __forceinline__ bool rutil_valid_dir(f3 d) { return norm2(d) > 0.9999f; }

__device__ __forceinline__ void frame(f3 e1, f3& e2, f3& e3) {  // Sampling.cpp:12-18
    if (fabs_(e1.x) > LO_1EM6) e2 = normalized(cross(mk(0, 1, 0), e1));  // (double)|e1.x| > 0.000001
    else e2 = normalized(cross(mk(1, 0, 0), e1));
    e3 = normalized(cross(e1, e2));
}

// Cos_Sampling (Sampling.cpp:7-28)
__device__ inline f3 cos_sampling(MTCursor& m, uint32_t* ovf, f3 n, float* pdf) {
    f3 e2, e3;
    frame(n, e2, e3);
    float ph = (float)(real1(m, ovf) * 2.0 * (double)SPPM_PI);
    float th = (float)real1(m, ovf);
    float th2 = sqrt_narrowed(th);
    f3 a = n * (float)__builtin_sqrt(1.000000001 - (double)th);
    f3 b = (e2 * (float)cos((double)ph)) * th2;
    f3 c = (e3 * (float)sin((double)ph)) * th2;
    f3 r = normalized((a + b) + c);
    *pdf = smax(SPPM_SMIN, dot(r, n) / SPPM_PI);
    return r;
}

// GGX_Sampling (Sampling.cpp:123-157)
__device__ inline f3 ggx_sampling(MTCursor& m, uint32_t* ovf, f3 n, f3 in, float alpha, float* pdf) {
    float phi = (float)(real1(m, ovf) * 2.0 * (double)SPPM_PI);
    float e = (float)real1(m, ovf);
    float cth = (float)__builtin_sqrt((1.0 - (double)e) / (double)(e * (alpha * alpha - 1.0f) + 1.0f));
    f3 e2, e3;
    frame(n, e2, e3);
    float cph = (float)cos((double)phi);
    float sph = (float)sin((double)phi);
    float sth = sqrtf_(1.0f - cth * cth);
    f3 mN = normalized(((e3 * sth) * cph + (e2 * sth) * sph) + n * cth);
    if ((double)dot(mN, in) > 0.0) { *pdf = SPPM_SMIN; return invalid_dir(); }
    f3 r = reflect_sampling(mN, in);
    if ((double)dot(r, n) < 0.0) { *pdf = SPPM_SMIN; return invalid_dir(); }
    float wm_N = dot(mN, n);
    float wo_wm = dot(r, mN);
    double q = (double)(1.0f - (1.0f - alpha * alpha) * wm_N * wm_N);
    *pdf = (float)((double)(alpha * alpha) / ((double)SPPM_PI * (q * q)) * (double)wm_N /
                   (double)(4.0f * wo_wm));
    return r;
}
// GGX_Sampling_PDF (Sampling.cpp:158-180)
__device__ inline float ggx_sampling_pdf(f3 out, f3 n, f3 in, float a) {
    if (rutil_is_reflect(in, out, n)) {
        f3 h = normalized((-in) + out);
        float h_n = dot(h, n);
        if (h_n < 0.0f) return SPPM_SMIN;
        float wi_h = -dot(in, h);
        if (wi_h < 0.0f) return SPPM_SMIN;
        float wo_h = dot(out, h);
        if (wo_h < 0.0f) return SPPM_SMIN;
        float jc = 1.0f / (4.0f * wo_h);
        return (a * a / SPPM_PI * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) * h_n * jc;
    }
    return SPPM_SMIN;
}
// GGXGlass_Sampling (Sampling.cpp:181-248)
__device__ inline f3 ggxglass_sampling(MTCursor& m, uint32_t* ovf, f3 n, f3 in, float alpha, float ii,
                                       float io, float* pdf) {
    const float SV = 1e-12f;
    f3 NN = dot(in, n) < 0 ? n : -n;
    *pdf = 1.0f;
    float phi = (float)(real1(m, ovf) * 2.0 * (double)SPPM_PI);
    float e = (float)real1(m, ovf);
    float cth = (float)__builtin_sqrt((1.0000001 - (double)e) /
                                      ((double)e * ((double)(alpha * alpha) - 0.99999) + 1.0));
    f3 e2, e3;
    frame(NN, e2, e3);
    float cph = (float)cos((double)phi);
    float sph = (float)sin((double)phi);
    float sth = sqrtf_(1.00000001f - cth * cth);
    f3 mN = normalized(((e3 * sth) * cph + (e2 * sth) * sph) + NN * cth);
    float wi_wm = -dot(in, mN);
    float wm_N = dot(mN, NN);
    float wi_N = -dot(in, NN);
    if (wi_N < SV) { *pdf = SPPM_SMIN; return invalid_dir(); }
    if (wi_wm < SV) { *pdf = SPPM_SMIN; return invalid_dir(); }
    if (wm_N < SV) { *pdf = SPPM_SMIN; return invalid_dir(); }
    float refl = fresnel_reflectance(mN, in, ii, io);
    f3 r;
    if ((double)refl > real1(m, ovf)) {
        *pdf = refl;
        r = reflect_sampling(mN, in);
        float wo_wm = dot(r, mN);
        if (wo_wm < SV) return invalid_dir();
        if (dot(r, NN) * dot(in, NN) > SV) { *pdf = SPPM_SMIN; return invalid_dir(); }
        float t = 1.0f - (1.0f - alpha * alpha) * wm_N * wm_N;
        *pdf = refl * (wm_N / (4.0f * wo_wm) * alpha * alpha / (SPPM_PI * (t * t)));
    } else {
        r = rutil_refract(in, mN, ii, io);
        float wo_wm = -dot(r, mN);
        float wo_N = -dot(r, NN);
        if (wi_N * wo_N < SV || wo_N < SV || wo_wm < SV) { *pdf = SPPM_SMIN; return invalid_dir(); }
        double s = (double)(ii * wi_wm + io * wo_wm);
        float t = 1.0f - (1.0f - alpha * alpha) * wm_N * wm_N;
        double val = (double)((1.0f - refl) * wm_N) * ((double)(io * io * wo_wm) / (s * s));
        val = val * (double)alpha;
        val = val * (double)alpha;
        *pdf = (float)(val / (double)(SPPM_PI * (t * t)));
    }
    return r;
}
// GGXGlass_Sampling_PDF (Sampling.cpp:249-288)
__device__ inline float ggxglass_sampling_pdf(f3 out, f3 n, f3 in, float a, float ii, float io) {
    if (rutil_is_reflect(in, out, n)) {
        f3 h = normalized((-in) + out);
        float h_n = dot(h, n);
        if (h_n < 0.0f) return SPPM_SMIN;
        float wi_h = -dot(in, h);
        if (wi_h < 0.0f) return SPPM_SMIN;
        float wo_h = dot(out, h);
        if (wo_h < 0.0f) return SPPM_SMIN;
        float refl = fresnel_reflectance(h, in, ii, io);
        float jc = 1.0f / (4.0f * wo_h);
        return refl * h_n * (a * a / SPPM_PI * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) * jc;
    }
    // GetMicrofacetNormal_FromRefraction (RayUtility.cpp:21-33)
    f3 h;
    {
        float c = fabs_(dot(in, out));
        float eta = io / ii;
        float isq = eta * eta - 2.0f * eta * c + 1.0f;
        if (isq < HI_1EM8) h = mk(0, 0, 0);  // (double)isq < 1e-8
        else {
            float ss = 1.0f / sqrtf_(isq);
            h = normalized(out * (-eta * ss) + in * ss);
            if (!(io > ii)) h = -h;
        }
    }
    float h_n = dot(h, n);
    if (h_n < 0.0f) return SPPM_SMIN;
    float wi_h = -dot(in, h);
    if (wi_h < 0.0f) return SPPM_SMIN;
    float wo_h = -dot(out, h);
    if (wo_h < 0.0f) return SPPM_SMIN;
    float refl = fresnel_reflectance(h, in, ii, io);
    float jc = io * io * wo_h * powf(ii * wi_h + io * wo_h, -2.0f);
    return (1.0f - refl) * h_n * (a * a / SPPM_PI * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) * jc;
}

// --------------------------------------------------------------------------
// Surfaces (material.h).  FUNCTION_PARAM (material.h:12-34)
// --------------------------------------------------------------------------
struct FParam {
    f3 in, out, n;
    float from, to, rough, absorb, refl;
};

__device__ __forceinline__ bool opaque_reflect(const FParam& p) {  // material.h:139-141
    return (dot(p.in, -p.n) > LO_1EM4 && dot(p.out, p.n) > LO_1EM4);  // (double)... > 1e-4
}

__device__ inline float brdf_rough_conductor(const FParam& p) {  // material.h:227-274
    const float sv = 1e-10f;
    if (dot(p.in, p.n) > -sv || dot(p.out, p.n) < sv) return 0.0f;
    float eta = p.to / p.from;
    float k = p.absorb;
    float alpha = p.rough;
    f3 h = normalized((-p.in) + p.out);
    float ccs = dot(h, p.n);
    if (ccs < sv || dot(h, p.in) > -sv || dot(h, p.out) < sv) return 0.0f;
    float q = (1.0001f - (1.0f - alpha * alpha) * ccs * ccs);
    float D = smax(SPPM_SMIN, (alpha * alpha) / (SPPM_PI * (q * q)));
    float lc = smax(0.0f, smin(0.9999f, dot(p.in, -p.n)));
    float lv = smax(0.0f, smin(0.9999f, dot(p.out, p.n)));
    if (lc < SPPM_SMIN || lv < SPPM_SMIN) return 0.0f;
    float L_l = (float)((-1.0 + __builtin_sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lc * lc)) - 1.0))) * 0.5);
    float L_v = (float)((-1.0 + __builtin_sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lv * lv)) - 1.0))) * 0.5);
    float G = 1.0f / ((1.0f + L_l) * (1.0f + L_v));
    float fcos = dot(p.n, -p.in);
    float temp = eta * eta - k * k - (1.0f - fcos * fcos);
    float apb = sqrtf_(temp * temp + k * k * eta * eta);
    float a = (float)__builtin_sqrt((double)(apb + temp) * 0.5);
    float tm1 = apb + fcos * fcos;
    float tm2 = a * 2.0f * fcos;
    float rs2 = (tm1 - tm2) / (tm1 + tm2);
    float tm3 = (float)((double)(apb * fcos * fcos) + (double)(1.0f - fcos * fcos) * (1.0 - (double)(fcos * fcos)));
    float tm4 = tm2 * (1.0f - fcos * fcos);
    float rp2 = rs2 * (tm3 - tm4) / (tm3 + tm4);
    float F = (float)(0.5 * (double)(rp2 + rs2));
    float fx = (4.0f * dot(p.in, -p.n) * dot(p.out, p.n));
    return (F * G * D) / fx;
}

__device__ inline float brdf_ggx_glass(const FParam& p) {  // material.h:314-390
    const float SV = 1e-8f;
    const bool isReflect = dot(p.n, p.in) * dot(p.n, p.out) < 0;
    const float a = p.rough;
    const float Ii = p.from, Io = p.to;
    if (isReflect) {
        f3 h = normalized((-p.in) + p.out);
        float n_h = dot(p.n, h);
        if (n_h < 0) return 0.0f;
        float D = (a * a) * powf(1.0f - (1.0f - a * a) * n_h * n_h, -2.0f) / SPPM_PI;
        float L_l = (-1.0f + sqrtf_(1.0001f + a * a * (powf(dot(p.n, -p.in), -2.0f) - 1.0f))) * 0.5f;
        float L_v = (-1.0f + sqrtf_(1.0001f + a * a * (powf(dot(p.n, p.out), -2.0f) - 1.0f))) * 0.5f;
        float G = 1.0f / (L_l + L_v + 1.0f);
        float F = fresnel_reflectance(h, p.in, Ii, Io);
        float wo_n = smax(0.0f, smin(1.0f, dot(p.out, h)));
        float wi_n = smax(0.0f, smin(1.0f, -dot(p.in, h)));
        if (wo_n < SV || wi_n < SV) return 0.0f;
        return F * G * D / (4.0f * wo_n * wi_n);
    }
    const float eta = Io / Ii;
    if (fabs_(eta - 1.0f) < SV) return 1.0f;
    if ((double)dot(p.in, p.out) < 0.0) return 0.0f;
    f3 h;  // EstimateN (material.h:297-309)
    {
        float c = fabs_(dot(p.in, p.out));
        float et = Io / Ii;
        float isq = et * et - 2.0f * et * c + 1.0f;
        if (isq < SPPM_SMIN) h = mk(0, 0, 0);
        else {
            float ss = 1.0f / sqrtf_(isq);
            h = normalized(p.out * (-et * ss) + p.in * ss);
            if (!(Io > Ii)) h = -h;
        }
    }
    if (norm(h) < SV) return 0.0f;
    float n_h = dot(h, p.n);
    if (n_h < SV) return 0.0f;
    float D = (a * a) * powf(1.0f - (1.0f - a * a) * n_h * n_h, -2.0f) / SPPM_PI;
    float L_l = (-1.0f + sqrtf_(1.0001f + a * a * (powf(dot(p.n, -p.in), -2.0f) - 1.0f))) * 0.5f;
    float L_v = (-1.0f + sqrtf_(1.0001f + a * a * (powf(dot(p.n, -p.out), -2.0f) - 1.0f))) * 0.5f;
    float G = 1.0f / (L_l + L_v + 1.0f);
    float F = fresnel_reflectance(h, p.in, Ii, Io);
    float wi_wm = -dot(p.in, h);
    float wo_wm = -dot(p.out, h);
    float wi_n = -dot(p.in, p.n);
    float wo_n = -dot(p.out, p.n);
    if (wo_wm < SV || wi_wm < SV) return 0.0f;
    if (wi_n < SV || wo_n < SV) return 0.0f;
    return (wi_wm * wo_wm) / (wi_n * wo_n) * (Io * Io) * (1.0f - F) * G * D * powf(Ii * wo_wm + Io * wi_wm, -2.0f);
}

// GlossTransmitSurface::ReflectByN (material.h:692-695): mirror r about the
// plane of n (its unused normalised tangent is not formed).
__device__ __forceinline__ f3 reflect_by_n(f3 r, f3 n) { return r - n * (2.0f * dot(n, r)); }

__device__ inline float brdf_gloss_transmit(const FParam& p) {  // material.h:719-756
    const float power = dot(p.n, p.in) * dot(p.n, p.out) < 0.0f ? p.refl : 1.0f - p.refl;
    const f3 N = dot(p.n, p.in) < 0 ? p.n : -p.n;
    f3 out = p.out;
    if (dot(out, N) < 0) out = reflect_by_n(out, N);
    const float sv = 1e-10f;
    if (dot(p.in, N) > -sv || dot(out, N) < sv) return 0.0f;
    const float alpha = p.rough;
    const f3 h = normalized((-p.in) + out);
    const float ccs = dot(h, N);
    if (ccs < sv || dot(h, p.in) > -sv || dot(h, out) < sv) return 0.0f;
    const float q = (1.0001f - (1.0f - alpha * alpha) * ccs * ccs);
    const float D = smax(SPPM_SMIN, (alpha * alpha) / (SPPM_PI * (q * q)));
    const float lc = smax(0.0f, smin(0.9999f, dot(p.in, -N)));
    const float lv = smax(0.0f, smin(0.9999f, dot(out, N)));
    if (lc < SPPM_SMIN || lv < SPPM_SMIN) return 0.0f;
    const float L_l =
        (float)((-1.0 + __builtin_sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lc * lc)) - 1.0))) * 0.5);
    const float L_v =
        (float)((-1.0 + __builtin_sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lv * lv)) - 1.0))) * 0.5);
    const float G = 1.0f / ((1.0f + L_l) * (1.0f + L_v));
    const float fx = (4.0f * dot(p.in, -N) * dot(out, N));
    return power * (G * D) / fx;
}

__device__ inline float surf_brdf(int kind, const FParam& p) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:  // material.h:168-173
        return opaque_reflect(p) ? 1.0f / SPPM_PI : 0.0f;
    case SPPM_MAT_METAL: {  // material.h:404-415
        if (!opaque_reflect(p)) return 0.0f;
        f3 ref = reflect_sampling(p.n, p.in);
        if (dot(ref, p.out) > 1.0f - 1.0f / SPPM_DELTA_PDF)
            return (float)((double)SPPM_DELTA_PDF / ((double)dot(p.out, p.n) + 0.00001));
        return 0.0f;
    }
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return brdf_rough_conductor(p);
    case SPPM_MAT_GGX_GLASS:
        return brdf_ggx_glass(p);
    case SPPM_MAT_GLASS: {  // material.h:499-508
        f3 NN = dot(p.n, p.in) < 0 ? p.n : -p.n;
        float refl = fresnel_reflectance(NN, p.in, p.from, p.to);
        double denom = (double)fabs_(dot(p.out, NN)) + 0.0001;
        if (dot(p.n, p.in) * dot(p.n, p.out) < 0.0f) return (float)((double)(SPPM_DELTA_PDF * refl) / denom);
        return (float)((double)(SPPM_DELTA_PDF * (1.0f - refl)) / denom);
    }
    case SPPM_MAT_BIMETAL: {  // material.h:437-449
        const f3 N = dot(p.n, p.in) < 0 ? p.n : -p.n;
        if (!opaque_reflect(p)) return 0.0f;
        f3 ref = reflect_sampling(N, p.in);
        if (dot(ref, p.out) > 1.0f - 1.0f / SPPM_DELTA_PDF)
            return (float)((double)SPPM_DELTA_PDF / ((double)dot(p.out, N) + 0.00001));
        return 0.0f;
    }
    case SPPM_MAT_DIFFUSE_TRANSMIT:  // material.h:666-674
        return dot(p.n, p.in) * dot(p.n, p.out) < 0.0f ? p.refl / SPPM_PI : (1.0f - p.refl) / SPPM_PI;
    case SPPM_MAT_GLOSS_TRANSMIT:
        return brdf_gloss_transmit(p);
    }
    return 0.0f;
}

// Roughconductor::BRDF (material.h:227-274) for the gather, where the value
// only scales a hit point's flux (SPPM.cpp:440) and never decides a path:
// the same terms in FP32 with hardware reciprocals, square roots and
// inverse square roots (about 1 ulp each) instead of the reference's FP64
// pieces and correctly rounded divisions.  Relative error ~1e-6, far inside
// the flux tolerance; photon and eye paths keep brdf_rough_conductor.
__device__ inline float brdf_rough_conductor_flux(const FParam& p) {
    const float sv = 1e-10f;
    const float cin = dot(p.in, p.n), cout = dot(p.out, p.n);
    if (cin > -sv || cout < sv) return 0.0f;
    const float eta = p.to * __builtin_amdgcn_rcpf(p.from);
    const float k = p.absorb;
    const float alpha = p.rough, a2 = alpha * alpha;
    const f3 hv = (-p.in) + p.out;
    const f3 h = hv * __builtin_amdgcn_rsqf(norm2(hv));
    const float ccs = dot(h, p.n);
    if (ccs < sv || dot(h, p.in) > -sv || dot(h, p.out) < sv) return 0.0f;
    const float q = 1.0001f - (1.0f - a2) * ccs * ccs;
    const float D = smax(SPPM_SMIN, a2 * __builtin_amdgcn_rcpf(SPPM_PI * (q * q)));
    const float lc = smax(0.0f, smin(0.9999f, -cin));
    const float lv = smax(0.0f, smin(0.9999f, cout));
    if (lc < SPPM_SMIN || lv < SPPM_SMIN) return 0.0f;
    const float L_l = 0.5f * (__builtin_amdgcn_sqrtf(1.0f + a2 * (__builtin_amdgcn_rcpf(lc * lc) - 1.0f)) - 1.0f);
    const float L_v = 0.5f * (__builtin_amdgcn_sqrtf(1.0f + a2 * (__builtin_amdgcn_rcpf(lv * lv) - 1.0f)) - 1.0f);
    const float G = __builtin_amdgcn_rcpf((1.0f + L_l) * (1.0f + L_v));
    const float fcos = -cin;
    const float s2 = 1.0f - fcos * fcos;
    const float temp = eta * eta - k * k - s2;
    const float apb = __builtin_amdgcn_sqrtf(temp * temp + k * k * eta * eta);
    const float a = __builtin_amdgcn_sqrtf((apb + temp) * 0.5f);
    const float tm1 = apb + fcos * fcos;
    const float tm2 = a * 2.0f * fcos;
    const float rs2 = (tm1 - tm2) * __builtin_amdgcn_rcpf(tm1 + tm2);
    const float tm3 = apb * fcos * fcos + s2 * s2;
    const float tm4 = tm2 * s2;
    const float rp2 = rs2 * (tm3 - tm4) * __builtin_amdgcn_rcpf(tm3 + tm4);
    const float F = 0.5f * (rp2 + rs2);
    return (F * G * D) * __builtin_amdgcn_rcpf(4.0f * fcos * cout);
}

// BRDF of a gather term (flux only): the rough conductor in fast FP32, the
// other surfaces as on the paths.
__device__ __forceinline__ float gather_brdf(int kind, const FParam& p) {
    if (kind == SPPM_MAT_ROUGH_CONDUCTOR) return brdf_rough_conductor_flux(p);
    return surf_brdf(kind, p);
}

__device__ inline float surf_pdf(int kind, const FParam& p) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:  // material.h:174-176
        return smax(SPPM_SMIN, dot(p.out, p.n) / SPPM_PI);
    case SPPM_MAT_METAL: {  // material.h:416-424
        f3 ref = reflect_sampling(p.n, p.in);
        return (dot(ref, p.out) > 1.0f - 1.0f / SPPM_DELTA_PDF) ? SPPM_DELTA_PDF : SPPM_SMIN;
    }
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return ggx_sampling_pdf(p.out, p.n, p.in, p.rough);
    case SPPM_MAT_GGX_GLASS:
        return ggxglass_sampling_pdf(p.out, p.n, p.in, p.rough, p.from, p.to);
    case SPPM_MAT_GLASS: {  // material.h:509-531
        f3 NN = dot(p.n, p.in) < 0 ? p.n : -p.n;
        float refl = fresnel_reflectance(NN, p.in, p.from, p.to);
        if (dot(p.n, p.in) * dot(p.n, p.out) < 0.0f) {
            f3 ref = reflect_sampling(NN, p.in);
            return (fabs_(dot(ref, p.out)) > LO_0P99) ? SPPM_DELTA_PDF * refl : SPPM_SMIN;  // (double)|.| > 0.99
        }
        f3 rac = refract_sampling(NN, p.in, p.from, p.to);
        return (fabs_(dot(rac, p.out)) > LO_0P99) ? SPPM_DELTA_PDF * (1.0f - refl) : SPPM_SMIN;
    }
    case SPPM_MAT_BIMETAL: {  // material.h:450-461
        const f3 N = dot(p.n, p.in) < 0 ? p.n : -p.n;
        f3 ref = reflect_sampling(N, p.in);
        return (dot(ref, p.out) > 1.0f - 1.0f / SPPM_DELTA_PDF) ? SPPM_DELTA_PDF : SPPM_SMIN;
    }
    case SPPM_MAT_DIFFUSE_TRANSMIT: {  // material.h:675-683
        const float c = (float)__builtin_fabs((double)dot(p.out, p.n) + 0.00001);
        return dot(p.n, p.in) * dot(p.n, p.out) < 0.0f ? p.refl * c / SPPM_PI : (1.0f - p.refl) * c / SPPM_PI;
    }
    case SPPM_MAT_GLOSS_TRANSMIT:  // material.h:757-765
        if (dot(p.n, p.in) * dot(p.n, p.out) < 0.0f) return p.refl * ggx_sampling_pdf(p.out, p.n, p.in, p.rough);
        return (1.0f - p.refl) * ggx_sampling_pdf(reflect_by_n(p.out, p.n), p.n, p.in, p.rough);
    }
    return SPPM_SMIN;
}

__device__ inline f3 glass_refract(float from, float to, f3 in, f3 n) {  // material.h:470-477
    float eta = from / to;
    f3 NN = (dot(in, n) > 0) ? -n : n;
    float LN = -dot(in, NN);
    float sq = smax(0.00000001f, 1.0f - eta * eta * (1.0f - LN * LN));
    return normalized((-NN) * sqrtf_(sq) - ((-in) - NN * LN) * eta);
}

__device__ inline f3 surf_sample(int kind, const FParam& p, MTCursor& m, uint32_t* ovf, float* pdf) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:
        return cos_sampling(m, ovf, p.n, pdf);
    case SPPM_MAT_METAL:
        *pdf = SPPM_DELTA_PDF;
        return reflect_sampling(p.n, p.in);
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return ggx_sampling(m, ovf, p.n, p.in, p.rough, pdf);
    case SPPM_MAT_GGX_GLASS:
        return ggxglass_sampling(m, ovf, p.n, p.in, p.rough, p.from, p.to, pdf);
    case SPPM_MAT_GLASS: {  // material.h:483-498
        f3 NN = dot(p.n, p.in) < 0 ? p.n : -p.n;
        float refl = fresnel_reflectance(NN, p.in, p.from, p.to);
        if ((double)refl > real1(m, ovf)) {
            *pdf = refl;
            return normalized(p.in - NN * (2.0f * dot(NN, p.in)));
        }
        *pdf = 1.0f - refl;
        float eta = p.from / p.to;
        float LN = -dot(p.in, NN);
        float sq = (float)(1.0 - (double)(eta * eta) * (1.0 - (double)(LN * LN)));
        if (sq < HI_1EM6) return normalized(p.in - p.n * (2.0f * dot(p.n, p.in)));  // (double)sq < 0.000001
        return glass_refract(p.from, p.to, p.in, p.n);
    }
    case SPPM_MAT_BIMETAL:  // material.h:428-436
        *pdf = SPPM_DELTA_PDF;
        return reflect_sampling(dot(p.n, p.in) < 0 ? p.n : -p.n, p.in);
    case SPPM_MAT_DIFFUSE_TRANSMIT: {  // material.h:651-665: reflect about -n, transmit about n
        const f3 n = dot(p.n, p.in) > 0 ? p.n : -p.n;
        if ((double)p.refl > real1(m, ovf)) {
            f3 r = cos_sampling(m, ovf, -n, pdf);
            *pdf *= p.refl;
            return r;
        }
        f3 r = cos_sampling(m, ovf, n, pdf);
        *pdf *= (1.0f - p.refl);
        return r;
    }
    case SPPM_MAT_GLOSS_TRANSMIT: {  // material.h:698-717
        const f3 n = dot(p.n, p.in) < 0 ? p.n : -p.n;
        if ((double)p.refl > real1(m, ovf)) {
            f3 r = ggx_sampling(m, ovf, n, p.in, p.rough, pdf);
            *pdf *= p.refl;
            return r;
        }
        f3 r = reflect_by_n(ggx_sampling(m, ovf, n, p.in, p.rough, pdf), n);
        *pdf *= (1.0f - p.refl);
        return r;
    }
    }
    *pdf = 1.0f;
    return invalid_dir();
}

// --------------------------------------------------------------------------
// Textures: LinearSampler::Sample (TextureSampler.h:60-79) on RGBA32F
// --------------------------------------------------------------------------
__device__ inline float4 tex_bilinear(const float4* __restrict__ data, int W, int H, float u, float v) {
    int ix = (int)((float)W * u);
    int iy = (int)((float)H * v);
    int x1 = abs(ix) % W;
    int y1 = abs(iy) % H;
    if (x1 < 0) x1 = 0;  // only for |ix| == INT_MIN (u outside the int range)
    if (y1 < 0) y1 = 0;
    int x2 = min(W - 1, x1 + 1);
    int y2 = min(H - 1, y1 + 1);
    float4 t00 = data[x1 + y1 * W];
    float4 t10 = data[x2 + y1 * W];
    float4 t01 = data[x1 + y2 * W];
    float4 t11 = data[x2 + y2 * W];
    float w = fabs_(u * (float)W) - (float)abs(ix);
    float h = fabs_(v * (float)H) - (float)abs(iy);
    float omw = (float)(1.0 - (double)w);
    float omh = (float)(1.0 - (double)h);
    float4 up, dn, r;
    up.x = omw * t00.x + w * t10.x; up.y = omw * t00.y + w * t10.y;
    up.z = omw * t00.z + w * t10.z; up.w = omw * t00.w + w * t10.w;
    dn.x = omw * t01.x + w * t11.x; dn.y = omw * t01.y + w * t11.y;
    dn.z = omw * t01.z + w * t11.z; dn.w = omw * t01.w + w * t11.w;
    r.x = omh * up.x + h * dn.x; r.y = omh * up.y + h * dn.y;
    r.z = omh * up.z + h * dn.z; r.w = omh * up.w + h * dn.w;
    return r;
}

}  // namespace sppm
