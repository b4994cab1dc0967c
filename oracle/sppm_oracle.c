/*
 * sppm_oracle.c -- CPU restatement of the FriendsRender R SPPM pass.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; never part of the product).
 * See sppm_oracle.h for what pins each part.
 *
 * Conventions that matter for bit-exactness, mirroring the reference:
 *  - PTUtility::Vec3 op order (Vec.h:179-243): dot = (x*x'+y*y')+z*z',
 *    cross component order, normalized() divides each component by the
 *    norm, norm = (float)sqrt((double)norm2).
 *  - Unqualified sqrt/cos/sin/exp on float resolve to the double libm
 *    function and the result is narrowed (verified with g++ 11 on the
 *    reference headers); std::sqrt(float)/std::pow(float,float) stay float.
 *    g++ folds std::pow(float, 2.0f) to x*x; std::pow(float, -2.0f) calls powf.
 *  - Comparisons of float against double literals are done in double.
 *  - Compile with -O2 -ffp-contract=off (no FMA contraction).
 */
#include "sppm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PIf 3.14159f           /* PTUtility::PI, Vec.h:10 */
#define RAY_OFFSET 1e-4f       /* Primitive::Ray::RAY_OFFSET, Primitive.h:31 */
#define SMIN 1e-16f            /* Sampling::sMin, Sampling.h:8 */
#define DELTA_PDF 1.0f         /* Sampling::DeltaPDF, Sampling.h:7 */
#define NUM_SLOTS 32           /* std::deque<Photon>(32), SPPM.cpp:753 */
#define LIGHT_TABLE_SAMPLES 1000 /* LightSampler::NumSample, LightSampler.h:17 */

/* ======================================================================
 * Vec3 (Vec.h:168-323)
 * ====================================================================== */
typedef struct { float x, y, z; } v3;
static inline v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
/* float * Vec3 and Vec3 * float (Vec.h:303-308): s*x per component */
static inline v3 vscale(v3 a, float s) { return V3(s * a.x, s * a.y, s * a.z); }
static inline v3 vdivs(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vnorm2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float vnorm(v3 a) { return (float)sqrt((double)vnorm2(a)); }
static inline v3 vnormalized(v3 a) {
    float d = vnorm(a);
    return V3(a.x / d, a.y / d, a.z / d);
}
static inline float vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline v3 vload(const float* p) { return V3(p[0], p[1], p[2]); }
static inline void vstore(float* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
/* std::max(a,b) = (a<b)?b:a and std::min(a,b) = (b<a)?b:a (argument order matters for NaN) */
static inline float fmaxf_(float a, float b) { return (a < b) ? b : a; }
static inline float fminf_(float a, float b) { return (b < a) ? b : a; }

/* ======================================================================
 * RandomMT: std::mt19937 + uniform_real_distribution<double>(0,1)
 * (MT.h:7-44; libstdc++ generate_canonical<double,53> with a 32-bit engine:
 *  two draws, sum = x0 + x1*2^32 in double, / 2^64, clamped below 1)
 * ====================================================================== */
void orc_mt_seed(orc_mt* m, uint32_t seed) {
    m->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        m->mt[i] = 1812433253u * (m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) + (uint32_t)i;
    m->idx = 624;
}
static void mt_twist(orc_mt* m) {
    const uint32_t UP = 0x80000000u, LO = 0x7fffffffu, A = 0x9908b0dfu;
    uint32_t* x = m->mt;
    int k;
    for (k = 0; k < 624 - 397; ++k) {
        uint32_t y = (x[k] & UP) | (x[k + 1] & LO);
        x[k] = x[k + 397] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
    }
    for (; k < 623; ++k) {
        uint32_t y = (x[k] & UP) | (x[k + 1] & LO);
        x[k] = x[k + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
    }
    uint32_t y = (x[623] & UP) | (x[0] & LO);
    x[623] = x[396] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
    m->idx = 0;
}
uint32_t orc_mt_next(orc_mt* m) {
    if (m->idx >= 624) mt_twist(m);
    uint32_t z = m->mt[m->idx++];
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}
/* Scripted draws (pinning tests only): while a script is installed on the
 * calling thread, every genrand64_real1 / real2 returns the next scripted
 * value instead of the engine's output, so a test can see which draw feeds
 * which variable.  Past the end of the script the draws return 0.5 and the
 * overrun is counted. */
static _Thread_local const double* g_script = NULL;
static _Thread_local int g_script_n = 0, g_script_pos = 0;
void orc_script_set(const double* draws, int n) {
    g_script = n >= 0 ? draws : NULL;
    g_script_n = n > 0 ? n : 0;
    g_script_pos = 0;
}
int orc_script_used(void) { return g_script_pos; }
double orc_mt_real1(orc_mt* m) {
    if (g_script) {
        const int k = g_script_pos++;
        return k < g_script_n ? g_script[k] : 0.5;
    }
    double sum = (double)orc_mt_next(m);
    sum += (double)orc_mt_next(m) * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;  /* * (1.0 - 0.0) + 0.0 */
}
/* genrand64_real2: min(real1, 1.0 - DBL_MIN) == real1 (1 - DBL_MIN rounds to 1) */
static inline double mt_real2(orc_mt* m) {
    double r = orc_mt_real1(m);
    return r < (1.0 - DBL_MIN) ? r : (1.0 - DBL_MIN);
}
void orc_mt_real1_seq(uint32_t seed, int n, double* out) {
    orc_mt m;
    orc_mt_seed(&m, seed);
    for (int i = 0; i < n; ++i) out[i] = orc_mt_real1(&m);
}
void orc_mt_raw_seq(uint32_t seed, int n, uint32_t* out) {
    orc_mt m;
    orc_mt_seed(&m, seed);
    for (int i = 0; i < n; ++i) out[i] = orc_mt_next(&m);
}
/* UniformSample::Sample (CameraSampler.h:25-27): Vec3(real1(), real1(), real1())
 * with g++ evaluating constructor arguments right to left: z first. */
static v3 uniform_sampler_draw(orc_mt* m) {
    double z = orc_mt_real1(m);
    double y = orc_mt_real1(m);
    double x = orc_mt_real1(m);
    return V3((float)x, (float)y, (float)z);
}
void orc_uniform_sampler_seq(int id, int ncalls, float* out) {
    orc_mt m;
    orc_mt_seed(&m, (uint32_t)id);
    for (int i = 0; i < ncalls; ++i) vstore(out + 3 * i, uniform_sampler_draw(&m));
}

/* QRandSequence::HaltonCore<B>::Sample (QRandSequence.h:7-23) */
float orc_halton(int B, int n) {
    float result = 0.0f;
    if (n <= 0) return result;
    int keta = (int)(log2((double)n) / log2((double)B) + 1);
    for (int i = 0; i < keta; ++i) {
        result = (float)((double)result + (double)(n % B) * pow((double)B, (double)(-(i + 1))));
        n /= B;
    }
    return result;
}
/* Hammersley::Sample (QRandSequence.h:35-46) */
void orc_hammersley(int nseq, int i, float* out3) {
    out3[0] = (float)i / (float)nseq;
    out3[1] = orc_halton(2, i);
    out3[2] = orc_halton(3, i);
}

/* ======================================================================
 * TableSampler (TableSampler.h:44-134): alias table in the reference's
 * deque order.
 * ====================================================================== */
typedef struct { float v; int id; } tpair;
void orc_table_build(const float* values, int n, float* prob, int32_t* left, int32_t* right,
                     float* avg_out) {
    if (n <= 0) return;
    float avg = 0.0f;
    for (int i = 0; i < n; ++i) avg += fmaxf_(0.0f, values[i]);
    avg /= (float)n;
    /* deques: bigger and smaller may each grow to at most 2n entries */
    tpair* bigger = (tpair*)malloc(sizeof(tpair) * (size_t)(2 * n + 2));
    tpair* smaller = (tpair*)malloc(sizeof(tpair) * (size_t)(2 * n + 2));
    int nb = 0, ns = 0;
    for (int i = 0; i < n; ++i) {
        float value = fmaxf_(0.0f, values[i]);
        tpair p = {value / avg, i};
        if (value >= avg) bigger[nb++] = p;
        else smaller[ns++] = p;
    }
    int IS = 0, IB = 0;
    for (int i = 0; i < n; ++i) {
        prob[i] = 1.0f; left[i] = -1; right[i] = -1;
        if (IS < ns) {
            left[i] = smaller[IS].id;
            prob[i] = smaller[IS].v;
            IS++;
            if (IB < nb) {
                right[i] = bigger[IB].id;
                bigger[IB].v -= (1.0f - prob[i]);
                if (bigger[IB].v < 1.0f) smaller[ns++] = bigger[IB++];
            } else {
                if (ns == IS) {
                    right[i] = -1;
                    prob[i] = 1.0f;
                } else {
                    right[i] = smaller[IS].id;
                }
            }
        } else {
            left[i] = bigger[IB].id;
            prob[i] = 1.0f;
            bigger[IB].v -= 1.0f;
            if (bigger[IB].v < 1.0f) smaller[ns++] = bigger[IB++];
        }
    }
    free(bigger);
    free(smaller);
    if (avg_out) *avg_out = avg;
}
/* TableSampler::Sample(RandomMT&) (TableSampler.h:126-129) */
static int table_sample(const float* prob, const int32_t* left, const int32_t* right, int n,
                        orc_mt* m) {
    size_t bin = (size_t)((double)n * mt_real2(m));
    double r = orc_mt_real1(m);
    return (r < (double)prob[bin]) ? left[bin] : right[bin];
}
void orc_table_sample_seq(const float* prob, const int32_t* left, const int32_t* right, int n,
                          uint32_t seed, int ndraw, int32_t* out) {
    orc_mt m;
    orc_mt_seed(&m, seed);
    for (int i = 0; i < ndraw; ++i) out[i] = table_sample(prob, left, right, n, &m);
}

/* ======================================================================
 * Textures: LinearSampler::Sample (TextureSampler.h:60-79), TColor = Vec4
 * ====================================================================== */
void orc_bilinear(const sppm_texture* tex, float u, float v, float* out4) {
    int W = tex->width, H = tex->height;
    int x1 = abs((int)((float)W * u)) % W;
    int y1 = abs((int)((float)H * v)) % H;
    int x2 = (W - 1) < (x1 + 1) ? (W - 1) : (x1 + 1);
    int y2 = (H - 1) < (y1 + 1) ? (H - 1) : (y1 + 1);
    const float* t00 = tex->rgba + 4 * (size_t)(x1 + y1 * W);
    const float* t10 = tex->rgba + 4 * (size_t)(x2 + y1 * W);
    const float* t01 = tex->rgba + 4 * (size_t)(x1 + y2 * W);
    const float* t11 = tex->rgba + 4 * (size_t)(x2 + y2 * W);
    float w = fabsf(u * (float)W) - (float)abs((int)((float)W * u));
    float h = fabsf(v * (float)H) - (float)abs((int)((float)H * v));
    float omw = (float)(1.0 - (double)w);
    float omh = (float)(1.0 - (double)h);
    for (int c = 0; c < 4; ++c) {
        float up = omw * t00[c] + w * t10[c];
        float down = omw * t01[c] + w * t11[c];
        out4[c] = omh * up + h * down;
    }
}
/* GetNormalFromNormalmap (TextureUtility.cpp:161-164) */
static v3 normal_from_map(const sppm_texture* tex, float u, float v) {
    float n[4];
    orc_bilinear(tex, u, v, n);
    return vnormalized(V3(2.0f * n[0] - 1.0f, 2.0f * n[1] - 1.0f, 2.0f * n[2] - 1.0f));
}

/* Matrix3 (Matrix.h:150-300): cofactor inverse */
static float det2(float a, float b, float c, float d) { return a * d - b * c; }
static float cofactor3(const float* m, int row, int col) {
    float s[4];
    int k = 0;
    for (int i = 0; i < 3; ++i) {
        if (i == row) continue;
        for (int j = 0; j < 3; ++j) {
            if (j == col) continue;
            s[k++] = m[3 * i + j];
        }
    }
    return det2(s[0], s[1], s[2], s[3]) * (float)(((row + col) % 2) ? -1 : 1);
}
void orc_matrix3_inv(const float* m, float* out) {
    float C[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = cofactor3(m, i, j);
    /* CofactorMatrix() = C^T */
    float adj[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) adj[3 * i + j] = C[3 * j + i];
    /* Determinant: m00*C00 + m10*C10 + m20*C20 */
    float det = m[0] * C[0] + m[3] * C[3] + m[6] * C[6];
    float ix = 1.0f / det;
    for (int i = 0; i < 9; ++i) out[i] = ix * adj[i];
}

/* ======================================================================
 * Sampling.cpp / RayUtility.cpp helpers
 * ====================================================================== */
static const v3 INVALID_DIR = {1e-3f, 1e-3f, 1e-3f};  /* RUtil::INVALID_DIR */

static inline v3 reflect_sampling(v3 n, v3 in) {  /* Sampling.cpp:293-295 */
    return vnormalized(vsub(in, vscale(n, 2.0f * vdot(n, in))));
}
static float fresnel_reflectance(v3 n, v3 in, float ii, float io) {  /* :297-313 */
    float n1n2 = ii / io;
    float cosA = fabsf(vdot(n, in));
    float sinA = sqrtf(1.0000001f - cosA * cosA);
    float sinB = n1n2 * sinA;
    if ((double)sinB > 0.999999) return 1.0f;
    float cosB = sqrtf(1.0000001f - sinB * sinB);
    float rp = (io * cosA - ii * cosB) / (io * cosA + ii * cosB);
    float rs = (ii * cosA - io * cosB) / (ii * cosA + io * cosB);
    return (0.5f * rs * rs) + (0.5f * rp * rp);
}
static v3 refract_sampling(v3 n, v3 in, float ii, float io) {  /* Sampling.cpp:329-337 */
    float eta = ii / io;
    v3 NN = (vdot(in, n) > 0) ? vneg(n) : n;
    float LN = -vdot(in, NN);
    float sq = fmaxf_(0.00000001f, 1.0f - eta * eta * (1.0f - LN * LN));
    return vnormalized(vsub(vscale(vneg(NN), sqrtf(sq)), vscale(vsub(vneg(in), vscale(NN, LN)), eta)));
}
static v3 rutil_refract(v3 in, v3 n, float from, float into) {  /* RayUtility.cpp:10-17 */
    v3 NN = (vdot(in, n) < 0) ? n : vneg(n);
    float eta = from / into;
    float LN = -vdot(in, NN);
    float sq = fmaxf_(0.0f, 1.0f - eta * eta * (1.0f - LN * LN));
    return vnormalized(vsub(vscale(vneg(NN), sqrtf(sq)), vscale(vsub(vneg(in), vscale(NN, LN)), eta)));
}
static inline v3 microfacet_from_reflection(v3 in, v3 out) {  /* RayUtility.cpp:18-20 */
    return vnormalized(vadd(vneg(in), out));
}
static v3 microfacet_from_refraction(v3 in, v3 out, float from, float into) {  /* :21-33 */
    float c = fabsf(vdot(in, out));
    float eta = into / from;
    float isq = eta * eta - 2.0f * eta * c + 1.0f;
    if ((double)isq < 1e-8) return V3(0, 0, 0);
    float ss = 1.0f / sqrtf(isq);
    v3 h = vnormalized(vadd(vscale(out, -eta * ss), vscale(in, ss)));
    return (into > from) ? h : vneg(h);
}
static inline int rutil_is_reflect(v3 in, v3 out, v3 n) {  /* :34-36 */
    return vdot(n, in) * vdot(n, out) < 0.0f;
}
static inline int rutil_valid_dir(v3 d) { return vnorm2(d) > 0.9999f; }  /* :38-40 */

/* orthonormal frame used by Cos/GGX sampling (Sampling.cpp:12-18) */
static void frame(v3 e1, v3* e2, v3* e3) {
    if ((double)fabsf(e1.x) > 0.000001) *e2 = vnormalized(vcross(V3(0, 1, 0), e1));
    else *e2 = vnormalized(vcross(V3(1, 0, 0), e1));
    *e3 = vnormalized(vcross(e1, *e2));
}

/* Cos_Sampling (Sampling.cpp:7-28) */
static v3 cos_sampling(orc_mt* m, v3 n, float* pdf) {
    v3 e2, e3;
    frame(n, &e2, &e3);
    float ph = (float)(orc_mt_real1(m) * 2.0 * (double)PIf);
    float th = (float)orc_mt_real1(m);
    float th2 = (float)sqrt((double)th);
    v3 a = vscale(n, (float)sqrt(1.000000001 - (double)th));
    v3 b = vscale(vscale(e2, (float)cos((double)ph)), th2);
    v3 c = vscale(vscale(e3, (float)sin((double)ph)), th2);
    v3 r = vnormalized(vadd(vadd(a, b), c));
    if (pdf) *pdf = fmaxf_(SMIN, vdot(r, n) / PIf);
    return r;
}

/* GGX_Sampling (Sampling.cpp:123-157) */
static v3 ggx_sampling(orc_mt* m, v3 n, v3 in, float alpha, float* pdf) {
    float phi = (float)(orc_mt_real1(m) * 2.0 * (double)PIf);
    float e = (float)orc_mt_real1(m);
    float cth = (float)sqrt((1.0 - (double)e) / (double)(e * (alpha * alpha - 1.0f) + 1.0f));
    v3 e2, e3;
    frame(n, &e2, &e3);
    float cph = (float)cos((double)phi);
    float sph = (float)sin((double)phi);
    float sth = sqrtf(1.0f - cth * cth);
    v3 mN = vnormalized(vadd(vadd(vscale(vscale(e3, sth), cph), vscale(vscale(e2, sth), sph)),
                             vscale(n, cth)));
    if ((double)vdot(mN, in) > 0.0) { *pdf = SMIN; return INVALID_DIR; }
    v3 r = reflect_sampling(mN, in);
    if ((double)vdot(r, n) < 0.0) { *pdf = SMIN; return INVALID_DIR; }
    float wm_N = vdot(mN, n);
    float wo_wm = vdot(r, mN);
    /* std::pow(float, 2) with an int exponent promotes to double (x*x) */
    double q = (double)(1.0f - (1.0f - alpha * alpha) * wm_N * wm_N);
    *pdf = (float)((double)(alpha * alpha) / ((double)PIf * (q * q)) * (double)wm_N /
                   (double)(4.0f * wo_wm));
    return r;
}
/* GGX_Sampling_PDF (Sampling.cpp:158-180) */
static float ggx_sampling_pdf(v3 out, v3 n, v3 in, float a) {
    if (rutil_is_reflect(in, out, n)) {
        v3 h = microfacet_from_reflection(in, out);
        float h_n = vdot(h, n);
        if (h_n < 0.0f) return SMIN;
        float wi_h = -vdot(in, h);
        if (wi_h < 0.0f) return SMIN;
        float wo_h = vdot(out, h);
        if (wo_h < 0.0f) return SMIN;
        float jc = 1.0f / (4.0f * wo_h);
        return (a * a / PIf * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) * h_n * jc;
    }
    return SMIN;
}
/* GGXGlass_Sampling (Sampling.cpp:181-248) */
static v3 ggxglass_sampling(orc_mt* m, v3 n, v3 in, float alpha, float ii, float io, float* pdf) {
    const float SV = 1e-12f;
    v3 NN = vdot(in, n) < 0 ? n : vneg(n);
    *pdf = 1.0f;
    float phi = (float)(orc_mt_real1(m) * 2.0 * (double)PIf);
    float e = (float)orc_mt_real1(m);
    float cth = (float)sqrt((1.0000001 - (double)e) /
                            ((double)e * ((double)(alpha * alpha) - 0.99999) + 1.0));
    v3 e2, e3;
    frame(NN, &e2, &e3);
    float cph = (float)cos((double)phi);
    float sph = (float)sin((double)phi);
    float sth = sqrtf(1.00000001f - cth * cth);
    v3 mN = vnormalized(vadd(vadd(vscale(vscale(e3, sth), cph), vscale(vscale(e2, sth), sph)),
                             vscale(NN, cth)));
    float wi_wm = -vdot(in, mN);
    float wm_N = vdot(mN, NN);
    float wi_N = -vdot(in, NN);
    if (wi_N < SV) { *pdf = SMIN; return INVALID_DIR; }
    if (wi_wm < SV) { *pdf = SMIN; return INVALID_DIR; }
    if (wm_N < SV) { *pdf = SMIN; return INVALID_DIR; }
    float refl = fresnel_reflectance(mN, in, ii, io);
    v3 r;
    if ((double)refl > orc_mt_real1(m)) {
        *pdf = refl;
        r = reflect_sampling(mN, in);
        float wo_wm = vdot(r, mN);
        if (wo_wm < SV) return INVALID_DIR;
        if (vdot(r, NN) * vdot(in, NN) > SV) { *pdf = SMIN; return INVALID_DIR; }
        float t = 1.0f - (1.0f - alpha * alpha) * wm_N * wm_N;
        *pdf = refl * (wm_N / (4.0f * wo_wm) * alpha * alpha / (PIf * (t * t)));
    } else {
        r = rutil_refract(in, mN, ii, io);
        float wo_wm = -vdot(r, mN);
        float wo_N = -vdot(r, NN);
        if (wi_N * wo_N < SV || wo_N < SV || wo_wm < SV) { *pdf = SMIN; return INVALID_DIR; }
        /* std::pow(float, 2) (int exponent) -> double x*x; the product runs in double */
        double s = (double)(ii * wi_wm + io * wo_wm);
        float t = 1.0f - (1.0f - alpha * alpha) * wm_N * wm_N;
        double val = (double)((1.0f - refl) * wm_N) * ((double)(io * io * wo_wm) / (s * s));
        val = val * (double)alpha;
        val = val * (double)alpha;
        *pdf = (float)(val / (double)(PIf * (t * t)));
    }
    return r;
}
/* GGXGlass_Sampling_PDF (Sampling.cpp:249-288) */
static float ggxglass_sampling_pdf(v3 out, v3 n, v3 in, float a, float ii, float io) {
    if (rutil_is_reflect(in, out, n)) {
        v3 h = microfacet_from_reflection(in, out);
        float h_n = vdot(h, n);
        if (h_n < 0.0f) return SMIN;
        float wi_h = -vdot(in, h);
        if (wi_h < 0.0f) return SMIN;
        float wo_h = vdot(out, h);
        if (wo_h < 0.0f) return SMIN;
        float refl = fresnel_reflectance(h, in, ii, io);
        float jc = 1.0f / (4.0f * wo_h);
        return refl * h_n * (a * a / PIf * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) * jc;
    } else {
        v3 h = microfacet_from_refraction(in, out, ii, io);
        float h_n = vdot(h, n);
        if (h_n < 0.0f) return SMIN;
        float wi_h = -vdot(in, h);
        if (wi_h < 0.0f) return SMIN;
        float wo_h = -vdot(out, h);
        if (wo_h < 0.0f) return SMIN;
        float refl = fresnel_reflectance(h, in, ii, io);
        float jc = io * io * wo_h * powf(ii * wi_h + io * wo_h, -2.0f);
        return (1.0f - refl) * h_n * (a * a / PIf * powf(1.0f - (1.0f - a * a) * h_n * h_n, -2.0f)) *
               jc;
    }
}

/* ======================================================================
 * Surfaces (material.h)
 * ====================================================================== */
typedef struct {
    v3 in, out, n;
    float from, to, rough, absorb, refl;
} fparam;

static inline int opaque_reflect(const fparam* p) {  /* material.h:139-141 */
    return ((double)vdot(p->in, vneg(p->n)) > 1e-4 && (double)vdot(p->out, p->n) > 1e-4);
}

static float brdf_rough_conductor(const fparam* p) {  /* material.h:227-274 */
    const float sv = 1e-10f;
    if (vdot(p->in, p->n) > -sv || vdot(p->out, p->n) < sv) return 0.0f;
    float i_to = p->to, i_from = p->from, k = p->absorb;
    float eta = i_to / i_from;
    float alpha = p->rough;
    v3 h = vnormalized(vadd(vneg(p->in), p->out));
    float ccs = vdot(h, p->n);
    if (ccs < sv || vdot(h, p->in) > -sv || vdot(h, p->out) < sv) return 0.0f;
    float q = (1.0001f - (1.0f - alpha * alpha) * ccs * ccs);
    float D = fmaxf_(SMIN, (alpha * alpha) / (PIf * (q * q)));
    float lc = fmaxf_(0.0f, fminf_(0.9999f, vdot(p->in, vneg(p->n))));
    float lv = fmaxf_(0.0f, fminf_(0.9999f, vdot(p->out, p->n)));
    if (lc < SMIN || lv < SMIN) return 0.0f;
    /* (-1.0f + std::sqrt(1.0f + a*a*(1.0f/(l*l) - 1.0))) * 0.5 runs in double */
    float L_l = (float)((-1.0 + sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lc * lc)) - 1.0))) * 0.5);
    float L_v = (float)((-1.0 + sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lv * lv)) - 1.0))) * 0.5);
    float G = 1.0f / ((1.0f + L_l) * (1.0f + L_v));
    float fcos = vdot(p->n, vneg(p->in));
    float temp = eta * eta - k * k - (1.0f - fcos * fcos);
    float apb = sqrtf(temp * temp + k * k * eta * eta);
    float a = (float)sqrt((double)(apb + temp) * 0.5);
    float tm1 = apb + fcos * fcos;
    float tm2 = a * 2.0f * fcos;
    float rs2 = (tm1 - tm2) / (tm1 + tm2);
    float tm3 = (float)((double)(apb * fcos * fcos) + (double)(1.0f - fcos * fcos) * (1.0 - (double)(fcos * fcos)));
    float tm4 = tm2 * (1.0f - fcos * fcos);
    float rp2 = rs2 * (tm3 - tm4) / (tm3 + tm4);
    float F = (float)(0.5 * (double)(rp2 + rs2));
    float fx = (4.0f * vdot(p->in, vneg(p->n)) * vdot(p->out, p->n));
    return (F * G * D) / fx;
}

static float brdf_ggx_glass(const fparam* p) {  /* material.h:314-390 */
    const float SV = 1e-8f;
    int isReflect = vdot(p->n, p->in) * vdot(p->n, p->out) < 0;
    float a = p->rough;
    float Ii = p->from, Io = p->to;
    if (isReflect) {
        v3 h = vnormalized(vadd(vneg(p->in), p->out));
        float n_h = vdot(p->n, h);
        if (n_h < 0) return 0.0f;
        float D = (a * a) * powf(1.0f - (1.0f - a * a) * n_h * n_h, -2.0f) / PIf;
        float L_l = (-1.0f + sqrtf(1.0001f + a * a * (powf(vdot(p->n, vneg(p->in)), -2.0f) - 1.0f))) * 0.5f;
        float L_v = (-1.0f + sqrtf(1.0001f + a * a * (powf(vdot(p->n, p->out), -2.0f) - 1.0f))) * 0.5f;
        float G = 1.0f / (L_l + L_v + 1.0f);
        float F = fresnel_reflectance(h, p->in, Ii, Io);
        float wo_n = fmaxf_(0.0f, fminf_(1.0f, vdot(p->out, h)));
        float wi_n = fmaxf_(0.0f, fminf_(1.0f, -vdot(p->in, h)));
        if (wo_n < SV || wi_n < SV) return 0.0f;
        return F * G * D / (4.0f * wo_n * wi_n);
    } else {
        float eta = Io / Ii;
        if (fabsf(eta - 1.0f) < SV) return 1.0f;
        if ((double)vdot(p->in, p->out) < 0.0) return 0.0f;
        /* EstimateN (material.h:297-309) */
        v3 h;
        {
            float c = fabsf(vdot(p->in, p->out));
            float et = Io / Ii;
            float isq = et * et - 2.0f * et * c + 1.0f;
            if (isq < SMIN) h = V3(0, 0, 0);
            else {
                float ss = 1.0f / sqrtf(isq);
                h = vnormalized(vadd(vscale(p->out, -et * ss), vscale(p->in, ss)));
                if (!(Io > Ii)) h = vneg(h);
            }
        }
        if (vnorm(h) < SV) return 0.0f;
        float n_h = vdot(h, p->n);
        if (n_h < SV) return 0.0f;
        float D = (a * a) * powf(1.0f - (1.0f - a * a) * n_h * n_h, -2.0f) / PIf;
        float L_l = (-1.0f + sqrtf(1.0001f + a * a * (powf(vdot(p->n, vneg(p->in)), -2.0f) - 1.0f))) * 0.5f;
        float L_v = (-1.0f + sqrtf(1.0001f + a * a * (powf(vdot(p->n, vneg(p->out)), -2.0f) - 1.0f))) * 0.5f;
        float G = 1.0f / (L_l + L_v + 1.0f);
        float F = fresnel_reflectance(h, p->in, Ii, Io);
        float wi_wm = -vdot(p->in, h);
        float wo_wm = -vdot(p->out, h);
        float wi_n = -vdot(p->in, p->n);
        float wo_n = -vdot(p->out, p->n);
        if (wo_wm < SV || wi_wm < SV) return 0.0f;
        if (wi_n < SV || wo_n < SV) return 0.0f;
        return (wi_wm * wo_wm) / (wi_n * wo_n) * (Io * Io) * (1.0f - F) * G * D *
               powf(Ii * wo_wm + Io * wi_wm, -2.0f);
    }
}

/* GlossTransmitSurface::ReflectByN (material.h:692-695): r - 2 (n.r) n */
static v3 reflect_by_n(v3 r, v3 n) { return vsub(r, vscale(n, 2.0f * vdot(n, r))); }

/* GlossTransmitSurface::BRDF (material.h:719-756) */
static float brdf_gloss_transmit(const fparam* p) {
    float power = vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f ? p->refl : 1.0f - p->refl;
    v3 InDir = p->in;
    v3 Normal = vdot(p->n, InDir) < 0 ? p->n : vneg(p->n);
    v3 OutDir = p->out;
    if (vdot(OutDir, Normal) < 0) OutDir = reflect_by_n(OutDir, Normal);
    const float smallValue = 1e-10f;
    if (vdot(InDir, Normal) > -smallValue || vdot(OutDir, Normal) < smallValue) return 0.0f;
    float alpha = p->rough;
    v3 h = vnormalized(vadd(vneg(InDir), OutDir));
    float ccs = vdot(h, Normal);
    if (ccs < smallValue || vdot(h, InDir) > -smallValue || vdot(h, OutDir) < smallValue) return 0.0f;
    /* std::pow(float, 2.0f) is the float overload: the rounded square */
    float q = (1.0001f - (1.0f - alpha * alpha) * ccs * ccs);
    float D = fmaxf_(SMIN, (alpha * alpha) / (PIf * (q * q)));
    float lc = fmaxf_(0.0f, fminf_(0.9999f, vdot(InDir, vneg(Normal))));
    float lv = fmaxf_(0.0f, fminf_(0.9999f, vdot(OutDir, Normal)));
    if (lc < SMIN || lv < SMIN) return 0.0f;
    float L_l = (float)((-1.0 + sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lc * lc)) - 1.0))) * 0.5);
    float L_v = (float)((-1.0 + sqrt(1.0 + (double)(alpha * alpha) * ((double)(1.0f / (lv * lv)) - 1.0))) * 0.5);
    float G = 1.0f / ((1.0f + L_l) * (1.0f + L_v));
    float fx = (4.0f * vdot(InDir, vneg(Normal)) * vdot(OutDir, Normal));
    return power * (G * D) / fx;
}

static float surf_brdf(int kind, const fparam* p) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:  /* material.h:168-173 */
        return opaque_reflect(p) ? 1.0f / PIf : 0.0f;
    case SPPM_MAT_METAL: {  /* material.h:404-415 */
        if (!opaque_reflect(p)) return 0.0f;
        v3 ref = reflect_sampling(p->n, p->in);
        if (vdot(ref, p->out) > 1.0f - 1.0f / DELTA_PDF)
            return (float)((double)DELTA_PDF / ((double)vdot(p->out, p->n) + 0.00001));
        return 0.0f;
    }
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return brdf_rough_conductor(p);
    case SPPM_MAT_GGX_GLASS:
        return brdf_ggx_glass(p);
    case SPPM_MAT_GLASS: {  /* material.h:499-508 */
        v3 NN = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        float refl = fresnel_reflectance(NN, p->in, p->from, p->to);
        double denom = (double)fabsf(vdot(p->out, NN)) + 0.0001;
        if (vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f) return (float)((double)(DELTA_PDF * refl) / denom);
        return (float)((double)(DELTA_PDF * (1.0f - refl)) / denom);
    }
    case SPPM_MAT_BIMETAL: {  /* material.h:437-449: MetalSurface about the flipped normal */
        v3 N = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        if (!opaque_reflect(p)) return 0.0f;
        v3 ref = reflect_sampling(N, p->in);
        if (vdot(ref, p->out) > 1.0f - 1.0f / DELTA_PDF)
            return (float)((double)DELTA_PDF / ((double)vdot(p->out, N) + 0.00001));
        return 0.0f;
    }
    case SPPM_MAT_DIFFUSE_TRANSMIT:  /* material.h:666-674 */
        if (vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f) return p->refl / PIf;
        return (1.0f - p->refl) / PIf;
    case SPPM_MAT_GLOSS_TRANSMIT:
        return brdf_gloss_transmit(p);
    }
    return 0.0f;
}

static float surf_pdf(int kind, const fparam* p) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:
        return fmaxf_(SMIN, vdot(p->out, p->n) / PIf);
    case SPPM_MAT_METAL: {
        v3 ref = reflect_sampling(p->n, p->in);
        return (vdot(ref, p->out) > 1.0f - 1.0f / DELTA_PDF) ? DELTA_PDF : SMIN;
    }
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return ggx_sampling_pdf(p->out, p->n, p->in, p->rough);
    case SPPM_MAT_GGX_GLASS:
        return ggxglass_sampling_pdf(p->out, p->n, p->in, p->rough, p->from, p->to);
    case SPPM_MAT_GLASS: {  /* material.h:509-531 */
        v3 NN = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        float refl = fresnel_reflectance(NN, p->in, p->from, p->to);
        if (vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f) {
            v3 ref = reflect_sampling(NN, p->in);
            if ((double)fabsf(vdot(ref, p->out)) > 0.99) return DELTA_PDF * refl;
            return SMIN;
        } else {
            v3 rac = refract_sampling(NN, p->in, p->from, p->to);
            if ((double)fabsf(vdot(rac, p->out)) > 0.99) return DELTA_PDF * (1.0f - refl);
            return SMIN;
        }
    }
    case SPPM_MAT_BIMETAL: {  /* material.h:450-461 */
        v3 N = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        v3 ref = reflect_sampling(N, p->in);
        return (vdot(ref, p->out) > 1.0f - 1.0f / DELTA_PDF) ? DELTA_PDF : SMIN;
    }
    case SPPM_MAT_DIFFUSE_TRANSMIT: {  /* material.h:675-683: std::abs of a double */
        float c = (float)fabs((double)vdot(p->out, p->n) + 0.00001);
        if (vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f) return p->refl * c / PIf;
        return (1.0f - p->refl) * c / PIf;
    }
    case SPPM_MAT_GLOSS_TRANSMIT:  /* material.h:757-765 */
        if (vdot(p->n, p->in) * vdot(p->n, p->out) < 0.0f) return p->refl * ggx_sampling_pdf(p->out, p->n, p->in, p->rough);
        return (1.0f - p->refl) * ggx_sampling_pdf(reflect_by_n(p->out, p->n), p->n, p->in, p->rough);
    }
    return SMIN;
}

static v3 glass_reflect(v3 in, v3 n) { return vnormalized(vsub(in, vscale(n, 2.0f * vdot(n, in)))); }
static v3 glass_refract(float from, float to, v3 in, v3 n) {  /* material.h:470-477 */
    float eta = from / to;
    v3 NN = (vdot(in, n) > 0) ? vneg(n) : n;
    float LN = -vdot(in, NN);
    float sq = fmaxf_(0.00000001f, 1.0f - eta * eta * (1.0f - LN * LN));
    return vnormalized(vsub(vscale(vneg(NN), sqrtf(sq)), vscale(vsub(vneg(in), vscale(NN, LN)), eta)));
}

static v3 surf_sample(int kind, const fparam* p, orc_mt* m, float* pdf) {
    switch (kind) {
    case SPPM_MAT_DIFFUSE:
        return cos_sampling(m, p->n, pdf);
    case SPPM_MAT_METAL:
        *pdf = DELTA_PDF;
        return reflect_sampling(p->n, p->in);
    case SPPM_MAT_ROUGH_CONDUCTOR:
        return ggx_sampling(m, p->n, p->in, p->rough, pdf);
    case SPPM_MAT_GGX_GLASS:
        return ggxglass_sampling(m, p->n, p->in, p->rough, p->from, p->to, pdf);
    case SPPM_MAT_GLASS: {  /* material.h:483-498 */
        v3 NN = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        float refl = fresnel_reflectance(NN, p->in, p->from, p->to);
        if ((double)refl > orc_mt_real1(m)) {
            *pdf = refl;
            return glass_reflect(p->in, NN);
        }
        *pdf = 1.0f - refl;
        float eta = p->from / p->to;
        float LN = -vdot(p->in, NN);
        float sq = (float)(1.0 - (double)(eta * eta) * (1.0 - (double)(LN * LN)));
        if ((double)sq < 0.000001) return glass_reflect(p->in, p->n);
        return glass_refract(p->from, p->to, p->in, p->n);
    }
    case SPPM_MAT_BIMETAL: {  /* material.h:428-436 */
        v3 N = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        *pdf = DELTA_PDF;
        return reflect_sampling(N, p->in);
    }
    case SPPM_MAT_DIFFUSE_TRANSMIT: {  /* material.h:651-665 */
        v3 n = vdot(p->n, p->in) > 0 ? p->n : vneg(p->n);
        v3 rs;
        if ((double)p->refl > orc_mt_real1(m)) {
            rs = cos_sampling(m, vneg(n), pdf);
            *pdf *= p->refl;
        } else {
            rs = cos_sampling(m, n, pdf);
            *pdf *= (1.0f - p->refl);
        }
        return rs;
    }
    case SPPM_MAT_GLOSS_TRANSMIT: {  /* material.h:698-717 */
        v3 n = vdot(p->n, p->in) < 0 ? p->n : vneg(p->n);
        v3 rs;
        if ((double)p->refl > orc_mt_real1(m)) {
            rs = ggx_sampling(m, n, p->in, p->rough, pdf);
            *pdf *= p->refl;
        } else {
            rs = reflect_by_n(ggx_sampling(m, n, p->in, p->rough, pdf), n);
            *pdf *= (1.0f - p->refl);
        }
        return rs;
    }
    }
    *pdf = 1.0f;
    return INVALID_DIR;
}

static void to_fparam(const orc_fparam* a, fparam* b) {
    b->in = vload(a->in); b->out = vload(a->out); b->n = vload(a->n);
    b->from = a->from_index; b->to = a->to_index; b->rough = a->roughness;
    b->absorb = a->absorption; b->refl = a->reflectance;
}
float orc_bsdf_brdf(int kind, const orc_fparam* p) { fparam q; to_fparam(p, &q); return surf_brdf(kind, &q); }
float orc_bsdf_pdf(int kind, const orc_fparam* p) { fparam q; to_fparam(p, &q); return surf_pdf(kind, &q); }
/* batches for the tests' Monte Carlo integrals: PDF()/BRDF() at n directions
 * `out` (the rest of p fixed), and n samples from RandomMT(seed0 + j) */
void orc_bsdf_eval_batch(int kind, const orc_fparam* p, int n, const float* outdirs, float* pdf, float* brdf) {
    fparam q;
    to_fparam(p, &q);
    for (int j = 0; j < n; ++j) {
        q.out = vload(outdirs + 3 * (size_t)j);
        if (pdf) pdf[j] = surf_pdf(kind, &q);
        if (brdf) brdf[j] = surf_brdf(kind, &q);
    }
}
void orc_bsdf_sample_batch(int kind, const orc_fparam* p, uint32_t seed0, int n, float* dirs, float* pdf) {
    fparam q;
    to_fparam(p, &q);
    for (int j = 0; j < n; ++j) {
        orc_mt m;
        orc_mt_seed(&m, seed0 + (uint32_t)j);
        vstore(dirs + 3 * (size_t)j, surf_sample(kind, &q, &m, pdf + j));
    }
}
void orc_bsdf_sample(int kind, const orc_fparam* p, uint32_t seed, float* dir, float* pdf, int* ndraw) {
    fparam q;
    to_fparam(p, &q);
    orc_mt m;
    orc_mt_seed(&m, seed);
    v3 d = surf_sample(kind, &q, &m, pdf);
    vstore(dir, d);
    if (ndraw) *ndraw = (m.idx == 624) ? 0 : m.idx / 2;
}

/* ======================================================================
 * Scene
 * ====================================================================== */
typedef struct {
    int facet;
    v3 flux;
    int textured;
    /* textured: 1000-entry table (LightSampler.cpp:30-61) */
    v3* uvw;
    v3* L;
    float* prob;
    int32_t* left;
    int32_t* right;
    float avg;
} olight;

typedef struct { float bmin[3], bmax[3]; int left, right, prim; } obvh_node;

struct orc_scene {
    int nf;
    float* pos;   /* 9 per facet */
    float* nrm;
    float* uv;
    int32_t* attr;
    int nattr;
    sppm_attr* attrs;
    int ntex;
    sppm_texture* tex;
    float** texdata;
    int nlights;
    olight* lights;
    float* lprob;
    int32_t* lleft;
    int32_t* lright;
    float lavg;
    int32_t* facet_light;  /* GetAreaLightFromFacetID (LightSampler.cpp:98-106): light of a facet, -1 none */
    /* environment light: the converted map (LightSampler.cpp:22-26) */
    int has_ibl;
    sppm_texture ibl;
    float* ibl_data;
    /* acceleration for the oracle's own speed (results equal brute force) */
    obvh_node* nodes;
    int nnodes;
};

static float facet_area(const float* p) {  /* inoutFacet::CalculateArea, Normal_Facet.h:56-58 */
    v3 a = vsub(vload(p + 3), vload(p));
    v3 b = vsub(vload(p + 6), vload(p));
    return 0.5f * vnorm(vcross(a, b));
}

/* LightSampler::CreateAreaLight (LightSampler.cpp:30-66) */
static void create_area_light(olight* L, const orc_scene* s, int f, v3 radiance, const sppm_texture* tex) {
    const float* uv = s->uv + 9 * f;
    float area = facet_area(s->pos + 9 * f);
    L->facet = f;
    if (tex) {
        int N = LIGHT_TABLE_SAMPLES;
        L->uvw = (v3*)malloc(sizeof(v3) * N);
        L->L = (v3*)malloc(sizeof(v3) * N);
        v3 total = V3(0, 0, 0);
        for (int i = 0; i < N; ++i) {
            float w[3];
            orc_hammersley(N, i, w);
            if (w[1] > w[0]) { w[0] *= 0.5f; w[1] -= w[0]; }
            else { w[1] *= 0.5f; w[0] -= w[1]; }
            v3 uvw = V3(w[0], w[1], w[2]);
            v3 tuv = vadd(vadd(vscale(vload(uv), w[0]), vscale(vload(uv + 3), w[1])),
                          vscale(vload(uv + 6), 1.0f - w[0] - w[1]));
            float c[4];
            orc_bilinear(tex, tuv.x, tuv.y, c);
            v3 Lv = vmul(V3(c[0], c[1], c[2]), radiance);
            if (vnorm2(Lv) < 0.01f) Lv = V3(0, 0, 0);
            L->uvw[i] = uvw;
            L->L[i] = Lv;
            total = vadd(total, Lv);
        }
        if ((double)vnorm2(total) < 1e-6) {
            L->L[0] = vscale(V3(1, 1, 1), (float)1e-6);
            L->uvw[0] = V3(0, 0, 0);
        } else {
            total = vscale(total, (2.0f * PIf * area / (float)N));
        }
        L->flux = total;
        L->textured = 1;
        float* vals = (float*)malloc(sizeof(float) * N);
        for (int i = 0; i < N; ++i) vals[i] = vnorm(L->L[i]);
        L->prob = (float*)malloc(sizeof(float) * N);
        L->left = (int32_t*)malloc(sizeof(int32_t) * N);
        L->right = (int32_t*)malloc(sizeof(int32_t) * N);
        orc_table_build(vals, N, L->prob, L->left, L->right, &L->avg);
        free(vals);
    } else {
        L->flux = vscale(radiance, 2.0f * PIf * area);
        L->textured = 0;
        L->uvw = L->L = NULL;
        L->prob = NULL; L->left = L->right = NULL;
    }
}

/* ---- BVH for the oracle's own speed (binary, median split) ------------- */
static void tri_bounds(const float* p, float* bmin, float* bmax) {
    for (int k = 0; k < 3; ++k) {
        float a = p[k], b = p[3 + k], c = p[6 + k];
        float mn = a < b ? a : b; mn = mn < c ? mn : c;
        float mx = a > b ? a : b; mx = mx > c ? mx : c;
        bmin[k] = mn - 0.001f;
        bmax[k] = mx + 0.001f;
    }
}
static const float* g_sort_cent;
static int g_sort_axis;
static int cmp_cent(const void* a, const void* b) {
    float ca = g_sort_cent[3 * *(const int*)a + g_sort_axis];
    float cb = g_sort_cent[3 * *(const int*)b + g_sort_axis];
    if (ca < cb) return -1;
    if (ca > cb) return 1;
    return (*(const int*)a) - (*(const int*)b);
}
static int bvh_build_rec(orc_scene* s, int* ids, int n, const float* cent) {
    int me = s->nnodes++;
    obvh_node* nd = &s->nodes[me];
    float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < n; ++i) {
        float a[3], b[3];
        tri_bounds(s->pos + 9 * ids[i], a, b);
        for (int k = 0; k < 3; ++k) {
            if (a[k] < bmin[k]) bmin[k] = a[k];
            if (b[k] > bmax[k]) bmax[k] = b[k];
        }
    }
    memcpy(nd->bmin, bmin, sizeof bmin);
    memcpy(nd->bmax, bmax, sizeof bmax);
    if (n == 1) {
        nd->prim = ids[0];
        nd->left = nd->right = -1;
        return me;
    }
    int axis = 0;
    float ext = bmax[0] - bmin[0];
    for (int k = 1; k < 3; ++k)
        if (bmax[k] - bmin[k] > ext) { ext = bmax[k] - bmin[k]; axis = k; }
    g_sort_cent = cent;
    g_sort_axis = axis;
    qsort(ids, (size_t)n, sizeof(int), cmp_cent);
    int h = n / 2;
    nd->prim = -1;
    int l = bvh_build_rec(s, ids, h, cent);
    int r = bvh_build_rec(s, ids + h, n - h, cent);
    s->nodes[me].left = l;
    s->nodes[me].right = r;
    return me;
}

orc_scene* orc_scene_create(const sppm_scene_desc* d) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    int nf = d->nfacets;
    s->nf = nf;
    s->pos = (float*)malloc(sizeof(float) * 9 * (size_t)nf);
    s->nrm = (float*)malloc(sizeof(float) * 9 * (size_t)nf);
    s->uv = (float*)malloc(sizeof(float) * 9 * (size_t)nf);
    s->attr = (int32_t*)malloc(sizeof(int32_t) * (size_t)nf);
    memcpy(s->pos, d->pos, sizeof(float) * 9 * (size_t)nf);
    memcpy(s->nrm, d->nrm, sizeof(float) * 9 * (size_t)nf);
    memcpy(s->uv, d->uv, sizeof(float) * 9 * (size_t)nf);
    memcpy(s->attr, d->attr, sizeof(int32_t) * (size_t)nf);
    s->nattr = d->nattrs;
    s->attrs = (sppm_attr*)malloc(sizeof(sppm_attr) * (size_t)(d->nattrs > 0 ? d->nattrs : 1));
    memcpy(s->attrs, d->attrs, sizeof(sppm_attr) * (size_t)d->nattrs);
    s->ntex = d->ntextures;
    s->tex = (sppm_texture*)calloc((size_t)(d->ntextures > 0 ? d->ntextures : 1), sizeof(sppm_texture));
    s->texdata = (float**)calloc((size_t)(d->ntextures > 0 ? d->ntextures : 1), sizeof(float*));
    for (int i = 0; i < d->ntextures; ++i) {
        size_t n = 4 * (size_t)d->textures[i].width * (size_t)d->textures[i].height;
        s->texdata[i] = (float*)malloc(sizeof(float) * n);
        memcpy(s->texdata[i], d->textures[i].rgba, sizeof(float) * n);
        s->tex[i].width = d->textures[i].width;
        s->tex[i].height = d->textures[i].height;
        s->tex[i].rgba = s->texdata[i];
    }
    /* environment light: IBL::Init copies the map through 0.4 * c^2
     * (LightSampler.cpp:22-26, IBL.h:163-175); operator*(float, Vec4) */
    if (d->has_ibl && d->ibl_texture >= 0 && d->ibl_texture < d->ntextures) {
        const sppm_texture* t = &d->textures[d->ibl_texture];
        size_t n = 4 * (size_t)t->width * (size_t)t->height;
        s->ibl_data = (float*)malloc(sizeof(float) * n);
        for (size_t k = 0; k < n; k += 4) {
            for (int c = 0; c < 3; ++c) s->ibl_data[k + c] = 0.4f * (t->rgba[k + c] * t->rgba[k + c]);
            s->ibl_data[k + 3] = 0.4f * 0.0f;
        }
        s->ibl.width = t->width;
        s->ibl.height = t->height;
        s->ibl.rgba = s->ibl_data;
        s->has_ibl = 1;
    }
    /* Area lights, registered in facet order (Scene.cpp:145-154) */
    s->lights = (olight*)calloc((size_t)(nf > 0 ? nf : 1), sizeof(olight));
    s->nlights = 0;
    for (int f = 0; f < nf; ++f) {
        const sppm_attr* a = &s->attrs[s->attr[f]];
        v3 rad = vload(a->radiance);
        if (rad.x + rad.y + rad.z > 0.00001f) {
            const sppm_texture* t = a->tex_albedo >= 0 ? &s->tex[a->tex_albedo] : NULL;
            create_area_light(&s->lights[s->nlights++], s, f, rad, t);
        }
    }
    s->facet_light = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nf > 0 ? nf : 1));
    for (int f = 0; f < nf; ++f) s->facet_light[f] = -1;
    for (int i = 0; i < s->nlights; ++i) s->facet_light[s->lights[i].facet] = i;
    /* light pick table over |Flux| (LightSampler.cpp:6-28) */
    if (s->nlights > 0) {
        float* vals = (float*)malloc(sizeof(float) * s->nlights);
        for (int i = 0; i < s->nlights; ++i) vals[i] = vnorm(s->lights[i].flux);
        s->lprob = (float*)malloc(sizeof(float) * s->nlights);
        s->lleft = (int32_t*)malloc(sizeof(int32_t) * s->nlights);
        s->lright = (int32_t*)malloc(sizeof(int32_t) * s->nlights);
        orc_table_build(vals, s->nlights, s->lprob, s->lleft, s->lright, &s->lavg);
        free(vals);
    }
    /* oracle BVH */
    if (nf > 0) {
        s->nodes = (obvh_node*)malloc(sizeof(obvh_node) * (size_t)(2 * nf));
        s->nnodes = 0;
        int* ids = (int*)malloc(sizeof(int) * (size_t)nf);
        float* cent = (float*)malloc(sizeof(float) * 3 * (size_t)nf);
        for (int i = 0; i < nf; ++i) {
            ids[i] = i;
            for (int k = 0; k < 3; ++k)
                cent[3 * i + k] = (s->pos[9 * i + k] + s->pos[9 * i + 3 + k] + s->pos[9 * i + 6 + k]) / 3.0f;
        }
        bvh_build_rec(s, ids, nf, cent);
        free(ids);
        free(cent);
    }
    return s;
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    for (int i = 0; i < s->nlights; ++i) {
        free(s->lights[i].uvw); free(s->lights[i].L);
        free(s->lights[i].prob); free(s->lights[i].left); free(s->lights[i].right);
    }
    for (int i = 0; i < s->ntex; ++i) free(s->texdata[i]);
    free(s->texdata); free(s->tex); free(s->lights);
    free(s->lprob); free(s->lleft); free(s->lright); free(s->facet_light);
    free(s->pos); free(s->nrm); free(s->uv); free(s->attr); free(s->attrs); free(s->nodes);
    free(s->ibl_data);
    free(s);
}

int orc_scene_num_lights(const orc_scene* s) { return s->nlights; }
void orc_scene_light_info(const orc_scene* s, int32_t* facet, float* flux3, int32_t* textured,
                          float* prob, int32_t* left, int32_t* right, float* avg) {
    for (int i = 0; i < s->nlights; ++i) {
        facet[i] = s->lights[i].facet;
        vstore(flux3 + 3 * i, s->lights[i].flux);
        textured[i] = s->lights[i].textured;
        prob[i] = s->lprob[i]; left[i] = s->lleft[i]; right[i] = s->lright[i];
    }
    if (avg) *avg = s->lavg;
}

/* ---- closest hit ---------------------------------------------------------
 * Isc::Intersect(Ray, Triangle_ID) (Intersect.cpp:67-89) and the acceptance
 * rule of findClosestElement (bvh8.h:591) with ties -> lowest facet id. */
static inline int tri_intersect(const float* P, v3 org, v3 dir, float* t, float* u, float* v) {
    v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
    v3 e1 = vsub(p1, p0);
    v3 e2 = vsub(p2, p0);
    v3 ps = vcross(dir, e2);
    float a = vdot(e1, ps);
    if (a > -FLT_MIN && a < FLT_MIN) return 0;
    float f = 1.0f / a;
    v3 sv = vsub(org, p0);
    float uu = f * vdot(sv, ps);
    if ((double)uu < 0.0 || (double)uu > 1.0) return 0;
    v3 q = vcross(sv, e1);
    float vv = f * vdot(dir, q);
    if ((double)vv < 0.0 || (double)(uu + vv) > 1.0) return 0;
    *t = f * vdot(e2, q);
    *u = uu;
    *v = vv;
    return 1;
}
static inline void try_tri(const orc_scene* s, int f, v3 org, v3 dir, float tmin, float* bt, float* bu,
                           float* bv, int* bf) {
    float t, u, v;
    if (tri_intersect(s->pos + 9 * f, org, dir, &t, &u, &v)) {
        if (t > tmin && (t < *bt || (t == *bt && (*bf < 0 || f < *bf)))) {
            *bt = t; *bu = u; *bv = v; *bf = f;
        }
    }
}
static int slab(const obvh_node* n, v3 o, v3 id, float tmax) {
    float tn = -FLT_MAX, tf = FLT_MAX;
    for (int k = 0; k < 3; ++k) {
        float ok = vget(o, k), ik = vget(id, k);
        float t0 = (n->bmin[k] - ok) * ik, t1 = (n->bmax[k] - ok) * ik;
        if (t0 > t1) { float x = t0; t0 = t1; t1 = x; }
        if (t0 > tn) tn = t0;
        if (t1 < tf) tf = t1;
    }
    /* conservative: accept equality so ties are always visited */
    return tf >= tn && tn <= tmax && tf >= 0.0f;
}
int orc_closest_hit(const orc_scene* s, int mode, const float* org_, const float* dir_, float tmin,
                    float tmax, float* t, float* u, float* v) {
    v3 org = vload(org_), dir = vload(dir_);
    float bt = tmax, bu = 0, bv = 0;
    int bf = -1;
    if (mode == 0 || s->nnodes == 0) {
        for (int f = 0; f < s->nf; ++f) try_tri(s, f, org, dir, tmin, &bt, &bu, &bv, &bf);
    } else {
        /* inverse direction in double then rounded (exact-ish), guarded */
        v3 id;
        id.x = fabsf(dir.x) < 1e-16f ? (dir.x >= 0 ? 1e16f : -1e16f) : (float)(1.0 / (double)dir.x);
        id.y = fabsf(dir.y) < 1e-16f ? (dir.y >= 0 ? 1e16f : -1e16f) : (float)(1.0 / (double)dir.y);
        id.z = fabsf(dir.z) < 1e-16f ? (dir.z >= 0 ? 1e16f : -1e16f) : (float)(1.0 / (double)dir.z);
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const obvh_node* n = &s->nodes[stack[--sp]];
            if (!slab(n, org, id, bt)) continue;
            if (n->prim >= 0) {
                try_tri(s, n->prim, org, dir, tmin, &bt, &bu, &bv, &bf);
            } else {
                if (sp + 2 > 128) { /* fall back: brute force (never expected) */
                    for (int f = 0; f < s->nf; ++f) try_tri(s, f, org, dir, tmin, &bt, &bu, &bv, &bf);
                    break;
                }
                stack[sp++] = n->left;
                stack[sp++] = n->right;
            }
        }
    }
    /* hit iff best t < MaxT (bvh8.h:640); MaxT - 1e-10 == MaxT for FLT_MAX */
    if (bf < 0 || !((double)bt < (double)tmax - 1e-10)) {
        *t = FLT_MAX; *u = 0; *v = 0;
        return -1;
    }
    *t = bt; *u = bu; *v = bv;
    return bf;
}
void orc_closest_hit_batch(const orc_scene* s, int mode, int n, const float* org, const float* dir,
                           float tmin, float tmax, int32_t* facet, float* t, float* u, float* v) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        facet[i] = orc_closest_hit(s, mode, org + 3 * i, dir + 3 * i, tmin, tmax, t + i, u + i, v + i);
}

/* Scene::IsOccluded (Scene.cpp:246-281), recursive as written there: the
 * closest hit in [MinDist, MaxDist) occludes unless its albedo texture gives
 * 1 - alpha > translucency; then a new Ray (ctor normalises) starts
 * RAY_OFFSET behind the hit with MinT = max(MinDist - dist, RAY_OFFSET) and
 * MaxT = MaxDist - dist.  depth guards the recursion (returns 2 past it). */
static int occluded_rec(const orc_scene* s, v3 org, v3 dir, float mind, float maxd, float transl, int depth) {
    if (depth > 4096) return 2;
    float t, u, v;
    float o3[3] = {org.x, org.y, org.z}, d3[3] = {dir.x, dir.y, dir.z};
    int f = orc_closest_hit(s, 1, o3, d3, mind, maxd, &t, &u, &v);
    if (f < 0) return 0;
    const sppm_attr* a = &s->attrs[s->attr[f]];
    if (a->tex_albedo < 0) return 1;
    const float* UV = s->uv + 9 * f;
    v3 t0 = vload(UV), t1 = vload(UV + 3), t2 = vload(UV + 6);
    v3 uv = vadd(vadd(vscale(vsub(t1, t0), u), vscale(vsub(t2, t0), v)), t0);
    float c[4];
    orc_bilinear(&s->tex[a->tex_albedo], uv.x, uv.y, c);
    float tr = (float)(1.0 - (double)c[3]);
    if (!(tr > transl)) return 1;
    float dist = t + RAY_OFFSET;
    v3 pos = vadd(org, vscale(dir, dist));
    float mn = mind - dist;
    return occluded_rec(s, pos, vnormalized(dir), mn > RAY_OFFSET ? mn : RAY_OFFSET, maxd - dist, transl, depth + 1);
}
void orc_occluded_batch(const orc_scene* s, int n, const float* org, const float* dir, float tmin,
                        const float* tmax, float transl, int32_t* occluded) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        occluded[i] = occluded_rec(s, vload(org + 3 * i), vload(dir + 3 * i), tmin, tmax[i], transl, 0);
}

/* ======================================================================
 * Camera (Camera.h:34-129)
 * ====================================================================== */
typedef struct {
    int type, W, H;
    v3 pos, dir, e1, e2;
    float fov, PW, PH, LW, LH, aspect, F, D;
    /* ThinLensCamera's m_Table over the diaphragm image (dn pixels; 0: plain
     * lens sampling, m_HoleRatio > 0.999 or no image) */
    int dn;
    float *dprob, *du, *dv;
    int32_t *dleft, *dright;
} ocam;
static void cam_setup(const sppm_camera_desc* c, ocam* o) {
    o->dn = 0;
    o->dprob = o->du = o->dv = NULL;
    o->dleft = o->dright = NULL;
    o->type = c->type; o->W = c->width; o->H = c->height;
    o->pos = vload(c->pos);
    o->dir = vnormalized(vsub(vload(c->target), o->pos));
    o->fov = PIf * c->fov_deg / 360.0f;
    o->F = c->F; o->D = c->dist_to_lens; o->aspect = c->aspect;
    o->PW = (float)(2.0 * (double)c->dist_to_lens * tan((double)o->fov) * (double)c->aspect);
    o->PH = (float)(2.0 * (double)c->dist_to_lens * tan((double)o->fov));
    o->LW = o->PW; o->LH = o->PH;
    if ((double)o->dir.y > 0.99) {
        o->e1 = V3(1, 0, 0);
        o->e2 = V3(0, 0, 1);
    } else {
        o->e1 = vnormalized(vcross(V3(0, 1, 0), o->dir));
        o->e2 = vnormalized(vcross(o->dir, o->e1));
    }
}
/* returns org, dir (already through the Ray ctor's normalisation) and pdf */
static void cam_ray(const ocam* o, float u, float v, v3 smp, v3* org, v3* dir, float* pdf) {
    if (o->type == SPPM_CAM_PINHOLE) {  /* Camera.h:73-82 */
        v3 off = vdiv(vmul(smp, V3(o->PW, o->PH, 0)), V3((float)o->W, (float)o->H, 1.0f));
        v3 lens = o->pos;
        v3 img = vsub(o->pos, vscale(o->dir, o->D));
        img = vadd(img, vscale(vscale(o->e1, u - 0.5f + off.x), o->PW));
        img = vadd(img, vscale(vscale(o->e2, v - 0.5f + off.y), o->PH));
        *pdf = 1.0f;
        *org = img;
        *dir = vnormalized(vnormalized(vsub(lens, img)));
    } else {  /* Camera.h:103-129 */
        v3 luv = smp;
        if (o->dn > 0) {
            /* LensUV = m_Table.Sample(Sample(index).xy()).UV (Camera.h:109-111):
             * bin = m_NumElements (int) * x in float, truncated; left when
             * y < probLeftSample (TableSampler.h:130-133).  A bin index of n
             * (x rounded up to 1) or a right id of -1 (y = 1 on a bin of
             * probability 1) are out of the table in the reference: the last
             * bin and the left pixel are taken. */
            int bin = (int)((float)o->dn * smp.x);
            if (bin > o->dn - 1) bin = o->dn - 1;
            int id = (smp.y < o->dprob[bin]) ? o->dleft[bin] : o->dright[bin];
            if (id < 0) id = o->dleft[bin];
            luv = V3(o->du[id], o->dv[id], 0.0f);
        }
        float pdf_area = 1.0f / (o->LW * o->LH);
        v3 lens = vadd(vadd(o->pos, vscale(vscale(o->e1, luv.x - 0.5f), o->LW)),
                       vscale(vscale(o->e2, luv.y - 0.5f), o->LH));
        v3 center = o->pos;
        v3 img = vsub(o->pos, vscale(o->dir, o->D));
        img = vadd(img, vscale(vscale(o->e1, u - 0.5f), o->PW));
        img = vadd(img, vscale(vscale(o->e2, v - 0.5f), o->PH));
        float A = o->D;
        float B = (A * o->F) / (A - o->F);
        v3 target = vadd(img, vdivs(vscale(vsub(center, img), A + B), A));
        float cs = vdot(o->dir, vnormalized(vsub(lens, img)));
        float d2 = vnorm2(vsub(lens, img));
        *pdf = pdf_area * (d2 / o->F / o->F) / cs;
        *org = lens;
        *dir = vnormalized(vnormalized(vsub(target, lens)));
    }
}
static void cam_diaphragm_free(ocam* o) {
    free(o->dprob); free(o->du); free(o->dv); free(o->dleft); free(o->dright);
    o->dprob = o->du = o->dv = NULL;
    o->dleft = o->dright = NULL;
    o->dn = 0;
}
/* ThinLensCamera(..., DiaphragmImageFileName) (Camera.h:136-158): per pixel
 * i of the image, UV = ((i % W) / W, 1 - (i / W) / H) and L = 1 - red; a
 * pixel is a hole when red < 1e-6; m_HoleRatio = holes / (W H) in float and
 * more than 0.999 keeps plain lens sampling; else m_Table = alias table over
 * L.  Returns 0, or -1 when the image has no weight (the reference's table
 * would divide by a zero average). */
static int cam_diaphragm(ocam* o, int w, int h, const float* rgba) {
    cam_diaphragm_free(o);
    if (!rgba || w <= 0 || h <= 0) return 0;
    const int n = w * h;
    float* L = (float*)malloc(sizeof(float) * (size_t)n);
    o->du = (float*)malloc(sizeof(float) * (size_t)n);
    o->dv = (float*)malloc(sizeof(float) * (size_t)n);
    int holes = 0;
    for (int i = 0; i < n; ++i) {
        const float red = rgba[4 * (size_t)i];
        if ((double)red < 1e-6) ++holes;
        o->du[i] = (float)(i % w) / (float)w;
        o->dv[i] = (float)(1.0 - (double)((float)(i / w) / (float)h));
        L[i] = (float)(1.0 - (double)red);
    }
    const float ratio = (float)holes / (float)n;
    if ((double)ratio > 0.999) {
        free(L);
        cam_diaphragm_free(o);
        return 0;
    }
    o->dprob = (float*)malloc(sizeof(float) * (size_t)n);
    o->dleft = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    o->dright = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    float avg = 0.0f;
    orc_table_build(L, n, o->dprob, o->dleft, o->dright, &avg);
    free(L);
    if (!(avg > 0.0f)) {
        cam_diaphragm_free(o);
        return -1;
    }
    o->dn = n;
    return 0;
}
int orc_camera_ray_diaphragm(const sppm_camera_desc* c, int dw, int dh, const float* drgba, float img_u,
                             float img_v, const float* sample3, float* org, float* dir, float* pdf) {
    ocam o;
    cam_setup(c, &o);
    if (cam_diaphragm(&o, dw, dh, drgba)) return -1;
    v3 a, b;
    cam_ray(&o, img_u, img_v, vload(sample3), &a, &b, pdf);
    cam_diaphragm_free(&o);
    vstore(org, a);
    vstore(dir, b);
    return 0;
}
void orc_camera_ray(const sppm_camera_desc* c, float img_u, float img_v, const float* sample3,
                    float* org, float* dir, float* pdf) {
    ocam o;
    cam_setup(c, &o);
    v3 a, b;
    cam_ray(&o, img_u, img_v, vload(sample3), &a, &b, pdf);
    vstore(org, a);
    vstore(dir, b);
}

/* ======================================================================
 * SPPM pass
 * ====================================================================== */
typedef struct {  /* StochasticProgressivePhotonMapping::Photon (SPPM.h:128-143) */
    int valid;
    v3 pos, flux, indir;
    float index, pdf;
} ophoton;

typedef struct {  /* HitPoint fields the gather reads (SPPM.h:83-110) */
    int type;     /* 0 NONE, 1 HIT, 2 ONLY_DIRECTLIGHT */
    int facet;
    v3 pos, outdir, weight, normal, direct;
    int kind;
    float index, rough, absorb, refl;
} ohit;

typedef struct {  /* SharedHitPoint (SPPM.h:112-126) */
    v3 radiance, direct;
    int num_photon;
    float radius;
    int num_gather, num_direct;
} oshared;

struct orc_guide;
struct orc_render {
    const struct orc_guide* guide;  /* guided emission (NULL: the reference's plain emission) */
    const orc_scene* s;
    ocam cam;
    sppm_camera_desc camdesc;
    sppm_params p;
    int nthreads;
    int W, H;
    ophoton* slots;   /* N * NUM_SLOTS */
    int nslots_paths;
    ophoton* flat;
    int nflat;
    ohit* hits;
    oshared* shared;
    int32_t* counts;
    int64_t* nb_off;
    int32_t* nb_idx;
    int64_t nb_total;
    orc_mt camstream;   /* UniformSample's shared engine, persists across passes */
};

orc_render* orc_render_create(const orc_scene* s, const sppm_camera_desc* cam, const sppm_params* p,
                              int nthreads) {
    orc_render* r = (orc_render*)calloc(1, sizeof(orc_render));
    r->s = s;
    r->camdesc = *cam;
    cam_setup(cam, &r->cam);
    r->p = *p;
    r->nthreads = nthreads > 0 ? nthreads : 1;
    r->W = cam->width;
    r->H = cam->height;
    r->nslots_paths = p->photons_per_pass;
    r->slots = NULL;  /* allocated by the first photon pass */
    size_t np = (size_t)r->W * (size_t)r->H;
    r->hits = (ohit*)calloc(np, sizeof(ohit));
    r->shared = (oshared*)calloc(np, sizeof(oshared));
    r->counts = (int32_t*)calloc(np, sizeof(int32_t));
    /* initializeSharedPoints (SPPM.cpp:356-373) */
    for (size_t i = 0; i < np; ++i) r->shared[i].radius = p->first_radius;
    if (cam->sampler == SPPM_SAMPLER_UNIFORM) orc_mt_seed(&r->camstream, (uint32_t)cam->sampler_id);
    return r;
}
int orc_render_diaphragm(orc_render* r, int w, int h, const float* rgba) {
    return cam_diaphragm(&r->cam, w, h, rgba);
}
void orc_render_destroy(orc_render* r) {
    if (!r) return;
    cam_diaphragm_free(&r->cam);
    free(r->slots); free(r->flat); free(r->hits); free(r->shared); free(r->counts);
    free(r->nb_off); free(r->nb_idx);
    free(r);
}

typedef struct {
    v3 pos, N, uv;
    int facet, attr, kind;
    float albedo_[3];
    v3 albedo, radiance;
    float rough, index, absorb, refl;
    int opaque, is_sppm;
    float transl;
} geom;

/* calculateGeometry (SPPM.cpp:36-77) */
static void calc_geometry(const orc_scene* s, int f, float u, float v, geom* g) {
    const float* P = s->pos + 9 * f;
    const float* Nn = s->nrm + 9 * f;
    const float* UV = s->uv + 9 * f;
    v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
    g->pos = vadd(vadd(vscale(vsub(p1, p0), u), vscale(vsub(p2, p0), v)), p0);
    v3 n0 = vload(Nn), n1 = vload(Nn + 3), n2 = vload(Nn + 6);
    g->N = vnormalized(vadd(vadd(vscale(vsub(n1, n0), u), vscale(vsub(n2, n0), v)), n0));
    v3 t0 = vload(UV), t1 = vload(UV + 3), t2 = vload(UV + 6);
    g->uv = vadd(vadd(vscale(vsub(t1, t0), u), vscale(vsub(t2, t0), v)), t0);
    g->facet = f;
    const sppm_attr* a = &s->attrs[s->attr[f]];
    g->attr = s->attr[f];
    g->kind = a->kind;
    g->albedo = vload(a->albedo);
    g->radiance = vload(a->radiance);
    g->rough = a->roughness;
    g->index = a->index;
    g->absorb = a->extinction;
    g->refl = a->reflectance;
    g->opaque = a->if_opaque;
    g->is_sppm = a->is_sppm;
    g->transl = 0.0f;
    if ((double)vnorm2(g->radiance) > 1e-6) g->albedo = V3(1, 1, 1);
    if (a->tex_roughness >= 0) {
        float c[4];
        orc_bilinear(&s->tex[a->tex_roughness], g->uv.x, g->uv.y, c);
        g->rough = fminf_(0.8f, fmaxf_(0.05f, 0.5f * c[0]));
    }
    if (a->tex_albedo >= 0) {
        float c[4];
        orc_bilinear(&s->tex[a->tex_albedo], g->uv.x, g->uv.y, c);
        g->albedo = V3(c[0], c[1], c[2]);
        g->radiance = vmul(g->radiance, g->albedo);
        g->transl = (float)(1.0 - (double)c[3]);
    }
    if (a->tex_bump >= 0) {
        v3 nm = normal_from_map(&s->tex[a->tex_bump], g->uv.x, g->uv.y);
        float z = nm.z, y = nm.y;
        nm = vnormalized(V3(nm.x, z, y));
        v3 e2 = g->N;
        v3 e1 = (double)fabsf(e2.y) > 0.9 ? vcross(e2, V3(1, 0, 0)) : vcross(e2, V3(0, 1, 0));
        v3 e3 = vnormalized(vcross(e1, e2));
        /* CreateMatrixFromColumns({e1,e2,e3}): columns -> row-major m[i][j] = col_j[i] */
        float m[9] = {e1.x, e2.x, e3.x, e1.y, e2.y, e3.y, e1.z, e2.z, e3.z};
        float mi[9];
        orc_matrix3_inv(m, mi);
        v3 r0 = vload(mi), r1 = vload(mi + 3), r2 = vload(mi + 6);
        g->N = vnormalized(V3(vdot(r0, nm), vdot(r1, nm), vdot(r2, nm)));
    }
}

static inline int near1(float s_) { return (double)fabsf(s_ - 1.0f) < 0.00001; }  /* SPPM.h:27-29 */

/* LightSampler::SamplePointOnAreaLight (LightSampler.cpp:68-82) */
static void light_point(const orc_scene* s, const olight* L, orc_mt* m, v3* lpos, v3* rad, float* pdf_area) {
    const float* P = s->pos + 9 * L->facet;
    v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
    float area = facet_area(P);
    if (L->textured) {  /* LightSampler.cpp:70-76 */
        int k = table_sample(L->prob, L->left, L->right, LIGHT_TABLE_SAMPLES, m);
        v3 w = L->uvw[k];
        *lpos = vadd(vadd(vscale(p0, w.x), vscale(p1, w.y)), vscale(p2, 1.0f - w.x - w.y));
        *rad = L->L[k];
        float gp = vnorm(L->L[k]) / L->avg / (float)LIGHT_TABLE_SAMPLES;
        *pdf_area = (gp * (float)LIGHT_TABLE_SAMPLES) / area;
    } else {  /* PointOnTriangle (Sampling.cpp:339-349) */
        float u = (float)orc_mt_real1(m);
        float v = (float)orc_mt_real1(m);
        if (v > u) { u *= 0.5f; v -= u; }
        else { v *= 0.5f; u -= v; }
        *pdf_area = 2.0f / fabsf(vnorm(vcross(vsub(p1, p0), vsub(p2, p0))));
        *lpos = vadd(vadd(vscale(p0, u), vscale(p1, v)), vscale(p2, 1.0f - u - v));
        *rad = vdivs(L->flux, 2.0f * PIf * area);
    }
}
static v3 light_normal_of(const orc_scene* s, const olight* L);
/* LightSampler::GetAreaLightPDF (LightSampler.h:87) = TableSampler::GetPDF */
static float light_pick_pdf(const orc_scene* s, const olight* L) {
    return vnorm(L->flux) / s->lavg / (float)s->nlights;
}
/* (n0 + n1 + n2).normalized() (SPPM.cpp:465-467, 201) */
static v3 light_normal(const orc_scene* s, const olight* L) {
    const float* Ln = s->nrm + 9 * L->facet;
    return vnormalized(vadd(vadd(vload(Ln), vload(Ln + 3)), vload(Ln + 6)));
}
static v3 light_normal_of(const orc_scene* s, const olight* L) { return light_normal(s, L); }

/* ======================================================================
 * Guided emission (no counterpart in the reference: it replaces the
 * position x direction sampling of generateLightRay, SPPM.cpp:453-471, and
 * the PDF the photon carries, SPPM.cpp:527).  Restates the product's
 * mixture draw and density: with probability w the plain emission (light
 * pick, point on the light or table point, cosine lobe about the light
 * rectangle's normal), otherwise a proposal cell (32x32 light (u, v) x 16x16
 * hemi-octahedral direction (s, t)) by CDF search and a uniform point and
 * direction inside it (a table point of the position cell, by the plain
 * probabilities, for a textured light).  The photon's PDF is the mixture
 * density at the drawn (point, direction).
 * ====================================================================== */
#define G_GU 32
#define G_GV 32
#define G_GS 16
#define G_GT 16
#define G_CELLS (G_GU * G_GV * G_GS * G_GT)

struct orc_guide {
    v3 P0, e1, e2, n, t1, t2;
    float inv_e1, inv_e2, inv_area;
    v3 rad;
    float w;
    int textured;
    float* prob;  /* G_CELLS, unnormalised; total = cdf[G_CELLS - 1] */
    float* cdf;
    /* textured: table point (light li, entry k) -> position cell; points by cell */
    int32_t* pt_cell;  /* 2 * 1000, index li * 1000 + k */
    float* pcell_mass; /* G_GU * G_GV */
    int32_t* pc_start; /* G_GU * G_GV + 1 */
    int32_t* pc_light; /* light and entry of each point, in cell order */
    int32_t* pc_entry;
    float* pc_cum;     /* inclusive cumulative plain probability within the cell */
};

/* The light frame: exactly one rectangle made of two facets.  Corner P0 is
 * one end of the longest vertex-to-vertex distance (the diagonal), e1 and e2
 * run to the two other vertices, ordered so that e1 x e2 faces like the
 * facets' normal. */
orc_guide* orc_guide_create(const orc_scene* s, float mix_plain) {
    if (s->nlights != 2 || s->lights[0].textured != s->lights[1].textured) return NULL;
    v3 v[6];
    int nv = 0;
    for (int li = 0; li < 2; ++li) {
        const float* P = s->pos + 9 * s->lights[li].facet;
        for (int k = 0; k < 3; ++k) {
            v3 p = vload(P + 3 * k);
            int dup = 0;
            for (int q = 0; q < nv; ++q) dup = dup || vnorm2(vsub(p, v[q])) < 1e-12f;
            if (!dup) v[nv++] = p;
        }
    }
    if (nv != 4) return NULL;
    int a = 0, b = 1;
    float best = -1.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j)
            if (vnorm2(vsub(v[i], v[j])) > best) { best = vnorm2(vsub(v[i], v[j])); a = i; b = j; }
    int o[2], k = 0;
    for (int i = 0; i < 4; ++i)
        if (i != a && i != b) o[k++] = i;
    orc_guide* g = (orc_guide*)calloc(1, sizeof(orc_guide));
    g->P0 = v[a];
    v3 e1 = vsub(v[o[0]], g->P0), e2 = vsub(v[o[1]], g->P0);
    const v3 ln = light_normal_of(s, &s->lights[0]);
    if (vdot(vcross(e1, e2), ln) < 0.0f) { v3 t = e1; e1 = e2; e2 = t; }
    const float l1 = sqrtf(vnorm2(e1)), l2 = sqrtf(vnorm2(e2));
    g->e1 = e1;
    g->e2 = e2;
    g->n = vnormalized(vcross(e1, e2));
    g->t1 = vdivs(e1, l1);
    g->t2 = vcross(g->n, g->t1);
    g->inv_e1 = 1.0f / vnorm2(e1);
    g->inv_e2 = 1.0f / vnorm2(e2);
    g->inv_area = 1.0f / (l1 * l2);
    {
        const olight* L0 = &s->lights[0];
        const float area = facet_area(s->pos + 9 * L0->facet);
        g->rad = V3(L0->flux.x / (2.0f * PIf * area), L0->flux.y / (2.0f * PIf * area), L0->flux.z / (2.0f * PIf * area));
    }
    g->w = mix_plain;
    g->textured = s->lights[0].textured;
    g->prob = (float*)calloc(G_CELLS, sizeof(float));
    g->cdf = (float*)calloc(G_CELLS, sizeof(float));
    if (g->textured) {
        /* plain probability of point k of light li: P(pick li) * value_k /
         * (1000 * average) (LightSampler.cpp:68-76); its position cell */
        const int ncell = G_GU * G_GV;
        g->pt_cell = (int32_t*)malloc(sizeof(int32_t) * 2 * LIGHT_TABLE_SAMPLES);
        double* pi = (double*)malloc(sizeof(double) * 2 * LIGHT_TABLE_SAMPLES);
        for (int li = 0; li < 2; ++li) {
            const olight* L = &s->lights[li];
            const double pick = (double)vnorm(L->flux) / ((double)s->lavg * (double)s->nlights);
            const float* P = s->pos + 9 * L->facet;
            const v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
            for (int q = 0; q < LIGHT_TABLE_SAMPLES; ++q) {
                const v3 w = L->uvw[q];
                const float val = vnorm(L->L[q]);
                pi[li * LIGHT_TABLE_SAMPLES + q] = pick * (double)val / ((double)L->avg * 1000.0);
                const v3 pos = vadd(vadd(vscale(p0, w.x), vscale(p1, w.y)), vscale(p2, 1.0f - w.x - w.y));
                const v3 rel = vsub(pos, g->P0);
                const float u = fminf_(fmaxf_(vdot(rel, e1) * g->inv_e1, 0.0f), 0.99999994f);
                const float vv = fminf_(fmaxf_(vdot(rel, e2) * g->inv_e2, 0.0f), 0.99999994f);
                int iu = (int)(u * (float)G_GU), iv = (int)(vv * (float)G_GV);
                if (iu > G_GU - 1) iu = G_GU - 1;
                if (iv > G_GV - 1) iv = G_GV - 1;
                g->pt_cell[li * LIGHT_TABLE_SAMPLES + q] = iu * G_GV + iv;
            }
        }
        g->pcell_mass = (float*)calloc(ncell, sizeof(float));
        g->pc_start = (int32_t*)calloc(ncell + 1, sizeof(int32_t));
        g->pc_light = (int32_t*)malloc(sizeof(int32_t) * 2 * LIGHT_TABLE_SAMPLES);
        g->pc_entry = (int32_t*)malloc(sizeof(int32_t) * 2 * LIGHT_TABLE_SAMPLES);
        g->pc_cum = (float*)malloc(sizeof(float) * 2 * LIGHT_TABLE_SAMPLES);
        int npts = 0;
        for (int c = 0; c < ncell; ++c) {  /* cells in order, points in (light, entry) order */
            g->pc_start[c] = npts;
            double run = 0.0;
            for (int j = 0; j < 2 * LIGHT_TABLE_SAMPLES; ++j) {
                if (g->pt_cell[j] != c || !(pi[j] > 0.0)) continue;
                run += pi[j];
                g->pc_light[npts] = j / LIGHT_TABLE_SAMPLES;
                g->pc_entry[npts] = j % LIGHT_TABLE_SAMPLES;
                g->pc_cum[npts] = (float)run;
                ++npts;
            }
            g->pcell_mass[c] = npts > g->pc_start[c] ? g->pc_cum[npts - 1] : 0.0f;
        }
        g->pc_start[ncell] = npts;
        free(pi);
    }
    return g;
}
void orc_guide_destroy(orc_guide* g) {
    if (!g) return;
    free(g->prob); free(g->cdf); free(g->pt_cell); free(g->pcell_mass); free(g->pc_start);
    free(g->pc_light); free(g->pc_entry); free(g->pc_cum);
    free(g);
}
void orc_guide_set_tables(orc_guide* g, const float* prob, const float* cdf) {
    memcpy(g->prob, prob, sizeof(float) * G_CELLS);
    memcpy(g->cdf, cdf, sizeof(float) * G_CELLS);
}
void orc_render_set_guide(orc_render* r, const orc_guide* g) { r->guide = g; }
void orc_guide_frame(const orc_guide* g, float* out) {  /* P0, e1, e2, n, t1, t2 (18 floats) */
    vstore(out, g->P0); vstore(out + 3, g->e1); vstore(out + 6, g->e2);
    vstore(out + 9, g->n); vstore(out + 12, g->t1); vstore(out + 15, g->t2);
}
/* textured light: plain probability mass and point count of each position cell */
void orc_guide_point_cells(const orc_guide* g, float* mass, int32_t* start, int32_t* pt_cell) {
    if (!g->textured) return;
    if (mass) memcpy(mass, g->pcell_mass, sizeof(float) * G_GU * G_GV);
    if (start) memcpy(start, g->pc_start, sizeof(int32_t) * (G_GU * G_GV + 1));
    if (pt_cell) memcpy(pt_cell, g->pt_cell, sizeof(int32_t) * 2 * LIGHT_TABLE_SAMPLES);
}

static int cdf_search(const float* cdf, int lo, int hi, float x) {  /* first entry > x in [lo, hi] */
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

/* Mixture density of (table point (tl, tk) or position lpos, direction
 * outdir); also the emission cell and the emitted radiance. */
static float guide_density(const orc_scene* s, const orc_guide* G, int tl, int tk, v3 lpos, v3 outdir, int* cell_out,
                           v3* rad0) {
    const float total = G->cdf[G_CELLS - 1];
    const float lx = vdot(outdir, G->t1), ly = vdot(outdir, G->t2), lz = vdot(outdir, G->n);
    const float cs = lz > 0.0f ? lz : 0.0f;
    const float l1 = fabsf(lx) + fabsf(ly) + cs;
    const float ox = lx / l1, oy = ly / l1, oz = cs / l1;
    const float ss = ox + oy, tt = ox - oy;
    int is = (int)((ss + 1.0f) * (0.5f * (float)G_GS)), it = (int)((tt + 1.0f) * (0.5f * (float)G_GT));
    is = is < 0 ? 0 : (is > G_GS - 1 ? G_GS - 1 : is);
    it = it < 0 ? 0 : (it > G_GT - 1 ? G_GT - 1 : it);
    int pcell;
    if (G->textured) {
        pcell = G->pt_cell[tl * LIGHT_TABLE_SAMPLES + tk];
    } else {
        const v3 rel = vsub(lpos, G->P0);
        const float u = fminf(fmaxf(vdot(rel, G->e1) * G->inv_e1, 0.0f), 0.99999994f);
        const float v = fminf(fmaxf(vdot(rel, G->e2) * G->inv_e2, 0.0f), 0.99999994f);
        int iu = (int)(u * (float)G_GU), iv = (int)(v * (float)G_GV);
        if (iu > G_GU - 1) iu = G_GU - 1;
        if (iv > G_GV - 1) iv = G_GV - 1;
        pcell = iu * G_GV + iv;
    }
    const int cell = (pcell * G_GS + is) * G_GT + it;
    const float plen = sqrtf(ox * ox + oy * oy + oz * oz);
    /* direction density inside a cell: 1 / (cell area in (s, t)) * 2|P|^3, since d omega = ds dt / (2|P|^3) */
    const float dd = (float)(G_GS * G_GT / 4) * (2.0f * plen * plen * plen);
    const float pc_prob = G->prob[cell] / total;
    float q;
    if (G->textured) {
        const olight* L = &s->lights[tl];
        const float area = facet_area(s->pos + 9 * L->facet);
        const float gp = vnorm(L->L[tk]) / L->avg / 1000.0f;
        const float Dk = ((gp * 1000.0f) / area) * (vnorm(L->flux) / s->lavg / (float)s->nlights);
        const float mass = G->pcell_mass[pcell];
        const float prop = mass > 0.0f ? pc_prob * dd / mass : 0.0f;
        q = Dk * (G->w * cs / PIf + (1.0f - G->w) * prop);
        if (rad0) *rad0 = L->L[tk];
    } else {
        const float p_mlp = pc_prob * (float)(G_GU * G_GV) * G->inv_area * dd;
        const float p_plain = G->inv_area * cs / PIf;
        q = G->w * p_plain + (1.0f - G->w) * p_mlp;
        if (rad0) *rad0 = G->rad;
    }
    if (cell_out) *cell_out = cell;
    return q;
}
float orc_guide_density(const orc_scene* s, const orc_guide* g, int light, int point, const float* pos,
                        const float* dir, int32_t* cell) {
    int c = 0;
    const float q = guide_density(s, g, light, point, vload(pos), vload(dir), &c, NULL);
    if (cell) *cell = c;
    return q;
}

void orc_guide_density_batch(const orc_scene* s, const orc_guide* g, int n, const int32_t* light,
                             const int32_t* point, const float* pos, const float* dir, float* q) {
    for (int j = 0; j < n; ++j)
        q[j] = guide_density(s, g, light ? light[j] : 0, point ? point[j] : 0, vload(pos + 3 * (size_t)j),
                             vload(dir + 3 * (size_t)j), NULL, NULL);
}

/* The mixture draw: position, direction, PDF, radiance and cell. */
static void guided_start(const orc_scene* s, const orc_guide* G, orc_mt* m, v3* lpos_o, v3* dir_o, float* pdf_o,
                         v3* rad_o, int* cell_o) {
    const float total = G->cdf[G_CELLS - 1];
    v3 lpos, outdir;
    int tl = 0, tk = 0;
    if (orc_mt_real1(m) < (double)G->w) {
        /* plain: light pick, then PointOnTriangle or a table point, cosine lobe (SPPM.cpp:453-468) */
        const int li = table_sample(s->lprob, s->lleft, s->lright, s->nlights, m);
        const olight* L = &s->lights[li];
        const float* P = s->pos + 9 * L->facet;
        const v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
        if (G->textured) {
            const int k = table_sample(L->prob, L->left, L->right, LIGHT_TABLE_SAMPLES, m);
            const v3 w = L->uvw[k];
            lpos = vadd(vadd(vscale(p0, w.x), vscale(p1, w.y)), vscale(p2, 1.0f - w.x - w.y));
            tl = li;
            tk = k;
        } else {
            float u = (float)orc_mt_real1(m);
            float v = (float)orc_mt_real1(m);
            if (v > u) { u *= 0.5f; v -= u; }
            else { v *= 0.5f; u -= v; }
            lpos = vadd(vadd(vscale(p0, u), vscale(p1, v)), vscale(p2, 1.0f - u - v));
        }
        float pdf_dir = 1.0f;
        outdir = vnormalized(cos_sampling(m, G->n, &pdf_dir));
    } else {
        /* proposal: a cell by its CDF (x < total, so the cell has probability > 0) */
        const float xr = (float)orc_mt_real1(m) * total;
        union { float f; uint32_t u; } below;
        below.f = total;
        below.u -= 1u;
        const float x = fminf(xr, below.f);
        const int c = cdf_search(G->cdf, 0, G_CELLS - 1, x);
        const int it = c % G_GT, is = (c / G_GT) % G_GS, iv = (c / (G_GS * G_GT)) % G_GV, iu = c / (G_GV * G_GS * G_GT);
        if (G->textured) {
            const int pc = iu * G_GV + iv;
            const int a = G->pc_start[pc], e = G->pc_start[pc + 1];
            const float x2 = (float)orc_mt_real1(m) * G->pcell_mass[pc];
            const int q = e > a ? cdf_search(G->pc_cum, a, e - 1, x2) : a;
            tl = G->pc_light[q];
            tk = G->pc_entry[q];
            const olight* L = &s->lights[tl];
            const float* P = s->pos + 9 * L->facet;
            const v3 p0 = vload(P), p1 = vload(P + 3), p2 = vload(P + 6);
            const v3 w = L->uvw[tk];
            lpos = vadd(vadd(vscale(p0, w.x), vscale(p1, w.y)), vscale(p2, 1.0f - w.x - w.y));
        } else {
            const float u = ((float)iu + (float)orc_mt_real1(m)) / (float)G_GU;
            const float v = ((float)iv + (float)orc_mt_real1(m)) / (float)G_GV;
            lpos = vadd(vadd(G->P0, vscale(G->e1, u)), vscale(G->e2, v));
        }
        const float ss = -1.0f + ((float)is + (float)orc_mt_real1(m)) * (2.0f / (float)G_GS);
        const float tt = -1.0f + ((float)it + (float)orc_mt_real1(m)) * (2.0f / (float)G_GT);
        const float px = 0.5f * (ss + tt), py = 0.5f * (ss - tt);
        const float pz = fmaxf(0.0f, 1.0f - fabsf(px) - fabsf(py));
        outdir = vnormalized(vadd(vadd(vscale(G->t1, px), vscale(G->t2, py)), vscale(G->n, pz)));
    }
    v3 rad0;
    const float q = guide_density(s, G, tl, tk, lpos, outdir, cell_o, &rad0);
    const float cs = fmaxf(vdot(outdir, G->n), 0.0f);
    *rad_o = vscale(rad0, fminf_(1.0f, cs));
    *lpos_o = lpos;
    *dir_o = outdir;
    *pdf_o = q;
}

/* The start of guided photon path i of pass `step`: slot-0 position, direction, PDF, cell. */
void orc_guide_emit(const orc_render* r, int step, int path, float* pos, float* dir, float* pdf, int32_t* cell) {
    const int N = r->p.photons_per_pass;
    orc_mt m;
    orc_mt_seed(&m, (uint32_t)path + 135u + (uint32_t)step * (uint32_t)N);
    v3 lpos, d, rad;
    int c = 0;
    guided_start(r->s, r->guide, &m, &lpos, &d, pdf, &rad, &c);
    vstore(pos, vadd(lpos, vscale(r->guide->n, RAY_OFFSET)));
    vstore(dir, d);
    if (cell) *cell = c;
}

void orc_guide_emit_batch(const orc_render* r, int step, int first, int n, float* pos, float* dir, float* pdf,
                          int32_t* cell) {
    for (int j = 0; j < n; ++j)
        orc_guide_emit(r, step, first + j, pos + 3 * (size_t)j, dir + 3 * (size_t)j, pdf + j, cell ? cell + j : NULL);
}

/* createPhotonMap, one path (SPPM.cpp:518-672) */
static int photon_path(const orc_render* r, ophoton* slot, int i, int step) {
    const orc_scene* s = r->s;
    const int N = r->p.photons_per_pass;
    orc_mt m;
    orc_mt_seed(&m, (uint32_t)i + 135u + (uint32_t)step * (uint32_t)N);
    v3 rad, cur_org, cur_dir;
    float PDF;
    if (r->guide) {
        /* guided emission: the mixture draw and its density */
        v3 lpos;
        int cell;
        guided_start(s, r->guide, &m, &lpos, &cur_dir, &PDF, &rad, &cell);
        cur_org = vadd(lpos, vscale(r->guide->n, RAY_OFFSET));
    } else {
        /* generateLightRay (SPPM.cpp:453-471) */
        int li = table_sample(s->lprob, s->lleft, s->lright, s->nlights, &m);
        const olight* L = &s->lights[li];
        v3 lpos;
        float pdf_area;
        light_point(s, L, &m, &lpos, &rad, &pdf_area);
        pdf_area *= light_pick_pdf(s, L);
        v3 lnorm = light_normal(s, L);
        float pdf_dir = 1.0f;
        v3 outdir = cos_sampling(&m, lnorm, &pdf_dir);
        rad = vscale(rad, fminf_(1.0f, fmaxf_(0.0f, vdot(lnorm, outdir))));
        /* Ray(lightPos + n*RAY_OFFSET, outDir): ctor normalises */
        cur_org = vadd(lpos, vscale(lnorm, RAY_OFFSET));
        cur_dir = vnormalized(outdir);
        PDF = pdf_area * pdf_dir;
    }
    float cur_index = 1.0f;
    v3 next_org = cur_org, next_dir = cur_dir;
    float next_index = 1.0f;
    v3 T = V3(1, 1, 1);
    int depth = 0;
    /* emission photon (SPPM.cpp:530-540) */
    {
        ophoton ph;
        ph.flux = vscale(vmul(rad, T), PIf);
        ph.pdf = PDF;
        ph.indir = cur_dir;
        ph.pos = cur_org;
        ph.valid = 1;
        ph.index = cur_index;
        slot[0] = ph;
    }
    int pidx = 0, loopcount = 0;
    for (;;) {
        if (loopcount++ > 32 || pidx >= 32 - 1) break;
        if ((double)T.x < 1e-8 || (double)T.y < 1e-8 || (double)T.z < 1e-8) break;
        float t, hu, hv;
        int f = orc_closest_hit(s, 1, (const float*)&cur_org, (const float*)&cur_dir, RAY_OFFSET,
                                FLT_MAX, &t, &hu, &hv);
        if (f < 0) break;  /* escape: IBL colour only (none), TraceNext = false */
        geom g;
        calc_geometry(s, f, hu, hv, &g);
        v3 Nf = vdot(g.N, cur_dir) > 0 ? vneg(g.N) : g.N;
        next_org = g.pos;
        if ((double)vnorm2(g.radiance) > 0.00001) break;
        if (g.transl > 0.5f) {
            T = vscale(T, g.transl);
            cur_org = vadd(g.pos, vscale(cur_dir, RAY_OFFSET));
            continue;
        }
        /* russianRulette_PhotonMapping (SPPM.h:33-52) */
        float pdf_rr = 1.0f;
        if (depth > 16) break;
        {
            float a0 = g.albedo.x, a1 = g.albedo.y, a2 = g.albedo.z;
            pdf_rr = 0.98f * (a0 > a1 ? (a0 > a2 ? a0 : a2) : (a1 > a2 ? a1 : a2));
            if (orc_mt_real1(&m) > (double)pdf_rr) break;
        }
        PDF *= pdf_rr;
        float Ii, Io = 1.0f;
        if (near1(cur_index)) { Ii = 1.0f; Io = g.index; }
        else { Ii = cur_index; Io = 1.0f; }
        fparam fp = {cur_dir, next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
        if (g.is_sppm) {
            ophoton ph;
            ph.flux = vmul(vmul(rad, T), g.albedo);
            ph.pdf = PDF;
            ph.indir = cur_dir;
            ph.pos = g.pos;
            ph.valid = 1;
            ph.index = cur_index;
            slot[pidx++] = ph;
        }
        float pdf_s = 1.0f;
        next_dir = surf_sample(g.kind, &fp, &m, &pdf_s);
        if (!rutil_valid_dir(next_dir)) break;
        if (g.opaque) {
            if ((double)vdot(next_dir, Nf) < 1e-6) break;
        } else {
            int refl = rutil_is_reflect(cur_dir, next_dir, Nf);
            next_index = refl ? cur_index : Io;
        }
        fp.out = next_dir;
        pdf_s = surf_pdf(g.kind, &fp);
        v3 outN = vdot(Nf, next_dir) > 0 ? Nf : vneg(Nf);
        next_org = vadd(next_org, vscale(outN, RAY_OFFSET));
        float brdf = surf_brdf(g.kind, &fp);
        float cs = fmaxf_(0.0f, fminf_(1.0f, vdot(next_dir, outN)));
        if ((double)brdf < 1e-8) break;
        if (pdf_s < 0) break;
        T = vmul(T, vscale(vscale(g.albedo, cs), brdf));
        PDF *= pdf_s;
        cur_org = next_org; cur_dir = next_dir; cur_index = next_index;
        depth++;
    }
    ophoton term;
    memset(&term, 0, sizeof term);
    term.valid = 0; term.pdf = 1.0f; term.index = 1.0f;
    const int k = pidx;
    slot[++pidx] = term;
    return k;
}

static void ensure_slots(orc_render* r) {
    if (r->slots) return;
    size_t n = (size_t)r->nslots_paths * NUM_SLOTS;
    r->slots = (ophoton*)calloc(n, sizeof(ophoton));
    for (size_t i = 0; i < n; ++i) {  /* Photon() : PDF 1, Index 1 */
        r->slots[i].pdf = 1.0f;
        r->slots[i].index = 1.0f;
    }
}

int orc_photon_pass(orc_render* r, int step) {
    int N = r->p.photons_per_pass;
    if (r->s->nlights == 0) return 0;
    ensure_slots(r);
#pragma omp parallel for schedule(dynamic, 256) num_threads(r->nthreads)
    for (int i = 0; i < N; ++i) photon_path(r, r->slots + (size_t)i * NUM_SLOTS, i, step);
    return N;
}

int orc_flatten(orc_render* r) {
    int N = r->p.photons_per_pass;
    ensure_slots(r);
    size_t cnt = 0;
    for (int i = 0; i < N; ++i) {
        const ophoton* sl = r->slots + (size_t)i * NUM_SLOTS;
        for (int k = 0; k < NUM_SLOTS && sl[k].valid; ++k) cnt++;
    }
    free(r->flat);
    r->flat = (ophoton*)malloc(sizeof(ophoton) * (cnt > 0 ? cnt : 1));
    size_t o = 0;
    for (int i = 0; i < N; ++i) {
        const ophoton* sl = r->slots + (size_t)i * NUM_SLOTS;
        for (int k = 0; k < NUM_SLOTS && sl[k].valid; ++k) r->flat[o++] = sl[k];
    }
    r->nflat = (int)cnt;
    return (int)cnt;
}
void orc_read_photons(const orc_render* r, float* pos, float* flux, float* indir, float* index,
                      float* pdf) {
    for (int i = 0; i < r->nflat; ++i) {
        vstore(pos + 3 * i, r->flat[i].pos);
        vstore(flux + 3 * i, r->flat[i].flux);
        vstore(indir + 3 * i, r->flat[i].indir);
        index[i] = r->flat[i].index;
        pdf[i] = r->flat[i].pdf;
    }
}

/* getHitpointInfo (SPPM.cpp:98-306) with no IBL: after the hit point is made
 * the rest of the path can only add IBL light (zero), so it stops there. */
/* IBL::GetUVFromDirection, SPHERE mapping (IBL.h:90-103): unqualified acos
 * on a float binds to the double function; u, v are narrowed to float.
 * IBL::GetIBLColorFromDirection (IBL.h:122-129) fetches bilinearly. */
static void ibl_uv(v3 dir, float* u_, float* v_) {
    float u = 0.0f, v = 0.0f;
    if (!((double)fabsf(dir.x) < 1e-6 && (double)fabsf(dir.z) < 1e-6)) {
        const v3 xz = vnormalized(V3(dir.x, 0.0f, dir.z));
        const double a = acos((double)xz.x);
        u = (float)((xz.z >= 0 ? a : (double)(2.0f * PIf) - a) / (double)(2.0f * PIf));
        v = (float)(acos((double)dir.y) / (double)PIf);
    }
    *u_ = u;
    *v_ = v;
}
static v3 ibl_color(const orc_scene* s, v3 dir) {
    float u, v;
    ibl_uv(dir, &u, &v);
    float c[4];
    orc_bilinear(&s->ibl, u, v, c);
    return V3(c[0], c[1], c[2]);
}

void orc_ibl_uv(const float* dir, float* uv) { ibl_uv(vload(dir), uv, uv + 1); }
void orc_ibl_color(const orc_scene* s, const float* dir, float* out3) {
    const v3 c = s->has_ibl ? ibl_color(s, vload(dir)) : V3(0, 0, 0);
    vstore(out3, c);
}

static void eye_path(orc_render* r, int pix, int step, v3 smp, ohit* hp) {
    const orc_scene* s = r->s;
    const int W = r->W, H = r->H;
    orc_mt m;
    orc_mt_seed(&m, (uint32_t)pix + (uint32_t)W * (uint32_t)H * (uint32_t)step);
    int w = pix % W, h = pix / W;
    float u = (float)w / (float)W, v = (float)h / (float)H;
    v3 cur_org, cur_dir;
    float campdf = 1.0f;
    cam_ray(&r->cam, u, v, smp, &cur_org, &cur_dir, &campdf);
    float cur_index = 1.0f;
    v3 next_org = cur_org, next_dir = cur_dir;
    float next_index = 1.0f;
    v3 T = V3(1, 1, 1);
    float PDF = 1.0f;
    int depth = 0;
    memset(hp, 0, sizeof *hp);
    hp->type = 0;
    hp->facet = -1;
    hp->weight = V3(1, 1, 1);
    hp->outdir = V3(0, 1, 0);
    v3 direct = V3(0, 0, 0);
    /* getHitpointInfo's compiled-out next-event estimation (SPPM.cpp:144-168,
     * 187-229), enabled by eye_nee; keep = SPPM_NEE_KEEP.  eye_nee = 3 exists
     * only here, as the checker of the MIS estimator: SPPM_NEE_KEEP without
     * the light samples, emitter hits at full weight (plain BSDF sampling). */
    const int nee = r->p.eye_nee != SPPM_NEE_OFF;
    const int keep = r->p.eye_nee >= SPPM_NEE_KEEP;
    const int light_samples = r->p.eye_nee != 3;
    v3 prev_pos = V3(0, 0, 0);  /* TraceParam::PrevGmt.Position, .param.IfDeltaSurface, pdf_prevDir */
    int prev_delta = 0;
    float pdf_prev = 0.0f;
    for (;;) {
        if (depth > 6) break;
        else if (depth > 4) {
            float rus = (float)exp((double)(4 - depth));
            if (orc_mt_real1(&m) > (double)rus) break;
            PDF *= rus;
        }
        float t, hu, hv;
        int f = orc_closest_hit(s, 1, (const float*)&cur_org, (const float*)&cur_dir, RAY_OFFSET,
                                FLT_MAX, &t, &hu, &hv);
        if (f < 0) {
            /* escape: the environment's colour (SPPM.cpp:290-300) */
            if (s->has_ibl) direct = vadd(direct, vdivs(vmul(ibl_color(s, cur_dir), T), PDF));
            if (hp->type == 0) hp->type = 2;
            break;
        }
        geom g;
        calc_geometry(s, f, hu, hv, &g);
        v3 Nf = vdot(g.N, cur_dir) > 0 ? vneg(g.N) : g.N;
        next_org = g.pos;
        if (g.transl > 0.5f) {
            T = vscale(T, g.transl);
            cur_org = vadd(g.pos, vscale(cur_dir, RAY_OFFSET));
            continue;
        }
        if (nee && (double)vnorm2(g.radiance) > 1e-6) {  /* SPPM.cpp:145-167 */
            if (hp->type != 1) {
                float mis = 1.0f;
                if (depth > 0 && !prev_delta && light_samples) {
                    int li = s->facet_light[f];
                    float pdf_nee = 0.0f;  /* a default AreaLight has |Flux| = 0 */
                    if (li >= 0) {
                        const olight* L = &s->lights[li];
                        /* SampleAreaLight_pdf (LightSampler.cpp:84-96) */
                        float ratio = 1.0f;
                        if (L->textured)
                            ratio = (vnorm(g.radiance) / L->avg / (float)LIGHT_TABLE_SAMPLES) * (float)LIGHT_TABLE_SAMPLES;
                        float tri = ratio / facet_area(s->pos + 9 * f);
                        float cosL = fmaxf_(0.0001f, fminf_(1.0f, vdot(Nf, vneg(cur_dir))));
                        pdf_nee = (light_pick_pdf(s, L) * tri) * vnorm2(vsub(prev_pos, g.pos)) / cosL;
                    }
                    float pt = pdf_prev;
                    mis = (pt * pt) / ((pdf_nee * pdf_nee) + (pt * pt));
                }
                direct = vadd(direct, vdivs(vmul(vmul(vscale(g.albedo, mis), g.radiance), T), PDF));
                if (keep && hp->type == 0) hp->type = 2;
            }
            break;
        }
        if (g.is_sppm && hp->type != 1) {  /* makeHitPoint (SPPM.cpp:79-96, 171-174) */
            if (!(nee && keep)) direct = V3(0, 0, 0);  /* the new HitPoint's DirectLight */
            hp->type = 1;
            hp->facet = f;
            hp->pos = g.pos;
            hp->outdir = vneg(cur_dir);
            hp->weight = vdivs(T, PDF);
            hp->normal = Nf;
            hp->kind = g.kind;
            hp->index = g.index;
            hp->rough = g.rough;
            hp->absorb = g.absorb;
            hp->refl = g.refl;
            /* the path continues after its hit point only to reach the
             * environment; without one nothing later changes the hit point */
            if (!s->has_ibl) break;
        }
        float Ii, Io = 1.0f;
        if (near1(cur_index)) { Ii = 1.0f; Io = g.index; }
        else { Ii = cur_index; Io = 1.0f; }
        fparam fp = {cur_dir, next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
        const int delta = s->attrs[s->attr[f]].if_delta != 0;
        if (nee && light_samples && !delta && hp->type != 1) {  /* SPPM.cpp:190-226 */
            int li = table_sample(s->lprob, s->lleft, s->lright, s->nlights, &m);
            const olight* L = &s->lights[li];
            float pdf_sel = light_pick_pdf(s, L);
            v3 lpos, rad;
            float pdf_area;
            light_point(s, L, &m, &lpos, &rad, &pdf_area);
            v3 gpos = g.pos;
            v3 lnorm = light_normal(s, L);
            v3 ldir = vnormalized(vnormalized(vsub(lpos, gpos)));  /* Ray's ctor normalises again */
            float dist = vnorm(vsub(lpos, gpos));
            float maxd = (float)((double)dist - (double)RAY_OFFSET * 2.0);
            if (occluded_rec(s, gpos, ldir, RAY_OFFSET, maxd, 0.5f, 0) == 0) {
                float cosL = fminf_(1.0f, fmaxf_(0.00001f, vdot(lnorm, vneg(ldir))));
                float cosE = fabsf(vdot(Nf, ldir));
                fp.out = ldir;
                float brdf = fmaxf_(0.0f, surf_brdf(g.kind, &fp));
                /* Radiance * cosL * pow(d, -2) / (pdf_area * pdf_sel) * cosE * Albedo * BRDF */
                v3 dlr = vscale(vmul(vscale(vdivs(vscale(vscale(rad, cosL), powf(dist, -2.0f)), pdf_area * pdf_sel), cosE),
                                     g.albedo), brdf);
                float pn = (pdf_area * pdf_sel) * (dist * dist) / cosL;
                float pt = surf_pdf(g.kind, &fp);
                float w = (pn * pn) / ((pn * pn) + (pt * pt));
                direct = vadd(direct, vdivs(vmul(vscale(dlr, w), T), PDF));
            }
        }
        float pdf_s = 1.0f;
        next_dir = surf_sample(g.kind, &fp, &m, &pdf_s);
        if (!rutil_valid_dir(next_dir)) break;
        if (g.opaque) {
            if ((double)vdot(next_dir, Nf) < 1e-6) break;
        } else {
            int refl = !(vdot(next_dir, Nf) < 0.0f);
            next_index = refl ? cur_index : Io;
        }
        fp.out = next_dir;
        pdf_s = surf_pdf(g.kind, &fp);
        v3 outN = vdot(Nf, next_dir) > 0 ? Nf : vneg(Nf);
        next_org = vadd(next_org, vscale(outN, RAY_OFFSET));
        float brdf = surf_brdf(g.kind, &fp);
        float cs = fmaxf_(0.0f, fminf_(1.0f, vdot(next_dir, outN)));
        if ((double)brdf < 1e-8) break;
        if (pdf_s < 0) break;
        T = vmul(T, vscale(vscale(g.albedo, cs), brdf));
        PDF *= pdf_s;
        pdf_prev = pdf_s;
        prev_pos = g.pos;
        prev_delta = delta;
        cur_org = next_org; cur_dir = next_dir; cur_index = next_index;
        depth++;
    }
    /* createHitpoints: PixelWeight /= pdf, DirectLight /= pdf (SPPM.cpp:346-347) */
    hp->weight = vdivs(hp->weight, campdf);
    hp->direct = vdivs(direct, campdf);
}

static v3 camera_sample(orc_render* r, int step) {
    const sppm_camera_desc* c = &r->camdesc;
    if (c->sampler == SPPM_SAMPLER_UNIFORM) return uniform_sampler_draw(&r->camstream);
    if (c->sampler == SPPM_SAMPLER_HALTON)
        return V3(orc_halton(2, step), orc_halton(3, step), orc_halton(5, step));
    float h[3];
    orc_hammersley(c->sampler_nsample, step, h);
    return V3(h[0], h[1], h[2]);
}

void orc_eye_pass(orc_render* r, int step) {
    int np = r->W * r->H;
    v3* smp = (v3*)malloc(sizeof(v3) * (size_t)np);
    /* the shared sampler is consumed in pixel order (OMP_NUM_THREADS=1 semantics) */
    for (int i = 0; i < np; ++i) smp[i] = camera_sample(r, step);
#pragma omp parallel for schedule(dynamic, 64) num_threads(r->nthreads)
    for (int i = 0; i < np; ++i) eye_path(r, i, step, smp[i], &r->hits[i]);
    free(smp);
}
void orc_read_hitpoints(const orc_render* r, int32_t* type, int32_t* facet, float* pos, float* weight,
                        float* outdir, float* normal) {
    int np = r->W * r->H;
    for (int i = 0; i < np; ++i) {
        const ohit* h = &r->hits[i];
        type[i] = h->type;
        facet[i] = h->facet;
        vstore(pos + 3 * i, h->pos);
        vstore(weight + 3 * i, h->weight);
        vstore(outdir + 3 * i, h->outdir);
        vstore(normal + 3 * i, h->normal);
    }
}

/* ---- gather: applyHitPointInfo (SPPM.cpp:395-451) ------------------------
 * Candidate photons come from a uniform grid (cell >= max radius, one
 * extra ring of cells); the neighbour set is defined only by the
 * reference's two tests (closed box, strict radius), so it equals the
 * kd-tree's set. */
typedef struct { int64_t* cell_start; int32_t* order; int nx, ny, nz; float cell; float bmin[3]; } ogrid;

static void grid_build(const orc_render* r, float cell, ogrid* g) {
    float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = 0; i < r->nflat; ++i)
        for (int k = 0; k < 3; ++k) {
            float x = vget(r->flat[i].pos, k);
            if (x < bmin[k]) bmin[k] = x;
            if (x > bmax[k]) bmax[k] = x;
        }
    int n[3];
    for (int k = 0; k < 3; ++k) {
        double ext = (double)bmax[k] - (double)bmin[k];
        double c = ext / cell;
        n[k] = (int)(c < 1024 ? c : 1024) + 1;
        g->bmin[k] = bmin[k];
    }
    /* bound total cells */
    while ((int64_t)n[0] * n[1] * n[2] > (1 << 24)) {
        cell *= 2.0f;
        for (int k = 0; k < 3; ++k) n[k] = (int)(((double)bmax[k] - (double)bmin[k]) / cell) + 1;
    }
    g->cell = cell;
    g->nx = n[0]; g->ny = n[1]; g->nz = n[2];
    int64_t nc = (int64_t)n[0] * n[1] * n[2];
    g->cell_start = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    int32_t* key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(r->nflat > 0 ? r->nflat : 1));
    for (int i = 0; i < r->nflat; ++i) {
        int c[3];
        for (int k = 0; k < 3; ++k) {
            int ci = (int)(((double)vget(r->flat[i].pos, k) - (double)g->bmin[k]) / (double)cell);
            if (ci < 0) ci = 0;
            if (ci >= n[k]) ci = n[k] - 1;
            c[k] = ci;
        }
        key[i] = (int32_t)(((int64_t)c[2] * n[1] + c[1]) * n[0] + c[0]);
        g->cell_start[key[i] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) g->cell_start[c + 1] += g->cell_start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc);
    memcpy(fill, g->cell_start, sizeof(int64_t) * (size_t)nc);
    g->order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(r->nflat > 0 ? r->nflat : 1));
    for (int i = 0; i < r->nflat; ++i) g->order[fill[key[i]]++] = i;  /* ascending within cell */
    free(fill);
    free(key);
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

void orc_gather(orc_render* r, int record) {
    int np = r->W * r->H;
    const float alpha = r->p.alpha;
    float rmax = 0.0f;
    for (int i = 0; i < np; ++i)
        if (r->shared[i].radius > rmax) rmax = r->shared[i].radius;
    ogrid g;
    memset(&g, 0, sizeof g);
    int have = r->nflat > 0;
    if (have) grid_build(r, rmax > 0 ? rmax : 1e-3f, &g);
    /* pass 1: neighbour lists (indices ascending) */
    int32_t** lists = (int32_t**)calloc((size_t)np, sizeof(int32_t*));
#pragma omp parallel for schedule(dynamic, 64) num_threads(r->nthreads)
    for (int v = 0; v < np; ++v) {
        const ohit* hp = &r->hits[v];
        r->counts[v] = 0;
        if (hp->type != 1 || !have) continue;
        const oshared* s = &r->shared[v];
        v3 c = hp->pos;
        float R = s->radius;
        v3 bbmin = vsub(c, vscale(V3(1, 1, 1), R));
        v3 bbmax = vadd(c, vscale(V3(1, 1, 1), R));
        float R2 = R * R;
        int lo[3], hi[3];
        int nn[3] = {g.nx, g.ny, g.nz};
        for (int k = 0; k < 3; ++k) {
            lo[k] = (int)floor(((double)vget(bbmin, k) - (double)g.bmin[k]) / (double)g.cell) - 1;
            hi[k] = (int)floor(((double)vget(bbmax, k) - (double)g.bmin[k]) / (double)g.cell) + 1;
            if (lo[k] < 0) lo[k] = 0;
            if (hi[k] > nn[k] - 1) hi[k] = nn[k] - 1;
        }
        int cap = 64, cnt = 0;
        int32_t* L = (int32_t*)malloc(sizeof(int32_t) * cap);
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    int64_t cell = ((int64_t)z * g.ny + y) * g.nx + x;
                    for (int64_t q = g.cell_start[cell]; q < g.cell_start[cell + 1]; ++q) {
                        int pi = g.order[q];
                        v3 pp = r->flat[pi].pos;
                        /* isIncludePoint (kdtree.h:78-88): closed box */
                        if (bbmin.x > pp.x || bbmax.x < pp.x) continue;
                        if (bbmin.y > pp.y || bbmax.y < pp.y) continue;
                        if (bbmin.z > pp.z || bbmax.z < pp.z) continue;
                        if (!(vnorm2(vsub(pp, c)) < R2)) continue;
                        if (cnt == cap) { cap *= 2; L = (int32_t*)realloc(L, sizeof(int32_t) * cap); }
                        L[cnt++] = pi;
                    }
                }
        qsort(L, (size_t)cnt, sizeof(int32_t), cmp_i32);
        lists[v] = L;
        r->counts[v] = cnt;
    }
    /* pass 2: accumulate in ascending photon order, progressive update */
#pragma omp parallel for schedule(dynamic, 64) num_threads(r->nthreads)
    for (int v = 0; v < np; ++v) {
        const ohit* hp = &r->hits[v];
        if (hp->type == 0) continue;
        oshared* s = &r->shared[v];
        /* DirectLight is always zero without an IBL: the 1e-8 test never passes */
        if (hp->type == 2) continue;
        v3 add = V3(0, 0, 0);
        int M = r->counts[v];
        for (int q = 0; q < M; ++q) {
            const ophoton* p = &r->flat[lists[v][q]];
            float Ii = p->index;
            float Io = near1(p->index) ? hp->index : 1.0f;
            v3 nrm = vdot(p->indir, hp->normal) < 0 ? hp->normal : vneg(hp->normal);
            fparam fp = {p->indir, hp->outdir, nrm, Ii, Io, hp->rough, hp->absorb, hp->refl};
            float b = surf_brdf(hp->kind, &fp);
            add = vadd(add, vdivs(vmul(vscale(hp->weight, b), p->flux), p->pdf));
        }
        if (M > 0) {
            float num = (float)s->num_photon + alpha * (float)M;
            float den = (float)(s->num_photon + M);
            s->radiance = vdivs(vscale(vadd(s->radiance, add), num), den);
            s->radius *= sqrtf(num / den);
            s->num_photon = (int)((float)s->num_photon + (float)M * alpha);
            s->num_gather++;
        }
    }
    if (record) {
        free(r->nb_off);
        free(r->nb_idx);
        r->nb_off = (int64_t*)malloc(sizeof(int64_t) * ((size_t)np + 1));
        r->nb_off[0] = 0;
        for (int v = 0; v < np; ++v) r->nb_off[v + 1] = r->nb_off[v] + r->counts[v];
        r->nb_total = r->nb_off[np];
        r->nb_idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(r->nb_total > 0 ? r->nb_total : 1));
        for (int v = 0; v < np; ++v)
            if (r->counts[v]) memcpy(r->nb_idx + r->nb_off[v], lists[v], sizeof(int32_t) * r->counts[v]);
    }
    for (int v = 0; v < np; ++v) free(lists[v]);
    free(lists);
    free(g.cell_start);
    free(g.order);
    /* DirectLight of every hit point that is not NONE (SPPM.cpp:404-418):
     * norm2() > 1e-8 compares the float norm against a double */
    for (int v = 0; v < np; ++v) {
        const ohit* hp = &r->hits[v];
        if (hp->type == 0) continue;
        if ((double)vnorm2(hp->direct) > 1e-8) {
            r->shared[v].direct = vadd(r->shared[v].direct, hp->direct);
            r->shared[v].num_direct++;
        }
    }
}

void orc_read_direct(const orc_render* r, float* hit_direct, float* shared_direct, int32_t* num_direct) {
    size_t np = (size_t)r->W * (size_t)r->H;
    for (size_t i = 0; i < np; ++i) {
        if (hit_direct) vstore(hit_direct + 3 * i, r->hits[i].direct);
        if (shared_direct) vstore(shared_direct + 3 * i, r->shared[i].direct);
        if (num_direct) num_direct[i] = r->shared[i].num_direct;
    }
}
void orc_read_counts(const orc_render* r, int32_t* m) {
    memcpy(m, r->counts, sizeof(int32_t) * (size_t)r->W * (size_t)r->H);
}
int64_t orc_neighbour_total(const orc_render* r) { return r->nb_total; }
void orc_read_neighbours(const orc_render* r, int64_t* offsets, int32_t* idx) {
    size_t np = (size_t)r->W * (size_t)r->H;
    memcpy(offsets, r->nb_off, sizeof(int64_t) * (np + 1));
    if (r->nb_total) memcpy(idx, r->nb_idx, sizeof(int32_t) * (size_t)r->nb_total);
}
void orc_read_shared(const orc_render* r, float* radiance, float* radius, int32_t* num_photon,
                     int32_t* num_gather) {
    size_t np = (size_t)r->W * (size_t)r->H;
    for (size_t i = 0; i < np; ++i) {
        vstore(radiance + 3 * i, r->shared[i].radiance);
        radius[i] = r->shared[i].radius;
        num_photon[i] = r->shared[i].num_photon;
        num_gather[i] = r->shared[i].num_gather;
    }
}
void orc_write_shared(orc_render* r, const float* radiance, const float* radius,
                      const int32_t* num_photon, const int32_t* num_gather) {
    size_t np = (size_t)r->W * (size_t)r->H;
    for (size_t i = 0; i < np; ++i) {
        r->shared[i].radiance = vload(radiance + 3 * i);
        r->shared[i].radius = radius[i];
        r->shared[i].num_photon = num_photon[i];
        r->shared[i].num_gather = num_gather[i];
    }
}

/* reconstructImageFromSharedPoints (SPPM.cpp:473-493) */
void orc_reconstruct(const orc_render* r, int32_t total_photons, int32_t passes, float* rgba) {
    size_t np = (size_t)r->W * (size_t)r->H;
    for (size_t i = 0; i < np; ++i) {
        const oshared* s = &r->shared[i];
        v3 dl = vdivs(s->direct, (float)((float)s->num_direct + 0.0001f));
        float den = ((float)total_photons * ((float)s->num_gather + 0.0001f) / (0.0001f + (float)passes)) *
                    PIf * s->radius * s->radius;
        v3 L = vadd(dl, vdivs(s->radiance, den));
        const float b = 1.2f;
        rgba[4 * i + 0] = powf(L.x, 1.0f / b);
        rgba[4 * i + 1] = powf(L.y, 1.0f / b);
        rgba[4 * i + 2] = powf(L.z, 1.0f / b);
        rgba[4 * i + 3] = 1.0f;
    }
}

/* Loop body of RenderImage_FixRenderTime (SPPM.cpp:771-841): photon pass
 * with Step = numstep+1, then the eye pass with step = (incremented)
 * numstep, i.e. the same value. */
int orc_full_pass(orc_render* r, int pass, int record) {
    int N = r->p.photons_per_pass;
    ensure_slots(r);
    for (int i = 0; i < N; ++i) {  /* photons[i][0] = Photon() */
        ophoton* s0 = r->slots + (size_t)i * NUM_SLOTS;
        memset(s0, 0, sizeof *s0);
        s0->pdf = 1.0f; s0->index = 1.0f;
    }
    int emitted = orc_photon_pass(r, pass + 1);
    orc_flatten(r);
    orc_eye_pass(r, pass + 1);
    orc_gather(r, record);
    return emitted;
}


/* ======================================================================
 * Spot checks
 * ====================================================================== */
void orc_uniform_sampler_at(int id, int n, const int64_t* calls, float* out) {
    orc_mt m;
    orc_mt_seed(&m, (uint32_t)id);
    int64_t pos = 0; /* index of the next call */
    for (int j = 0; j < n; ++j) {
        int64_t skip_words = (calls[j] - pos) * 6;
        /* discard whole twist blocks without tempering */
        while (skip_words > 0) {
            if (m.idx >= 624) mt_twist(&m);
            int64_t avail = 624 - m.idx;
            if (skip_words >= avail) {
                skip_words -= avail;
                m.idx = 624;
            } else {
                m.idx += (int)skip_words;
                skip_words = 0;
            }
        }
        vstore(out + 3 * j, uniform_sampler_draw(&m));
        pos = calls[j] + 1;
    }
}

void orc_eye_pixel(const orc_render* r, int step, int pixel, const float* sample3, int32_t* type,
                   int32_t* facet, float* pos, float* weight, float* normal, float* outdir) {
    ohit h;
    eye_path((orc_render*)r, pixel, step, vload(sample3), &h);
    *type = h.type;
    *facet = h.facet;
    vstore(pos, h.pos);
    vstore(weight, h.weight);
    vstore(normal, h.normal);
    vstore(outdir, h.outdir);
}

int orc_photon_path(const orc_render* r, int step, int path, float* pos, float* flux, float* indir,
                    float* index, float* pdf, int32_t* nlist) {
    /* run the path on a private slot row that starts invalid */
    ophoton* row = (ophoton*)calloc(NUM_SLOTS, sizeof(ophoton));
    int k = photon_path(r, row, path, step);
    const ophoton* sl = row;
    int n = 0;
    while (n < NUM_SLOTS && sl[n].valid) ++n;
    if (nlist) *nlist = n;
    for (int k = 0; k < n; ++k) {
        vstore(pos + 3 * k, sl[k].pos);
        vstore(flux + 3 * k, sl[k].flux);
        vstore(indir + 3 * k, sl[k].indir);
        index[k] = sl[k].index;
        pdf[k] = sl[k].pdf;
    }
    free(row);
    return k;
}

void orc_read_path_counts(const orc_render* r, int32_t* counts) {
    int N = r->p.photons_per_pass;
    for (int i = 0; i < N; ++i) {
        const ophoton* sl = r->slots + (size_t)i * NUM_SLOTS;
        int k = 0;
        while (k < NUM_SLOTS && sl[k].valid) ++k;
        counts[i] = k;
    }
}
