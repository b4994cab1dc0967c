// sppm_host_math.h -- host-side scene setup arithmetic (camera basis, light
// registration, alias tables, Hammersley points, bilinear fetch).  Same float
// operation order as the reference so the device sees identical tables.
// Compiled with -ffp-contract=off.
#pragma once
#include <math.h>
#include <stdint.h>

#include <cstdlib>
#include <vector>

namespace sppm_host {

constexpr float PI = 3.14159f;  // PTUtility::PI, Vec.h:10

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return v3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float norm2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float norm(V3 a) { return (float)sqrt((double)norm2(a)); }
inline V3 normalized(V3 a) {
    float d = norm(a);
    return v3(a.x / d, a.y / d, a.z / d);
}
inline V3 ld(const float* p) { return v3(p[0], p[1], p[2]); }
inline float smax(float a, float b) { return (a < b) ? b : a; }

// HaltonCore<B>::Sample (QRandSequence.h:7-23)
inline float halton(int B, int n) {
    float result = 0.0f;
    if (n <= 0) return result;
    int keta = (int)(log2((double)n) / log2((double)B) + 1);
    for (int i = 0; i < keta; ++i) {
        result = (float)((double)result + (double)(n % B) * pow((double)B, (double)(-(i + 1))));
        n /= B;
    }
    return result;
}

// TableSampler<T>::Init (TableSampler.h:44-120): alias table, reference deque order
inline void build_alias(const std::vector<float>& values, std::vector<float>& prob, std::vector<int>& left,
                        std::vector<int>& right, float& avg) {
    const int n = (int)values.size();
    prob.assign(n, 1.0f);
    left.assign(n, -1);
    right.assign(n, -1);
    avg = 0.0f;
    if (n <= 0) return;
    for (int i = 0; i < n; ++i) avg += smax(0.0f, values[i]);
    avg /= (float)n;
    struct P { float v; int id; };
    std::vector<P> bigger, smaller;
    bigger.reserve(2 * n);
    smaller.reserve(2 * n);
    for (int i = 0; i < n; ++i) {
        float value = smax(0.0f, values[i]);
        P p{value / avg, i};
        if (value >= avg) bigger.push_back(p);
        else smaller.push_back(p);
    }
    size_t IS = 0, IB = 0;
    for (int i = 0; i < n; ++i) {
        if (IS < smaller.size()) {
            left[i] = smaller[IS].id;
            prob[i] = smaller[IS].v;
            IS++;
            if (IB < bigger.size()) {
                right[i] = bigger[IB].id;
                bigger[IB].v -= (1.0f - prob[i]);
                if (bigger[IB].v < 1.0f) smaller.push_back(bigger[IB++]);
            } else {
                if (smaller.size() == IS) {
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
            if (bigger[IB].v < 1.0f) smaller.push_back(bigger[IB++]);
        }
    }
}

// LinearSampler::Sample (TextureSampler.h:60-79) on RGBA32F
inline void bilinear(const float* rgba, int W, int H, float u, float v, float* out) {
    int ix = (int)((float)W * u), iy = (int)((float)H * v);
    int x1 = std::abs(ix) % W, y1 = std::abs(iy) % H;
    int x2 = (W - 1) < (x1 + 1) ? (W - 1) : (x1 + 1);
    int y2 = (H - 1) < (y1 + 1) ? (H - 1) : (y1 + 1);
    const float* t00 = rgba + 4 * (size_t)(x1 + y1 * W);
    const float* t10 = rgba + 4 * (size_t)(x2 + y1 * W);
    const float* t01 = rgba + 4 * (size_t)(x1 + y2 * W);
    const float* t11 = rgba + 4 * (size_t)(x2 + y2 * W);
    float w = fabsf(u * (float)W) - (float)std::abs(ix);
    float h = fabsf(v * (float)H) - (float)std::abs(iy);
    float omw = (float)(1.0 - (double)w), omh = (float)(1.0 - (double)h);
    for (int c = 0; c < 4; ++c) {
        float up = omw * t00[c] + w * t10[c];
        float dn = omw * t01[c] + w * t11[c];
        out[c] = omh * up + h * dn;
    }
}

}  // namespace sppm_host
