// ref_harness.cpp -- drives the parts of the FriendsRender R reference that
// compile as shipped, so the oracle can be checked against the real code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile.ref from the reference
// sources where they lie (/root/reference/SPPM/SPPM, read-only); output goes
// to oracle/_ref/libsppm_ref.so.  Nothing of the reference is copied here:
// this file only includes its headers and links its .cpp files.
//
// The full reference does not compile with g++ 11 or clang 22: Primitive.h
// brace-initialises mSIMD::Double4/Float8 from a braced list (MSVC-only), and
// every TU that includes Primitive.h fails.  The pieces below do not include
// it: MT.h, CameraSampler.h, QRandSequence.h, TableSampler.h, TextureSampler.h,
// TextureUtility.cpp (+stb), Matrix.h/.cpp, Vec.h/.cpp.
//
// TableSampler keeps its alias table private; `private` is redefined to
// `public` for that header only, to read the table without changing it.
#include <cstdint>
#include <cstring>
#include <vector>

#include "Vec.h"
#include "MT.h"
#include "QRandSequence.h"
#include "CameraSampler.h"
#define private public
#include "TableSampler.h"
#undef private
#include "Texture.h"
#include "TextureUtility.h"
#include "TextureSampler.h"
#include "Matrix.h"

using PTUtility::Vec2;
using PTUtility::Vec3;
using PTUtility::Vec4;

extern "C" {

// RandomMT(seed).genrand64_real1() x n            (MT.h:17-32)
void ref_mt_real1_seq(int seed, int n, double* out) {
    RandomMT m(seed);
    for (int i = 0; i < n; ++i) out[i] = m.genrand64_real1();
}

// UniformSample(ID).Sample(k) x ncalls            (CameraSampler.h:19-31)
void ref_uniform_sampler_seq(int id, int ncalls, float* out) {
    UniformSample s(id);
    for (int i = 0; i < ncalls; ++i) {
        Vec3 v = s.Sample(i);
        out[3 * i + 0] = v[0];
        out[3 * i + 1] = v[1];
        out[3 * i + 2] = v[2];
    }
}

// HaltonCore<B>::Sample(n) for B in {2,3,5}      (QRandSequence.h:7-23)
float ref_halton(int base, int n) {
    switch (base) {
    case 2: return QRandSequence::HaltonCore<2>::Sample(n);
    case 3: return QRandSequence::HaltonCore<3>::Sample(n);
    case 5: return QRandSequence::HaltonCore<5>::Sample(n);
    }
    return -1.0f;
}

// Hammersley(N).Sample(i)                          (QRandSequence.h:35-46)
void ref_hammersley(int nseq, int i, float* out3) {
    QRandSequence::Hammersley h(nseq);
    Vec3 v = h.Sample(i);
    out3[0] = v[0]; out3[1] = v[1]; out3[2] = v[2];
}

// TableSampler<float>::Init with f = identity       (TableSampler.h:44-120)
void ref_table_build(const float* values, int n, float* prob, int32_t* left, int32_t* right,
                     float* avg) {
    TableSampler<float> t;
    t.Init(values, n, [](const float& x) { return x; });
    for (int i = 0; i < n; ++i) {
        prob[i] = t.m_Table[i].probLeftSample;
        left[i] = t.m_Table[i].ID_Left;
        right[i] = t.m_Table[i].ID_Right;
    }
    *avg = t.m_Average;
}

// GetPDF of each element                          (TableSampler.h:122-124)
void ref_table_pdf(const float* values, int n, float* pdf) {
    TableSampler<float> t;
    t.Init(values, n, [](const float& x) { return x; });
    for (int i = 0; i < n; ++i) pdf[i] = t.GetPDF(values[i]);
}

// Sample(RandomMT(seed)) x ndraw -> element index  (TableSampler.h:126-129)
struct IdxVal {
    int idx;
    float v;
};
void ref_table_sample_seq(const float* values, int n, int seed, int ndraw, int32_t* out) {
    std::vector<IdxVal> d(n);
    for (int i = 0; i < n; ++i) d[i] = IdxVal{i, values[i]};
    TableSampler<IdxVal> t;
    t.Init(d.data(), n, [](const IdxVal& x) { return x.v; });
    RandomMT m(seed);
    for (int i = 0; i < ndraw; ++i) out[i] = t.Sample(m).idx;
}

// TextureUtility::GetColor_UV on an RGBA32F image   (TextureUtility.cpp:120-123)
void ref_bilinear(int w, int h, const float* rgba, float u, float v, float* out4) {
    Texture2D tex;
    tex.Init(w, h);
    for (int i = 0; i < w * h; ++i)
        tex.GetRawData()[i] = TColor(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]);
    TColor c = TextureUtility::GetColor_UV(tex, u, v);
    for (int k = 0; k < 4; ++k) out4[k] = c[k];
}

// Same, many lookups on one texture
void ref_bilinear_batch(int w, int h, const float* rgba, int n, const float* uv, float* out4) {
    Texture2D tex;
    tex.Init(w, h);
    for (int i = 0; i < w * h; ++i)
        tex.GetRawData()[i] = TColor(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]);
    for (int j = 0; j < n; ++j) {
        TColor c = TextureUtility::GetColor_UV(tex, uv[2 * j], uv[2 * j + 1]);
        for (int k = 0; k < 4; ++k) out4[4 * j + k] = c[k];
    }
}

// TextureUtility::GetNormalFromNormalmap          (TextureUtility.cpp:161-164)
void ref_normal_from_map(int w, int h, const float* rgba, float u, float v, float* out3) {
    Texture2D tex;
    tex.Init(w, h);
    for (int i = 0; i < w * h; ++i)
        tex.GetRawData()[i] = TColor(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]);
    Vec3 n = TextureUtility::GetNormalFromNormalmap(tex, u, v);
    out3[0] = n[0]; out3[1] = n[1]; out3[2] = n[2];
}

// Matrix3(a).inv()                                 (Matrix.h:180-211)
void ref_matrix3_inv(const float* m, float* out) {
    std::array<float, 9> a;
    for (int i = 0; i < 9; ++i) a[i] = m[i];
    PTUtility::Matrix3 M(a);
    PTUtility::Matrix3 I = M.inv();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[3 * i + j] = I[i][j];
}

// CreateMatrixFromColumns({e1,e2,e3}).inv() * n  (SPPM.cpp:68-74 bump step)
void ref_bump_rotate(const float* e1, const float* e2, const float* e3, const float* n, float* out3) {
    PTUtility::Matrix3 mat = PTUtility::CreateMatrixFromColumns(
        {Vec3(e1[0], e1[1], e1[2]), Vec3(e2[0], e2[1], e2[2]), Vec3(e3[0], e3[1], e3[2])}).inv();
    Vec3 r = (mat * Vec3(n[0], n[1], n[2])).normalized();
    out3[0] = r[0]; out3[1] = r[1]; out3[2] = r[2];
}

}  // extern "C"
