// FriendsRenderSPPM.cpp -- the host facade: OBJ/MTL scene ingestion, the
// material plug-in, camera/sampler records and the render entry points, all
// on top of the C ABI in include/sppm_amd.h.  Host-only C++17.
#include "FriendsRenderSPPM.h"

#include <math.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "../../include/sppm_amd.h"

namespace PTUtility {
namespace Material {
int DiffuseSurface::Kind() const { return SPPM_MAT_DIFFUSE; }
int MetalSurface::Kind() const { return SPPM_MAT_METAL; }
int Roughconductor::Kind() const { return SPPM_MAT_ROUGH_CONDUCTOR; }
int GGXGlass::Kind() const { return SPPM_MAT_GGX_GLASS; }
int GlassSurface::Kind() const { return SPPM_MAT_GLASS; }
int BiMetalSurface::Kind() const { return SPPM_MAT_BIMETAL; }
int DiffuseTransmitSurface::Kind() const { return SPPM_MAT_DIFFUSE_TRANSMIT; }
int GlossTransmitSurface::Kind() const { return SPPM_MAT_GLOSS_TRANSMIT; }
}  // namespace Material

// Default plug-in (SceneMaterial.h:13-16): a diffuse surface, other
// parameters left at their SURFACE_PARAM defaults.
Material::Surface* SceneMaterialManager::SetParamtoAttribute_p(Attribute* at, const std::string&) const {
    at->m_Surface = new Material::DiffuseSurface;
    return at->m_Surface;
}

void SceneMaterialManager::SetParamtoAttribute(Attribute* at, const std::string& MaterialName) const {
    Material::Surface* s = SetParamtoAttribute_p(at, MaterialName);
    if (s) m_Materials.push_back(std::unique_ptr<Material::Surface>(s));
}
}  // namespace PTUtility

using PTUtility::Attribute;
using PTUtility::Vec3;

// ---- Texture2D ---------------------------------------------------------------
void Texture2D::Init(int width, int height) {
    m_Width = width;
    m_Height = height;
    m_Data.assign((size_t)width * height * 4, 0.0f);
}

std::array<float, 4> Texture2D::GetPixel(int x, int y) const {
    const float* p = &m_Data[4 * ((size_t)x + (size_t)y * m_Width)];
    return {p[0], p[1], p[2], p[3]};
}

namespace TextureUtility {

static bool read_token(std::istream& in, std::string& tok) {
    tok.clear();
    int c;
    while ((c = in.get()) != EOF) {
        if (c == '#') {
            while ((c = in.get()) != EOF && c != '\n') {}
            continue;
        }
        if (!isspace(c)) {
            tok.push_back((char)c);
            break;
        }
    }
    while ((c = in.peek()) != EOF && !isspace(c)) tok.push_back((char)in.get());
    return !tok.empty();
}

bool LoadTextureFromFile(Texture2D& tex, const char* fileName) {
    std::ifstream in(fileName, std::ios::binary);
    if (!in) return false;
    std::string magic, sw, sh, sm;
    if (!read_token(in, magic) || !read_token(in, sw) || !read_token(in, sh) || !read_token(in, sm)) return false;
    int w = atoi(sw.c_str()), h = atoi(sh.c_str());
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1 << 28)) return false;
    in.get();  // the single whitespace byte before the raster
    tex.Init(w, h);
    float* d = tex.Data();
    if (magic == "P6") {
        if (atoi(sm.c_str()) != 255) return false;
        std::vector<unsigned char> raw((size_t)w * h * 3);
        if (!in.read((char*)raw.data(), (std::streamsize)raw.size())) return false;
        // 8-bit texel -> byte / 255.0f, alpha 1 (TextureUtility.cpp:12-35)
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            for (int c = 0; c < 3; ++c) d[4 * i + c] = (float)raw[3 * i + c] / 255.0f;
            d[4 * i + 3] = 1.0f;
        }
        return true;
    }
    if (magic == "PF" || magic == "Pf") {
        const int nc = magic == "PF" ? 3 : 1;
        const bool little = atof(sm.c_str()) < 0.0;
        std::vector<float> raw((size_t)w * h * nc);
        if (!in.read((char*)raw.data(), (std::streamsize)(raw.size() * 4))) return false;
        const uint16_t probe = 1;
        const bool host_little = *(const unsigned char*)&probe == 1;
        if (little != host_little) {
            for (float& f : raw) {
                unsigned char* b = (unsigned char*)&f;
                std::swap(b[0], b[3]);
                std::swap(b[1], b[2]);
            }
        }
        // PFM stores rows bottom to top
        for (int y = 0; y < h; ++y) {
            const float* src = &raw[(size_t)(h - 1 - y) * w * nc];
            for (int x = 0; x < w; ++x) {
                float* t = &d[4 * ((size_t)x + (size_t)y * w)];
                for (int c = 0; c < 3; ++c) t[c] = src[(size_t)x * nc + (nc == 3 ? c : 0)];
                t[3] = 1.0f;
            }
        }
        return true;
    }
    return false;
}

bool ExportPFM(const Texture2D& tex, const char* fileName) {
    std::ofstream out(fileName, std::ios::binary);
    if (!out) return false;
    const int w = tex.GetWidth(), h = tex.GetHeight();
    out << "PF\n" << w << " " << h << "\n-1.0\n";
    std::vector<float> row((size_t)w * 3);
    for (int y = h - 1; y >= 0; --y) {
        for (int x = 0; x < w; ++x) {
            auto p = tex.GetPixel(x, y);
            row[3 * (size_t)x + 0] = p[0];
            row[3 * (size_t)x + 1] = p[1];
            row[3 * (size_t)x + 2] = p[2];
        }
        out.write((const char*)row.data(), (std::streamsize)(row.size() * 4));
    }
    return (bool)out;
}

void SetColor(Texture2D& tex, float r, float g, float b, float a) {
    float* d = tex.Data();
    for (size_t i = 0; i < (size_t)tex.GetWidth() * tex.GetHeight(); ++i) {
        d[4 * i + 0] = r;
        d[4 * i + 1] = g;
        d[4 * i + 2] = b;
        d[4 * i + 3] = a;
    }
}

}  // namespace TextureUtility

// ---- samplers and cameras ---------------------------------------------------------
int UniformSample::Kind() const { return SPPM_SAMPLER_UNIFORM; }
int HaltonSample::Kind() const { return SPPM_SAMPLER_HALTON; }
int HammersleySample::Kind() const { return SPPM_SAMPLER_HAMMERSLEY; }
int PinholeCamera::Type() const { return SPPM_CAM_PINHOLE; }
int ThinLensCamera::Type() const { return SPPM_CAM_THINLENS; }

ThinLensCamera::ThinLensCamera(int ScreenWidth, int ScreenHeight, const Vec3& CameraPos, const Vec3& Target,
                               float Fov, float F, float DistanceToLens, float AspectRatio,
                               iCameraSsmpler* CameraSampler, const char* DiaphragmImageFileName)
    : iCamera(ScreenWidth, ScreenHeight, CameraPos, Target, Fov, F, DistanceToLens, AspectRatio, CameraSampler) {
    if (DiaphragmImageFileName != nullptr) {
        if (!TextureUtility::LoadTextureFromFile(_DiaphragmImage, DiaphragmImageFileName))
            throw std::invalid_argument(std::string("ThinLensCamera: cannot load diaphragm image ") +
                                        DiaphragmImageFileName);
        _HasDiaphragm = true;
    }
}

// ---- Scene: OBJ / MTL ingestion (Scene.cpp:13-218) ------------------------------
namespace {

struct MtlRecord {
    std::string name;
    float kd[3] = {0, 0, 0};
    float ke[3] = {0, 0, 0};
    std::string map_kd, map_d, map_bump;
};

std::string dir_of(const std::string& path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

bool parse_floats(std::istringstream& ls, float* out, int n) {
    for (int i = 0; i < n; ++i) {
        std::string t;
        if (!(ls >> t)) return false;
        char* end = nullptr;
        out[i] = strtof(t.c_str(), &end);
        if (end == t.c_str()) return false;
    }
    return true;
}

bool load_mtl(const std::string& path, std::vector<MtlRecord>& mats, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open material library " + path;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            mats.emplace_back();
            ls >> mats.back().name;
            continue;
        }
        if (mats.empty()) continue;
        MtlRecord& m = mats.back();
        if (key == "Kd") {
            if (!parse_floats(ls, m.kd, 3)) { err = "bad Kd in " + path; return false; }
        } else if (key == "Ke") {
            if (!parse_floats(ls, m.ke, 3)) { err = "bad Ke in " + path; return false; }
        } else if (key == "map_Kd") {
            ls >> m.map_kd;
        } else if (key == "map_d") {
            ls >> m.map_d;
        } else if (key == "map_bump" || key == "bump") {
            ls >> m.map_bump;
        }
    }
    return true;
}

// OBJ index (1-based, negative = relative to the end) -> 0-based or -1
int obj_index(const std::string& s, size_t count) {
    if (s.empty()) return -1;
    long v = strtol(s.c_str(), nullptr, 10);
    if (v > 0) return (int)(v - 1);
    if (v < 0) return (int)((long)count + v);
    return -2;
}

}  // namespace

int Scene::textureId(const std::string& path) {
    for (size_t i = 0; i < m_TextureNames.size(); ++i)
        if (m_TextureNames[i] == path) return (int)i;
    std::unique_ptr<Texture2D> t(new Texture2D);
    if (!TextureUtility::LoadTextureFromFile(*t, path.c_str())) return -1;
    m_Textures.push_back(std::move(t));
    m_TextureNames.push_back(path);
    return (int)m_Textures.size() - 1;
}

void Scene::ClearScene() {
    m_Pos.clear();
    m_Nrm.clear();
    m_UV.clear();
    m_Attr.clear();
    m_Attributes.clear();
    m_TexIds.clear();
    m_Textures.clear();
    m_TextureNames.clear();
    m_Error.clear();
    m_IBL = -1;
    ++m_Version;
}

bool Scene::CreateScene(const char* OBJ_FileName, const PTUtility::SceneMaterialManager* sceneMaterial,
                        Texture2D* IBLTexture) {
    ClearScene();
    PTUtility::SceneMaterialManager defmat;
    const PTUtility::SceneMaterialManager* smat = sceneMaterial ? sceneMaterial : &defmat;
    if (!OBJ_FileName) {
        m_Error = "NULL OBJ file name";
        return false;
    }
    const std::string base = dir_of(OBJ_FileName);
    std::ifstream in(OBJ_FileName);
    if (!in) {
        m_Error = std::string("cannot open ") + OBJ_FileName;
        return false;
    }
    std::vector<float> V, VN, VT;
    std::vector<MtlRecord> mats;
    int cur_mat = -1;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "v") {
            float p[3];
            if (!parse_floats(ls, p, 3)) { m_Error = "bad vertex at line " + std::to_string(lineno); return false; }
            V.insert(V.end(), p, p + 3);
        } else if (key == "vn") {
            float p[3];
            if (!parse_floats(ls, p, 3)) { m_Error = "bad normal at line " + std::to_string(lineno); return false; }
            VN.insert(VN.end(), p, p + 3);
        } else if (key == "vt") {
            float p[2];
            if (!parse_floats(ls, p, 2)) { m_Error = "bad texcoord at line " + std::to_string(lineno); return false; }
            VT.insert(VT.end(), p, p + 2);
        } else if (key == "mtllib") {
            std::string f;
            ls >> f;
            if (!load_mtl(base + f, mats, m_Error)) return false;
        } else if (key == "usemtl") {
            std::string n;
            ls >> n;
            cur_mat = -1;
            for (size_t i = 0; i < mats.size(); ++i)
                if (mats[i].name == n) cur_mat = (int)i;
            if (cur_mat < 0) { m_Error = "unknown material " + n + " at line " + std::to_string(lineno); return false; }
        } else if (key == "f") {
            if (cur_mat < 0) { m_Error = "face without material at line " + std::to_string(lineno); return false; }
            std::vector<std::array<int, 3>> corners;
            std::string t;
            while (ls >> t) {
                std::array<int, 3> c{-1, -1, -1};
                size_t s1 = t.find('/');
                std::string a = t.substr(0, s1), b, cc;
                if (s1 != std::string::npos) {
                    size_t s2 = t.find('/', s1 + 1);
                    b = t.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                    if (s2 != std::string::npos) cc = t.substr(s2 + 1);
                }
                c[0] = obj_index(a, V.size() / 3);
                c[1] = obj_index(b, VT.size() / 2);
                c[2] = obj_index(cc, VN.size() / 3);
                if (c[0] < 0 || c[0] >= (int)(V.size() / 3) || c[1] < -1 || c[1] >= (int)(VT.size() / 2) ||
                    c[2] < -1 || c[2] >= (int)(VN.size() / 3)) {
                    m_Error = "face index out of range at line " + std::to_string(lineno);
                    return false;
                }
                corners.push_back(c);
            }
            if (corners.size() < 3) { m_Error = "face with < 3 vertices at line " + std::to_string(lineno); return false; }
            // polygons: fan (0, k, k+1)
            for (size_t k = 1; k + 1 < corners.size(); ++k) {
                const std::array<int, 3>* tri[3] = {&corners[0], &corners[k], &corners[k + 1]};
                bool have_n = true;
                for (auto* c : tri) have_n = have_n && (*c)[2] >= 0;
                for (int j = 0; j < 3; ++j)
                    for (int e = 0; e < 3; ++e) m_Pos.push_back(V[3 * (size_t)(*tri[j])[0] + e]);
                const float* P = &m_Pos[m_Pos.size() - 9];
                if (have_n) {
                    for (int j = 0; j < 3; ++j)
                        for (int e = 0; e < 3; ++e) m_Nrm.push_back(VN[3 * (size_t)(*tri[j])[2] + e]);
                } else {
                    // face normal (p1 - p2).normalized().cross((p2 - p0).normalized()) (Scene.cpp:121-126)
                    float a[3], b[3];
                    for (int e = 0; e < 3; ++e) {
                        a[e] = P[3 + e] - P[6 + e];
                        b[e] = P[6 + e] - P[e];
                    }
                    float da = (float)sqrt((double)(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]));
                    float db = (float)sqrt((double)(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
                    for (int e = 0; e < 3; ++e) {
                        a[e] = a[e] / da;
                        b[e] = b[e] / db;
                    }
                    float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
                    for (int j = 0; j < 3; ++j) m_Nrm.insert(m_Nrm.end(), n, n + 3);
                }
                for (int j = 0; j < 3; ++j) {
                    int ti = (*tri[j])[1];
                    m_UV.push_back(ti >= 0 ? VT[2 * (size_t)ti] : 0.0f);
                    m_UV.push_back(ti >= 0 ? VT[2 * (size_t)ti + 1] : 0.0f);
                    m_UV.push_back(0.0f);
                }
                m_Attr.push_back(cur_mat);
            }
        }
        // o, g, s and other statements do not change the facet records
    }
    if (m_Attr.empty()) {
        m_Error = "no faces in the OBJ file";
        return false;
    }
    // attributes: Index 1, Albedo = Kd, Radiance = Ke, then the plug-in (Scene.cpp:58-76)
    for (const MtlRecord& m : mats) {
        std::unique_ptr<Attribute> at(new Attribute);
        at->m_SurfaceParam.Index = 1.0f;
        at->m_SurfaceParam.Albedo = Vec3(m.kd[0], m.kd[1], m.kd[2]);
        at->m_SurfaceParam.Radiance = Vec3(m.ke[0], m.ke[1], m.ke[2]);
        at->_AttributeName = m.name;
        smat->SetParamtoAttribute(at.get(), m.name);
        if (!at->m_Surface) {
            m_Error = "material manager gave no surface for " + m.name;
            return false;
        }
        std::array<int32_t, 3> ids{-1, -1, -1};
        const std::string* paths[3] = {&m.map_kd, &m.map_d, &m.map_bump};
        std::string* names[3] = {&at->m_AlbedoTextureFileName, &at->m_RoughnessTextureFileName,
                                 &at->m_BumpMapTextureFileName};
        for (int r = 0; r < 3; ++r) {
            if (paths[r]->size() <= 2) continue;  // Scene.cpp:41 ignores names of <= 2 characters
            const std::string full = base + *paths[r];
            ids[r] = textureId(full);
            if (ids[r] < 0) {
                m_Error = "cannot load texture " + full + " (binary PPM P6 or PFM expected)";
                return false;
            }
            *names[r] = full;
        }
        m_TexIds.push_back(ids);
        m_Attributes.push_back(std::move(at));
    }
    // environment light (Scene.cpp:214 hands it to the light sampler): a copy
    // of the map; the core converts it as IBL::Init does
    if (IBLTexture != nullptr) {
        if (IBLTexture->GetWidth() <= 0 || IBLTexture->GetHeight() <= 0) {
            m_Error = "empty IBL texture";
            return false;
        }
        std::unique_ptr<Texture2D> t(new Texture2D);
        t->Init(IBLTexture->GetWidth(), IBLTexture->GetHeight());
        memcpy(t->Data(), IBLTexture->Data(),
               sizeof(float) * 4 * (size_t)IBLTexture->GetWidth() * (size_t)IBLTexture->GetHeight());
        m_Textures.push_back(std::move(t));
        m_TextureNames.push_back("(environment map)");
        m_IBL = (int)m_Textures.size() - 1;
    }
    ++m_Version;
    return true;
}

void Scene::AttributeRecord(int i, void* out) const {
    const Attribute& a = *m_Attributes[i];
    const auto& p = a.m_SurfaceParam;
    sppm_attr r;
    memset(&r, 0, sizeof(r));
    r.kind = a.m_Surface->Kind();
    r.index = p.Index;
    r.if_delta = p.IfDeltaSurface;
    r.if_opaque = p.IfOpeque;
    for (int c = 0; c < 3; ++c) {
        r.albedo[c] = p.Albedo[c];
        r.radiance[c] = p.Radiance[c];
    }
    r.roughness = p.Roughness;
    r.reflectance = p.Reflectance;
    r.extinction = p.Extinction;
    r.is_sppm = p.isSPPMSurface;
    r.tex_albedo = m_TexIds[i][0];
    r.tex_roughness = m_TexIds[i][1];
    r.tex_bump = m_TexIds[i][2];
    memcpy(out, &r, sizeof(r));
}

float Scene::Diagonal() const {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < m_Pos.size(); i += 3)
        for (int e = 0; e < 3; ++e) {
            lo[e] = fminf(lo[e], m_Pos[i + e]);
            hi[e] = fmaxf(hi[e], m_Pos[i + e]);
        }
    double s = 0.0;
    for (int e = 0; e < 3; ++e) s += (double)(hi[e] - lo[e]) * (hi[e] - lo[e]);
    return (float)sqrt(s);
}

// ---- Render ----------------------------------------------------------------------
namespace Render {

static void check(int rc, const char* what) {
    if (rc != SPPM_OK) {
        const char* msg = sppm_last_error();
        throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "unknown error") + " (status " +
                                 std::to_string(rc) + ")");
    }
}

struct StochasticProgressivePhotonMapping::Impl {
    sppm_ctx* ctx = nullptr;
    const Scene* scene = nullptr;
    uint64_t version = 0;
    int bvh_width = 0;
    ~Impl() {
        if (ctx) sppm_destroy(ctx);
    }
};

StochasticProgressivePhotonMapping::StochasticProgressivePhotonMapping() : m_Impl(new Impl) {}
StochasticProgressivePhotonMapping::~StochasticProgressivePhotonMapping() {}

void StochasticProgressivePhotonMapping::prepare(const Scene& scene, const iCamera& camera, int photonsPerPass) {
    Impl& I = *m_Impl;
    if (!I.ctx) check(sppm_create(&I.ctx, m_Device, nullptr), "sppm_create");
    if (I.scene != &scene || I.version != scene.Version() || I.bvh_width != m_BVHWidth) {
        std::vector<sppm_attr> attrs(scene.NumAttributes());
        for (int i = 0; i < scene.NumAttributes(); ++i) scene.AttributeRecord(i, &attrs[i]);
        std::vector<sppm_texture> tex(scene.NumTextures());
        for (int i = 0; i < scene.NumTextures(); ++i) {
            tex[i].width = scene.GetTexture(i).GetWidth();
            tex[i].height = scene.GetTexture(i).GetHeight();
            tex[i].rgba = scene.GetTexture(i).Data();
        }
        sppm_scene_desc d;
        d.nfacets = scene.NumFacets();
        d.pos = scene.Positions().data();
        d.nrm = scene.Normals().data();
        d.uv = scene.UVs().data();
        d.attr = scene.FacetAttributes().data();
        d.nattrs = (int32_t)attrs.size();
        d.attrs = attrs.data();
        d.ntextures = (int32_t)tex.size();
        d.textures = tex.data();
        d.has_ibl = scene.IBLTextureId() >= 0 ? 1 : 0;
        d.ibl_texture = scene.IBLTextureId();
        check(sppm_scene_upload(I.ctx, &d, m_BVHWidth), "sppm_scene_upload");
        I.scene = &scene;
        I.version = scene.Version();
        I.bvh_width = m_BVHWidth;
    }
    sppm_camera_desc c;
    memset(&c, 0, sizeof(c));
    c.type = camera.Type();
    c.width = camera._ScreenWidth;
    c.height = camera._ScreenHeight;
    for (int e = 0; e < 3; ++e) {
        c.pos[e] = camera._CameraPos[e];
        c.target[e] = camera._Target[e];
    }
    c.fov_deg = camera._FovDegree;
    c.F = camera._F;
    c.dist_to_lens = camera._DistanceToLens;
    c.aspect = camera._AspectRatio;
    if (!camera._CameraSampler) throw std::invalid_argument("camera has no sampler");
    c.sampler = camera._CameraSampler->Kind();
    c.sampler_id = camera._CameraSampler->m_ID;
    c.sampler_nsample = camera._CameraSampler->m_NumSample;
    check(sppm_camera_set(I.ctx, &c), "sppm_camera_set");
    if (const Texture2D* d = camera.Diaphragm())
        check(sppm_camera_diaphragm(I.ctx, d->GetWidth(), d->GetHeight(), d->Data()), "sppm_camera_diaphragm");
    sppm_params p;
    memset(&p, 0, sizeof(p));
    p.alpha = m_Alpha;
    p.first_radius = m_FirstRadius;
    p.photons_per_pass = photonsPerPass;
    p.bvh_width = m_BVHWidth;
    p.eye_nee = m_EyeNEE;
    check(sppm_params_set(I.ctx, &p), "sppm_params_set");
    check(sppm_begin(I.ctx), "sppm_begin");
    m_Stats = RenderStatistics();
}

void StochasticProgressivePhotonMapping::finish(Texture2D& renderTarget, int64_t totalPhotons, int passes) {
    Impl& I = *m_Impl;
    int64_t hp, nb, st, ps;
    check(sppm_counters(I.ctx, &hp, &nb, &st, nullptr, &ps), "sppm_counters");
    m_Stats.hit_points = hp;
    m_Stats.neighbours = nb;
    m_Stats.photons_stored = st;
    check(sppm_reconstruct(I.ctx, totalPhotons, passes, renderTarget.Data()), "sppm_reconstruct");
}

static void check_target(const Texture2D& rt, const iCamera& cam) {
    if (rt.GetWidth() != cam._ScreenWidth || rt.GetHeight() != cam._ScreenHeight)
        throw std::invalid_argument("render target size differs from the camera's screen size");
}

void StochasticProgressivePhotonMapping::RenderImage_FixPasses(Texture2D& renderTarget, const Scene& scene,
                                                               const iCamera& camera, int NumPhotonsPerStep,
                                                               int Passes) {
    check_target(renderTarget, camera);
    TextureUtility::SetColor(renderTarget, 0, 0, 0, 1);
    if (Passes <= 0) return;
    prepare(scene, camera, NumPhotonsPerStep);
    auto t0 = std::chrono::steady_clock::now();
    // passes are queued without a host round trip; the last one reports stats
    sppm_pass_stats s;
    for (int i = 0; i < Passes; ++i) check(sppm_pass(m_Impl->ctx, i + 1 == Passes ? &s : nullptr), "sppm_pass");
    const int64_t total = (int64_t)s.photons_emitted * Passes;
    m_Stats.photons_emitted = total;
    m_Stats.flags = s.flags;
    m_Stats.passes = Passes;
    m_Stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finish(renderTarget, total, Passes);
}

void StochasticProgressivePhotonMapping::RenderImage_FixRenderTime(Texture2D& renderTarget, const Scene& scene,
                                                                   const iCamera& camera, int NumPhotonsPerStep,
                                                                   int RenderTimeSecond) {
    check_target(renderTarget, camera);
    TextureUtility::SetColor(renderTarget, 0, 0, 0, 1);
    prepare(scene, camera, NumPhotonsPerStep);
    auto t0 = std::chrono::steady_clock::now();
    const double budget = (double)RenderTimeSecond;
    int64_t total = 0;
    int passes = 0;
    double avg = 0.0;
    while (true) {
        auto ts = std::chrono::steady_clock::now();
        sppm_pass_stats s;
        check(sppm_pass(m_Impl->ctx, &s), "sppm_pass");  // stats make the pass synchronous
        total += s.photons_emitted;
        m_Stats.photons_emitted += s.photons_emitted;
        m_Stats.flags |= s.flags;
        ++passes;
        auto te = std::chrono::steady_clock::now();
        const double dt = std::chrono::duration<double>(te - ts).count();
        avg = (avg * (passes - 1) + dt) / passes;
        const double used = std::chrono::duration<double>(te - t0).count();
        if (used + avg > budget) break;
    }
    m_Stats.passes = passes;
    m_Stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finish(renderTarget, total, passes);
}

void StochasticProgressivePhotonMapping::RenderImage_FixStep(Texture2D& renderTarget, const Scene& scene,
                                                             const iCamera& camera, int NumPhotons, int Step) {
    check_target(renderTarget, camera);
    TextureUtility::SetColor(renderTarget, 0, 0, 0, 1);
    if (Step <= 0) return;
    prepare(scene, camera, NumPhotons / Step);
    auto t0 = std::chrono::steady_clock::now();
    sppm_pass_stats s;
    for (int i = 0; i < Step; ++i) check(sppm_pass(m_Impl->ctx, i + 1 == Step ? &s : nullptr), "sppm_pass");
    m_Stats.photons_emitted = (int64_t)s.photons_emitted * Step;
    m_Stats.flags = s.flags;
    m_Stats.passes = Step;
    m_Stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finish(renderTarget, NumPhotons, Step);
}

}  // namespace Render
