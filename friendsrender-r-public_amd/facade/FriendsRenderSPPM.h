// FriendsRenderSPPM.h -- C++ host facade over the SPPM pass core C ABI
// (include/sppm_amd.h).  It keeps the class names, constructor signatures and
// render entry points of the FriendsRender R host code so a program written
// against the reference switches by changing its includes:
//
//   Render::StochasticProgressivePhotonMapping   SPPM/SPPM/SPPM.h:14-188
//     RenderImage_FixRenderTime / RenderImage_FixStep  SPPM.h:184-185
//   Scene::CreateScene(obj, SceneMaterialManager*, IBL) Scene.h:70, Scene.cpp:13-218
//   PTUtility::SceneMaterialManager::SetParamtoAttribute_p  SceneMaterial.h:9-33
//   PTUtility::Attribute / SURFACE_PARAM      Normal_Facet.h:86-105, material.h:35-90
//   PTUtility::Material::{DiffuseSurface, MetalSurface, Roughconductor,
//                         GGXGlass, GlassSurface, BiMetalSurface,
//                         DiffuseTransmitSurface, GlossTransmitSurface}
//                                             material.h
//   iCamera / PinholeCamera / ThinLensCamera  Camera.h:11-160
//   iCameraSsmpler / UniformSample / HaltonSample / HammersleySample
//                                             CameraSampler.h:7-55
//   Texture2D                                 Texture.h:10-42
//
// Differences, each an error rather than a silent change:
//   * the environment light (IBL texture argument) lights escaping eye
//     paths as the reference's does (SPHERE mapping, colour 0.4 c^2); like
//     the reference, it emits no photons;
//   * a ThinLensCamera diaphragm image that cannot be read throws (the
//     reference goes on with an empty image);
//   * textures referenced by an MTL file are read from binary PPM (P6, 8-bit,
//     stored as byte/255 like the reference's loader) or PFM (float) files;
//   * surface kinds outside the five above make CreateScene return false.
// Device work runs on one HIP device; the renderer owns the device state.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace PTUtility {

struct Vec3 {
    float v[3];
    Vec3() : v{0.0f, 0.0f, 0.0f} {}
    Vec3(float x, float y, float z) : v{x, y, z} {}
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float operator[](int i) const { return v[i]; }
    float& operator[](int i) { return v[i]; }
    static Vec3 Ones() { return Vec3(1.0f, 1.0f, 1.0f); }
    static Vec3 Zero() { return Vec3(); }
    Vec3 operator*(float s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
};

namespace Material {

// Surface kinds the pass core implements (material.h).  The objects carry no
// state; the kind selects the device BSDF.
struct Surface {
    virtual ~Surface() {}
    virtual int Kind() const = 0;
};
struct DiffuseSurface : Surface { int Kind() const override; };
struct MetalSurface : Surface { int Kind() const override; };
struct Roughconductor : Surface { int Kind() const override; };
struct GGXGlass : Surface { int Kind() const override; };
struct GlassSurface : Surface { int Kind() const override; };
struct BiMetalSurface : Surface { int Kind() const override; };
struct DiffuseTransmitSurface : Surface { int Kind() const override; };
struct GlossTransmitSurface : Surface { int Kind() const override; };

// SURFACE_PARAM with the reference's defaults (material.h:77-89).
struct SURFACE_PARAM {
    float Index = 1.0f;
    bool IfDeltaSurface = false;
    bool IfOpeque = true;
    Vec3 Albedo;
    float Roughness = 1.0f;
    float RoughnessU = 0.0f;
    float RoughnessV = 0.0f;
    Vec3 Radiance;
    float Reflectance = 1.0f;
    float Extinction = 1.0f;
    bool isSPPMSurface = false;
    bool Bidirectional = false;
};

}  // namespace Material

// Attribute (Normal_Facet.h:86-105): surface object + parameters + texture
// file names for map_Kd / map_d / map_bump.
struct Attribute {
    Material::Surface* m_Surface = nullptr;
    Material::SURFACE_PARAM m_SurfaceParam;
    std::string m_AlbedoTextureFileName;
    std::string m_RoughnessTextureFileName;
    std::string m_BumpMapTextureFileName;
    std::string _AttributeName = "Default";
};

// Material plug-in (SceneMaterial.h:9-33).  Override SetParamtoAttribute_p
// to set the surface kind and parameters for each MTL material name; the
// manager owns the surfaces it returns and must outlive the Scene.
class SceneMaterialManager {
    mutable std::vector<std::unique_ptr<Material::Surface>> m_Materials;

protected:
    virtual Material::Surface* SetParamtoAttribute_p(Attribute* at, const std::string& MaterialName) const;

public:
    void SetParamtoAttribute(Attribute* at, const std::string& MaterialName) const;
    virtual ~SceneMaterialManager() {}
};

}  // namespace PTUtility

// RGBA32F image (Texture.h:10-42), texel (x, y) at 4 * (x + y * width).
class Texture2D {
    int m_Width = 0, m_Height = 0;
    std::vector<float> m_Data;

public:
    void Init(int width, int height);
    int GetWidth() const { return m_Width; }
    int GetHeight() const { return m_Height; }
    float* Data() { return m_Data.data(); }
    const float* Data() const { return m_Data.data(); }
    std::array<float, 4> GetPixel(int x, int y) const;
};

namespace TextureUtility {
// 8-bit binary PPM (P6) -> byte / 255, alpha 1; or PFM (PF/Pf) floats.
bool LoadTextureFromFile(Texture2D& tex, const char* fileName);
// PFM (RGB float, bottom-to-top rows as the format specifies).
bool ExportPFM(const Texture2D& tex, const char* fileName);
void SetColor(Texture2D& tex, float r, float g, float b, float a);
}  // namespace TextureUtility

// Camera sample streams (CameraSampler.h).  They only carry the kind and
// parameters: the samples themselves are generated on the device.
class iCameraSsmpler {
public:
    virtual ~iCameraSsmpler() {}
    virtual int Kind() const = 0;
    int m_ID = 0;
    int m_NumSample = 1;
};
class UniformSample : public iCameraSsmpler {
public:
    explicit UniformSample(int ID) { m_ID = ID; }
    int Kind() const override;
};
class HaltonSample : public iCameraSsmpler {
public:
    explicit HaltonSample(int ID) { m_ID = ID; }
    int Kind() const override;
};
class HammersleySample : public iCameraSsmpler {
public:
    HammersleySample(int ID, int NumSample) { m_ID = ID; m_NumSample = NumSample; }
    int Kind() const override;
};

// Cameras (Camera.h:11-160): the constructor arguments are kept and passed
// to the device, which derives the basis and image plane as Camera.h:34-53.
struct iCamera {
    int _ScreenWidth, _ScreenHeight;
    PTUtility::Vec3 _CameraPos, _Target;
    float _FovDegree, _F, _DistanceToLens, _AspectRatio;
    iCameraSsmpler* _CameraSampler;
    iCamera(int ScreenWidth, int ScreenHeight, const PTUtility::Vec3& CameraPos, const PTUtility::Vec3& Target,
            float Fov, float F, float DistanceToLens, float AspectRatio, iCameraSsmpler* CameraSampler)
        : _ScreenWidth(ScreenWidth), _ScreenHeight(ScreenHeight), _CameraPos(CameraPos), _Target(Target),
          _FovDegree(Fov), _F(F), _DistanceToLens(DistanceToLens), _AspectRatio(AspectRatio),
          _CameraSampler(CameraSampler) {}
    virtual ~iCamera() {}
    virtual int Type() const = 0;
    // the thin lens's diaphragm image (NULL: none)
    virtual const Texture2D* Diaphragm() const { return nullptr; }
};
struct PinholeCamera : iCamera {
    using iCamera::iCamera;
    int Type() const override;
};
struct ThinLensCamera : iCamera {
    ThinLensCamera(int ScreenWidth, int ScreenHeight, const PTUtility::Vec3& CameraPos,
                   const PTUtility::Vec3& Target, float Fov, float F, float DistanceToLens, float AspectRatio,
                   iCameraSsmpler* CameraSampler, const char* DiaphragmImageFileName);
    int Type() const override;
    const Texture2D* Diaphragm() const override { return _HasDiaphragm ? &_DiaphragmImage : nullptr; }

private:
    Texture2D _DiaphragmImage;
    bool _HasDiaphragm = false;
};

// Scene (Scene.h:55-93): facets, attributes and textures on the host; the
// renderer uploads them (and builds the closest-hit structure on the GPU)
// when it first renders the scene.
class Scene {
public:
    // CreateScene as Scene.cpp:13-218: OBJ + MTL (v, vn, vt, f, usemtl,
    // mtllib; newmtl, Kd, Ke, map_Kd, map_d, map_bump), attribute i =
    // MTL material i with Index 1, Albedo = Kd, Radiance = Ke, then the
    // material manager (a default one when null).  Texture paths are
    // relative to the OBJ file's directory.  Returns false on error (message
    // in LastError()).
    bool CreateScene(const char* OBJ_FileName, const PTUtility::SceneMaterialManager* sceneMaterial,
                     Texture2D* IBLTexture);
    void ClearScene();
    const std::string& LastError() const { return m_Error; }

    int NumFacets() const { return (int)m_Attr.size(); }
    // Facet records as uploaded: 9 floats each for positions, normals, UVs.
    const std::vector<float>& Positions() const { return m_Pos; }
    const std::vector<float>& Normals() const { return m_Nrm; }
    const std::vector<float>& UVs() const { return m_UV; }
    const std::vector<int32_t>& FacetAttributes() const { return m_Attr; }
    int NumAttributes() const { return (int)m_Attributes.size(); }
    const PTUtility::Attribute& GetAttribute(int i) const { return *m_Attributes[i]; }
    // Attribute i as the C ABI record (kind, parameters, texture ids).
    void AttributeRecord(int i, void* sppm_attr_out) const;
    int NumTextures() const { return (int)m_Textures.size(); }
    // texture id of the environment map copied from CreateScene's IBL argument, or -1
    int IBLTextureId() const { return m_IBL; }
    const Texture2D& GetTexture(int i) const { return *m_Textures[i]; }
    float Diagonal() const;
    uint64_t Version() const { return m_Version; }

private:
    std::vector<float> m_Pos, m_Nrm, m_UV;
    std::vector<int32_t> m_Attr;
    std::vector<std::unique_ptr<PTUtility::Attribute>> m_Attributes;
    std::vector<std::array<int32_t, 3>> m_TexIds;  // albedo, roughness, bump per attribute
    std::vector<std::unique_ptr<Texture2D>> m_Textures;
    std::vector<std::string> m_TextureNames;
    std::string m_Error;
    uint64_t m_Version = 0;
    int m_IBL = -1;
    int textureId(const std::string& path);
};

namespace Render {

// Per-pass statistics of the last render (sppm_pass_stats summed/kept).
struct RenderStatistics {
    int passes = 0;
    int64_t photons_emitted = 0;
    int64_t photons_stored = 0;
    int64_t hit_points = 0;
    int64_t neighbours = 0;
    int flags = 0;
    double seconds = 0.0;
};

class StochasticProgressivePhotonMapping {
public:
    StochasticProgressivePhotonMapping();
    ~StochasticProgressivePhotonMapping();
    StochasticProgressivePhotonMapping(const StochasticProgressivePhotonMapping&) = delete;
    StochasticProgressivePhotonMapping& operator=(const StochasticProgressivePhotonMapping&) = delete;

    // SPPM.cpp:719-858.  Passes run until the time budget would be exceeded
    // (at least one pass); each pass emits NumPhotonsPerStep photons with
    // step = pass + 1.  The reference's photon reduction near the deadline
    // (SPPM.cpp:783-795) is not applied: every pass is full size.
    void RenderImage_FixRenderTime(Texture2D& renderTarget, const Scene& scene, const iCamera& camera,
                                   int NumPhotonsPerStep, int RenderTimeSecond);
    // Step passes of NumPhotons / Step photons each with the loop body of
    // RenderImage_FixRenderTime, reconstructed with (NumPhotons, Step) as
    // SPPM.cpp:682-717 does.  (The reference's own FixStep indexes an empty
    // photon buffer, so its photon pass is undefined.)
    void RenderImage_FixStep(Texture2D& renderTarget, const Scene& scene, const iCamera& camera, int NumPhotons,
                             int Step);
    // Fixed pass count variant of RenderImage_FixRenderTime (the loop of
    // SPPM.cpp:770-854 run Passes times, then reconstructed with
    // TotalPhoton = Passes * NumPhotonsPerStep).
    void RenderImage_FixPasses(Texture2D& renderTarget, const Scene& scene, const iCamera& camera,
                               int NumPhotonsPerStep, int Passes);

    // Constants of SPPM.cpp:20-21 (defaults 0.9 and 0.015) and run options.
    void SetAlpha(float alpha) { m_Alpha = alpha; }
    void SetFirstRadius(float r) { m_FirstRadius = r; }
    void SetBVHWidth(int w) { m_BVHWidth = w; }
    void SetDevice(int device) { m_Device = device; }
    // Next-event estimation on eye paths, the reference's compiled-out
    // blocks of getHitpointInfo (SPPM.cpp:144-168, 187-229): SPPM_NEE_OFF
    // (default, as the reference is built), SPPM_NEE_AS_WRITTEN, SPPM_NEE_KEEP.
    void SetEyeNEE(int mode) { m_EyeNEE = mode; }
    const RenderStatistics& GetStatistics() const { return m_Stats; }

private:
    struct Impl;
    std::unique_ptr<Impl> m_Impl;
    float m_Alpha = 0.9f;
    float m_FirstRadius = 0.015f;
    int m_BVHWidth = 8;
    int m_Device = 0;
    int m_EyeNEE = 0;
    RenderStatistics m_Stats;
    void prepare(const Scene& scene, const iCamera& camera, int photonsPerPass);
    void finish(Texture2D& renderTarget, int64_t totalPhotons, int passes);
};

}  // namespace Render
