// render_obj -- renders an OBJ/MTL scene with the SPPM pass core through the
// C++ facade, the way the reference's RenderScene does (main.cpp:117-167):
// a SceneMaterialManager subclass maps MTL material names to surfaces, the
// scene is created from the OBJ file, a camera with a UniformSample stream is
// built, and RenderImage_* renders into a Texture2D that is written out.
//
// usage: render_obj SCENE.obj OUT.pfm [options]
//   --size W H            screen size (default 256 256)
//   --photons N           photons per pass (default 100000)
//   --passes P            fixed pass count (default 16)
//   --seconds S           time budget instead of a pass count (RenderImage_FixRenderTime)
//   --fixstep             RenderImage_FixStep(N * P photons, P steps)
//   --pinhole | --thinlens A F   camera kind; thin lens takes DistanceToLens A and F
//   --pos x y z --target x y z --fov deg
//   --alpha a --radius r  SPPM constants (default 0.9 / 0.015, SPPM.cpp:20-21)
//   --bvh W               BVH width 2, 4 or 8
//   --sampler-id ID       UniformSample seed ID
//   --ibl FILE            environment map (binary PPM or PFM), Scene::CreateScene's IBL argument
//   --diaphragm FILE      thin-lens diaphragm image (binary PPM or PFM), ThinLensCamera's last argument
//   --nee M               eye-path next-event estimation: 0 off, 1 as written, 2 kept
//   --dump-scene FILE     write the parsed facet records and attribute table
//                         (no rendering, no GPU)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/sppm_amd.h"
#include "FriendsRenderSPPM.h"

using namespace PTUtility;

// Material names of the generated test scenes, mapped like main.cpp:14-114.
class TestSceneMaterial : public SceneMaterialManager {
protected:
    Material::Surface* SetParamtoAttribute_p(Attribute* at, const std::string& name) const override {
        auto& p = at->m_SurfaceParam;
        if (name == "mirror_sphere") {
            at->m_Surface = new Material::MetalSurface();
            p.IfDeltaSurface = true;
            p.isSPPMSurface = false;
        } else if (name == "conductor_sphere" || name == "displaced_mesh") {
            at->m_Surface = new Material::Roughconductor();
            p.Roughness = name == "conductor_sphere" ? 0.3f : 0.4f;
            p.Index = 1.0f;
            p.Extinction = 20.0f;
            p.isSPPMSurface = true;
        } else if (name == "glass_sphere") {
            at->m_Surface = new Material::GGXGlass();
            p.Roughness = 0.1f;
            p.Index = 1.5f;
            p.IfOpeque = false;
            p.isSPPMSurface = false;
        } else if (name == "smooth_glass") {
            at->m_Surface = new Material::GlassSurface();
            p.Index = 1.5f;
            p.IfDeltaSurface = true;
            p.IfOpeque = false;
            p.isSPPMSurface = false;
        } else {
            at->m_Surface = new Material::DiffuseSurface;
            p.isSPPMSurface = true;
            p.IfOpeque = true;
        }
        return at->m_Surface;
    }
};

static int dump_scene(const Scene& sc, const char* path) {
    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    int32_t hdr[3] = {sc.NumFacets(), sc.NumAttributes(), sc.NumTextures()};
    fwrite(hdr, 4, 3, f);
    fwrite(sc.Positions().data(), 4, sc.Positions().size(), f);
    fwrite(sc.Normals().data(), 4, sc.Normals().size(), f);
    fwrite(sc.UVs().data(), 4, sc.UVs().size(), f);
    fwrite(sc.FacetAttributes().data(), 4, sc.FacetAttributes().size(), f);
    for (int i = 0; i < sc.NumAttributes(); ++i) {
        sppm_attr a;
        sc.AttributeRecord(i, &a);
        fwrite(&a, sizeof a, 1, f);
    }
    for (int i = 0; i < sc.NumTextures(); ++i) {
        const Texture2D& t = sc.GetTexture(i);
        int32_t wh[2] = {t.GetWidth(), t.GetHeight()};
        fwrite(wh, 4, 2, f);
        fwrite(t.Data(), 4, (size_t)wh[0] * wh[1] * 4, f);
    }
    return fclose(f) == 0 ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s SCENE.obj OUT.pfm [options]\n", argv[0]);
        return 2;
    }
    const char* obj = argv[1];
    const char* out = argv[2];
    int W = 256, H = 256, photons = 100000, passes = 16, seconds = -1, bvh = 8, sid = 0, nee = 0;
    bool thin = false, fixstep = false;
    float A = 1.0f, F = 1.0f, fov = 45.0f, alpha = 0.9f, radius = 0.015f;
    float pos[3] = {0.5f, 0.5f, 2.2f}, tgt[3] = {0.5f, 0.5f, 0.5f};
    const char* dump = nullptr;
    const char* ibl = nullptr;
    const char* diaphragm = nullptr;
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](int n) {
            if (i + n >= argc) {
                fprintf(stderr, "%s needs %d value(s)\n", a.c_str(), n);
                exit(2);
            }
        };
        if (a == "--size") { need(2); W = atoi(argv[++i]); H = atoi(argv[++i]); }
        else if (a == "--photons") { need(1); photons = atoi(argv[++i]); }
        else if (a == "--passes") { need(1); passes = atoi(argv[++i]); }
        else if (a == "--seconds") { need(1); seconds = atoi(argv[++i]); }
        else if (a == "--fixstep") { fixstep = true; }
        else if (a == "--pinhole") { thin = false; }
        else if (a == "--thinlens") { need(2); thin = true; A = (float)atof(argv[++i]); F = (float)atof(argv[++i]); }
        else if (a == "--pos") { need(3); for (float& v : pos) v = (float)atof(argv[++i]); }
        else if (a == "--target") { need(3); for (float& v : tgt) v = (float)atof(argv[++i]); }
        else if (a == "--fov") { need(1); fov = (float)atof(argv[++i]); }
        else if (a == "--alpha") { need(1); alpha = (float)atof(argv[++i]); }
        else if (a == "--radius") { need(1); radius = (float)atof(argv[++i]); }
        else if (a == "--bvh") { need(1); bvh = atoi(argv[++i]); }
        else if (a == "--ibl") { need(1); ibl = argv[++i]; }
        else if (a == "--diaphragm") { need(1); diaphragm = argv[++i]; }
        else if (a == "--nee") { need(1); nee = atoi(argv[++i]); }
        else if (a == "--sampler-id") { need(1); sid = atoi(argv[++i]); }
        else if (a == "--dump-scene") { need(1); dump = argv[++i]; }
        else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }

    TestSceneMaterial material;
    Scene scene;
    Texture2D iblTex;
    if (ibl && !TextureUtility::LoadTextureFromFile(iblTex, ibl)) {
        fprintf(stderr, "cannot read the environment map %s\n", ibl);
        return 1;
    }
    if (!scene.CreateScene(obj, &material, ibl ? &iblTex : nullptr)) {
        fprintf(stderr, "CreateScene failed: %s\n", scene.LastError().c_str());
        return 1;
    }
    if (dump) return dump_scene(scene, dump);

    UniformSample sampler(sid);
    const Vec3 P(pos[0], pos[1], pos[2]), T(tgt[0], tgt[1], tgt[2]);
    const float aspect = W / (float)H;
    try {
        Texture2D target;
        target.Init(W, H);
        Render::StochasticProgressivePhotonMapping render;
        render.SetAlpha(alpha);
        render.SetFirstRadius(radius);
        render.SetBVHWidth(bvh);
        render.SetEyeNEE(nee);
        PinholeCamera pin(W, H, P, T, fov, 1.0f, 1.0f, aspect, &sampler);
        ThinLensCamera lens(W, H, P, T, fov, F, A, aspect, &sampler, diaphragm);
        const iCamera& cam = thin ? (const iCamera&)lens : (const iCamera&)pin;
        if (seconds >= 0)
            render.RenderImage_FixRenderTime(target, scene, cam, photons, seconds);
        else if (fixstep)
            render.RenderImage_FixStep(target, scene, cam, photons * passes, passes);
        else
            render.RenderImage_FixPasses(target, scene, cam, photons, passes);
        const auto& s = render.GetStatistics();
        printf("passes %d photons %lld stored %lld hit_points %lld neighbours %lld flags %d seconds %.3f\n", s.passes,
               (long long)s.photons_emitted, (long long)s.photons_stored, (long long)s.hit_points,
               (long long)s.neighbours, s.flags, s.seconds);
        if (!TextureUtility::ExportPFM(target, out)) {
            fprintf(stderr, "cannot write %s\n", out);
            return 1;
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "render failed: %s\n", e.what());
        return 1;
    }
    return 0;
}
