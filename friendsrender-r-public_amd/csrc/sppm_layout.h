// sppm_layout.h -- HBM layout of the SPPM pass core, shared by kernels and host.
//
// Everything is structure-of-arrays so a wavefront of 64 lanes touching
// consecutive paths/pixels/photons issues coalesced 16-byte loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sppm_amd.h"

namespace sppm {

// ---- closest-hit structure ------------------------------------------------
// Wide BVH node of width W (2, 4, 8): child boxes as 6 rows of W floats
// (min x, min y, min z, max x, max y, max z), then W child references,
// padded to a multiple of 16 bytes.  Child reference: >= 0 internal node,
// LEAF_BIT | (start << 3) | (count - 1) leaf range of 1-8 triangles in the
// leaf-triangle array, EMPTY_CHILD empty slot (its box is inverted).
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t EMPTY_CHILD = 0xffffffffu;
constexpr int TRACE_LDS = 16;  // traversal stack entries per lane kept in LDS (deeper ones spill)
// LDS words of a 64-lane block's traversal stack: the node references, then
// their boxes' entry distances
constexpr int TRACE_STACK_WORDS = 2 * TRACE_LDS * 64;
__host__ __device__ constexpr int node_floats(int W) { return ((7 * W + 3) / 4) * 4; }

// Leaf triangle t (sorted order): 3 float4 = (v0.xyz, facet id bits),
// (e1 = v1 - v0, 0), (e2 = v2 - v0, 0).  e1/e2 are exactly the reference's
// p.m_V[1].m_Pos - p.m_V[0].m_Pos (Intersect.cpp:69-70).

struct DevAttr {       // SURFACE_PARAM subset + surface kind + texture roles
    int kind;
    float index;
    int opaque;
    int is_sppm;
    float albedo[3];
    float rough;
    float radiance[3];
    float refl;
    float absorb;
    int tex_albedo, tex_rough, tex_bump;
    float radiance_norm2;
    int delta;  // IfDeltaSurface
};

struct TexDesc {
    int width, height;
    long long offset;  // first texel in the texel pool
};

struct DevLight {      // LightSampler::AreaLight (LightSampler.h:20-42)
    int facet;
    int textured;
    int tab_off;       // first entry of this light's 1000-entry table
    float flux_norm;   // |Flux|, the pick-table value
    float flux[3];
    float area;        // inoutFacet::CalculateArea
    float normal[3];   // (n0+n1+n2).normalized()
    float tab_avg;     // TableSampler::m_Average of the 1000-entry table
    float p[9];        // facet vertices P0, P1, P2
};

struct SceneDev {
    int width;                        // BVH width
    int nnodes;
    const float* __restrict__ nodes;  // nnodes * node_floats(width)
    int ntri;
    const float4* __restrict__ tri;   // 3 * ntri
    int nfacets;
    const float* __restrict__ facet_nrm;   // 9 per facet
    const float2* __restrict__ facet_uv;   // 3 per facet
    const int* __restrict__ facet_attr;
    const DevAttr* __restrict__ attrs;
    const TexDesc* __restrict__ tex;
    const float4* __restrict__ texels;
    // lights
    int nlights;
    float lpick_avg;
    const float* __restrict__ lpick_prob;
    const int* __restrict__ lpick_left;
    const int* __restrict__ lpick_right;
    const DevLight* __restrict__ lights;
    const float4* __restrict__ ltab_uvw;  // (u, v, w, |L|)
    const float4* __restrict__ ltab_L;    // (L.xyz, 0)
    const float* __restrict__ ltab_prob;
    const int* __restrict__ ltab_left;
    const int* __restrict__ ltab_right;
    const int* __restrict__ facet_light;  // area light of each facet, -1 none
    // environment light: texture id of the converted map (0.4 c^2), -1 none
    int ibl_tex;
};

// ---- camera (iCamera, Camera.h:11-53), precomputed on the host ------------
struct CamDev {
    int type, W, H;
    float pos[3], dir[3], e1[3], e2[3];
    float PW, PH, LW, LH, F, D;
    int sampler;         // SPPM_SAMPLER_*
    float fixed_smp[3];  // Halton / Hammersley sample of this pass (index = step)
    // thin-lens diaphragm (Camera.h:106-111): alias table over the aperture
    // image's pixels, dia_n bins (0: plain lens sampling); bin k keeps its
    // left-pick probability and the lens (u, v) of its left and right pixels
    int dia_n;
    const float* dia_prob;
    const float4* dia_uv;
};

// ---- photon slots: std::vector<std::deque<Photon>> (SPPM.cpp:753) --------
// Persistent [slot][path] arrays.  A path stores at most 17 surface photons
// (depth 0..16) and its terminator lands at slot <= 18, so 19 slots hold
// every slot the reference ever touches.
constexpr int NSLOT = 19;

struct SlotsDev {
    int npaths;
    float4* pos;     // (pos.xyz, index (IOR))       [NSLOT][npaths]
    float4* flux;    // (flux.xyz, pdf)
    float4* dir;     // (indir.xyz, 0)
    uint8_t* valid;  // Photon::Valid               [NSLOT][npaths]
    int* count;      // flattened photons of path i this pass
};

// ---- flattened + cell-sorted photons (kdtree<Photon> replacement) --------
struct GridDev {
    int nx, ny, nz;
    float cell, inv_cell;
    float bmin[3];
    int ncells;
    const int* __restrict__ cell_start;  // ncells + 1
    const float4* __restrict__ pos;      // (pos.xyz, index)   cell order
    const float4* __restrict__ flux;     // (flux.xyz, pdf)
    const float4* __restrict__ dir;      // (indir.xyz, flat index bits)
};

// ---- hit points (HitPoint, SPPM.h:83-110) --------------------------------
struct HitDev {
    int np;
    float4* pos;     // (pos.xyz, type as float: 0 NONE, 1 HIT, 2 ONLY_DIRECTLIGHT)
    float4* nrm;     // (face-forwarded normal.xyz, IOR)
    float4* out;     // (-ray dir.xyz, roughness)
    float4* weight;  // (PixelWeight.xyz, reflectance)
    int2* mat;       // (kind, facet id)
    float* absorb;   // extinction
    float4* direct;  // (DirectLight.xyz, 0): the environment seen by escaping eye paths
};

// ---- SharedHitPoint (SPPM.h:112-126) --------------------------------------
struct SharedDev {
    float4* rad;   // (Radiance.xyz, Radius)
    int2* cnt;     // (NumPhoton, NumGather)
    float4* direct;  // (DirectLight.xyz, NumDirectLight as int bits)
};

}  // namespace sppm
