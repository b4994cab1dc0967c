/*
 * sppm_amd.h -- C ABI of the MI355X-native SPPM pass core.
 *
 * This is the drop-in boundary for the per-pass loop of FriendsRender R
 * (SPPM/SPPM/SPPM.cpp:770-854).  Every entry point is plain C: opaque
 * context handle, plain pointers and sizes, int status (0 = ok, <0 = error,
 * message in sppm_last_error()).  Device memory is owned by the context;
 * host buffers are owned by the caller.  A context is bound to one HIP
 * stream (two private side streams join it inside each pass) and is not
 * thread-safe; distinct contexts are independent.
 *
 * Reference interfaces replaced (file:line under SPPM/SPPM/):
 *   Scene::CreateScene ............... Scene.cpp:13-218   -> sppm_scene_upload
 *   ThinLensCamera / PinholeCamera ... Camera.h:72-160    -> sppm_camera_set
 *   UniformSample/Halton/Hammersley .. CameraSampler.h:19-55 (sampler kind)
 *   constants ppm_Alpha, FirstRadius . SPPM.cpp:20-21     -> sppm_params
 *   initializeSharedPoints ........... SPPM.cpp:356-373   -> sppm_begin
 *   loop body resets ................. SPPM.cpp:771-779   -> sppm_pass
 *   createPhotonMap .................. SPPM.cpp:495-680   -> sppm_pass_photons
 *   photon flatten + kdtree::CreateTree SPPM.cpp:812-828, kdtree.h:417-431
 *                                                          -> sppm_pass_build
 *   createHitpoints .................. SPPM.cpp:308-353   -> sppm_pass_eye
 *   applyHitPointInfo ................ SPPM.cpp:395-451   -> sppm_pass_gather
 *   reconstructImageFromSharedPoints . SPPM.cpp:473-493   -> sppm_reconstruct
 *   N_BVH<...>::ClosestHit ........... bvh8.h:709-712     -> sppm_trace_closest
 */
#ifndef SPPM_AMD_H
#define SPPM_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPPM_ABI_VERSION 6  /* 2: sppm_guide_params gained hidden, layers, fp16; 3: sppm_params.eye_nee;
                               4: sppm_guide_params.train_interval, sppm_params.step_offset;
                               5: sppm_params.photon_kernel, wave_bounces, wave_lanes,
                                  sppm_guide_params.target;
                               6: sppm_tune_result gained bvh_builder ... photons_suggested */

/* ---- status codes -------------------------------------------------- */
#define SPPM_OK 0
#define SPPM_E_ARG -1        /* bad argument / size */
#define SPPM_E_HIP -2        /* a HIP runtime call failed */
#define SPPM_E_STATE -3      /* call out of order (no scene, no camera, ...) */
#define SPPM_E_NOMEM -4      /* device allocation failed */
#define SPPM_E_SCENE -5      /* scene validation failed */
#define SPPM_E_NUMERIC -6    /* NaN/Inf detected in per-pass flux */
#define SPPM_E_OVERFLOW -7   /* a fixed-capacity device structure overflowed */

/* ---- material kinds (subclasses of PTUtility::Material::Surface) ------- */
#define SPPM_MAT_DIFFUSE 0          /* DiffuseSurface   material.h:163-177 */
#define SPPM_MAT_METAL 1            /* MetalSurface     material.h:396-425 */
#define SPPM_MAT_ROUGH_CONDUCTOR 2  /* Roughconductor   material.h:222-278 */
#define SPPM_MAT_GGX_GLASS 3        /* GGXGlass         material.h:280-394 */
#define SPPM_MAT_GLASS 4            /* GlassSurface     material.h:465-532 */
#define SPPM_MAT_BIMETAL 5          /* BiMetalSurface   material.h:427-463 */
#define SPPM_MAT_DIFFUSE_TRANSMIT 6 /* DiffuseTransmitSurface material.h:646-683 (Reflectance = reflected share) */
#define SPPM_MAT_GLOSS_TRANSMIT 7   /* GlossTransmitSurface   material.h:685-766 (Reflectance, Roughness) */

/* Attribute = SURFACE_PARAM (material.h:35-90) + surface kind + texture
 * roles map_Kd / map_d / map_bump (Scene.cpp:72-75). */
typedef struct sppm_attr {
    int32_t kind;          /* SPPM_MAT_* */
    float index;           /* SURFACE_PARAM::Index (IOR) */
    int32_t if_delta;      /* IfDeltaSurface (informational) */
    int32_t if_opaque;     /* IfOpeque */
    float albedo[3];       /* Albedo */
    float roughness;       /* Roughness */
    float radiance[3];     /* Radiance (emission, Ke) */
    float reflectance;     /* Reflectance */
    float extinction;      /* Extinction (conductor k) */
    int32_t is_sppm;       /* isSPPMSurface: hit points / photons stored here */
    int32_t tex_albedo;    /* texture id or -1 (map_Kd) */
    int32_t tex_roughness; /* texture id or -1 (map_d) */
    int32_t tex_bump;      /* texture id or -1 (map_bump, normal map) */
} sppm_attr;

/* RGBA32F texture, row-major, texel (x,y) at rgba[4*(x + y*width)]
 * (Texture2D, Texture.h:10-42; loader convention TextureUtility.cpp:12-35). */
typedef struct sppm_texture {
    int32_t width, height;
    const float* rgba;
} sppm_texture;

/* Flattened facet list (inoutFacet, Normal_Facet.h:16-59).  Facet i is
 * triangle i of the closest-hit structure and its FacetID (Scene.cpp:141-142).
 * pos/nrm/uv: 9 floats per facet (vertex 0,1,2 xyz; uv.z is 0). */
typedef struct sppm_scene_desc {
    int32_t nfacets;
    const float* pos;
    const float* nrm;
    const float* uv;
    const int32_t* attr;    /* attribute id per facet, 0 <= id < nattrs */
    int32_t nattrs;
    const sppm_attr* attrs;
    int32_t ntextures;
    const sppm_texture* textures;
    /* Environment light (IBL.h, LightSampler.cpp:22-26; Scene::CreateScene's
     * IBLTexture): texture id of the map, used with the SPHERE mapping and the
     * colour conversion 0.4 * c^2 the reference applies; eye paths that escape
     * add it to the hit point's DirectLight (SPPM.cpp:290-300).  Photons are
     * not emitted from it, as in the reference. */
    int32_t has_ibl;        /* 0: no environment light */
    int32_t ibl_texture;    /* texture id when has_ibl */
} sppm_scene_desc;

/* ---- camera ------------------------------------------------------------ */
#define SPPM_CAM_PINHOLE 0    /* PinholeCamera  Camera.h:72-87 */
#define SPPM_CAM_THINLENS 1   /* ThinLensCamera Camera.h:89-160 (diaphragm: sppm_camera_diaphragm) */

#define SPPM_SAMPLER_UNIFORM 0    /* UniformSample(ID): one shared mt19937 stream */
#define SPPM_SAMPLER_HALTON 1     /* HaltonSample */
#define SPPM_SAMPLER_HAMMERSLEY 2 /* HammersleySample(ID, NumSample) */

typedef struct sppm_camera_desc {
    int32_t type;              /* SPPM_CAM_* */
    int32_t width, height;     /* ScreenWidth, ScreenHeight */
    float pos[3];              /* CameraPos */
    float target[3];           /* Target */
    float fov_deg;             /* Fov (full angle in degrees; half angle = PI*Fov/360) */
    float F;                   /* focal length F */
    float dist_to_lens;        /* DistanceToLens */
    float aspect;              /* AspectRatio */
    int32_t sampler;           /* SPPM_SAMPLER_* */
    int32_t sampler_id;        /* UniformSample seed ID */
    int32_t sampler_nsample;   /* Hammersley NumSample */
} sppm_camera_desc;

/* ---- parameters (SPPM.cpp:20-21 constants + run shape) ------------------ */
typedef struct sppm_params {
    float alpha;               /* ppm_Alpha; reference 0.9 */
    float first_radius;        /* FirstRadius; reference 0.015 */
    int32_t photons_per_pass;  /* NumPhotonsPerStep */
    int32_t bvh_width;         /* 2, 4 or 8 (closest hit is width-independent) */
    int32_t block_size;        /* threads per block for path kernels (0 = default) */
    int32_t check_numeric;     /* 1: fail a pass with SPPM_E_NUMERIC on NaN/Inf flux */
    int32_t record_neighbours; /* 1: neighbour lists from a separate recording gather kernel;
                                  2: the production gather itself appends its (hit point,
                                  photon) pairs, at most max_neighbours per pass (debug/parity) */
    int32_t max_neighbours;    /* capacity of the pair buffer in mode 2 (entries) */
    int32_t eye_nee;           /* SPPM_NEE_*: next-event estimation on eye paths */
    int32_t step_offset;       /* sppm_pass runs step = passes done + 1 + step_offset (the reference's
                                  step, SPPM.cpp:798, 836, which seeds every photon and eye path):
                                  a second render with another offset draws independent streams */
    int32_t photon_kernel;     /* SPPM_PHOTON_*: how the photon pass (createPhotonMap) runs */
    int32_t wave_bounces;      /* wavefront: bounces traced from ray queues before the last ones
                                  finish one lane per path (0 = default) */
    int32_t wave_lanes;        /* wavefront: lanes per ray of the closest-hit kernel, 2, 4 or 8 and
                                  at most the BVH width (0 = default) */
} sppm_params;

/* Photon pass kernels (the stored photons are identical either way).
 *   SPPM_PHOTON_MEGAKERNEL  (default, 0) one persistent kernel, one lane per
 *                           path from emission to its end;
 *   SPPM_PHOTON_WAVEFRONT   queues of rays between an emission kernel, a
 *                           closest-hit kernel with wave_lanes lanes per ray
 *                           (coalesced wide-node fetch; a ray's group takes the
 *                           next ray as soon as its ray ends) and a shading
 *                           kernel, wave_bounces times, then one lane per path.
 *                           Faster per ray on full queues, but every bounce
 *                           launch waits for its slowest ray: on the
 *                           full-material workload a whole pass takes longer. */
#define SPPM_PHOTON_MEGAKERNEL 0
#define SPPM_PHOTON_WAVEFRONT 1

/* Next-event estimation with MIS on eye paths: the two blocks of
 * getHitpointInfo that the reference compiles out with #if 0
 * (SPPM.cpp:144-168 at emitter hits, SPPM.cpp:187-229 at surfaces before the
 * hit point; shadow rays through Scene::IsOccluded with translucency 0.5).
 *   SPPM_NEE_OFF         the reference as compiled (default);
 *   SPPM_NEE_AS_WRITTEN  both blocks enabled exactly as written: eye paths end
 *                        at emitters, and the direct light they estimate is
 *                        kept only by paths that leave the scene before making
 *                        a hit point (makeHitPoint resets DirectLight, and a
 *                        path ending on an emitter keeps type NONE);
 *   SPPM_NEE_KEEP        the same estimates, kept: a hit point starts from
 *                        the DirectLight gathered before it, and a path ending
 *                        on an emitter with no hit point counts as
 *                        ONLY_DIRECTLIGHT.  This adds the light that reaches
 *                        the eye through surfaces without hit points
 *                        (emitters seen directly or through glossy and
 *                        specular non-SPPM surfaces), which photons never
 *                        carry. */
#define SPPM_NEE_OFF 0
#define SPPM_NEE_AS_WRITTEN 1
#define SPPM_NEE_KEEP 2

/* Per-pass statistics. */
typedef struct sppm_pass_stats {
    int32_t photons_emitted;   /* createPhotonMap return value */
    int32_t photons_stored;    /* flattened photon count (SPPM.cpp:812-822) */
    int32_t hit_points;        /* hit points of type HIT */
    int64_t neighbours;        /* sum over hit points of in-radius photons (M) */
    float mean_radius;         /* mean shared radius of the hit points after the update */
    int32_t flags;             /* bit0 NaN/Inf radiance; bit1 rng cursor overflow; bit2 traversal
                                  stack overflow; bit3 photon slot overflow; bit4 a stored photon
                                  had pdf <= 0 or non-finite flux; bit5 a hit point weight was
                                  non-finite (bits 4-5 follow the reference's arithmetic) */
    float ms_photons, ms_build, ms_eye, ms_gather;  /* phase times (hipEvent) */
} sppm_pass_stats;

typedef struct sppm_ctx sppm_ctx;

/* ---- lifecycle ---------------------------------------------------------- */
int sppm_abi_version(void);
const char* sppm_last_error(void);
/* device: HIP device ordinal; stream: hipStream_t or NULL for a private one */
int sppm_create(sppm_ctx** out, int device, void* stream);
int sppm_destroy(sppm_ctx* ctx);

/* ---- setup (Scene::CreateScene, camera ctor, constants) ------------------ */
/* bvh_width: 2, 4 or 8, optionally | SPPM_BVH_LBVH.  The closest-hit
 * structure is built on the GPU: binned-SAH top-down by default (replaces
 * N_BVH::Build3, bvh8.h:135-189), or with SPPM_BVH_LBVH a radix tree over
 * Morton codes; either is collapsed to `bvh_width`-wide nodes. */
#define SPPM_BVH_LBVH 0x100
int sppm_scene_upload(sppm_ctx* ctx, const sppm_scene_desc* scene, int bvh_width);
/* Clears the thin-lens diaphragm (sppm_camera_diaphragm); the render must be
 * begun again (sppm_begin). */
int sppm_camera_set(sppm_ctx* ctx, const sppm_camera_desc* cam);
/* ThinLensCamera's diaphragm image (the DiaphragmImageFileName argument,
 * Camera.h:136-158, and the lens sampling of Camera.h:106-111): the image the
 * reference loads from that file, width x height RGBA float32 in the
 * texture's storage order (pixel i = x + y * width).  Lens points are drawn
 * with the camera sampler's first two numbers from an alias table over the
 * pixels weighted by 1 - red; pixel i sits at lens (u, v) = ((i % width) /
 * width, 1 - (i / width) / height).  When more than 99.9 % of the pixels are
 * open (red < 1e-6) the lens keeps its plain sampling, as in the reference.
 * rgba = NULL clears it.  Thin-lens cameras only; call after sppm_camera_set
 * and before sppm_begin. */
int sppm_camera_diaphragm(sppm_ctx* ctx, int32_t width, int32_t height, const float* rgba);
int sppm_params_set(sppm_ctx* ctx, const sppm_params* params);
/* initializeSharedPoints + photon slot allocation; resets the pass counter. */
int sppm_begin(sppm_ctx* ctx);

/* ---- per pass ------------------------------------------------------------- */
/* One full pass (SPPM.cpp:770-843): resets, photon pass (step = pass+1),
 * gather-structure build, eye pass (step = pass+1), gather + update.  The
 * eye pass runs on a side stream beside the photon pass and the build; when
 * the previous call was also sppm_pass and emission is not guided, the
 * photon pass starts on a second side stream beside the previous pass's
 * gather.  Work queued on the context stream afterwards sees the whole
 * pass. */
int sppm_pass(sppm_ctx* ctx, sppm_pass_stats* stats);
/* Split steps of the same pass, for tests, one after another on the context
 * stream.  Order: photons, build, eye, gather; they do not advance the pass
 * counter (steps are explicit, as in createPhotonMap / createHitpoints). */
int sppm_pass_photons(sppm_ctx* ctx, int step, int32_t* emitted);
int sppm_pass_build(sppm_ctx* ctx, int32_t* stored);
int sppm_pass_eye(sppm_ctx* ctx, int step);
int sppm_pass_gather(sppm_ctx* ctx);
/* Synchronise the context's stream and its side streams. */
int sppm_sync(sppm_ctx* ctx);

/* reconstructImageFromSharedPoints: rgba = W*H*4 floats. total_photons and
 * passes as the reference passes TotalPhoton and numstep; total_photons is
 * 64-bit here (the reference's int overflows beyond 2^31 emitted photons)
 * and is converted to float as the reference converts its int. */
int sppm_reconstruct(sppm_ctx* ctx, int64_t total_photons, int32_t passes, float* rgba);

/* ---- debug readback (parity tests) ---------------------------------------- */
/* Flattened photon list of the last build, in reference order:
 * pos/flux/indir 3 floats each, index (IOR) and pdf 1 float each. */
int sppm_read_photons(sppm_ctx* ctx, int32_t cap, float* pos, float* flux, float* indir,
                      float* index, float* pdf, int32_t* count);
/* Hit points of the last eye pass (index == pixel): type (0 NONE, 1 HIT,
 * 2 ONLY_DIRECTLIGHT), facet id (-1 if none), position, weight, out dir,
 * normal (face-forwarded), 3 floats each. */
int sppm_read_hitpoints(sppm_ctx* ctx, int32_t* type, int32_t* facet, float* pos,
                        float* weight, float* outdir, float* normal);
/* Shared hit points: radiance 3f, radius, num_photon (int N), num_gather. */
int sppm_read_shared(sppm_ctx* ctx, float* radiance, float* radius, int32_t* num_photon,
                     int32_t* num_gather);
/* DirectLight of the last eye pass's hit points (3 floats each; 0 without an
 * environment light), and SharedHitPoint::DirectLight / NumDirectLight
 * (SPPM.h:112-126, accumulated as SPPM.cpp:415-418).  Any may be NULL. */
int sppm_read_direct(sppm_ctx* ctx, float* hit_direct, float* shared_direct, int32_t* num_direct);
/* Flattened photon count of each path of the last photon pass (N entries):
 * surface photons stored + the stale slot the reference keeps, or 1 for the
 * emission photon (SPPM.cpp:530-540, 672, 812-822). */
int sppm_read_path_counts(sppm_ctx* ctx, int32_t* counts);
/* The photon grid of the last build, the gather structure that takes the place
 * of kdtree<Photon>::CreateTree (R/kdtree.h:417-431); the reference keeps its
 * tree in host memory and has no read-back of it.  dims = {nx, ny, nz,
 * ncells, nphotons}, geom = {bmin x, y, z, cell edge}.  cell_start (ncells + 1
 * entries, when not NULL): the photons of cell c = (z * ny + y) * nx + x sit at
 * grid positions [cell_start[c], cell_start[c + 1]).  order (nphotons entries,
 * when not NULL): the flattened photon index at each grid position (inside a
 * cell in increasing order).  Capacities too small: SPPM_E_ARG. */
int sppm_read_grid(sppm_ctx* ctx, int32_t* dims, float* geom, int64_t cell_cap, int32_t* cell_start,
                   int64_t photon_cap, int32_t* order);
/* Per-hit-point in-radius photon count M of the last gather. */
int sppm_read_counts(sppm_ctx* ctx, int32_t* m);
/* Neighbour lists of the last gather (requires params.record_neighbours):
 * CSR offsets (W*H+1) and flattened photon indices, each list sorted ascending. */
int sppm_read_neighbours(sppm_ctx* ctx, int64_t cap, int64_t* offsets, int32_t* idx,
                         int64_t* total);

/* ---- closest hit (N_BVH::ClosestHit) on a batch of rays ------------------- */
/* org/dir: n*3 floats (dir used as given), tmin/tmax per batch.  Outputs
 * facet id (-1 miss), t, u, v.  Ties in t go to the lowest facet id. */
int sppm_trace_closest(sppm_ctx* ctx, int32_t n, const float* org, const float* dir,
                       float tmin, float tmax, int32_t* facet, float* t, float* u, float* v);
/* The same query without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (serving loops, code under HIP
 * graph capture, harnesses that forbid hipMalloc).  sppm_trace_closest
 * allocates device staging for the rays and hits on every call; this entry
 * point stages them in one fixed device array per GPU instead (2.5 MiB,
 * SPPM_TRACE_STATIC_MAX_RAYS rays, held until the process ends), and calls
 * no allocating routine.  More rays than that: SPPM_E_ARG.  All calls on one
 * GPU share that GPU's array: calls that overlap on one GPU (from several
 * threads, or contexts on other host threads) race. */
#define SPPM_TRACE_STATIC_MAX_RAYS 65536
int sppm_trace_closest_static(sppm_ctx* ctx, int32_t n, const float* org, const float* dir, float tmin,
                              float tmax, int32_t* facet, float* t, float* u, float* v);
/* Scene::IsOccluded (Scene.cpp:246-281) for a batch of rays: whether a
 * surface lies in [tmin, tmax[i]) along ray i, passing through surfaces whose
 * albedo texture alpha gives 1 - alpha > translucency (the call site
 * SPPM.cpp:205 passes 0.5 and a per-ray light distance).  occluded[i] = 0 / 1;
 * SPPM_E_OVERFLOW past 256 translucent layers on one ray. */
int sppm_trace_occluded(sppm_ctx* ctx, int32_t n, const float* org, const float* dir, float tmin,
                        const float* tmax, float translucency, int32_t* occluded);

/* Cumulative counters since sppm_begin: valid hit points gathered (type HIT),
 * in-radius photons summed over hit points (sum of M), photons stored
 * (flattened list lengths), candidate photons the gather tested (the
 * photons of the grid cells its query boxes cover; 0 in record mode),
 * passes.  Any pointer may be NULL. */
/* Closest-hit queries traced since sppm_begin: by the photon passes
 * (Scene::GetOneFacetFromRay calls of createPhotonMap, SPPM.cpp:555) and by
 * the eye passes (getHitpointInfo, SPPM.cpp:125, plus shadow-ray layers when
 * eye_nee is on). */
int sppm_ray_counts(sppm_ctx* ctx, int64_t* photon_rays, int64_t* eye_rays);
/* Of those, the closest hits where another facet is hit at exactly the same
 * distance.  On them the reference keeps the last triangle its traversal
 * tested (N_BVH::findClosestElement's `best >= t`, bvh8.h:591), which depends
 * on its tree; this core keeps the lowest facet id, so these are the only
 * rays where the two rules can pick different facets. */
int sppm_tie_counts(sppm_ctx* ctx, int64_t* photon_ties, int64_t* eye_ties);
int sppm_counters(sppm_ctx* ctx, int64_t* hit_points, int64_t* neighbours, int64_t* photons_stored,
                  int64_t* candidates, int64_t* passes);

/* Closest-hit structure of the uploaded scene: facets, area lights, BVH
 * width, wide nodes, depth in wide levels, traversal stack entries per lane
 * (LDS part) and the global-memory stack entries per lane behind it. */
typedef struct sppm_scene_info {
    int32_t nfacets, nlights, bvh_width, bvh_nodes, bvh_depth, stack_lds, stack_global;
} sppm_scene_info;
int sppm_scene_stats(sppm_ctx* ctx, sppm_scene_info* info);

/* ---- learned photon-emission proposal --------------------------------------
 * No counterpart in the reference: it replaces the position x direction
 * sampling of generateLightRay (SPPM.cpp:453-471) and the PDF the photon
 * carries (SPPM.cpp:527).  Photons are drawn from the mixture
 * q = w * plain + (1 - w) * proposal and carry PDF = q, so the estimate stays
 * consistent.  The proposal is a softmax over 32x32 light (u, v) cells x 16x16
 * hemi-octahedral direction cells whose logits come from an MLP of Fourier
 * features of the cell centre (64 -> hidden x layers -> 1, ReLU), trained on
 * the GPU (MFMA in bf16 or fp16 with float32 accumulation, Adam)
 * from the photons' deposits on visible hit points: per cell, the sum over the
 * photons emitted from it of (hit points that gather the photon) x (photon
 * luminance / PDF).  The photon pass of a pass samples the proposal trained
 * two passes before (published by the previous pass's build), so it can run
 * beside the previous pass's gather and training.
 * Scenes: exactly one untextured rectangular light made of two facets. */
#define SPPM_GUIDE_CELLS 262144
#define SPPM_GUIDE_NPARAMS 16705 /* parameters of the default network (hidden 64, 4 layers) */
typedef struct sppm_guide_params {
    int32_t enable;        /* 0: plain emission */
    float mix_plain;       /* w, probability of the plain emission, in [0.25, 1] */
    float learning_rate;   /* Adam step size */
    int32_t train_steps;   /* Adam steps after each pass's gather */
    uint32_t seed;         /* weight initialisation */
    float history_decay;   /* training histogram = decay * previous + this pass's deposits, in [0, 1) */
    int32_t hidden;        /* units per hidden layer: 64 or 128 (0: 64) */
    int32_t layers;        /* hidden layers: 4 or 6 (0: 4) */
    int32_t fp16;          /* 1: fp16 MFMA, 0: bf16 */
    int32_t train_interval; /* train after every k-th pass (0, 1: every pass); the deposits of the
                               passes between gather into the histogram.  Captured passes need 0 or 1. */
    int32_t target;        /* SPPM_GUIDE_TARGET_*: what the proposal is trained toward */
} sppm_guide_params;
/* Training targets of the proposal (the cross-entropy's weight per cell).
 *   SPPM_GUIDE_TARGET_MEAN  the histogram of the paths' deposits (hit points
 *                           reached x luminance / PDF): proportional to
 *                           p_c E[c | cell], p_c the plain emission's cell
 *                           probability and c a path's contribution;
 *   SPPM_GUIDE_TARGET_RMS   sqrt(p_c S_c), S_c the histogram of deposit^2 x
 *                           q / p (an estimate of p_c E_p[c^2 | cell]): the
 *                           cell proposal that minimises the per-path variance
 *                           (for emission inside a cell as the plain one). */
#define SPPM_GUIDE_TARGET_MEAN 0
#define SPPM_GUIDE_TARGET_RMS 1
/* Parameter count: 64*hidden + hidden + (layers-1)*(hidden*hidden + hidden) + hidden + 1. */
/* Enable (after sppm_scene_upload) or disable guided emission; re-initialises the weights. */
int sppm_guide_set(sppm_ctx* ctx, const sppm_guide_params* params);
/* Read (out) and/or replace (in, resets Adam) the float32 weights
 * (W1[in][out], b1, ..., W_layers, b_layers, w_out, b_out). */
int sppm_guide_weights(sppm_ctx* ctx, float* out, const float* in);
/* Loss and gradient of the current weights for a deposit histogram
 * (SPPM_GUIDE_CELLS floats; NULL: the last pass's histogram). */
int sppm_guide_gradient(sppm_ctx* ctx, const float* hist, float* loss, float* grad);
/* `steps` Adam steps on a histogram (NULL: the last pass's), then the proposal is rebuilt. */
int sppm_guide_train(sppm_ctx* ctx, const float* hist, int steps);
/* Current proposal probabilities per cell and the training histogram (either may be NULL). */
int sppm_guide_read(sppm_ctx* ctx, float* prob, float* hist);
/* The tables the next guided photon pass samples: the proposal per cell
 * (unnormalised) and its inclusive prefix sum (either may be NULL). */
int sppm_guide_tables(sppm_ctx* ctx, float* prob, float* cdf);
/* The emission cell (index into the 262,144 cells) of every photon path of
 * the last guided pass, whichever mixture component drew it: N entries. */
int sppm_guide_path_cells(sppm_ctx* ctx, int32_t* cells);

/* ---- graph-captured passes, batches of scenes, autotuner --------------------
 * No counterpart in the reference (it renders one scene per process,
 * main.cpp:117-167).  A captured pass replays every kernel of sppm_pass from
 * graphs of its steps (photon pass, eye pass, camera words, build, gather),
 * launched on the same streams with the same events as an eager pass, so a
 * replay's photon pass runs beside the previous gather; a batch captures one
 * pass of several contexts (each set up with its own scene, camera and params
 * and begun) into one graph whose branches run concurrently.  Capture right after sppm_begin; needs the
 * UniformSample camera stream and record_neighbours 0.  After capture,
 * sppm_pass replays the graph and the split pass steps are refused;
 * sppm_begin drops the graph. */
int sppm_graph_capture(sppm_ctx* ctx);
typedef struct sppm_batch sppm_batch;
int sppm_batch_create(sppm_batch** out, int device);
int sppm_batch_add(sppm_batch* batch, sppm_ctx* ctx);
/* How sppm_batch_capture lays out the batch's pass (before capture):
 *   SPPM_BATCH_MULTI     one launch per phase for all contexts: the photon
 *                        passes, the camera words, the eye passes and the
 *                        gathers' chunk kernels of every scene run as one
 *                        persistent launch each (per BVH width / guided / NEE
 *                        group), the per-scene setup kernels as graph branches;
 *   SPPM_BATCH_BRANCHES  each context's pass as its own graph branch (the
 *                        phases of different scenes overlap);
 *   SPPM_BATCH_AUTO      (default) MULTI when the contexts average fewer than
 *                        SPPM_BATCH_AUTO_PHOTONS photons per pass (many small
 *                        passes, whose launch tails the shared launches
 *                        remove), else BRANCHES (passes that fill the device
 *                        by themselves overlap better).
 * Every mode renders each scene exactly as its own passes would. */
#define SPPM_BATCH_MULTI 0
#define SPPM_BATCH_BRANCHES 1
#define SPPM_BATCH_AUTO 2
#define SPPM_BATCH_AUTO_PHOTONS 2000000
int sppm_batch_mode(sppm_batch* batch, int mode);
/* the mode in use (after sppm_batch_capture: the one SPPM_BATCH_AUTO chose) */
int sppm_batch_get_mode(sppm_batch* batch, int32_t* mode);
int sppm_batch_capture(sppm_batch* batch);
/* one pass of every context in the batch (one graph launch; asynchronous) */
int sppm_batch_pass(sppm_batch* batch);
int sppm_batch_sync(sppm_batch* batch);
/* destroys the batch (not its contexts) */
int sppm_batch_destroy(sppm_batch* batch);

/* Autotuner: times `passes` passes (after one warm-up pass) of each
 * configuration in stages and keeps the fastest of each stage (the context
 * is left begun with the result):
 *   1. closest-hit structure: builder {binned SAH, LBVH} x BVH width {2, 4, 8};
 *   2. path-kernel grid {device-filling, 3/4, 1/2} (blocks of 64 lanes);
 *   3. photon kernel {megakernel, wavefront}.
 * None of these changes a pass's result.  Then it times photons/s at half,
 * the same and twice photons_per_pass with the chosen configuration and
 * reports the best as photons_suggested, without applying it (the photon
 * count per pass is part of the estimate).  Call after sppm_camera_set and
 * sppm_params_set; `scene` is the scene to (re)upload. */
typedef struct sppm_tune_result {
    int32_t bvh_width;     /* chosen BVH width */
    int32_t grid_blocks;   /* chosen path-kernel grid (blocks of 64 lanes) */
    float ms_per_pass;     /* chosen configuration */
    float trial_ms[6];     /* stage 1: width 2, 4, 8 with SAH, then with LBVH */
    int32_t bvh_builder;   /* chosen builder: 0 binned SAH, SPPM_BVH_LBVH */
    int32_t photon_kernel; /* chosen SPPM_PHOTON_* */
    float grid_ms[3];      /* stage 2: grid full, 3/4, 1/2 */
    float kernel_ms[2];    /* stage 3: megakernel, wavefront */
    float photons_per_s[3];     /* half, same, twice photons_per_pass */
    int32_t photons_suggested;  /* the fastest of those photon counts per pass */
} sppm_tune_result;
int sppm_autotune(sppm_ctx* ctx, const sppm_scene_desc* scene, int passes, sppm_tune_result* out);

/* ---- timing helpers for the benchmark -------------------------------------- */
/* Average duration (ms) of the named pass phase over the `sppm_pass` and
 * split-step calls since sppm_begin or the last reset, measured with
 * hipEvents recorded around the phase's launches on the stream it runs on
 * (in sppm_pass, phases on different streams overlap).  name: "photons", "build",
 * "eye", "gather", "guide" (training, guided emission only); "reset" clears
 * the averages. */
int sppm_kernel_ms(sppm_ctx* ctx, const char* name, float* ms);

/* ---- checkpoint / resume ------------------------------------------------------
 * No counterpart in the reference, whose only persistent state between passes
 * is SharedHitPoint (SPPM.h:112-126) and the photon slot array
 * (SPPM.cpp:753), never saved.  A checkpoint holds what a pass carries to the
 * next: photon slots, shared hit points, radius statistic, counters, camera
 * stream state, pass counter and, when guided, the proposal's weights, Adam
 * state and histogram.  Load it into a context set up the same way (scene,
 * camera, parameters, guided emission) and begun; the render continues as if
 * never interrupted.  The caller keeps the emitted-photon total for
 * sppm_reconstruct (passes x photons per pass). */
int sppm_checkpoint_size(sppm_ctx* ctx, int64_t* bytes);
int sppm_checkpoint_save(sppm_ctx* ctx, void* buf, int64_t bytes);
int sppm_checkpoint_load(sppm_ctx* ctx, const void* buf, int64_t bytes);

/* ---- host utility ----------------------------------------------------------- */
/* mt19937 jump-ahead on the host (no counterpart in the reference, whose
 * UniformSample stream, CameraSampler.h:19-27, is consumed sequentially; the
 * context splits that stream the same way).  `state` is the window of the
 * 624 most recent raw (untempered) words, oldest first -- for a freshly
 * seeded std::mt19937 the 624 seed words; the next output is the tempered
 * x[624] = x[397] ^ mix(x[0], x[1]).  `out` receives the window `steps` words
 * further on (the low 31 bits of out[0] may differ from sequential
 * generation; no later output reads them).  No GPU involved. */
int sppm_mt19937_jump(const uint32_t* state, uint64_t steps, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* SPPM_AMD_H */
