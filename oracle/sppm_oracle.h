/*
 * sppm_oracle.h -- CPU restatement of the FriendsRender R SPPM pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function restates the reference algorithm (file:line under
 * SPPM/SPPM/ of tabochans/FriendsRender-R-public) with the same float/double
 * operation order, so that integer outputs (closest-hit facet ids, photon
 * lists, neighbour sets, counts) are bit-exact with a strict-FP build of the
 * reference (no FMA contraction, IEEE division/sqrt).
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   RNG (mt19937 + generate_canonical) .. C++ KAT + reference MT.h built in oracle/_ref
 *   camera sampler stream order ......... reference CameraSampler.h built in oracle/_ref
 *   alias table (TableSampler) .......... reference TableSampler.h built in oracle/_ref
 *   Hammersley / Halton ................. reference QRandSequence.h built in oracle/_ref
 *   bilinear texture fetch .............. reference TextureSampler.h built in oracle/_ref
 *   Matrix3 inverse (bump TBN) .......... reference Matrix.h/.cpp built in oracle/_ref
 *   closest hit ......................... brute force == BVH, analytic hits
 *   BSDFs ............................... closed forms, MC pdf normalisation,
 *                                         reflection/refraction laws
 *   camera .............................. closed form (central ray, pinhole geometry)
 *   thin-lens diaphragm ................. one-open-pixel images land on that pixel's
 *                                         lens UV; pick frequencies == 1 - red;
 *                                         open images == plain lens sampling
 *   neighbour sets / update / reconstruct  brute force + hand-built recurrences
 *   BiMetal / DiffuseTransmit / GlossTransmit  closed forms, density normalisation,
 *                                         mirror symmetry, agreement with Metal/Roughconductor
 *   environment light (SPHERE map, 0.4c^2)  closed-form directions, conversion, DL sums
 *   IsOccluded .......................... rays whose answer follows from a scene's layout
 *   eye-path NEE / MIS .................. closed form (camera filled by the light) and
 *                                         equality of means with plain BSDF sampling
 *                                         (eye_nee = 3, an oracle-only checker mode)
 *   draw order (which draw feeds which variable, per routine and along the
 *   eye and photon paths) ............... scripted draws with closed-form
 *                                         expectations (orc_script_set)
 *   GGX dielectric ....................... PDF mass over the sphere = the
 *                                         microfacet mass it covers, sampled
 *                                         density == PDF(), BRDF closed forms
 *                                         at normal incidence
 *   guided emission ..................... mixture density integrates to 1,
 *                                         sample frequencies == density,
 *                                         unbiased against plain emission
 */
#ifndef SPPM_ORACLE_H
#define SPPM_ORACLE_H

#include <stdint.h>
#include "../include/sppm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: RandomMT (MT.h:7-44) ---------------------------------------- */
typedef struct orc_mt { uint32_t mt[624]; int idx; } orc_mt;
void orc_mt_seed(orc_mt* m, uint32_t seed);
uint32_t orc_mt_next(orc_mt* m);
double orc_mt_real1(orc_mt* m);
/* n successive genrand64_real1() of RandomMT(seed) */
void orc_mt_real1_seq(uint32_t seed, int n, double* out);
/* n successive raw 32-bit outputs of std::mt19937(seed) */
void orc_mt_raw_seq(uint32_t seed, int n, uint32_t* out);
/* UniformSample(ID)::Sample() x n calls (CameraSampler.h:25-27), 3 floats per call */
void orc_uniform_sampler_seq(int id, int ncalls, float* out);
/* QRandSequence::HaltonCore<B>::Sample(n) and Hammersley(N).Sample(i) */
float orc_halton(int base, int n);
void orc_hammersley(int nseq, int i, float* out3);

/* ---- alias table: TableSampler<T>::Init (TableSampler.h:44-120) ---------- */
/* values f(t_i) for n elements; outputs prob[n], left[n], right[n], *avg. */
void orc_table_build(const float* values, int n, float* prob, int32_t* left, int32_t* right,
                     float* avg);
/* TableSampler::Sample(RandomMT&) for `ndraw` samples from RandomMT(seed). */
void orc_table_sample_seq(const float* prob, const int32_t* left, const int32_t* right, int n,
                          uint32_t seed, int ndraw, int32_t* out);

/* ---- textures (TextureSampler.h:60-79, TextureUtility.cpp:161-164) ------- */
void orc_bilinear(const sppm_texture* tex, float u, float v, float* out4);
/* Matrix3 inverse through cofactors (Matrix.h:180-211); m, out row-major 9 */
void orc_matrix3_inv(const float* m, float* out);

/* ---- BSDFs (material.h, Sampling.cpp) ------------------------------------ */
typedef struct orc_fparam {
    float in[3], out[3], n[3];
    float from_index, to_index, roughness, absorption, reflectance;
} orc_fparam;
float orc_bsdf_brdf(int kind, const orc_fparam* p);
float orc_bsdf_pdf(int kind, const orc_fparam* p);
/* Draws from RandomMT(seed) ; returns direction and *pdf; *ndraw = doubles used */
void orc_bsdf_sample(int kind, const orc_fparam* p, uint32_t seed, float* dir, float* pdf,
                     int* ndraw);

/* PDF() and BRDF() at n outgoing directions (either output may be NULL), and
 * n samples drawn from RandomMT(seed0 + j), for Monte Carlo checks */
void orc_bsdf_eval_batch(int kind, const orc_fparam* p, int n, const float* outdirs, float* pdf, float* brdf);
void orc_bsdf_sample_batch(int kind, const orc_fparam* p, uint32_t seed0, int n, float* dirs, float* pdf);

/* ---- scene ------------------------------------------------------------- */
typedef struct orc_scene orc_scene;
orc_scene* orc_scene_create(const sppm_scene_desc* d);
void orc_scene_destroy(orc_scene* s);
int orc_scene_num_lights(const orc_scene* s);
/* light pick table and per-light data, for tests */
void orc_scene_light_info(const orc_scene* s, int32_t* facet, float* flux3, int32_t* textured,
                          float* prob, int32_t* left, int32_t* right, float* avg);
/* closest hit, brute force over all facets (mode 0) or BVH (mode 1).
 * Returns facet id or -1; min t then min facet id. */
int orc_closest_hit(const orc_scene* s, int mode, const float* org, const float* dir, float tmin,
                    float tmax, float* t, float* u, float* v);
void orc_closest_hit_batch(const orc_scene* s, int mode, int n, const float* org, const float* dir,
                           float tmin, float tmax, int32_t* facet, float* t, float* u, float* v);
/* Scene::IsOccluded per ray: 0 / 1 (2 if more than 4096 translucent layers) */
void orc_occluded_batch(const orc_scene* s, int n, const float* org, const float* dir, float tmin,
                        const float* tmax, float transl, int32_t* occluded);

/* ---- camera ------------------------------------------------------------ */
/* One SampleCameraRay with a given sampler value (3 floats) -> org, dir, pdf */
void orc_camera_ray(const sppm_camera_desc* c, float img_u, float img_v, const float* sample3,
                    float* org, float* dir, float* pdf);
/* The same with ThinLensCamera's diaphragm image (dw x dh RGBA, NULL: none);
 * -1 when the image has no open pixel */
int orc_camera_ray_diaphragm(const sppm_camera_desc* c, int dw, int dh, const float* drgba, float img_u,
                             float img_v, const float* sample3, float* org, float* dir, float* pdf);

/* ---- the SPPM pass ------------------------------------------------------ */
typedef struct orc_render orc_render;
orc_render* orc_render_create(const orc_scene* s, const sppm_camera_desc* cam,
                              const sppm_params* p, int nthreads);
void orc_render_destroy(orc_render* r);
/* ThinLensCamera's diaphragm image for the render's camera (NULL: none) */
int orc_render_diaphragm(orc_render* r, int w, int h, const float* rgba);
/* createPhotonMap(step): returns processed photons */
int orc_photon_pass(orc_render* r, int step);
/* flatten slots (SPPM.cpp:812-822); returns stored count */
int orc_flatten(orc_render* r);
void orc_read_photons(const orc_render* r, float* pos, float* flux, float* indir, float* index,
                      float* pdf);
/* createHitpoints(step) */
void orc_eye_pass(orc_render* r, int step);
void orc_read_hitpoints(const orc_render* r, int32_t* type, int32_t* facet, float* pos,
                        float* weight, float* outdir, float* normal);
/* applyHitPointInfo; record=1 keeps neighbour lists */
void orc_gather(orc_render* r, int record);
void orc_read_counts(const orc_render* r, int32_t* m);
int64_t orc_neighbour_total(const orc_render* r);
void orc_read_neighbours(const orc_render* r, int64_t* offsets, int32_t* idx);
/* IBL::GetIBLColorFromDirection of the scene's converted environment map (0 without one). */
void orc_ibl_color(const orc_scene* s, const float* dir, float* out3);
/* IBL::GetUVFromDirection, SPHERE mapping: (u, v) of a direction. */
void orc_ibl_uv(const float* dir, float* uv);
/* DirectLight of the hit points and the shared hit points (IBL escapes). */
void orc_read_direct(const orc_render* r, float* hit_direct, float* shared_direct, int32_t* num_direct);
void orc_read_shared(const orc_render* r, float* radiance, float* radius, int32_t* num_photon,
                     int32_t* num_gather);
/* set shared state (for hand-built update tests) */
void orc_write_shared(orc_render* r, const float* radiance, const float* radius,
                      const int32_t* num_photon, const int32_t* num_gather);
void orc_reconstruct(const orc_render* r, int32_t total_photons, int32_t passes, float* rgba);
/* one full pass in reference order (SPPM.cpp:771-841); pass = 0-based index */
int orc_full_pass(orc_render* r, int pass, int record);

/* ---- spot checks at full size (no full-pass run needed) ------------------- */
/* UniformSample(id) draws number calls[j] (0-based, ascending) -> out 3 floats each */
void orc_uniform_sampler_at(int id, int n, const int64_t* calls, float* out);
/* createHitpoints for one pixel of pass `step` with the given camera sample:
 * type, facet, pos[3], weight[3], normal[3], outdir[3] */
void orc_eye_pixel(const orc_render* r, int step, int pixel, const float* sample3, int32_t* type,
                   int32_t* facet, float* pos, float* weight, float* normal, float* outdir);
/* createPhotonMap for one path of pass `step`, on a private slot array that
 * starts invalid: returns the number of surface photons stored (k) and writes
 * the path's new photons (emission photon when k == 0) */
int orc_photon_path(const orc_render* r, int step, int path, float* pos, float* flux, float* indir,
                    float* index, float* pdf, int32_t* nlist);
/* flattened photon count of each path of the last photon pass (N entries) */
void orc_read_path_counts(const orc_render* r, int32_t* counts);

/* ---- scripted draws (pinning tests) ---------------------------------------
 * While installed on the calling thread, every RandomMT draw returns the next
 * of `n` scripted values (0.5 past the end); n < 0 removes the script.
 * orc_script_used() = draws taken since it was installed. */
void orc_script_set(const double* draws, int n);
int orc_script_used(void);

/* ---- guided emission (no reference counterpart; replaces generateLightRay's
 * sampling, SPPM.cpp:453-471, and the photon PDF, SPPM.cpp:527) ------------ */
typedef struct orc_guide orc_guide;
/* NULL unless the scene's only light is one rectangle of two facets */
orc_guide* orc_guide_create(const orc_scene* s, float mix_plain);
void orc_guide_destroy(orc_guide* g);
/* proposal over 262,144 cells (unnormalised) and its inclusive prefix sum */
void orc_guide_set_tables(orc_guide* g, const float* prob, const float* cdf);
/* photon passes of r emit from the mixture (g NULL: plain emission) */
void orc_render_set_guide(orc_render* r, const orc_guide* g);
/* P0, e1, e2, n, t1, t2 of the light frame (18 floats) */
void orc_guide_frame(const orc_guide* g, float* out);
/* textured light: plain mass per position cell (1024), CSR starts (1025), cell per table point (2000) */
void orc_guide_point_cells(const orc_guide* g, float* mass, int32_t* start, int32_t* pt_cell);
/* mixture density at (point, dir); textured lights: (light, point) is the table point, pos ignored */
float orc_guide_density(const orc_scene* s, const orc_guide* g, int light, int point, const float* pos,
                        const float* dir, int32_t* cell);
/* the start of guided path `path` of pass `step`: slot-0 position, direction, PDF, cell */
void orc_guide_emit(const orc_render* r, int step, int path, float* pos, float* dir, float* pdf, int32_t* cell);
/* batches of the two (light / point may be NULL for an untextured light; cell may be NULL) */
void orc_guide_density_batch(const orc_scene* s, const orc_guide* g, int n, const int32_t* light,
                             const int32_t* point, const float* pos, const float* dir, float* q);
void orc_guide_emit_batch(const orc_render* r, int step, int first, int n, float* pos, float* dir, float* pdf,
                          int32_t* cell);

#ifdef __cplusplus
}
#endif
#endif
