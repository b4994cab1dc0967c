// sppm_pass.h -- launch shapes and launchers of the per-pass kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_guide.h"
#include "sppm_layout.h"

namespace sppm {

constexpr int PATH_BLOCK = 64;     // lanes per block of the path kernels (1 wave)
// Waves per SIMD the path kernels are compiled for: caps them at 128 VGPRs
// (they would take 134-175 and run 2-3 waves per SIMD); traversal is latency
// bound, and the fourth wave is worth more than the spared registers.
constexpr int PATH_WAVES_PER_SIMD = 4;
constexpr int GATHER_WAVES = 4096; // persistent waves of the chunked gather
constexpr int GATHER_CHUNK_CAP = 65536;  // chunk budget per pass (partial sums buffer)

// Grid parameters, computed on the device each pass from the largest radius.
struct GridParams {
    int nx, ny, nz, ncells;
    float cell, inv_cell;
    float bmin[3];
    int nphotons;
};

// Cumulative device counters (since sppm_begin).
struct Counters {
    unsigned long long hit_points;
    unsigned long long neighbours;
    unsigned long long stored;
    unsigned long long candidates;  // photons the balanced gather tested (union boxes of its tiles)
    unsigned long long photon_rays; // closest-hit queries of the photon pass
    unsigned long long eye_rays;    // closest-hit queries of the eye pass (shadow-ray layers included)
    // closest hits with another facet at exactly the same t (the rule "lowest
    // facet id" and the reference's "last tested" may differ on them); each
    // sits two words after its ray counter
    unsigned long long photon_ties;
    unsigned long long eye_ties;
};

// Camera words of one pass: K segments of L words (K * L = 6 * pixels);
// states holds K windows of 624 raw words (then K more as scratch), phi the
// jump polynomial of (K-1)L.
void launch_camera_words(uint32_t* states, const uint32_t* phi, int K, long long L, uint32_t* words, hipStream_t st);
// The camera words of several streams in two launches: job s has K segments,
// blocks kstart[s] .. kstart[s + 1] - 1 (nseg = kstart[njobs] in all).
struct CamJob {
    uint32_t* states;
    const uint32_t* phi;
    int K;
    long long L;
    uint32_t* words;
};
void launch_camera_words_multi(const CamJob* jobs, const int* kstart, int njobs, int nseg, hipStream_t st);
// Traversal stack spill area of one path kernel: blocks * 64 lanes,
// cap - TRACE_LDS entries each.  Kernels that may run at the same time (the
// photon and eye passes) take separate areas.
struct TraceSpill {
    uint32_t* buf;
    int blocks;
    int cap;
};
int trace_grid_blocks(int width);
struct GuideEmit;
// step: the pass step (SPPM.cpp:798), or, when step_dev is given, *step_dev + 1
// rays: a 64-bit counter the launch adds its closest-hit queries to; the
// exact-tie count goes to rays[2]
void launch_photon(const SceneDev& S, const SlotsDev& slots, int N, int step, const int* step_dev,
                   const TraceSpill& sp, int* next_path, const GuideEmit* guide, uint32_t* flags,
                   unsigned long long* rays, hipStream_t st);
void launch_add_i32(int* p, int v, hipStream_t st);
void launch_zero_i32(int* p, int n, hipStream_t st);
// pdep: per flattened photon, the luminance its gathering hit points received (fixed point, guide_fix)
void launch_deposit_clear(unsigned long long* pdep, const int* nstored, hipStream_t st);
// dst[0 .. N-1] = off, dst[N] = *nstored
void launch_copy_offsets(const int* off, const int* nstored, int N, int* dst, hipStream_t st);
// step: the pass step, or *step_dev + 1 when step_dev is given (graph replays)
// ratio (optional): q / p_plain per path; with it the deposit is d^2 ratio
void launch_guide_deposit(const int* off, const int* nstored, const unsigned long long* pdep, const int* path_cell,
                          const float* ratio, int N, int step, const int* step_dev, unsigned long long* hist,
                          hipStream_t st);
// nee: SPPM_NEE_* (the compiled-out next-event estimation of getHitpointInfo)
void launch_eye(const SceneDev& S, const CamDev& C, const HitDev& H, int step, const int* step_dev,
                const TraceSpill& sp, const uint32_t* words, uint32_t* flags, unsigned long long* rays, int* queue,
                int nee, hipStream_t st);
// Closest hits in (tmin, tmax) for n rays (t, u, v may be NULL).
void launch_trace(const SceneDev& S, int n, const float* org, const float* dir, float tmin, float tmax,
                  const TraceSpill& sp, int* facet, float* t, float* u, float* v, uint32_t* flags, hipStream_t st);
// Scene::IsOccluded for a batch of rays: occluded[i] = 0 / 1.
void launch_occluded(const SceneDev& S, int n, const float* org, const float* dir, float tmin, const float* tmax,
                     float transl,
                     const TraceSpill& sp, int* occluded, uint32_t* flags, hipStream_t st);
// Sort buffers of the grid build, photon_cap entries each (hist:
// radix_hist_entries(photon_cap)): cell key and flattened index of every
// flattened photon.
struct PhotonSortBuffers {
    uint32_t *keys, *vals, *tmp_k, *tmp_v, *hist;
};
// The flattened photons of a pass in reference order, copied out of the
// slots by the build's first step so that the next photon pass may run while
// the rest of the build reads them: photon j is the 48-byte record
// rec[3j .. 3j+2] = (pos, flux, dir with the flattened index in dir.w), so
// the cell-order copy reads one record per photon instead of three scattered
// vectors (3 * photon_cap float4).
struct StagedPhotons {
    float4* rec;
};
// off: exclusive scan of the slots' per-path counts
void launch_photon_stage(const SlotsDev& slots, const int* off, const StagedPhotons& sp, hipStream_t st);
constexpr int PHOTON_KEY_BITS = 22;  // cell keys are below the cell capacity 2^22
// Photon grid: cell size, keys, stable sort by cell, cell starts
// (cell_cap + 1 entries; gaps: cell_gap_entries(cell_cap); ngap: 1 int) and
// the cell-sorted records.
void launch_build_grid(const StagedPhotons& sp, const int* nstored, Counters* ctr, GridParams* gp,
                       const unsigned long long* rstat, const float* bmin, const float* ext, int cell_cap,
                       int* cell_start, int4* gaps, int* ngap, const PhotonSortBuffers& sb, int photon_cap,
                       float4* gpos, float4* gflux, float4* gdir, hipStream_t st);
size_t cell_gap_entries(int cell_cap);
void launch_photon_flat(const SlotsDev& slots, const int* off, float4* fpos, float4* fflux, float4* fdir,
                        hipStream_t st);
void launch_gather(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int W, int Hh,
                   float alpha, int mode, int* counts, const int* list_off, int* lists, unsigned long long* rstat,
                   Counters* ctr, uint32_t* flags, hipStream_t st);
struct TileRange;
// Sort buffers of the hit-point groups (np entries each; hist as radix_hist_entries(np)).
struct GroupBuffers {
    uint32_t *keys, *vals, *tmp_k, *tmp_v, *hist;
    int* nvalid;  // valid hit points
    int* chunkq;  // striped chunk counters of the gather (STRIPE_INTS)
    unsigned long long* stat;  // 3 per group: radius sum, active hit points, neighbours
};
// Neighbour pairs (hit point, flattened photon index) that the production
// gather appends when recording (parity tests): at most cap, *count may run
// past cap (the overflow is reported by the reader).
struct PairRecord {
    int2* pairs;
    unsigned long long* count;
    long long cap;
};
void launch_gather_balanced(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int np,
                            const float* bmin, const float* ext, const GroupBuffers& gb, float alpha, TileRange* tiles,
                            int cap, int* chunkp, int* nch, int* chunk_off, int* scan_tmp, float4* partial,
                            int* counts, unsigned long long* rstat, Counters* ctr, uint32_t* flags,
                            unsigned long long* pdep,
                            const PairRecord& rec, hipStream_t st);
// The same in three parts: planning (hit-point groups, tiles, chunk lists;
// returns the group order), the chunk kernel, and the per-group sums.
const uint32_t* launch_gather_plan(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp,
                                   int np, const float* bmin, const float* ext, const GroupBuffers& gb,
                                   TileRange* tiles, int cap, int* chunkp, int* nch, int* chunk_off, int* scan_tmp,
                                   int* counts, Counters* ctr, const PairRecord& rec, hipStream_t st);
void launch_gather_run(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int np,
                       const uint32_t* order, const GroupBuffers& gb, const TileRange* tiles, const int* chunkp,
                       const int* chunk_off, float4* partial, unsigned long long* pdep, const PairRecord& rec,
                       hipStream_t st);
void launch_gather_finish(const HitDev& H, const SharedDev& SH, int np, const uint32_t* order, const GroupBuffers& gb,
                          float alpha, const int* chunk_off, const float4* partial, int* counts,
                          unsigned long long* rstat, Counters* ctr, uint32_t* flags, hipStream_t st);
int gather_group_count(int np);
int gather_tile_count(int W, int Hh);
size_t gather_tile_bytes();
void launch_set_u64x2(unsigned long long* p, unsigned long long a, unsigned long long b, hipStream_t st);

// ---- multi-scene launches ----------------------------------------------------
// One launch runs the photon pass, the eye pass or the gather's chunk kernel
// of every scene of a batch: the per-scene arguments of the one-scene
// launches, one entry per scene in device memory.  Waves start on the scene
// that owns their block (bstart: njobs + 1 block offsets, in proportion to
// the scenes' work) and then help the others.  The scenes' work counters
// (next_path, queue) are zeroed before the launch, as for the one-scene
// kernels.  step_dev: the scene's pass counter (the step is *step_dev + 1),
// or NULL for `step`.
struct PhotonJob {
    SceneDev S;
    SlotsDev slots;
    int N, step;
    const int* step_dev;
    int* next_path;
    GuideEmit G;
    uint32_t* flags;
    unsigned long long* rays;
};
struct EyeJob {
    SceneDev S;
    CamDev C;
    HitDev H;
    int step, keep;
    const int* step_dev;
    const uint32_t* words;
    int* queue;
    uint32_t* flags;
    unsigned long long* rays;
};
struct GatherJob {
    HitDev H;
    SharedDev SH;
    GridDev G;
    const GridParams* gp;
    const uint32_t* order;
    const int* nvalid;
    const TileRange* tiles;
    int ntiles;
    const int* chunk_off;
    const int* chunkp;
    int* queue;
    float4* partial;
    unsigned long long* pdep;
    PairRecord rec;
};
// width: the scenes' common BVH width; sp: a spill area for sp.blocks blocks
void launch_photon_multi(int width, bool guided, const PhotonJob* jobs, int njobs, const int* bstart,
                         const TraceSpill& sp, hipStream_t st);
void launch_eye_multi(int width, bool nee, const EyeJob* jobs, int njobs, const int* bstart, const TraceSpill& sp,
                      hipStream_t st);
void launch_gather_chunks_multi(const GatherJob* jobs, int njobs, const int* bstart, hipStream_t st);
int gather_wave_count();
// the group order launch_gather_plan returns for these buffers (no launch)
const uint32_t* gather_order(const GroupBuffers& gb, int np);
// Adds the hit points' DirectLight into the shared hit points (environment light only).
void launch_direct_light(const HitDev& H, const SharedDev& SH, int np, hipStream_t st);
void launch_reconstruct(const SharedDev& SH, int np, long long total, int passes, float4* out, hipStream_t st);
void launch_init_shared(const SharedDev& SH, int np, float r0, hipStream_t st);
void launch_init_slots(const SlotsDev& s, hipStream_t st);
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t st);

}  // namespace sppm
