// sppm_pass.hip -- the per-pass kernels of the SPPM core on gfx950.
//
// One pass (SPPM.cpp:770-843):
//   k_camw_emit/jump shared UniformSample stream words for this pass (K
//                    segments, each one wave, jump-ahead between passes)
//   k_photon<W>      createPhotonMap: persistent, one lane per photon path at
//                    a time, writes slots
//   scan(count)      flatten offsets (SPPM.cpp:812-822 order)
//   k_grid_params    cell size from the mean radius (device-side)
//   k_photon_keys    per flattened photon: cell key, flattened index, slot
//   radix sort       stable, by cell key
//   k_cell_starts    cell starts from the sorted keys (k_fill_gaps: long
//                    runs of empty cells)
//   k_photon_sorted  cell-sorted photon records (kdtree::CreateTree replacement)
//   k_eye<W>         createHitpoints: one lane per pixel
//   k_hit_keys ... k_gather_finish, k_group_stats
//                    applyHitPointInfo + progressive update over Morton
//                    groups of 64 hit points, chunked candidate lists
//   k_gather         the same with neighbour-list recording (tests)
//   k_reconstruct    reconstructImageFromSharedPoints
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_guide.h"
#include "sppm_kernels.h"
#include "sppm_pass.h"
#include "sppm_photon.h"
#include "sppm_stripes.h"
#include "sppm_trace.h"

namespace sppm {

// ---------------------------------------------------------------------------
// Shared camera sampler stream: UniformSample(ID) owns one mt19937 for the
// whole render (CameraSampler.h:19-31); pixel i of a pass draws 3 doubles in
// pixel order, 6 words per pixel.  The block of a pass (6 * pixels words) is
// cut into K segments of L words; segment k's state window (where it
// starts) comes from the host for the first pass.
//   k_camw_emit  one wave per segment emits the segment's tempered words and
//                stores the window after it;
//   k_camw_jump  one workgroup per segment jumps that window over the other
//                K - 1 segments (t^((K-1)L) mod P applied as a sum of shifted
//                windows): the start of its segment in the next pass.
// Raw word x[m], m >= 624, is mix(x[m-624], x[m-623], x[m-227]), so 227
// consecutive new words depend only on earlier ones, and a ring of 1024 raw
// words holds everything either kernel reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t temper(uint32_t z) {
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}
__device__ __forceinline__ uint32_t mt_mix(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}
constexpr int CAMW_RING = 1024;       // raw words kept (>= 624 + 227, and >= 227 + 624 for the jump's reads)
constexpr int CAMW_DEG = 19937;       // degree of the characteristic polynomial
constexpr int CAMW_JUMP_THREADS = 640;  // 624 output words of the jumped window

// One wave per segment: x[624 + q] -> out[q] for q < L, 227 words per step
// in four rounds of 64 lanes.  A step reads only words of earlier steps, and
// the wave's LDS accesses are executed in order, so no barrier is needed.
// segment seg of a stream: states, words and ends as for launch_camera_words
__device__ __forceinline__ void camw_emit(const uint32_t* __restrict__ states, long long L,
                                          uint32_t* __restrict__ words, uint32_t* __restrict__ ends, int seg,
                                          uint32_t* ring) {
    const int lane = threadIdx.x;
    const uint32_t* st = states + (size_t)seg * 624;
    for (int k = lane; k < 624; k += 64) ring[k] = st[k];
    __builtin_amdgcn_wave_barrier();
    uint32_t* out = words + (size_t)seg * (size_t)L;
    const long long mend = 624 + L;
    for (long long m0 = 624; m0 < mend; m0 += 227) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int idx = r * 64 + lane;
            const long long m = m0 + idx;
            if (idx < 227 && m < mend) {
                const uint32_t x = mt_mix(ring[(m - 624) & (CAMW_RING - 1)], ring[(m - 623) & (CAMW_RING - 1)],
                                          ring[(m - 227) & (CAMW_RING - 1)]);
                ring[m & (CAMW_RING - 1)] = x;
                out[m - 624] = temper(x);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    // the window after the segment: x[L .. L + 623]
    uint32_t* e = ends + (size_t)seg * 624;
    for (int j = lane; j < 624; j += 64) e[j] = ring[(L + j) & (CAMW_RING - 1)];
}
__global__ void __launch_bounds__(64) k_camw_emit(const uint32_t* __restrict__ states, long long L,
                                                  uint32_t* __restrict__ words, uint32_t* __restrict__ ends) {
    __shared__ uint32_t ring[CAMW_RING];
    camw_emit(states, L, words, ends, (int)blockIdx.x, ring);
}

// Jumped window: word t = XOR of y[i + t] over the coefficients phi_i = 1,
// y[k] = x[L + k].  The coefficients are taken 227 at a time; before a block
// [i0, i0 + 227) the words up to y[i0 + 226 + 623] are generated into the
// ring (227 at a time, one barrier), whose oldest slot then holds y[i0 - 175]
// or later, so the block's reads y[i0 ..] are intact.
__device__ __forceinline__ void camw_jump(const uint32_t* __restrict__ ends, const uint32_t* __restrict__ phi,
                                          uint32_t* __restrict__ states, int seg, uint32_t* ring) {
    const int t = threadIdx.x;
    const uint32_t* e = ends + (size_t)seg * 624;
    if (t < 624) ring[t] = e[t];  // y[k] at ring[k & 1023]
    __syncthreads();
    uint32_t acc = 0;
    int gen = 623;  // highest y generated
    for (int i0 = 0; i0 < CAMW_DEG; i0 += 227) {
        const int i1 = i0 + 227 < CAMW_DEG ? i0 + 227 : CAMW_DEG;
        const int need = i1 - 1 + 623;
        if (t < need - gen) {
            const int k = gen + 1 + t;
            ring[k & (CAMW_RING - 1)] = mt_mix(ring[(k - 624) & (CAMW_RING - 1)], ring[(k - 623) & (CAMW_RING - 1)],
                                               ring[(k - 227) & (CAMW_RING - 1)]);
        }
        gen = need;
        __syncthreads();
        if (t < 624) {
            for (int w = i0 >> 5; w <= (i1 - 1) >> 5; ++w) {
                uint32_t bits = phi[w];
                const int lo = i0 - w * 32, hi = i1 - w * 32;
                if (lo > 0) bits &= ~((1u << lo) - 1u);
                if (hi < 32) bits &= (1u << hi) - 1u;
                while (bits) {
                    const int b = __builtin_ctz(bits);
                    bits &= bits - 1u;
                    acc ^= ring[(w * 32 + b + t) & (CAMW_RING - 1)];
                }
            }
        }
        __syncthreads();
    }
    if (t < 624) states[(size_t)seg * 624 + t] = acc;
}
__global__ void __launch_bounds__(CAMW_JUMP_THREADS) k_camw_jump(const uint32_t* __restrict__ ends,
                                                                 const uint32_t* __restrict__ phi,
                                                                 uint32_t* __restrict__ states) {
    __shared__ uint32_t ring[CAMW_RING];
    camw_jump(ends, phi, states, (int)blockIdx.x, ring);
}

// The camera words of every scene of a batch: block b is segment b -
// kstart[s] of scene s, kstart[s] <= b < kstart[s + 1].
__device__ __forceinline__ int owner_of(const int* __restrict__ start, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}
__global__ void __launch_bounds__(64) k_camw_emit_multi(const CamJob* __restrict__ jobs,
                                                        const int* __restrict__ kstart, int njobs) {
    __shared__ uint32_t ring[CAMW_RING];
    const int s = owner_of(kstart, njobs, (int)blockIdx.x);
    const CamJob& J = jobs[s];
    camw_emit(J.states, J.L, J.words, J.states + (size_t)J.K * 624, (int)blockIdx.x - kstart[s], ring);
}
__global__ void __launch_bounds__(CAMW_JUMP_THREADS) k_camw_jump_multi(const CamJob* __restrict__ jobs,
                                                                       const int* __restrict__ kstart, int njobs) {
    __shared__ uint32_t ring[CAMW_RING];
    const int s = owner_of(kstart, njobs, (int)blockIdx.x);
    const CamJob& J = jobs[s];
    camw_jump(J.states + (size_t)J.K * 624, J.phi, J.states, (int)blockIdx.x - kstart[s], ring);
}

// Persistent photon pass: each lane traces one path at a time; when at least
// PHOTON_REGEN lanes of a wave have finished their paths, the wave claims the
// next path indices from the striped path counters (one atomic per claim),
// so lanes do not idle behind the longest path of their wave.  A path's result depends
// only on its index, not on the lane or order that traces it.  Every wave
// exits once the counter has passed N and its lanes are done.
// The photon paths [0, N) of one scene, traced by the calling wave until the
// scene's counter is exhausted and the wave's lanes are done.
template <int W, bool GUIDED>
__device__ __forceinline__ void photon_paths(const SceneDev& S, const SlotsDev& slots, int N, int step,
                                             int* __restrict__ next_path, const GuideEmit& G, const TraceStack& ts,
                                             uint32_t& ovf, uint32_t& fl, uint32_t& nray, uint32_t& ntie) {
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    PhotonPath P;
    bool alive = false, exhausted = false;
    StripeCursor cur = stripe_start();
    for (;;) {
        const unsigned long long idle = __ballot(!alive);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= PHOTON_REGEN) {
            int first = 0;
            const int got = stripe_claim(next_path, N, cur, nidle, &first);
            if (got == 0) exhausted = true;
            if (!alive) {
                const int r = __popcll(idle & below);
                if (r < got) {
                    const int i = first + r;
                    if constexpr (GUIDED) photon_start_guided(S, slots, G, N, step, i, P, &ovf);
                    else photon_start(S, slots, N, step, i, P, &ovf);
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (alive && !photon_bounce<W>(S, slots, P, ts, &ovf, &fl, nray, ntie)) {
            photon_finish(slots, P);
            alive = false;
        }
    }
}

template <int W, bool GUIDED>
__global__ void __launch_bounds__(PATH_BLOCK, PATH_WAVES_PER_SIMD) k_photon(SceneDev S, SlotsDev slots, int N, int step_arg,
                                                       const int* __restrict__ step_dev, int cap,
                                                       uint32_t* __restrict__ spill, int* __restrict__ next_path,
                                                       GuideEmit G, uint32_t* __restrict__ flags,
                                                       unsigned long long* __restrict__ rays) {
    // the pass step comes from the device pass counter when the pass is replayed from a graph
    const int step = step_dev ? *step_dev + 1 : step_arg;
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    uint32_t ovf = 0, fl = 0, nray = 0, ntie = 0;
    photon_paths<W, GUIDED>(S, slots, N, step, next_path, G, ts, ovf, fl, nray, ntie);
    if (ovf) fl |= 2u;
    if (fl) atomicOr(flags, fl);
    add_wave_count(rays, nray);
    add_wave_count(rays + 2, ntie);
}

// ---------------------------------------------------------------------------
// Multi-scene launches: one launch runs a phase of the pass for every scene
// of a batch.  Wave w starts on the scene that owns launch block w (blocks
// are dealt to the scenes in proportion to their work, bstart[s] <=
// blockIdx.x < bstart[s + 1]), works there until the scene's counters are
// exhausted, then visits the other scenes in order.  A wave that finds every
// stripe of a scene exhausted raises the scene's done flag (an unused word
// of its counter area, zeroed with it), so later visits cost one load.
// Every item's result depends only on the item, as in the one-scene
// kernels, so a batch renders each scene exactly as its own passes would.
// ---------------------------------------------------------------------------
constexpr int STRIPE_DONE = 1;  // word of a striped counter area that flags it exhausted
__device__ __forceinline__ int home_scene(const int* __restrict__ bstart, int n) {
    return owner_of(bstart, n, (int)blockIdx.x);
}
__device__ __forceinline__ bool scene_done(const int* ctr) {
    return __hip_atomic_load(ctr + STRIPE_DONE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ void scene_mark_done(int* ctr) {
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(ctr + STRIPE_DONE, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int W, bool GUIDED>
__global__ void __launch_bounds__(PATH_BLOCK, PATH_WAVES_PER_SIMD) k_photon_multi(const PhotonJob* __restrict__ jobs,
                                                                                int njobs,
                                                                                const int* __restrict__ bstart, int cap,
                                                                                uint32_t* __restrict__ spill) {
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    const int home = home_scene(bstart, njobs);
    for (int k = 0; k < njobs; ++k) {
        const int s = __builtin_amdgcn_readfirstlane(home + k < njobs ? home + k : home + k - njobs);
        const PhotonJob& J = jobs[s];
        if (scene_done(J.next_path)) continue;
        const int step = J.step_dev ? *J.step_dev + 1 : J.step;
        uint32_t ovf = 0, fl = 0, nray = 0, ntie = 0;
        photon_paths<W, GUIDED>(J.S, J.slots, J.N, step, J.next_path, J.G, ts, ovf, fl, nray, ntie);
        scene_mark_done(J.next_path);
        if (ovf) fl |= 2u;
        if (fl) atomicOr(J.flags, fl);
        add_wave_count(J.rays, nray);
        add_wave_count(J.rays + 2, ntie);
    }
}

// ---------------------------------------------------------------------------
// Grid over the stored photons.  Cell size = largest shared radius, so a
// query box [c - R, c + R] spans at most 3 cells per axis.  Cell coordinates
// are clamped to the scene box; the same monotone map is used for photons
// and query boxes, so no neighbour is ever missed.
// ---------------------------------------------------------------------------
__global__ void k_grid_params(GridParams* gp, const unsigned long long* __restrict__ rstat, float bmin_x, float bmin_y,
                              float bmin_z, float ext_x, float ext_y, float ext_z, int cell_cap,
                              const int* __restrict__ nstored, Counters* __restrict__ ctr, int* __restrict__ ngap) {
    *ngap = 0;
    // cell size = mean radius of the hit points after the previous gather;
    // any cell size gives the same neighbour sets (queries walk every cell
    // their box overlaps), this one keeps ~3 cells per query axis.
    const double mean = rstat[1] ? ((double)rstat[0] / 4294967296.0) / (double)rstat[1] : 1e-3;
    float cell = (float)mean > 1e-5f ? (float)mean : 1e-5f;
    float ext[3] = {ext_x, ext_y, ext_z};
    int n[3];
    for (;;) {
        long long tot = 1;
        for (int k = 0; k < 3; ++k) {
            float c = ext[k] / cell;
            n[k] = (int)fminf(c, 1.0e6f) + 1;
            tot *= n[k];
        }
        if (tot <= cell_cap) break;
        cell *= 1.25f;
    }
    gp->nx = n[0]; gp->ny = n[1]; gp->nz = n[2];
    gp->ncells = n[0] * n[1] * n[2];
    gp->cell = cell;
    gp->inv_cell = 1.0f / cell;
    gp->bmin[0] = bmin_x; gp->bmin[1] = bmin_y; gp->bmin[2] = bmin_z;
    gp->nphotons = *nstored;
    ctr->stored += (unsigned long long)*nstored;
}

__device__ __forceinline__ int cell_coord(float p, float lo, float inv, int n) {
    float c = floorf((p - lo) * inv);
    int ci = (int)fminf(fmaxf(c, 0.0f), (float)(n - 1));
    return ci;
}

__global__ void k_zero_i32(int* __restrict__ p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The pass's flattened photons (SPPM.cpp:812-822) copied out of the slots in
// reference order, each record's flattened index in dir.w: the rest of the
// build reads only these, so the next photon pass may rewrite the slots.
__global__ void __launch_bounds__(256) k_photon_stage(SlotsDev slots, const int* __restrict__ off,
                                                      StagedPhotons sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots.npaths) return;
    const int c = slots.count[i];
    const int base = off[i];
    const size_t NP = (size_t)slots.npaths;
    for (int s = 0; s < c; ++s) {
        const size_t o = (size_t)s * NP + i;
        float4 d = slots.dir[o];
        d.w = __int_as_float(base + s);
        float4* r = sp.rec + 3 * (size_t)(base + s);
        r[0] = slots.pos[o];
        r[1] = slots.flux[o];
        r[2] = d;
    }
}

// Cell key and flattened index of every staged photon.
__global__ void __launch_bounds__(256) k_photon_keys(const float4* __restrict__ rec, const int* __restrict__ nstored,
                                                     const GridParams* __restrict__ gp, PhotonSortBuffers sb) {
    const int n = *nstored;
    const GridParams g = *gp;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float4 p = rec[3 * (size_t)j];
        int cx = cell_coord(p.x, g.bmin[0], g.inv_cell, g.nx);
        int cy = cell_coord(p.y, g.bmin[1], g.inv_cell, g.ny);
        int cz = cell_coord(p.z, g.bmin[2], g.inv_cell, g.nz);
        sb.keys[j] = (uint32_t)((cz * g.ny + cy) * g.nx + cx);
        sb.vals[j] = (uint32_t)j;
    }
}

// Cell starts from the cell-sorted keys: sorted position i starts every cell
// in (keys[i-1], keys[i]] (position n the cells after the last key, up to
// ncells).  Runs of up to CELL_RUN_INLINE cells are written by the position's
// own thread, longer ones are queued for k_fill_gaps.  Every entry of
// cell_start[0..ncells] is written exactly once.
constexpr int CELL_RUN_INLINE = 64;
__global__ void __launch_bounds__(256) k_cell_starts(const uint32_t* __restrict__ keys,
                                                     const int* __restrict__ nstored,
                                                     const GridParams* __restrict__ gp, int* __restrict__ cell_start,
                                                     int4* __restrict__ gaps, int* __restrict__ ngap) {
    const int n = *nstored;
    const int nc = gp->ncells;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        const int a = i > 0 ? (int)keys[i - 1] : -1;
        const int b = i < n ? (int)keys[i] : nc;
        if (b <= a) continue;
        if (b - a <= CELL_RUN_INLINE) {
            for (int c = a + 1; c <= b; ++c) cell_start[c] = i;
        } else {
            const int q = atomicAdd(ngap, 1);
            gaps[q] = make_int4(a + 1, b, i, 0);
        }
    }
}
__global__ void __launch_bounds__(256) k_fill_gaps(const int4* __restrict__ gaps, const int* __restrict__ ngap,
                                                   int* __restrict__ cell_start) {
    const int ng = *ngap;
    for (int g = blockIdx.x; g < ng; g += gridDim.x) {
        const int4 r = gaps[g];
        for (int c = r.x + (int)threadIdx.x; c <= r.y; c += blockDim.x) cell_start[c] = r.z;
    }
}
size_t cell_gap_entries(int cell_cap) { return (size_t)(cell_cap + 1) / CELL_RUN_INLINE + 2; }

// Cell-sorted photon records: after the stable sort by cell key, position j
// holds flattened photon vals[j], so photons of a cell are in flattened
// (reference) order and every run lays them out the same way.  The record
// carries its flattened index in dir.w.
__global__ void __launch_bounds__(256) k_photon_sorted(StagedPhotons sp, const int* __restrict__ nstored,
                                                       const uint32_t* __restrict__ order,
                                                       float4* __restrict__ gpos, float4* __restrict__ gflux,
                                                       float4* __restrict__ gdir) {
    const int n = *nstored;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float4* r = sp.rec + 3 * (size_t)order[j];
        gpos[j] = r[0];
        gflux[j] = r[1];
        gdir[j] = r[2];
    }
}

// Flattened photon list in reference order (debug readback only)
__global__ void k_photon_flat(SlotsDev slots, const int* __restrict__ off, float4* __restrict__ fpos,
                              float4* __restrict__ fflux, float4* __restrict__ fdir) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots.npaths) return;
    const int c = slots.count[i];
    const int base = off[i];
    const size_t NP = (size_t)slots.npaths;
    for (int s = 0; s < c; ++s) {
        const size_t o = (size_t)s * NP + i;
        fpos[base + s] = slots.pos[o];
        fflux[base + s] = slots.flux[o];
        fdir[base + s] = slots.dir[o];
    }
}

// ---------------------------------------------------------------------------
// Camera (Camera.h:72-129)
// ---------------------------------------------------------------------------
__device__ inline void camera_ray(const CamDev& c, float u, float v, f3 smp, f3& org, f3& dir, float& pdf) {
    const f3 pos = mk(c.pos[0], c.pos[1], c.pos[2]);
    const f3 cd = mk(c.dir[0], c.dir[1], c.dir[2]);
    const f3 e1 = mk(c.e1[0], c.e1[1], c.e1[2]);
    const f3 e2 = mk(c.e2[0], c.e2[1], c.e2[2]);
    if (c.type == SPPM_CAM_PINHOLE) {
        f3 off = (smp * mk(c.PW, c.PH, 0.0f)) / mk((float)c.W, (float)c.H, 1.0f);
        f3 img = pos - cd * c.D;
        img = img + (e1 * (u - 0.5f + off.x)) * c.PW;
        img = img + (e2 * (v - 0.5f + off.y)) * c.PH;
        pdf = 1.0f;
        org = img;
        dir = normalized(normalized(pos - img));
    } else {
        float lu = smp.x, lv = smp.y;
        if (c.dia_n > 0) {
            // TableSampler::Sample(Vec2) (TableSampler.h:129-132): bin
            // n * x (float product, truncated), left pixel when y < its
            // probability.  x = 1 (a float-rounded canonical draw) would index
            // past the table in the reference; it takes the last bin here.
            const int k = min((int)((float)c.dia_n * smp.x), c.dia_n - 1);
            const float4 uv = c.dia_uv[k];
            const bool left = smp.y < c.dia_prob[k];
            lu = left ? uv.x : uv.z;
            lv = left ? uv.y : uv.w;
        }
        float pdf_area = 1.0f / (c.LW * c.LH);
        f3 lens = (pos + (e1 * (lu - 0.5f)) * c.LW) + (e2 * (lv - 0.5f)) * c.LH;
        f3 img = pos - cd * c.D;
        img = img + (e1 * (u - 0.5f)) * c.PW;
        img = img + (e2 * (v - 0.5f)) * c.PH;
        float A = c.D;
        float B = (A * c.F) / (A - c.F);
        f3 target = img + ((pos - img) * (A + B)) / A;
        float cs = dot(cd, normalized(lens - img));
        float d2 = norm2(lens - img);
        pdf = pdf_area * (d2 / c.F / c.F) / cs;
        org = lens;
        dir = normalized(normalized(target - lens));
    }
}

// ---------------------------------------------------------------------------
// createHitpoints + getHitpointInfo (SPPM.cpp:98-353), one lane per pixel.
// After the hit point is made the rest of the eye path can only add
// environment light (SPPM.cpp:290-300), so without an environment map (and
// without next-event estimation) the path stops there.
// ---------------------------------------------------------------------------
// IBL::GetUVFromDirection, SPHERE mapping (IBL.h:90-103; unqualified acos on a
// float is the double function, u and v narrowed to float), then the
// bilinear fetch of GetIBLColorFromDirection (IBL.h:122-129).
__device__ __forceinline__ f3 ibl_color(const SceneDev& S, f3 dir) {
    float u = 0.0f, v = 0.0f;
    if (!(fabs_(dir.x) < HI_1EM6 && fabs_(dir.z) < HI_1EM6)) {  // (double)|.| < 1e-6
        const f3 xz = normalized(mk(dir.x, 0.0f, dir.z));
        const double a = acos((double)xz.x);
        u = (float)((xz.z >= 0.0f ? a : (double)(2.0f * SPPM_PI) - a) / (double)(2.0f * SPPM_PI));
        v = (float)(acos((double)dir.y) / (double)SPPM_PI);
    }
    const TexDesc td = S.tex[S.ibl_tex];
    const float4 c = tex_bilinear(S.texels + td.offset, td.width, td.height, u, v);
    return mk(c.x, c.y, c.z);
}

// Scene::IsOccluded (Scene.cpp:246-281): the closest hit in [MinDist, MaxDist)
// occludes unless its surface's albedo texture makes it more translucent
// than the threshold (1 - alpha > translucency); then the ray goes on from
// just behind it with the range shortened, renormalised as Ray's constructor
// does.  The reference recurses; this loops, passing at most OCC_LAYERS
// surfaces (flag bit 2 beyond, counted as occluded).
constexpr int OCC_LAYERS = 256;
template <int W>
__device__ __forceinline__ bool occluded_ray(const SceneDev& S, f3 o, f3 d, float mn, float mx, float transl,
                                             const TraceStack& ts, uint32_t& fl) {
    for (int layer = 0;; ++layer) {
        if (layer == OCC_LAYERS) {
            fl |= 4u;
            return true;
        }
        HitRec h;
        if (!closest_hit<W>(S, o, d, mn, ts, h, &fl) || !((double)h.t < (double)mx - 1e-10)) return false;
        const DevAttr& a = S.attrs[S.facet_attr[h.facet]];
        if (a.tex_albedo < 0) return true;
        const float2* tv = S.facet_uv + 3 * (size_t)h.facet;
        const float2 t0 = tv[0], t1 = tv[1], t2 = tv[2];
        const float uvx = ((t1.x - t0.x) * h.u + (t2.x - t0.x) * h.v) + t0.x;
        const float uvy = ((t1.y - t0.y) * h.u + (t2.y - t0.y) * h.v) + t0.y;
        const TexDesc td = S.tex[a.tex_albedo];
        const float4 c = tex_bilinear(S.texels + td.offset, td.width, td.height, uvx, uvy);
        const float tr = (float)(1.0 - (double)c.w);
        if (!(tr > transl)) return true;
        const float dist = h.t + SPPM_RAY_OFFSET;
        o = o + d * dist;
        d = normalized(d);
        mn = fmaxf(mn - dist, SPPM_RAY_OFFSET);
        mx = mx - dist;
    }
}

// NEE: the compiled-out next-event estimation of getHitpointInfo
// (SPPM.cpp:144-168, 187-229); keep: keep what it estimates (SPPM_NEE_KEEP).
template <int W, bool NEE>
__device__ __forceinline__ void eye_path(const SceneDev& S, const CamDev& C, const HitDev& H, int step, int i,
                                         const uint32_t* __restrict__ words, const TraceStack& ts, uint32_t& fl,
                                         bool keep, uint32_t& nray, uint32_t& ntie) {
    uint32_t ovf = 0;
    MTCursor m;
    mt_seed(m, (uint32_t)i + (uint32_t)C.W * (uint32_t)C.H * (uint32_t)step);
    const int w = i % C.W, hh = i / C.W;
    const float u = (float)w / (float)C.W, v = (float)hh / (float)C.H;
    f3 smp;
    if (C.sampler == SPPM_SAMPLER_UNIFORM) {
        const uint32_t* wd = words + 6 * (size_t)i;
        // Vec3(real1(), real1(), real1()) with g++'s right-to-left argument order
        smp.z = (float)canonical(wd[0], wd[1]);
        smp.y = (float)canonical(wd[2], wd[3]);
        smp.x = (float)canonical(wd[4], wd[5]);
    } else {
        smp = mk(C.fixed_smp[0], C.fixed_smp[1], C.fixed_smp[2]);
    }
    f3 cur_org, cur_dir;
    float campdf = 1.0f;
    camera_ray(C, u, v, smp, cur_org, cur_dir, campdf);
    float cur_index = 1.0f;
    f3 next_org = cur_org, next_dir = cur_dir;
    float next_index = 1.0f;
    f3 T = mk(1.0f, 1.0f, 1.0f);
    float PDF = 1.0f;
    int depth = 0;
    int type = 0, facet = -1, kind = 0;
    f3 hpos = mk(0, 0, 0), hout = mk(0, 1, 0), hw = mk(1, 1, 1), hn = mk(0, 0, 0);
    float hidx = 1.0f, hrough = 1.0f, habs = 1.0f, hrefl = 1.0f;
    f3 direct = mk(0.0f, 0.0f, 0.0f);
    const bool ibl = S.ibl_tex >= 0;
    // TraceParam::PrevGmt (position, IfDeltaSurface) and pdf_prevDir (SPPM.h:78-80)
    f3 prev_pos = mk(0.0f, 0.0f, 0.0f);
    float pdf_prev = 0.0f;
    bool prev_delta = false;
    for (;;) {
        if (depth > 6) break;
        else if (depth > 4) {
            float rus = (float)exp((double)(4 - depth));
            if (real1(m, &ovf) > (double)rus) break;
            PDF *= rus;
        }
        HitRec h;
        ++nray;
        const bool got = closest_hit<W>(S, cur_org, cur_dir, SPPM_RAY_OFFSET, ts, h, &fl);
        ntie += h.tie ? 1u : 0u;
        if (!got) {
            // escape: the environment's colour (SPPM.cpp:290-300)
            if (ibl) direct = direct + (ibl_color(S, cur_dir) * T) / PDF;
            if (type == 0) type = 2;
            break;
        }
        Geom g;
        calc_geometry(S, h, g);
        f3 Nf = dot(g.N, cur_dir) > 0 ? -g.N : g.N;
        next_org = g.pos;
        if (g.transl > 0.5f) {
            T = T * g.transl;
            cur_org = g.pos + cur_dir * SPPM_RAY_OFFSET;
            continue;
        }
        if (NEE && norm2(g.radiance) > LO_1EM6) {  // (double)norm2 > 1e-6
            // an emitter ends the path; before the hit point its radiance is
            // DirectLight, MIS-weighted against the light sample of the
            // previous vertex unless that was a delta surface (SPPM.cpp:145-167)
            if (type != 1) {
                float mis = 1.0f;
                if (depth > 0 && !prev_delta) {
                    const int li = S.facet_light[h.facet];
                    float pdf_nee = 0.0f;  // GetAreaLightPDF of a facet that is no light: 0
                    if (li >= 0) {
                        const DevLight& L = S.lights[li];
                        // SampleAreaLight_pdf (LightSampler.cpp:84-96)
                        const float ratio = L.textured ? (norm(g.radiance) / L.tab_avg / 1000.0f) * 1000.0f : 1.0f;
                        const float tri = ratio / L.area;
                        const float cosL = smax(0.0001f, smin(1.0f, dot(Nf, -cur_dir)));
                        pdf_nee = (light_pick_pdf(S, L) * tri) * norm2(prev_pos - g.pos) / cosL;
                    }
                    mis = (pdf_prev * pdf_prev) / ((pdf_nee * pdf_nee) + (pdf_prev * pdf_prev));
                }
                direct = direct + (((g.albedo * mis) * g.radiance) * T) / PDF;
                if (keep && type == 0) type = 2;
            }
            break;
        }
        if (g.is_sppm && type != 1) {  // makeHitPoint (SPPM.cpp:79-96, 171-174)
            if (!(NEE && keep)) direct = mk(0.0f, 0.0f, 0.0f);  // the new HitPoint's DirectLight
            type = 1;
            facet = h.facet;
            hpos = g.pos;
            hout = -cur_dir;
            hw = T / PDF;
            hn = Nf;
            kind = g.kind;
            hidx = g.index;
            hrough = g.rough;
            habs = g.absorb;
            hrefl = g.refl;
            // the path goes on after its hit point only to reach the
            // environment; without one nothing later changes the hit point
            if (!ibl) break;
        }
        float Ii, Io = 1.0f;
        if (near1(cur_index)) { Ii = 1.0f; Io = g.index; }
        else { Ii = cur_index; Io = 1.0f; }
        FParam fp{cur_dir, next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
        bool delta = false;
        if (NEE) {
            delta = S.attrs[S.facet_attr[h.facet]].delta != 0;
            if (!delta && type != 1) {
                // a light sample from this vertex (SPPM.cpp:190-226): pick a
                // light, a point on it, and a shadow ray from the hit position
                const int li = table_sample(S.lpick_prob, S.lpick_left, S.lpick_right, S.nlights, m, &ovf);
                const DevLight& L = S.lights[li];
                const float psel = light_pick_pdf(S, L);
                f3 lpos, rad;
                float parea;
                light_point(S, L, m, &ovf, lpos, rad, parea);
                const f3 ln = mk(L.normal[0], L.normal[1], L.normal[2]);
                const f3 dv = lpos - g.pos;
                const f3 ldir = normalized(normalized(dv));  // normalised, then again by Ray's ctor
                const float dist = norm(dv);
                const float maxd = (float)((double)dist - (double)SPPM_RAY_OFFSET * 2.0);
                ++nray;
                if (!occluded_ray<W>(S, g.pos, ldir, SPPM_RAY_OFFSET, maxd, 0.5f, ts, fl)) {
                    const float cosL = smin(1.0f, smax(0.00001f, dot(ln, -ldir)));
                    const float cosE = fabs_(dot(Nf, ldir));
                    FParam fl2 = fp;
                    fl2.out = ldir;
                    const float b = smax(0.0f, surf_brdf(g.kind, fl2));
                    const float pp = parea * psel;
                    const f3 dlr = (((((rad * cosL) * powf(dist, -2.0f)) / pp) * cosE) * g.albedo) * b;
                    const float pn = pp * (dist * dist) / cosL;
                    const float pt = surf_pdf(g.kind, fl2);
                    const float w = (pn * pn) / ((pn * pn) + (pt * pt));
                    direct = direct + ((dlr * w) * T) / PDF;
                }
            }
        }
        float pdf_s = 1.0f;
        next_dir = surf_sample(g.kind, fp, m, &ovf, &pdf_s);
        if (!rutil_valid_dir(next_dir)) break;
        if (g.opaque) {
            if (dot(next_dir, Nf) < HI_1EM6) break;  // (double)dot < 1e-6
        } else {
            next_index = !(dot(next_dir, Nf) < 0.0f) ? cur_index : Io;
        }
        fp.out = next_dir;
        pdf_s = surf_pdf(g.kind, fp);
        f3 outN = dot(Nf, next_dir) > 0 ? Nf : -Nf;
        next_org = next_org + outN * SPPM_RAY_OFFSET;
        float brdf = surf_brdf(g.kind, fp);
        float cs = smax(0.0f, smin(1.0f, dot(next_dir, outN)));
        if (brdf < HI_1EM8) break;  // (double)brdf < 1e-8
        if (pdf_s < 0) break;
        T = T * ((g.albedo * cs) * brdf);
        PDF *= pdf_s;
        if (NEE) {
            pdf_prev = pdf_s;
            prev_pos = g.pos;
            prev_delta = delta;
        }
        cur_org = next_org;
        cur_dir = next_dir;
        cur_index = next_index;
        ++depth;
    }
    hw = hw / campdf;  // createHitpoints: PixelWeight /= pdf, DirectLight /= pdf (SPPM.cpp:346-347)
    if (ibl || NEE) {
        direct = direct / campdf;
        H.direct[i] = make_float4(direct.x, direct.y, direct.z, 0.0f);
    }
    if (type == 1 && (!isfinite(hw.x) || !isfinite(hw.y) || !isfinite(hw.z))) fl |= 32u;
    H.pos[i] = make_float4(hpos.x, hpos.y, hpos.z, (float)type);
    H.nrm[i] = make_float4(hn.x, hn.y, hn.z, hidx);
    H.out[i] = make_float4(hout.x, hout.y, hout.z, hrough);
    H.weight[i] = make_float4(hw.x, hw.y, hw.z, hrefl);
    H.mat[i] = make_int2(kind, facet);
    H.absorb[i] = habs;
    if (ovf) fl |= 2u;
}

// Eye pass: persistent waves take 64 consecutive pixels at a time from a
// counter (zeroed before the launch), so waves whose pixels have long paths
// (glass, mirrors) do not hold back the others; a pixel's result does not
// depend on the wave that traces it.
// The pixels of one scene, 64 at a time, until its counter is exhausted.
template <int W, bool NEE>
__device__ __forceinline__ void eye_pixels(const SceneDev& S, const CamDev& C, const HitDev& H, int step,
                                           const uint32_t* __restrict__ words, int* __restrict__ queue, bool keep,
                                           const TraceStack& ts, uint32_t& fl, uint32_t& nray, uint32_t& ntie) {
    const int lane = threadIdx.x;
    const int NPIX = C.W * C.H;
    StripeCursor cur = stripe_start();
    for (;;) {
        int first = 0;
        const int got = stripe_claim(queue, NPIX, cur, 64, &first);
        if (got == 0) break;
        if ((lane & 63) < got) eye_path<W, NEE>(S, C, H, step, first + (lane & 63), words, ts, fl, keep, nray, ntie);
    }
}

template <int W, bool NEE>
__global__ void __launch_bounds__(PATH_BLOCK, PATH_WAVES_PER_SIMD) k_eye(SceneDev S, CamDev C, HitDev H, int step_arg,
                                                    const int* __restrict__ step_dev, int cap,
                                                    uint32_t* __restrict__ spill, const uint32_t* __restrict__ words,
                                                    uint32_t* __restrict__ flags, int keep,
                                                    unsigned long long* __restrict__ rays, int* __restrict__ queue) {
    const int step = step_dev ? *step_dev + 1 : step_arg;
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    uint32_t fl = 0, nray = 0, ntie = 0;
    eye_pixels<W, NEE>(S, C, H, step, words, queue, keep != 0, ts, fl, nray, ntie);
    if (fl) atomicOr(flags, fl);
    add_wave_count(rays, nray);
    add_wave_count(rays + 2, ntie);
}

// The eye pass of every scene of a batch (see k_photon_multi).
template <int W, bool NEE>
__global__ void __launch_bounds__(PATH_BLOCK, PATH_WAVES_PER_SIMD) k_eye_multi(const EyeJob* __restrict__ jobs, int njobs,
                                                                             const int* __restrict__ bstart, int cap,
                                                                             uint32_t* __restrict__ spill) {
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    const int home = home_scene(bstart, njobs);
    for (int k = 0; k < njobs; ++k) {
        const int s = __builtin_amdgcn_readfirstlane(home + k < njobs ? home + k : home + k - njobs);
        const EyeJob& J = jobs[s];
        if (scene_done(J.queue)) continue;
        const int step = J.step_dev ? *J.step_dev + 1 : J.step;
        uint32_t fl = 0, nray = 0, ntie = 0;
        eye_pixels<W, NEE>(J.S, J.C, J.H, step, J.words, J.queue, J.keep != 0, ts, fl, nray, ntie);
        scene_mark_done(J.queue);
        if (fl) atomicOr(J.flags, fl);
        add_wave_count(J.rays, nray);
        add_wave_count(J.rays + 2, ntie);
    }
}

// ---------------------------------------------------------------------------
// applyHitPointInfo (SPPM.cpp:395-451).  Lane <-> pixel in 8x8 tiles so a
// wave gathers from a compact patch of the photon grid.
// mode 0: gather + update; 1: count only (M); 2: gather + update + lists.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tiled_pixel(int lin, int W, int H) {
    const int tiles_x = (W + 7) >> 3;
    int tile = lin >> 6, in = lin & 63;
    int tx = tile % tiles_x, ty = tile / tiles_x;
    int x = tx * 8 + (in & 7), y = ty * 8 + (in >> 3);
    if (x >= W || y >= H) return -1;
    return y * W + x;
}

// One neighbour candidate: the reference's two tests, then BRDF accumulation.
struct GatherLane {
    f3 c, bbmin, bbmax, N, outd, wgt, add;
    float R2, ior, rough, refl, absorb;
    int kind, M, lo;
};
__device__ __forceinline__ void gather_candidate(GatherLane& L, const float4 pp, const float4 pf, const float4 pd,
                                                 int mode, int* __restrict__ lists) {
    // isIncludePoint (kdtree.h:78-88) then the strict radius test (SPPM.cpp:432)
    if (L.bbmin.x > pp.x || L.bbmax.x < pp.x) return;
    if (L.bbmin.y > pp.y || L.bbmax.y < pp.y) return;
    if (L.bbmin.z > pp.z || L.bbmax.z < pp.z) return;
    const f3 d = mk(pp.x, pp.y, pp.z) - L.c;
    if (!(norm2(d) < L.R2)) return;
    ++L.M;
    if (mode == 1) return;
    const f3 indir = mk(pd.x, pd.y, pd.z);
    const float Io = near1(pp.w) ? L.ior : 1.0f;
    const f3 nrm = dot(indir, L.N) < 0 ? L.N : -L.N;
    FParam fp{indir, L.outd, nrm, pp.w, Io, L.rough, L.absorb, L.refl};
    const float b = surf_brdf(L.kind, fp);
    L.add = L.add + ((L.wgt * b) * mk(pf.x, pf.y, pf.z)) / pf.w;
    if (mode == 2) lists[L.lo++] = __float_as_int(pd.w);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wave per 8x8 pixel tile.  The tile's hit points usually lie on one
// small surface patch, so their query boxes overlap almost entirely: the
// wave walks the cells of the union box once, stages 64 photons at a time in
// LDS with coalesced loads, and every lane tests each staged photon against
// its own hit point (LDS broadcast reads).  Tiles whose union box is much
// larger than the lanes' own boxes (silhouettes, depth edges) walk their own
// cells per lane instead.  Either way each lane applies exactly the
// reference's box + radius tests, so the neighbour set is unchanged.
__global__ void __launch_bounds__(64) k_gather(HitDev H, SharedDev SH, GridDev G, const GridParams* __restrict__ gp,
                                               int W, int Hh, float alpha, int mode, int* __restrict__ counts,
                                               const int* __restrict__ list_off, int* __restrict__ lists,
                                               unsigned long long* __restrict__ rstat, Counters* __restrict__ ctr,
                                               uint32_t* __restrict__ flags) {
    __shared__ float4 s_pos[64], s_flux[64], s_dir[64];
    const int lane = threadIdx.x;
    const int p = tiled_pixel(blockIdx.x * 64 + lane, W, Hh);
    float4 hp = make_float4(0, 0, 0, 0), sr = make_float4(0, 0, 0, 0);
    if (p >= 0) {
        hp = H.pos[p];
        sr = SH.rad[p];
    }
    const bool active = p >= 0 && (int)hp.w == 1;
    GatherLane L;
    L.M = 0;
    L.lo = 0;
    L.add = mk(0.0f, 0.0f, 0.0f);
    const float R = sr.w;
    L.c = mk(hp.x, hp.y, hp.z);
    L.bbmin = L.c - mk(1.0f, 1.0f, 1.0f) * R;
    L.bbmax = L.c + mk(1.0f, 1.0f, 1.0f) * R;
    L.R2 = R * R;
    const GridParams g = *gp;
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1, z0 = 0, z1 = -1;
    if (active) {
        x0 = cell_coord(L.bbmin.x, g.bmin[0], g.inv_cell, g.nx); x1 = cell_coord(L.bbmax.x, g.bmin[0], g.inv_cell, g.nx);
        y0 = cell_coord(L.bbmin.y, g.bmin[1], g.inv_cell, g.ny); y1 = cell_coord(L.bbmax.y, g.bmin[1], g.inv_cell, g.ny);
        z0 = cell_coord(L.bbmin.z, g.bmin[2], g.inv_cell, g.nz); z1 = cell_coord(L.bbmax.z, g.bmin[2], g.inv_cell, g.nz);
        float4 hn = H.nrm[p], ho = H.out[p], hw = H.weight[p];
        L.N = mk(hn.x, hn.y, hn.z);
        L.ior = hn.w;
        L.outd = mk(ho.x, ho.y, ho.z);
        L.rough = ho.w;
        L.wgt = mk(hw.x, hw.y, hw.z);
        L.refl = hw.w;
        L.kind = H.mat[p].x;
        L.absorb = H.absorb[p];
        if (mode == 2) L.lo = list_off[p];
    }
    const uint64_t act = __ballot(active);
    if (act) {
        // union of the active lanes' cell ranges
        const int X0 = -wave_max_i(active ? -x0 : -0x7fffffff), X1 = wave_max_i(active ? x1 : -1);
        const int Y0 = -wave_max_i(active ? -y0 : -0x7fffffff), Y1 = wave_max_i(active ? y1 : -1);
        const int Z0 = -wave_max_i(active ? -z0 : -0x7fffffff), Z1 = wave_max_i(active ? z1 : -1);
        const int own = active ? (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1) : 0;
        const int own_max = wave_max_i(own);
        const long long ucells = (long long)(X1 - X0 + 1) * (Y1 - Y0 + 1) * (Z1 - Z0 + 1);
        if (ucells <= 8LL * own_max) {
            for (int z = Z0; z <= Z1; ++z)
                for (int y = Y0; y <= Y1; ++y) {
                    const int row = (z * g.ny + y) * g.nx;
                    const int qa = G.cell_start[row + X0];
                    const int qb = G.cell_start[row + X1 + 1];
                    const bool mine = active && z >= z0 && z <= z1 && y >= y0 && y <= y1;
                    for (int base = qa; base < qb; base += 64) {
                        const int q = base + lane;
                        if (q < qb) {
                            s_pos[lane] = G.pos[q];
                            s_flux[lane] = G.flux[q];
                            s_dir[lane] = G.dir[q];
                        }
                        __syncthreads();
                        const int n = min(64, qb - base);
                        if (mine)
                            for (int j = 0; j < n; ++j) gather_candidate(L, s_pos[j], s_flux[j], s_dir[j], mode, lists);
                        __syncthreads();
                    }
                }
        } else if (active) {
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y) {
                    const int row = (z * g.ny + y) * g.nx;
                    const int qa = G.cell_start[row + x0];
                    const int qb = G.cell_start[row + x1 + 1];
                    for (int q = qa; q < qb; ++q) gather_candidate(L, G.pos[q], G.flux[q], G.dir[q], mode, lists);
                }
        }
    }
    float rnew = p >= 0 ? sr.w : 0.0f;
    if (active && mode != 1 && L.M > 0) {
        // progressive update (SPPM.cpp:444-449); N is an int: N += alpha*M truncates
        int2 cn = SH.cnt[p];
        const float num = (float)cn.x + alpha * (float)L.M;
        const float den = (float)(cn.x + L.M);
        f3 r = mk(sr.x, sr.y, sr.z);
        r = ((r + L.add) * num) / den;
        const float R1 = sr.w * sqrtf_(num / den);
        SH.rad[p] = make_float4(r.x, r.y, r.z, R1);
        SH.cnt[p] = make_int2((int)((float)cn.x + (float)L.M * alpha), cn.y + 1);
        rnew = R1;
        if (!(r.x == r.x) || !(r.y == r.y) || !(r.z == r.z) || __builtin_isinf(r.x) || __builtin_isinf(r.y) ||
            __builtin_isinf(r.z))
            atomicOr(flags, 1u);
    }
    if (p >= 0 && counts) counts[p] = active ? L.M : 0;
    // wave totals: radius sum (next pass's cell size), hit points, neighbours
    unsigned long long rs = active ? (unsigned long long)((double)rnew * 4294967296.0) : 0ull;
    for (int off = 32; off > 0; off >>= 1) rs += __shfl_xor(rs, off, 64);
    const int th = wave_sum_i(active ? 1 : 0);
    const int tn = wave_sum_i(L.M);
    if (lane == 0 && mode != 1) {
        atomicAdd(&rstat[0], rs);
        atomicAdd(&rstat[1], (unsigned long long)th);
        atomicAdd(&ctr->hit_points, (unsigned long long)th);
        atomicAdd(&ctr->neighbours, (unsigned long long)tn);
    }
}

// ---------------------------------------------------------------------------
// Load-balanced gather for the production pass (no neighbour lists).
// Candidate counts per 8x8 tile vary by orders of magnitude (caustics), so:
//   k_gather_plan    per tile: union cell box of its hit points and its
//                    candidate count
//   k_chunk_size     candidates per chunk so that all chunks fit the buffer
//   k_tile_chunks    chunks per tile
//   scan             chunk offsets per tile
//   k_gather_chunks  persistent waves take chunks of one tile's candidates,
//                    stage them in LDS and write per-lane partial sums
//   k_gather_finish  per tile: partial sums in chunk order + the update
// ---------------------------------------------------------------------------
struct TileRange {
    int x0, x1, y0, y1, z0, z1, rowlen_pad, cand;
};

// Hit-point groups of the balanced gather: the valid hit points sorted by the
// Morton code of their position (1024 steps per scene-box axis),
// GATHER_GROUP per group.  World-space groups keep the union of their query
// boxes close to one query box; screen tiles do not (a thin lens spreads a
// tile's hit points over the circle of confusion).  The smaller the group,
// the closer its union box to each member's own box: a wave holds each hit
// point of a group in GATHER_PHASES lanes, and each of those lanes tests
// every GATHER_PHASES-th staged photon.
#ifndef SPPM_GATHER_GROUP
#define SPPM_GATHER_GROUP 32
#endif
constexpr int GATHER_GROUP = SPPM_GATHER_GROUP;
constexpr int GATHER_PHASES = 64 / GATHER_GROUP;
static_assert(GATHER_GROUP == 16 || GATHER_GROUP == 32 || GATHER_GROUP == 64, "group of 16, 32 or 64 hit points");
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ int quant10(float p, float lo, float ext) {
    float f = (p - lo) / ext * 1024.0f;
    int q = (int)f;
    return q < 0 ? 0 : (q > 1023 ? 1023 : q);
}
constexpr uint32_t HIT_KEY_INVALID = 1u << 30;  // above every 30-bit Morton code: invalid hit points sort last
__global__ void k_hit_keys(HitDev H, int np, float bx, float by, float bz, float ex, float ey, float ez,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float4 hp = H.pos[i];
    uint32_t key = HIT_KEY_INVALID;
    if ((int)hp.w == 1)
        key = spread10((uint32_t)quant10(hp.x, bx, ex)) | (spread10((uint32_t)quant10(hp.y, by, ey)) << 1) |
              (spread10((uint32_t)quant10(hp.z, bz, ez)) << 2);
    keys[i] = key;
    vals[i] = (uint32_t)i;
}
// Number of valid hit points = first position of HIT_KEY_INVALID in the
// sorted keys.  One wave probes 64 evenly spaced positions per round, so a
// million keys take about four rounds of dependent loads (one global counter
// bumped by every wave would serialise thousands of atomics on one address).
__global__ void __launch_bounds__(64) k_count_valid(const uint32_t* __restrict__ keys, int n, int* __restrict__ nvalid) {
    const int lane = threadIdx.x;
    int lo = 0, hi = n;  // keys[< lo] valid, keys[>= hi] invalid
    while (lo < hi) {
        const int step = (hi - lo + 63) / 64;
        const int q = lo + lane * step;
        const bool inr = q < hi;
        const bool ge = inr && keys[q] >= HIT_KEY_INVALID;
        const unsigned long long m = __ballot(ge || !inr);
        if (m == 0ull) {  // all 64 probes valid
            lo = lo + 63 * step + 1;
            continue;
        }
        const int k = __builtin_ctzll(m);
        const int qk = lo + k * step;
        const int nlo = k > 0 ? lo + (k - 1) * step + 1 : lo;
        hi = qk < hi ? qk : hi;
        lo = nlo;
    }
    if (lane == 0) *nvalid = lo;
}
__device__ __forceinline__ int group_pixel(const uint32_t* __restrict__ order, int nvalid, int lin) {
    return lin < nvalid ? (int)order[lin] : -1;
}

__device__ __forceinline__ void load_gather_lane(const HitDev& H, const SharedDev& SH, const GridParams& g, int p,
                                                 GatherLane& L, bool& active, float4& sr, int* rng) {
    float4 hp = make_float4(0, 0, 0, 0);
    sr = make_float4(0, 0, 0, 0);
    if (p >= 0) {
        hp = H.pos[p];
        sr = SH.rad[p];
    }
    active = p >= 0 && (int)hp.w == 1;
    L.M = 0;
    L.lo = 0;
    L.add = mk(0.0f, 0.0f, 0.0f);
    const float R = sr.w;
    L.c = mk(hp.x, hp.y, hp.z);
    L.bbmin = L.c - mk(1.0f, 1.0f, 1.0f) * R;
    L.bbmax = L.c + mk(1.0f, 1.0f, 1.0f) * R;
    L.R2 = R * R;
    rng[0] = 0; rng[1] = -1; rng[2] = 0; rng[3] = -1; rng[4] = 0; rng[5] = -1;
    if (active) {
        rng[0] = cell_coord(L.bbmin.x, g.bmin[0], g.inv_cell, g.nx); rng[1] = cell_coord(L.bbmax.x, g.bmin[0], g.inv_cell, g.nx);
        rng[2] = cell_coord(L.bbmin.y, g.bmin[1], g.inv_cell, g.ny); rng[3] = cell_coord(L.bbmax.y, g.bmin[1], g.inv_cell, g.ny);
        rng[4] = cell_coord(L.bbmin.z, g.bmin[2], g.inv_cell, g.nz); rng[5] = cell_coord(L.bbmax.z, g.bmin[2], g.inv_cell, g.nz);
        float4 hn = H.nrm[p], ho = H.out[p], hw = H.weight[p];
        L.N = mk(hn.x, hn.y, hn.z);
        L.ior = hn.w;
        L.outd = mk(ho.x, ho.y, ho.z);
        L.rough = ho.w;
        L.wgt = mk(hw.x, hw.y, hw.z);
        L.refl = hw.w;
        L.kind = H.mat[p].x;
        L.absorb = H.absorb[p];
    }
}

__global__ void __launch_bounds__(64) k_gather_plan(HitDev H, SharedDev SH, GridDev G, const GridParams* __restrict__ gp,
                                                    const uint32_t* __restrict__ order, const int* __restrict__ nvalidp,
                                                    TileRange* __restrict__ tiles) {
    const int lane = threadIdx.x;
    const int p = lane < GATHER_GROUP ? group_pixel(order, *nvalidp, blockIdx.x * GATHER_GROUP + lane) : -1;
    const GridParams g = *gp;
    GatherLane L;
    bool active;
    float4 sr;
    int r[6];
    load_gather_lane(H, SH, g, p, L, active, sr, r);
    TileRange T;
    T.x0 = -wave_max_i(active ? -r[0] : -0x7fffffff); T.x1 = wave_max_i(active ? r[1] : -1);
    T.y0 = -wave_max_i(active ? -r[2] : -0x7fffffff); T.y1 = wave_max_i(active ? r[3] : -1);
    T.z0 = -wave_max_i(active ? -r[4] : -0x7fffffff); T.z1 = wave_max_i(active ? r[5] : -1);
    int cand = 0;
    if (__ballot(active)) {
        const int ny = T.y1 - T.y0 + 1, nrows = ny * (T.z1 - T.z0 + 1);
        for (int k = lane; k < nrows; k += 64) {
            const int y = T.y0 + k % ny, z = T.z0 + k / ny;
            const int row = (z * g.ny + y) * g.nx;
            cand += G.cell_start[row + T.x1 + 1] - G.cell_start[row + T.x0];
        }
    } else {
        T.x0 = 0; T.x1 = -1; T.y0 = 0; T.y1 = -1; T.z0 = 0; T.z1 = -1;
    }
    cand = wave_sum_i(cand);
    T.rowlen_pad = 0;
    T.cand = cand;
    if (lane == 0) tiles[blockIdx.x] = T;
}

// chunk size = max(1024, ceil(total / cap)) rounded to a multiple of 64
__global__ void __launch_bounds__(256) k_chunk_size(const TileRange* __restrict__ tiles, int ntiles, int cap,
                                                    int* __restrict__ chunk, Counters* __restrict__ ctr) {
    __shared__ unsigned long long part[4];
    unsigned long long s = 0;
    for (int t = threadIdx.x; t < ntiles; t += 256) s += (unsigned long long)tiles[t].cand;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = part[0] + part[1] + part[2] + part[3];
        unsigned long long c = (tot + (unsigned long long)cap - 1) / (unsigned long long)cap;
        if (c < 1024) c = 1024;
        c = (c + 63) & ~63ull;
        *chunk = (int)c;
        ctr->candidates += tot;
    }
}

__global__ void k_tile_chunks(const TileRange* __restrict__ tiles, int ntiles, const int* __restrict__ chunk,
                              int* __restrict__ nch) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) { nch[t] = 0; return; }
    const int c = *chunk;
    nch[t] = (tiles[t].cand + c - 1) / c;
}

// The reference's two neighbour tests for one candidate, closed box
// (kdtree.h:78-88) and the strict radius (SPPM.cpp:432), evaluated without
// early exit.
__device__ __forceinline__ bool in_query_nb(const GatherLane& L, const float4 pp) {
    const bool box = (L.bbmin.x <= pp.x) & (pp.x <= L.bbmax.x) & (L.bbmin.y <= pp.y) & (pp.y <= L.bbmax.y) &
                     (L.bbmin.z <= pp.z) & (pp.z <= L.bbmax.z);
    const f3 d = mk(pp.x, pp.y, pp.z) - L.c;
    return box & (norm2(d) < L.R2);
}

// Persistent waves over the chunk list.  Per staged batch of 64 photons every
// lane runs the two tests against its own hit point and lists the staged
// photons that pass (LDS, one byte per entry, [entry][lane]); then every
// lane walks its own list and accumulates W * BRDF * Flux / PDF
// (SPPM.cpp:434-440).  Splitting the cheap tests from the BRDF evaluation
// keeps the lanes of a wave busy in the second loop instead of evaluating
// one lane's BRDF per candidate; each lane still adds its neighbours in
// ascending candidate order.
// The chunks of one scene's gather, claimed until its counters are exhausted.
__device__ __forceinline__ void gather_chunks(HitDev H, SharedDev SH, GridDev G,
                                              const GridParams* __restrict__ gp, const uint32_t* __restrict__ order,
                                              const int* __restrict__ nvalidp, const TileRange* __restrict__ tiles,
                                              int ntiles, const int* __restrict__ chunk_off,
                                              const int* __restrict__ chunkp, int* __restrict__ queue,
                                              float4* __restrict__ partial, unsigned long long* __restrict__ pdep,
                                              PairRecord rec) {
    // LDS of the loop (one wave per block; each kernel that inlines this
    // function gets its own arrays)
    __shared__ float4 s_pos[64], s_flux[64], s_dir[64];
    __shared__ unsigned char s_hit[64 * 64];
    __shared__ int s_cnt[64];  // hit points of this group gathering each staged photon (guided emission)
    const int lane = threadIdx.x;
    const int hl = lane % GATHER_GROUP;   // the lane's hit point in the group
    const int ph = lane / GATHER_GROUP;   // the lane tests staged photons ph, ph + GATHER_PHASES, ...
    const int total = chunk_off[ntiles];
    const int CH = *chunkp;
    const int nvalid = *nvalidp;
    const GridParams g = *gp;
    // chunks are claimed round-robin through the striped counters (zeroed
    // before the launch); every wave ends once every stripe is exhausted.  The partial
    // sums go to the chunk's own slot, so the result does not depend on which
    // wave took which chunk.
    StripeCursor cur = stripe_start();
    for (;;) {
        int ch = 0;
        if (!stripe_claim_rr(queue, total, cur, &ch)) break;
        // owning tile: last t with chunk_off[t] <= ch
        int lo = 0, hi = ntiles;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (chunk_off[mid] <= ch) lo = mid; else hi = mid;
        }
        const int t = lo;
        const TileRange T = tiles[t];
        const int p = group_pixel(order, nvalid, t * GATHER_GROUP + hl);
        GatherLane L;
        bool active;
        float4 sr;
        int r[6];
        load_gather_lane(H, SH, g, p, L, active, sr, r);
        const long long first = (long long)(ch - chunk_off[t]) * CH;
        const long long last = first + CH < T.cand ? first + CH : T.cand;
        const int ny = T.y1 - T.y0 + 1, nrows = ny * (T.z1 - T.z0 + 1);
        long long acc = 0;  // candidates before the current row
        for (int k = 0; k < nrows && acc < last; ++k) {
            const int y = T.y0 + k % ny, z = T.z0 + k / ny;
            const int row = (z * g.ny + y) * g.nx;
            const int qa = G.cell_start[row + T.x0];
            const int qb = G.cell_start[row + T.x1 + 1];
            const long long len = qb - qa;
            const long long s0 = first > acc ? first - acc : 0;
            const long long s1 = last - acc < len ? last - acc : len;
            acc += len;
            if (s0 >= s1) continue;
            const bool mine = active && z >= r[4] && z <= r[5] && y >= r[2] && y <= r[3];
            for (long long base = qa + s0; base < qa + s1; base += 64) {
                const long long q = base + lane;
                if (q < qa + s1) {
                    s_pos[lane] = G.pos[q];
                    s_flux[lane] = G.flux[q];
                    s_dir[lane] = G.dir[q];
                }
                if (pdep) s_cnt[lane] = 0;
                __syncthreads();
                const int n = (int)((qa + s1 - base) < 64 ? (qa + s1 - base) : 64);
                int nh = 0;
                if (mine) {
                    // four of this lane's staged photons per step, no branches: the list
                    // slot nh is written every time and kept only when the photon passes
                    // (nh <= the photons tested so far, so the slot stays below 64)
                    int j = ph;
                    for (; j + 3 * GATHER_PHASES < n; j += 4 * GATHER_PHASES) {
                        const float4 p0 = s_pos[j], p1 = s_pos[j + GATHER_PHASES];
                        const float4 p2 = s_pos[j + 2 * GATHER_PHASES], p3 = s_pos[j + 3 * GATHER_PHASES];
                        const bool h0 = in_query_nb(L, p0), h1 = in_query_nb(L, p1);
                        const bool h2 = in_query_nb(L, p2), h3 = in_query_nb(L, p3);
                        s_hit[nh * 64 + lane] = (unsigned char)j;
                        nh += h0;
                        s_hit[nh * 64 + lane] = (unsigned char)(j + GATHER_PHASES);
                        nh += h1;
                        s_hit[nh * 64 + lane] = (unsigned char)(j + 2 * GATHER_PHASES);
                        nh += h2;
                        s_hit[nh * 64 + lane] = (unsigned char)(j + 3 * GATHER_PHASES);
                        nh += h3;
                    }
                    for (; j < n; j += GATHER_PHASES) {
                        s_hit[nh * 64 + lane] = (unsigned char)j;
                        nh += in_query_nb(L, s_pos[j]);
                    }
                }
                L.M += nh;
                if (rec.pairs) {
                    // recording (parity tests): this lane's neighbours among the staged
                    // photons, appended at a wave-wide prefix of the lanes' counts
                    int incl = nh;
                    for (int off = 1; off < 64; off <<= 1) {
                        const int v = __shfl_up(incl, off, 64);
                        if (lane >= off) incl += v;
                    }
                    const int tot = __shfl(incl, 63, 64);
                    unsigned long long base = 0;
                    if (lane == 0 && tot > 0) base = atomicAdd(rec.count, (unsigned long long)tot);
                    base = __shfl(base, 0, 64) + (unsigned long long)(incl - nh);
                    for (int e = 0; e < nh; ++e)
                        if ((long long)(base + e) < rec.cap)
                            rec.pairs[base + e] = make_int2(p, __float_as_int(s_dir[s_hit[e * 64 + lane]].w));
                }
                const float on = dot(L.outd, L.N);
                for (int e = 0; e < nh; ++e) {
                    const int j = s_hit[e * 64 + lane];
                    const float4 pp = s_pos[j], pf = s_flux[j], pd = s_dir[j];
                    const f3 indir = mk(pd.x, pd.y, pd.z);
                    const float din = dot(indir, L.N);
                    float b;
                    if (L.kind == SPPM_MAT_DIFFUSE) {
                        // DiffuseSurface::BRDF (material.h:168-173) with n = din < 0 ? N : -N:
                        // dot(in, -n) = |din| and dot(out, n) = +-dot(out, N), exactly
                        // (negating a dot product's operand negates its rounded value)
                        const float co = din < 0 ? on : -on;
                        b = (fabs_(din) > LO_1EM4 && co > LO_1EM4) ? 1.0f / SPPM_PI : 0.0f;
                    } else {
                        const float Io = near1(pp.w) ? L.ior : 1.0f;
                        const f3 nrm = din < 0 ? L.N : -L.N;
                        FParam fp{indir, L.outd, nrm, pp.w, Io, L.rough, L.absorb, L.refl};
                        b = gather_brdf(L.kind, fp);
                    }
                    L.add = L.add + ((L.wgt * b) * mk(pf.x, pf.y, pf.z)) * __builtin_amdgcn_rcpf(pf.w);
                    // guided emission: per staged photon, the hit points that gather it
                    if (pdep) atomicAdd(&s_cnt[j], 1);
                }
                if (pdep) {
                    // training signal of the emission proposal: hit points reached x photon power
                    __syncthreads();
                    if (lane < n && s_cnt[lane] > 0) {
                        const float4 pf = s_flux[lane];
                        const float d = (float)s_cnt[lane] * (0.2126f * pf.x + 0.7152f * pf.y + 0.0722f * pf.z) / pf.w;
                        if (d > 0.0f && d < 3.0e38f) atomicAdd(&pdep[__float_as_int(s_dir[lane].w)], guide_fix(d));
                    }
                }
                __syncthreads();
            }
        }
        // the hit point's sums over its lanes, in phase order (deterministic)
        f3 tot = mk(0.0f, 0.0f, 0.0f);
        int Mt = 0;
#pragma unroll
        for (int q = 0; q < GATHER_PHASES; ++q) {
            const int src = hl + q * GATHER_GROUP;
            tot = tot + mk(__shfl(L.add.x, src, 64), __shfl(L.add.y, src, 64), __shfl(L.add.z, src, 64));
            Mt += __shfl(L.M, src, 64);
        }
        if (ph == 0)
            partial[(size_t)ch * GATHER_GROUP + hl] = make_float4(tot.x, tot.y, tot.z, __int_as_float(Mt));
    }
}

// The kernels fit 128 VGPRs, 4 waves per SIMD (one register more drops them
// to 3 waves and costs 10 % of the gather).  gather_chunks takes its
// arguments by value: with references the one-scene kernel needed 129.
constexpr int GATHER_WAVES_PER_SIMD = 4;
__global__ void __launch_bounds__(64) k_gather_chunks(HitDev H, SharedDev SH, GridDev G,
                                                      const GridParams* __restrict__ gp,
                                                      const uint32_t* __restrict__ order, const int* __restrict__ nvalidp,
                                                      const TileRange* __restrict__ tiles, int ntiles,
                                                      const int* __restrict__ chunk_off, const int* __restrict__ chunkp,
                                                      int* __restrict__ queue, float4* __restrict__ partial,
                                                      unsigned long long* __restrict__ pdep, PairRecord rec) {
    gather_chunks(H, SH, G, gp, order, nvalidp, tiles, ntiles, chunk_off, chunkp, queue, partial, pdep, rec);
}

// The gather chunks of every scene of a batch (see k_photon_multi).
__global__ void __launch_bounds__(64, GATHER_WAVES_PER_SIMD) k_gather_chunks_multi(const GatherJob* __restrict__ jobs, int njobs,
                                                            const int* __restrict__ bstart) {
    const int home = home_scene(bstart, njobs);
    for (int k = 0; k < njobs; ++k) {
        const int s = __builtin_amdgcn_readfirstlane(home + k < njobs ? home + k : home + k - njobs);
        const GatherJob& J = jobs[s];
        if (scene_done(J.queue)) continue;
        gather_chunks(J.H, J.SH, J.G, J.gp, J.order, J.nvalid, J.tiles, J.ntiles, J.chunk_off, J.chunkp, J.queue,
                      J.partial, J.pdep, J.rec);
        scene_mark_done(J.queue);
    }
}

__global__ void __launch_bounds__(64) k_gather_finish(HitDev H, SharedDev SH, const uint32_t* __restrict__ order,
                                                      const int* __restrict__ nvalidp, float alpha,
                                                      const int* __restrict__ chunk_off,
                                                      const float4* __restrict__ partial, int* __restrict__ counts,
                                                      unsigned long long* __restrict__ gstat,
                                                      uint32_t* __restrict__ flags) {
    const int lane = threadIdx.x;
    const int t = blockIdx.x;
    const int p = lane < GATHER_GROUP ? group_pixel(order, *nvalidp, t * GATHER_GROUP + lane) : -1;
    float4 hp = make_float4(0, 0, 0, 0), sr = make_float4(0, 0, 0, 0);
    if (p >= 0) {
        hp = H.pos[p];
        sr = SH.rad[p];
    }
    const bool active = p >= 0 && (int)hp.w == 1;
    f3 add = mk(0.0f, 0.0f, 0.0f);
    int M = 0;
#pragma unroll 4
    for (int c = chunk_off[t]; c < chunk_off[t + 1]; ++c) {
        if (lane < GATHER_GROUP) {
            float4 v = partial[(size_t)c * GATHER_GROUP + lane];
            add = add + mk(v.x, v.y, v.z);
            M += __float_as_int(v.w);
        }
    }
    float rnew = sr.w;
    if (active && M > 0) {
        int2 cn = SH.cnt[p];
        const float num = (float)cn.x + alpha * (float)M;
        const float den = (float)(cn.x + M);
        f3 r = mk(sr.x, sr.y, sr.z);
        r = ((r + add) * num) / den;
        const float R1 = sr.w * sqrtf_(num / den);
        SH.rad[p] = make_float4(r.x, r.y, r.z, R1);
        SH.cnt[p] = make_int2((int)((float)cn.x + (float)M * alpha), cn.y + 1);
        rnew = R1;
        if (!(r.x == r.x) || !(r.y == r.y) || !(r.z == r.z) || __builtin_isinf(r.x) || __builtin_isinf(r.y) ||
            __builtin_isinf(r.z))
            atomicOr(flags, 1u);
    }
    if (p >= 0 && counts) counts[p] = active ? M : 0;
    // mean radius of the hit points (fixed point, 2^-32 units) -> next cell size
    unsigned long long rs = active ? (unsigned long long)((double)rnew * 4294967296.0) : 0ull;
    unsigned long long rc = active ? 1ull : 0ull;
    unsigned long long tn = (unsigned long long)M;
    for (int off = 32; off > 0; off >>= 1) {
        rs += __shfl_xor(rs, off, 64);
        rc += __shfl_xor(rc, off, 64);
        tn += __shfl_xor(tn, off, 64);
    }
    // per-group sums, added up by k_group_stats (not one atomic per group on
    // the same four addresses)
    if (lane == 0) {
        gstat[3 * t] = rs;
        gstat[3 * t + 1] = rc;
        gstat[3 * t + 2] = tn;
    }
}

// Sums the per-group statistics of k_gather_finish: radius sum, active hit
// points (the next pass's cell size) and the counters.  One workgroup; the
// only writer of rstat and of these counters in the pass.
__global__ void __launch_bounds__(1024) k_group_stats(const unsigned long long* __restrict__ gstat, int ngroups,
                                                      unsigned long long* __restrict__ rstat,
                                                      Counters* __restrict__ ctr) {
    __shared__ unsigned long long part[3][16];
    unsigned long long rs = 0, rc = 0, tn = 0;
    for (int g = threadIdx.x; g < ngroups; g += 1024) {
        rs += gstat[3 * g];
        rc += gstat[3 * g + 1];
        tn += gstat[3 * g + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
        rs += __shfl_xor(rs, off, 64);
        rc += __shfl_xor(rc, off, 64);
        tn += __shfl_xor(tn, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = rs;
        part[1][w] = rc;
        part[2][w] = tn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0, c = 0;
        for (int k = 0; k < 16; ++k) {
            a += part[0][k];
            b += part[1][k];
            c += part[2][k];
        }
        rstat[0] += a;
        rstat[1] += b;
        ctr->hit_points += b;
        ctr->neighbours += c;
    }
}

// DirectLight of every hit point that is not NONE (SPPM.cpp:404-418); only
// an environment light makes it nonzero.  norm2() > 1e-8 compares the float
// norm against a double.
__global__ void k_direct_light(HitDev H, SharedDev SH, int np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if ((int)H.pos[i].w == 0) return;
    const float4 d = H.direct[i];
    const f3 dv = mk(d.x, d.y, d.z);
    if (!(norm2(dv) > LO_1EM8)) return;  // (double)norm2 > 1e-8
    float4 s = SH.direct[i];
    SH.direct[i] = make_float4(s.x + dv.x, s.y + dv.y, s.z + dv.z, __int_as_float(__float_as_int(s.w) + 1));
}
void launch_direct_light(const HitDev& H, const SharedDev& SH, int np, hipStream_t st) {
    hipLaunchKernelGGL(k_direct_light, dim3((np + 255) / 256), dim3(256), 0, st, H, SH, np);
}

// reconstructImageFromSharedPoints (SPPM.cpp:473-493)
__global__ void k_reconstruct(SharedDev SH, int np, long long total_photons, int passes, float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 sr = SH.rad[i];
    int2 cn = SH.cnt[i];
    const float4 d4 = SH.direct[i];
    const f3 dl = mk(d4.x, d4.y, d4.z) / (float)((float)__float_as_int(d4.w) + 0.0001f);
    float den = ((float)total_photons * ((float)cn.y + 0.0001f) / (0.0001f + (float)passes)) * SPPM_PI * sr.w * sr.w;
    f3 L = dl + mk(sr.x, sr.y, sr.z) / den;
    const float b = 1.2f;
    out[i] = make_float4(powf(L.x, 1.0f / b), powf(L.y, 1.0f / b), powf(L.z, 1.0f / b), 1.0f);
}

__global__ void k_init_shared(SharedDev SH, int np, float r0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    SH.rad[i] = make_float4(0.0f, 0.0f, 0.0f, r0);
    SH.cnt[i] = make_int2(0, 0);
    SH.direct[i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
}
__global__ void k_init_slots(SlotsDev s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)NSLOT * s.npaths) return;
    s.valid[i] = 0;
}
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// Batch closest hit (N_BVH::ClosestHit) for queries: hits are accepted in
// (tmin, tmax) (bvh8.h:591, 640: the MaxT test in double); a miss gives facet
// -1, t = FLT_MAX, u = v = 0.  t, u, v may be NULL.
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK) k_trace(SceneDev S, int n, const float* __restrict__ org,
                                                      const float* __restrict__ dir, float tmin, float tmax, int cap,
                                                      uint32_t* __restrict__ spill, int* __restrict__ facet,
                                                      float* __restrict__ t, float* __restrict__ u,
                                                      float* __restrict__ v, uint32_t* __restrict__ flags) {
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    uint32_t fl = 0;
    for (int i = blockIdx.x * PATH_BLOCK + lane; i < n; i += gridDim.x * PATH_BLOCK) {
        HitRec h;
        f3 o = mk(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        f3 d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        bool hit = closest_hit<W>(S, o, d, tmin, ts, h, &fl);
        hit = hit && (double)h.t < (double)tmax - 1e-10;
        facet[i] = hit ? h.facet : -1;
        if (t) t[i] = hit ? h.t : 3.40282347e38f;
        if (u) u[i] = hit ? h.u : 0.0f;
        if (v) v[i] = hit ? h.v : 0.0f;
    }
    if (fl) atomicOr(flags, fl);
}

// Scene::IsOccluded for a batch of rays (MaxDist per ray).
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK) k_occluded(SceneDev S, int n, const float* __restrict__ org,
                                                         const float* __restrict__ dir, float tmin,
                                                         const float* __restrict__ tmax, float transl, int cap,
                                                         uint32_t* __restrict__ spill,
                                                         int* __restrict__ occluded, uint32_t* __restrict__ flags) {
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    uint32_t fl = 0;
    for (int i = blockIdx.x * PATH_BLOCK + lane; i < n; i += gridDim.x * PATH_BLOCK) {
        const f3 o = mk(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const f3 d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        occluded[i] = occluded_ray<W>(S, o, d, tmin, tmax[i], transl, ts, fl) ? 1 : 0;
    }
    if (fl) atomicOr(flags, fl);
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
// Blocks of the path kernels: enough resident waves to fill the device (the
// traversal spill area is sized for this many lanes).
int trace_grid_blocks(int width) {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e;
    if (width == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photon<2, false>, PATH_BLOCK, 0);
    else if (width == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photon<4, false>, PATH_BLOCK, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_photon<8, false>, PATH_BLOCK, 0);
    if (e != hipSuccess || nb < 1) nb = 4;
    return cus * nb;
}

void launch_camera_words(uint32_t* states, const uint32_t* phi, int K, long long L, uint32_t* words, hipStream_t st) {
    uint32_t* ends = states + (size_t)K * 624;  // scratch: the windows after the segments
    hipLaunchKernelGGL(k_camw_emit, dim3(K), dim3(64), 0, st, states, L, words, ends);
    hipLaunchKernelGGL(k_camw_jump, dim3(K), dim3(CAMW_JUMP_THREADS), 0, st, ends, phi, states);
}
void launch_camera_words_multi(const CamJob* jobs, const int* kstart, int njobs, int nseg, hipStream_t st) {
    hipLaunchKernelGGL(k_camw_emit_multi, dim3(nseg), dim3(64), 0, st, jobs, kstart, njobs);
    hipLaunchKernelGGL(k_camw_jump_multi, dim3(nseg), dim3(CAMW_JUMP_THREADS), 0, st, jobs, kstart, njobs);
}

void launch_photon(const SceneDev& S, const SlotsDev& slots, int N, int step, const int* step_dev,
                   const TraceSpill& sp, int* next_path, const GuideEmit* guide, uint32_t* flags,
                   unsigned long long* rays, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3(STRIPE_INTS / 256), dim3(256), 0, st, next_path, STRIPE_INTS);
    dim3 g(sp.blocks), b(PATH_BLOCK);
    const GuideEmit G = guide ? *guide : GuideEmit{};
#define SPPM_PHOTON(WW, GG) \
    hipLaunchKernelGGL((k_photon<WW, GG>), g, b, 0, st, S, slots, N, step, step_dev, sp.cap, sp.buf, next_path, G, \
                       flags, rays)
    if (guide) {
        if (S.width == 2) SPPM_PHOTON(2, true);
        else if (S.width == 4) SPPM_PHOTON(4, true);
        else SPPM_PHOTON(8, true);
    } else {
        if (S.width == 2) SPPM_PHOTON(2, false);
        else if (S.width == 4) SPPM_PHOTON(4, false);
        else SPPM_PHOTON(8, false);
    }
#undef SPPM_PHOTON
}

template <bool NEE>
static void launch_eye_t(const SceneDev& S, const CamDev& C, const HitDev& H, int step, const int* step_dev,
                         const TraceSpill& sp, const uint32_t* words, uint32_t* flags, unsigned long long* rays,
                         int* queue, int keep, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3(STRIPE_INTS / 256), dim3(256), 0, st, queue, STRIPE_INTS);
    dim3 g(sp.blocks), b(PATH_BLOCK);
    if (S.width == 2)
        hipLaunchKernelGGL((k_eye<2, NEE>), g, b, 0, st, S, C, H, step, step_dev, sp.cap, sp.buf, words, flags, keep,
                           rays, queue);
    else if (S.width == 4)
        hipLaunchKernelGGL((k_eye<4, NEE>), g, b, 0, st, S, C, H, step, step_dev, sp.cap, sp.buf, words, flags, keep,
                           rays, queue);
    else
        hipLaunchKernelGGL((k_eye<8, NEE>), g, b, 0, st, S, C, H, step, step_dev, sp.cap, sp.buf, words, flags, keep,
                           rays, queue);
}
void launch_eye(const SceneDev& S, const CamDev& C, const HitDev& H, int step, const int* step_dev,
                const TraceSpill& sp, const uint32_t* words, uint32_t* flags, unsigned long long* rays, int* queue,
                int nee, hipStream_t st) {
    if (nee == SPPM_NEE_OFF) launch_eye_t<false>(S, C, H, step, step_dev, sp, words, flags, rays, queue, 0, st);
    else
        launch_eye_t<true>(S, C, H, step, step_dev, sp, words, flags, rays, queue, nee == SPPM_NEE_KEEP ? 1 : 0, st);
}

__global__ void k_add_i32(int* __restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}
void launch_add_i32(int* p, int v, hipStream_t st) { hipLaunchKernelGGL(k_add_i32, dim3(1), dim3(64), 0, st, p, v); }
void launch_zero_i32(int* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

void launch_trace(const SceneDev& S, int n, const float* org, const float* dir, float tmin, float tmax,
                  const TraceSpill& sp, int* facet, float* t, float* u, float* v, uint32_t* flags, hipStream_t st) {
    dim3 g(sp.blocks), b(PATH_BLOCK);
#define SPPM_TRACE(WW) \
    hipLaunchKernelGGL(k_trace<WW>, g, b, 0, st, S, n, org, dir, tmin, tmax, sp.cap, sp.buf, facet, t, u, v, flags)
    if (S.width == 2) SPPM_TRACE(2);
    else if (S.width == 4) SPPM_TRACE(4);
    else SPPM_TRACE(8);
#undef SPPM_TRACE
}

void launch_occluded(const SceneDev& S, int n, const float* org, const float* dir, float tmin, const float* tmax,
                     float transl,
                     const TraceSpill& sp, int* occluded, uint32_t* flags, hipStream_t st) {
    dim3 g(sp.blocks), b(PATH_BLOCK);
    if (S.width == 2)
        hipLaunchKernelGGL(k_occluded<2>, g, b, 0, st, S, n, org, dir, tmin, tmax, transl, sp.cap, sp.buf, occluded, flags);
    else if (S.width == 4)
        hipLaunchKernelGGL(k_occluded<4>, g, b, 0, st, S, n, org, dir, tmin, tmax, transl, sp.cap, sp.buf, occluded, flags);
    else
        hipLaunchKernelGGL(k_occluded<8>, g, b, 0, st, S, n, org, dir, tmin, tmax, transl, sp.cap, sp.buf, occluded, flags);
}

void launch_photon_stage(const SlotsDev& slots, const int* off, const StagedPhotons& sp, hipStream_t st) {
    hipLaunchKernelGGL(k_photon_stage, dim3((slots.npaths + 255) / 256), dim3(256), 0, st, slots, off, sp);
}

void launch_build_grid(const StagedPhotons& sp, const int* nstored, Counters* ctr, GridParams* gp,
                       const unsigned long long* rstat, const float* bmin, const float* ext, int cell_cap,
                       int* cell_start, int4* gaps, int* ngap, const PhotonSortBuffers& sb, int photon_cap,
                       float4* gpos, float4* gflux, float4* gdir, hipStream_t st) {
    hipLaunchKernelGGL(k_grid_params, dim3(1), dim3(1), 0, st, gp, rstat, bmin[0], bmin[1], bmin[2], ext[0],
                       ext[1], ext[2], cell_cap, nstored, ctr, ngap);
    hipLaunchKernelGGL(k_photon_keys, dim3(2048), dim3(256), 0, st, sp.rec, nstored, gp, sb);
    // stable sort of the flattened photons by cell: cell keys are below
    // cell_cap = 2^22, the flattened order is kept inside a cell
    const SortResult sr =
        radix_sort_pairs(sb.keys, sb.vals, sb.tmp_k, sb.tmp_v, sb.hist, photon_cap, PHOTON_KEY_BITS, st, nstored);
    hipLaunchKernelGGL(k_cell_starts, dim3(2048), dim3(256), 0, st, sr.keys, nstored, gp, cell_start, gaps, ngap);
    hipLaunchKernelGGL(k_fill_gaps, dim3(1024), dim3(256), 0, st, gaps, ngap, cell_start);
    hipLaunchKernelGGL(k_photon_sorted, dim3(2048), dim3(256), 0, st, sp, nstored, sr.vals, gpos, gflux, gdir);
}

void launch_photon_flat(const SlotsDev& slots, const int* off, float4* fpos, float4* fflux, float4* fdir,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_photon_flat, dim3((slots.npaths + 255) / 256), dim3(256), 0, st, slots, off, fpos, fflux,
                       fdir);
}

void launch_gather(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int W, int Hh,
                   float alpha, int mode, int* counts, const int* list_off, int* lists, unsigned long long* rstat,
                   Counters* ctr, uint32_t* flags, hipStream_t st) {
    int tiles = ((W + 7) >> 3) * ((Hh + 7) >> 3);
    hipLaunchKernelGGL(k_gather, dim3(tiles), dim3(64), 0, st, H, SH, G,
                       gp, W, Hh, alpha, mode, counts, list_off, lists, rstat, ctr, flags);
}

const uint32_t* launch_gather_plan(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp,
                                   int np, const float* bmin, const float* ext, const GroupBuffers& gb,
                                   TileRange* tiles, int cap, int* chunkp, int* nch, int* chunk_off, int* scan_tmp,
                                   int* counts, Counters* ctr, const PairRecord& rec, hipStream_t st) {
    const int ntiles = gather_group_count(np);
    // hit-point groups: valid hit points in Morton order, GATHER_GROUP per group
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, gb.nvalid, 1);  // valid count
    hipLaunchKernelGGL(k_zero_i32, dim3(STRIPE_INTS / 256), dim3(256), 0, st, gb.chunkq, STRIPE_INTS);
    if (counts) hipLaunchKernelGGL(k_zero_i32, dim3((np + 255) / 256), dim3(256), 0, st, counts, np);
    hipLaunchKernelGGL(k_hit_keys, dim3((np + 255) / 256), dim3(256), 0, st, H, np, bmin[0], bmin[1], bmin[2], ext[0],
                       ext[1], ext[2], gb.keys, gb.vals);
    const SortResult sr = radix_sort_pairs(gb.keys, gb.vals, gb.tmp_k, gb.tmp_v, gb.hist, np, 31, st);
    hipLaunchKernelGGL(k_count_valid, dim3(1), dim3(64), 0, st, sr.keys, np, gb.nvalid);
    const uint32_t* order = sr.vals;  // == gather_order(gb, np)
    hipLaunchKernelGGL(k_gather_plan, dim3(ntiles), dim3(64), 0, st, H, SH, G, gp, order, gb.nvalid, tiles);
    hipLaunchKernelGGL(k_chunk_size, dim3(1), dim3(256), 0, st, tiles, ntiles, cap, chunkp, ctr);
    hipLaunchKernelGGL(k_tile_chunks, dim3((ntiles + 1 + 255) / 256), dim3(256), 0, st, tiles, ntiles, chunkp, nch);
    exclusive_scan(nch, chunk_off, ntiles + 1, nullptr, scan_tmp, st);
    if (rec.pairs) launch_set_u64x2(rec.count, 0ull, 0ull, st);
    return order;
}
void launch_gather_run(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int np,
                       const uint32_t* order, const GroupBuffers& gb, const TileRange* tiles, const int* chunkp,
                       const int* chunk_off, float4* partial, unsigned long long* pdep, const PairRecord& rec,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_gather_chunks, dim3(GATHER_WAVES), dim3(64), 0, st, H, SH, G, gp, order, gb.nvalid, tiles,
                       gather_group_count(np), chunk_off, chunkp, gb.chunkq, partial, pdep, rec);
}
void launch_gather_finish(const HitDev& H, const SharedDev& SH, int np, const uint32_t* order, const GroupBuffers& gb,
                          float alpha, const int* chunk_off, const float4* partial, int* counts,
                          unsigned long long* rstat, Counters* ctr, uint32_t* flags, hipStream_t st) {
    const int ntiles = gather_group_count(np);
    hipLaunchKernelGGL(k_gather_finish, dim3(ntiles), dim3(64), 0, st, H, SH, order, gb.nvalid, alpha, chunk_off,
                       partial, counts, gb.stat, flags);
    hipLaunchKernelGGL(k_group_stats, dim3(1), dim3(1024), 0, st, gb.stat, ntiles, rstat, ctr);
}
void launch_gather_balanced(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int np,
                            const float* bmin, const float* ext, const GroupBuffers& gb, float alpha, TileRange* tiles,
                            int cap, int* chunkp, int* nch, int* chunk_off, int* scan_tmp, float4* partial,
                            int* counts, unsigned long long* rstat, Counters* ctr, uint32_t* flags,
                            unsigned long long* pdep,
                            const PairRecord& rec, hipStream_t st) {
    const uint32_t* order = launch_gather_plan(H, SH, G, gp, np, bmin, ext, gb, tiles, cap, chunkp, nch, chunk_off,
                                               scan_tmp, counts, ctr, rec, st);
    launch_gather_run(H, SH, G, gp, np, order, gb, tiles, chunkp, chunk_off, partial, pdep, rec, st);
    launch_gather_finish(H, SH, np, order, gb, alpha, chunk_off, partial, counts, rstat, ctr, flags, st);
}

void launch_photon_multi(int width, bool guided, const PhotonJob* jobs, int njobs, const int* bstart,
                         const TraceSpill& sp, hipStream_t st) {
    dim3 g(sp.blocks), b(PATH_BLOCK);
#define SPPM_PM(WW, GG) hipLaunchKernelGGL((k_photon_multi<WW, GG>), g, b, 0, st, jobs, njobs, bstart, sp.cap, sp.buf)
    if (guided) {
        if (width == 2) SPPM_PM(2, true);
        else if (width == 4) SPPM_PM(4, true);
        else SPPM_PM(8, true);
    } else {
        if (width == 2) SPPM_PM(2, false);
        else if (width == 4) SPPM_PM(4, false);
        else SPPM_PM(8, false);
    }
#undef SPPM_PM
}
void launch_eye_multi(int width, bool nee, const EyeJob* jobs, int njobs, const int* bstart, const TraceSpill& sp,
                      hipStream_t st) {
    dim3 g(sp.blocks), b(PATH_BLOCK);
#define SPPM_EM(WW, NN) hipLaunchKernelGGL((k_eye_multi<WW, NN>), g, b, 0, st, jobs, njobs, bstart, sp.cap, sp.buf)
    if (nee) {
        if (width == 2) SPPM_EM(2, true);
        else if (width == 4) SPPM_EM(4, true);
        else SPPM_EM(8, true);
    } else {
        if (width == 2) SPPM_EM(2, false);
        else if (width == 4) SPPM_EM(4, false);
        else SPPM_EM(8, false);
    }
#undef SPPM_EM
}
void launch_gather_chunks_multi(const GatherJob* jobs, int njobs, const int* bstart, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_chunks_multi, dim3(GATHER_WAVES), dim3(64), 0, st, jobs, njobs, bstart);
}
int gather_wave_count() { return GATHER_WAVES; }
const uint32_t* gather_order(const GroupBuffers& gb, int np) {
    return radix_sort_result(gb.keys, gb.vals, gb.tmp_k, gb.tmp_v, np, 31).vals;
}
int gather_tile_count(int W, int Hh) { return ((W + 7) >> 3) * ((Hh + 7) >> 3); }
int gather_group_count(int np) { return (np + GATHER_GROUP - 1) / GATHER_GROUP; }
size_t gather_tile_bytes() { return sizeof(TileRange); }

void launch_reconstruct(const SharedDev& SH, int np, long long total, int passes, float4* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reconstruct, dim3((np + 255) / 256), dim3(256), 0, st, SH, np, total, passes, out);
}
void launch_init_shared(const SharedDev& SH, int np, float r0, hipStream_t st) {
    hipLaunchKernelGGL(k_init_shared, dim3((np + 255) / 256), dim3(256), 0, st, SH, np, r0);
}
void launch_init_slots(const SlotsDev& s, hipStream_t st) {
    size_t n = (size_t)NSLOT * s.npaths;
    hipLaunchKernelGGL(k_init_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s);
}
__global__ void k_set_u64x2(unsigned long long* p, unsigned long long a, unsigned long long b) {
    p[0] = a;
    p[1] = b;
}
void launch_set_u64x2(unsigned long long* p, unsigned long long a, unsigned long long b, hipStream_t st) {
    hipLaunchKernelGGL(k_set_u64x2, dim3(1), dim3(1), 0, st, p, a, b);
}
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, p, v);
}

}  // namespace sppm

namespace sppm {
// ---- guided emission: training records ----------------------------------
// Deposit per flattened photon -> per path -> the histogram of
// the paths' emission cells (float atomics: the summation order, and with it
// the last bits of the histogram, vary from run to run).
// Per path: its flattened photons' deposits (offsets of this pass's build;
// the path's last photon ends at the next path's offset, the last path's at
// the stored count) into the histogram cell it was emitted from (the block of
// path_cell the photon pass of this step wrote).  The slot counts are not
// read: the next pass's photon pass may already be rewriting them.
// ratio (optional, q / p_plain per path): deposit d^2 q / p instead of d (the
// RMS target's histogram).
__global__ void k_guide_deposit(const int* __restrict__ off, const int* __restrict__ nstored,
                                const unsigned long long* __restrict__ pdep, const int* __restrict__ path_cell,
                                const float* __restrict__ ratio, int N, int step_arg,
                                const int* __restrict__ step_dev, unsigned long long* __restrict__ hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int step = step_dev ? *step_dev + 1 : step_arg;
    const int o = off[i], e = i + 1 < N ? off[i + 1] : *nstored;
    unsigned long long sum = 0;
    for (int k = o; k < e; ++k) sum += pdep[k];
    float d = (float)sum * GUIDE_FIX_INV;
    const size_t slot = (size_t)(step & 1) * (size_t)N + i;
    if (ratio) d = d * d * ratio[slot];
    if (d > 0.0f && d < 3.0e38f) atomicAdd(&hist[path_cell[slot]], guide_fix(d));
}
__global__ void k_zero_u64_dev(unsigned long long* __restrict__ p, const int* __restrict__ n) {
    const int cnt = *n;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) p[k] = 0ull;
}
// The offsets and stored count of this pass's photons (dst: N + 1 entries),
// for readers that run while the next pass stages its own.
__global__ void k_copy_offsets(const int* __restrict__ off, const int* __restrict__ nstored, int N,
                               int* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = off[i];
    if (i == N) dst[N] = *nstored;
}
void launch_copy_offsets(const int* off, const int* nstored, int N, int* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_offsets, dim3((N + 1 + 255) / 256), dim3(256), 0, st, off, nstored, N, dst);
}
void launch_deposit_clear(unsigned long long* pdep, const int* nstored, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64_dev, dim3(2048), dim3(256), 0, st, pdep, nstored);
}
void launch_guide_deposit(const int* off, const int* nstored, const unsigned long long* pdep, const int* path_cell,
                          const float* ratio, int N, int step, const int* step_dev, unsigned long long* hist,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_guide_deposit, dim3((N + 255) / 256), dim3(256), 0, st, off, nstored, pdep, path_cell, ratio,
                       N, step, step_dev, hist);
}
}  // namespace sppm
