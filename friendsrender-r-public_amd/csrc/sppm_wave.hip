// sppm_wave.hip -- the photon pass (createPhotonMap, SPPM.cpp:495-680) as
// wavefront kernels over queues of rays.
//
//   k_pw_emit       one lane per path: emission (plain or guided), the
//                   slot-0 photon, the path state and the path's first ray
//                   (queue 0 holds path i at entry i)
//
// A ray queue is STRIPES partitions of P = ceil(N / STRIPES) entries, each
// with its own size and claim counter (256 bytes apart: atomics on one
// address serialise).  Partition s of queue 0 holds paths [s P, (s + 1) P);
// shading appends a path's next ray to the same partition of the next queue,
// which therefore never holds more than P entries.
//   k_pw_trace<W,L> persistent closest hit: L lanes per ray.  At an inner
//                   node the L lanes of a ray split the W child boxes, so a
//                   ray's group fetches the W-wide node in one coalesced
//                   read; at a leaf they split its triangles.  A group takes
//                   its next ray from a wave-local pool (refilled with one
//                   atomic per chunk) as soon as its ray ends, so lanes do
//                   not wait for the longest ray of the wave.
//   k_pw_shade      one lane per queued ray: the loop body after the trace
//                   (photon_shade), the checks before the next trace, and
//                   either the path's next ray appended to the next queue or
//                   the path's end
//   k_pw_resume<W>  after the wavefront bounces, the one-lane-per-path loop
//                   of the megakernel for the paths still alive
//
// A path's arithmetic is the megakernel's (the same device functions), and
// the closest hit keeps its rule (nearest t, ties to the lowest facet id), so
// the stored photons do not depend on which kernel, lane or queue position
// handled a ray.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_photon.h"
#include "sppm_stripes.h"
#include "sppm_wave.h"

namespace sppm {

constexpr int WT_BLOCK = 256;              // threads per block of the trace kernel
constexpr int WT_WAVES = WT_BLOCK / 64;
constexpr int WT_LDS = 32;                 // stack entries per ray kept in LDS (deeper ones spill)
#ifndef SPPM_WT_MIN_WAVES
#define SPPM_WT_MIN_WAVES 6
#endif
constexpr int WT_MIN_WAVES = SPPM_WT_MIN_WAVES;  // waves per SIMD the trace kernel is compiled for (<= 80 VGPRs)
constexpr int HIT_TIE = 0x40000000;        // hit record: tie flag beside the leaf-triangle slot
constexpr float F_MAX = 3.40282347e38f;    // Ray::m_MaxT

// ---- path state ----------------------------------------------------------
// MT cursor: ia = 1 + k and ib = 397 + k hold for every cursor (mt_seed
// starts at k = 0, ia = 1, ib = 397 and mt_next advances all three).
__device__ __forceinline__ void path_store(const PathStateDev& Q, const PhotonPath& P) {
    const int i = P.i;
    Q.rng[i] = make_uint4(P.m.a0, P.m.a1, P.m.b, P.m.k);
    Q.org[i] = make_float4(P.cur_org.x, P.cur_org.y, P.cur_org.z, P.cur_index);
    Q.dir[i] = make_float4(P.cur_dir.x, P.cur_dir.y, P.cur_dir.z, P.next_index);
    Q.ndir[i] = make_float4(P.next_dir.x, P.next_dir.y, P.next_dir.z, P.PDF);
    Q.rad[i] = make_float4(P.rad.x, P.rad.y, P.rad.z,
                           __uint_as_float((uint32_t)P.pidx | ((uint32_t)P.loopcount << 8) | ((uint32_t)P.depth << 16)));
    Q.thr[i] = make_float4(P.T.x, P.T.y, P.T.z, 0.0f);
}
__device__ __forceinline__ void path_load(const PathStateDev& Q, int i, PhotonPath& P) {
    P.i = i;
    const uint4 m = Q.rng[i];
    P.m.a0 = m.x; P.m.a1 = m.y; P.m.b = m.z; P.m.k = m.w;
    P.m.ia = 1u + m.w; P.m.ib = 397u + m.w;
    const float4 o = Q.org[i], d = Q.dir[i], n = Q.ndir[i], r = Q.rad[i], t = Q.thr[i];
    P.cur_org = mk(o.x, o.y, o.z); P.cur_index = o.w;
    P.cur_dir = mk(d.x, d.y, d.z); P.next_index = d.w;
    P.next_dir = mk(n.x, n.y, n.z); P.PDF = n.w;
    P.rad = mk(r.x, r.y, r.z);
    const uint32_t c = __float_as_uint(r.w);
    P.pidx = (int)(c & 0xffu); P.loopcount = (int)((c >> 8) & 0xffu); P.depth = (int)(c >> 16);
    P.T = mk(t.x, t.y, t.z);
}
__device__ __forceinline__ void ray_store(float4* __restrict__ rays, int q, const PhotonPath& P) {
    rays[2 * (size_t)q] = make_float4(P.cur_org.x, P.cur_org.y, P.cur_org.z, __int_as_float(P.i));
    rays[2 * (size_t)q + 1] = make_float4(P.cur_dir.x, P.cur_dir.y, P.cur_dir.z, 0.0f);
}

// ---- emission ------------------------------------------------------------
// A fresh path always passes the checks before its first trace (loop count
// 0, no photon stored yet, throughput 1), so queue 0 is the identity.
template <bool GUIDED>
__global__ void __launch_bounds__(256) k_pw_emit(SceneDev S, SlotsDev slots, int N, int step_arg,
                                                 const int* __restrict__ step_dev, GuideEmit G, PathStateDev Q,
                                                 float4* __restrict__ rays, int* __restrict__ qctl, int nctl, int Pq,
                                                 uint32_t* __restrict__ flags) {
    const int step = step_dev ? *step_dev + 1 : step_arg;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // queue sizes and claim counters: queue 0's partitions full, the rest empty
    for (int k = i; k < 2 * nctl * STRIPE_INTS; k += gridDim.x * blockDim.x) {
        int v = 0;
        if (k < STRIPE_INTS && k % STRIPE_PAD == 0) {
            const int left = N - (k / STRIPE_PAD) * Pq;
            v = left < 0 ? 0 : (left < Pq ? left : Pq);
        }
        qctl[k] = v;
    }
    if (i >= N) return;
    uint32_t ovf = 0;
    PhotonPath P;
    if constexpr (GUIDED) photon_start_guided(S, slots, G, N, step, i, P, &ovf);
    else photon_start(S, slots, N, step, i, P, &ovf);
    photon_pretrace(P);
    path_store(Q, P);
    ray_store(rays, i, P);
    if (ovf) atomicOr(flags, 2u);
}

// ---- closest hit, L lanes per ray -----------------------------------------
// The result equals closest_hit<W>'s: accepted hits (t > tmin) are kept by
// nearest t, then lowest facet id, and the tie flag says that two facets were
// hit at the final t; neither depends on the order in which nodes or
// triangles are tested.  Child boxes are tested as closest_hit does.
// At an inner node lane c of a ray's group tests children c*CPL .. c*CPL +
// CPL - 1 (CPL = W / L, one vector load per row of the node); the hit
// children are ranked nearest first (equal distances by child slot), rank 0
// is descended into and the others are pushed far-to-near.  At a leaf lane c
// tests triangles c, c + L, ...
// Stack: entry s of a ray in LDS for s < WT_LDS, else in the spill area at
// spill[(s - WT_LDS) * gstride + group]; entries past cap are dropped and
// flagged (bit 2), and the ray ends when it would pop one, as closest_hit.
// A wave claims queue entries in chunks (one atomic per chunk, through the
// partitions' counters) sized so that a wave claims a few times per launch:
// large chunks for a full queue, one ray per group for a short one.
template <int W, int L>
__global__ void __launch_bounds__(WT_BLOCK, WT_MIN_WAVES) k_pw_trace(
    SceneDev S, const float4* __restrict__ rays, const int* __restrict__ qcnt, int* __restrict__ fetch, int P,
    float4* __restrict__ hits, float tmin, int cap, uint32_t* __restrict__ spill, uint32_t* __restrict__ flags,
    unsigned long long* __restrict__ counts) {
    constexpr int G = 64 / L;     // rays per wave
    constexpr int CPL = W / L;    // children per lane at an inner node
    constexpr int NF = node_floats(W);
    __shared__ uint32_t stk[WT_WAVES * G][WT_LDS + 1];  // +1: a group's entries start in another bank
    const int lane = threadIdx.x & 63;
    const int c = lane & (L - 1);
    const int gl = lane & ~(L - 1);  // the group's first lane
    const int grp = (int)(threadIdx.x >> 6) * G + lane / L;
    uint32_t* const ls = &stk[grp][0];
    const size_t gstride = (size_t)gridDim.x * (WT_WAVES * G);
    uint32_t* const gs = spill + (size_t)blockIdx.x * (WT_WAVES * G) + grp;
    const uint64_t gmask = ((L == 64) ? ~0ull : ((1ull << L) - 1ull)) << gl;
    const uint64_t below_g = (1ull << gl) - 1ull;
    StripeCursor cur = stripe_start();
    const int per_part = (int)gridDim.x * WT_WAVES / STRIPES;
    int chunk = qcnt[cur.s * STRIPE_PAD] / ((per_part > 0 ? per_part : 1) * 2);
    chunk = chunk < G ? G : (chunk > 4 * G ? 4 * G : chunk);
    uint32_t fl = 0, nray = 0, ntie = 0;
    bool active = false, exhausted = false, nomore = false;
    int pbase = 0, pend = 0;  // the wave's pool of queue entries [pbase, pend)
    int q = 0, sp = 0;
    uint32_t ref = 0;
    f3 org = mk(0.0f, 0.0f, 0.0f), dir = org, inv = org, oi = org;
    float best = F_MAX, bu = 0.0f, bv = 0.0f;
    int bfacet = 0x7fffffff, bslot = -1;
    bool tie = false;
    for (;;) {
        // groups whose ray has ended take the next entries of the pool
        const uint64_t im = __ballot(!active && c == 0);
        const int nidle = __popcll(im);
        if (nidle > 0 && !exhausted) {
            const int avail = pend - pbase;
            const bool need = avail < nidle;
            int nb = 0, got = 0;
            if (need && !nomore) {
                got = part_claim(qcnt, fetch, P, cur, chunk, &nb);
                if (got == 0) nomore = true;
            }
            if (!active) {
                const int r = __popcll(im & below_g);
                const int qq = r < avail ? pbase + r : (r - avail < got ? nb + (r - avail) : -1);
                if (qq >= 0) {
                    const float4 a = rays[2 * (size_t)qq], b = rays[2 * (size_t)qq + 1];
                    org = mk(a.x, a.y, a.z);
                    dir = mk(b.x, b.y, b.z);
                    const float BIG = 1e16f;
                    inv.x = fabs_(dir.x) < 1e-16f ? (dir.x > 0.0f ? BIG : -BIG) : 1.0f / dir.x;
                    inv.y = fabs_(dir.y) < 1e-16f ? (dir.y > 0.0f ? BIG : -BIG) : 1.0f / dir.y;
                    inv.z = fabs_(dir.z) < 1e-16f ? (dir.z > 0.0f ? BIG : -BIG) : 1.0f / dir.z;
                    oi = mk(org.x * inv.x, org.y * inv.y, org.z * inv.z);
                    best = F_MAX; bu = 0.0f; bv = 0.0f; bfacet = 0x7fffffff; bslot = -1; tie = false;
                    sp = 0;
                    ref = 0;  // root
                    q = qq;
                    active = true;
                }
            }
            if (need) {  // the old pool is used up; the new one starts after the entries just taken
                const int used = nidle - avail < got ? nidle - avail : got;
                pbase = nb + used;
                pend = nb + got;
            } else {
                pbase += nidle;
            }
            if (nomore && pbase >= pend) exhausted = true;
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (active) {
            bool pop = false;
            if (ref & LEAF_BIT) {
                const int start = (int)((ref & ~LEAF_BIT) >> 3);
                const int count = (int)(ref & 7u) + 1;
                // this lane's triangles, kept by closest_hit's rule
                float lt = __int_as_float(0x7f800000), lu = 0.0f, lv = 0.0f;
                int lf = 0x7fffffff, lsl = -1;
                bool ltie = false;
                for (int k = c; k < count; k += L) {
                    const int s = start + k;
                    float t, u, v;
                    if (tri_test(S.tri, s, org, dir, t, u, v)) {
                        const int fid = __float_as_int(S.tri[3 * s].w);
                        if (t > tmin) {
                            ltie = t == lt ? (ltie || fid != lf) : (t < lt ? false : ltie);
                            if (t < lt || (t == lt && fid < lf)) {
                                lt = t; lu = u; lv = v; lf = fid; lsl = s;
                            }
                        }
                    }
                }
                // the group's nearest (t, facet)
                float tm = lt;
                int fm = lf;
#pragma unroll
                for (int off = 1; off < L; off <<= 1) {
                    const float t2 = __shfl_xor(tm, off, 64);
                    const int f2 = __shfl_xor(fm, off, 64);
                    if (t2 < tm || (t2 == tm && f2 < fm)) {
                        tm = t2;
                        fm = f2;
                    }
                }
                const uint64_t at = __ballot(lsl >= 0 && lt == tm) & gmask;
                const uint64_t att = __ballot(lsl >= 0 && lt == tm && ltie) & gmask;
                const uint64_t win = __ballot(lsl >= 0 && lt == tm && lf == fm) & gmask;
                if (win) {
                    const int wl = __ffsll((unsigned long long)win) - 1;
                    const float wu = __shfl(lu, wl, 64), wv = __shfl(lv, wl, 64);
                    const int ws = __shfl(lsl, wl, 64);
                    const bool tb = __popcll(at) >= 2 || att != 0ull;
                    if (tm < best) {
                        best = tm; bfacet = fm; bu = wu; bv = wv; bslot = ws; tie = tb;
                    } else if (tm == best) {
                        if (fm < bfacet) {
                            bfacet = fm; bu = wu; bv = wv; bslot = ws;
                        }
                        tie = true;
                    }
                }
                pop = true;
            } else {
                const float* nd = S.nodes + (size_t)ref * NF + c * CPL;
                float d[CPL];
                uint32_t cr[CPL];
                bool hit[CPL];
                uint64_t hm[CPL];
                int nh = 0;
#pragma unroll
                for (int k = 0; k < CPL; ++k) {
                    const float bx0 = nd[0 * W + k], by0 = nd[1 * W + k], bz0 = nd[2 * W + k];
                    const float bx1 = nd[3 * W + k], by1 = nd[4 * W + k], bz1 = nd[5 * W + k];
                    cr[k] = __float_as_uint(nd[6 * W + k]);
                    const float tx0 = fmaf(bx0, inv.x, -oi.x), tx1 = fmaf(bx1, inv.x, -oi.x);
                    const float ty0 = fmaf(by0, inv.y, -oi.y), ty1 = fmaf(by1, inv.y, -oi.y);
                    const float tz0 = fmaf(bz0, inv.z, -oi.z), tz1 = fmaf(bz1, inv.z, -oi.z);
                    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
                    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                    hit[k] = (cr[k] != EMPTY_CHILD) && (tf >= tn) && (tn <= best) && (tf >= 0.0f);
                    d[k] = hit[k] ? tn : 3.0e38f;
                    hm[k] = __ballot(hit[k]) & gmask;
                    nh += __popcll(hm[k]);
                }
                if (nh == 0) {
                    pop = true;
                } else {
                    // rank of each of this lane's children among the group's hit
                    // children: (distance, child slot) ascending
                    int rank[CPL];
#pragma unroll
                    for (int k = 0; k < CPL; ++k) rank[k] = 0;
#pragma unroll
                    for (int j = 0; j < L; ++j) {
#pragma unroll
                        for (int kk = 0; kk < CPL; ++kk) {
                            const float dj = __shfl(d[kk], gl + j, 64);
                            const bool hj = ((hm[kk] >> (gl + j)) & 1ull) != 0ull;
                            const int sj = j * CPL + kk;
#pragma unroll
                            for (int k = 0; k < CPL; ++k)
                                rank[k] += (hj && (dj < d[k] || (dj == d[k] && sj < c * CPL + k))) ? 1 : 0;
                        }
                    }
                    bool mine0 = false;
                    uint32_t near_local = 0;
#pragma unroll
                    for (int k = 0; k < CPL; ++k)
                        if (hit[k] && rank[k] == 0) {
                            mine0 = true;
                            near_local = cr[k];
                        }
                    const uint64_t nm = __ballot(mine0) & gmask;
                    const uint32_t nearest = __shfl(near_local, __ffsll((unsigned long long)nm) - 1, 64);
#pragma unroll
                    for (int k = 0; k < CPL; ++k) {
                        if (hit[k] && rank[k] > 0) {
                            const int s = sp + (nh - 1 - rank[k]);
                            if (s < WT_LDS) ls[s] = cr[k];
                            else if (s < cap) gs[(size_t)(s - WT_LDS) * gstride] = cr[k];
                            else fl |= 4u;
                        }
                    }
                    sp += nh - 1;
                    ref = nearest;
                }
            }
            if (pop) {
                if (sp == 0 || sp > cap) {
                    // the ray is done (or would pop a dropped entry)
                    if (c == 0) {
                        const bool h = bslot >= 0 && best < F_MAX;
                        hits[q] = h ? make_float4(best, bu, bv, __int_as_float(bslot | (tie ? HIT_TIE : 0)))
                                    : make_float4(F_MAX, 0.0f, 0.0f, __int_as_float(-1));
                        ++nray;
                        ntie += (h && tie) ? 1u : 0u;
                    }
                    active = false;
                } else {
                    --sp;
                    ref = sp < WT_LDS ? ls[sp] : gs[(size_t)(sp - WT_LDS) * gstride];
                }
            }
        }
    }
    if (fl) atomicOr(flags, fl);
    add_wave_count(counts, nray);
    add_wave_count(counts + 2, ntie);
}

// ---- shading ---------------------------------------------------------------
__device__ __forceinline__ bool hit_from_record(const SceneDev& S, float4 r, HitRec& h) {
    const int code = __float_as_int(r.w);
    if (code < 0) return false;
    h.slot = code & ~HIT_TIE;
    h.tie = (code & HIT_TIE) != 0;
    h.t = r.x; h.u = r.y; h.v = r.z;
    h.facet = __float_as_int(S.tri[3 * h.slot].w);
    return true;
}

// Block b shades partition b mod STRIPES (the grid is a multiple of
// STRIPES) and appends the continuing paths to the same partition of the
// next queue.
__global__ void __launch_bounds__(256) k_pw_shade(SceneDev S, SlotsDev slots, PathStateDev Q,
                                                  const float4* __restrict__ rays_in, const float4* __restrict__ hits,
                                                  const int* __restrict__ cnt_in, float4* __restrict__ rays_out,
                                                  int* __restrict__ cnt_out, int Pq, uint32_t* __restrict__ flags) {
    const int part = (int)blockIdx.x % STRIPES;
    const int kb = (int)blockIdx.x / STRIPES, nkb = (int)gridDim.x / STRIPES;
    const int nq = cnt_in[part * STRIPE_PAD];
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t ovf = 0, fl = 0;
    // wave-uniform trip count: every lane of a wave reaches the ballot
    for (int base = kb * (int)blockDim.x + (int)(threadIdx.x & ~63u); base < nq; base += nkb * (int)blockDim.x) {
        const int qi = part * Pq + base + lane;
        bool cont = false;
        PhotonPath P;
        if (base + lane < nq) {
            path_load(Q, __float_as_int(rays_in[2 * (size_t)qi].w), P);
            HitRec h;
            cont = hit_from_record(S, hits[qi], h) && photon_shade(S, slots, P, h, &ovf, &fl) && photon_pretrace(P);
            if (!cont) photon_finish(slots, P);
        }
        const uint64_t cm = __ballot(cont);
        if (cm) {
            const int leader = __ffsll((unsigned long long)cm) - 1;
            int o = 0;
            if (lane == leader) o = atomicAdd(&cnt_out[part * STRIPE_PAD], __popcll(cm));
            o = __shfl(o, leader, 64);
            if (cont) {
                path_store(Q, P);
                ray_store(rays_out, part * Pq + o + __popcll(cm & below), P);
            }
        }
    }
    if (ovf) fl |= 2u;
    if (fl) atomicOr(flags, fl);
}

// ---- the last bounces: one lane per path ---------------------------------
// k_photon's loop over the paths of a queue: a lane takes a queued path with
// its state, traces the queued ray (the checks before it were made by
// k_pw_shade) and continues the path to its end.
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK, PATH_WAVES_PER_SIMD) k_pw_resume(
    SceneDev S, SlotsDev slots, PathStateDev Q, const float4* __restrict__ rays, const int* __restrict__ qcnt,
    int* __restrict__ fetch, int Pq, int cap, uint32_t* __restrict__ spill, uint32_t* __restrict__ flags,
    unsigned long long* __restrict__ counts) {
    __shared__ uint32_t stk[TRACE_STACK_WORDS];
    const int lane = threadIdx.x;
    const TraceStack ts{stk + lane, spill + (size_t)blockIdx.x * PATH_BLOCK + lane, (int)gridDim.x * PATH_BLOCK, cap};
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t ovf = 0, fl = 0, nray = 0, ntie = 0;
    PhotonPath P;
    bool alive = false, fresh = false, exhausted = false;
    StripeCursor cur = stripe_start();
    for (;;) {
        const unsigned long long idle = __ballot(!alive);
        const int nidle = __popcll(idle);
        if (!exhausted && nidle >= PHOTON_REGEN) {
            int first = 0;
            const int got = part_claim(qcnt, fetch, Pq, cur, nidle, &first);
            if (got == 0) exhausted = true;
            if (!alive) {
                const int r = __popcll(idle & below);
                if (r < got) {
                    path_load(Q, __float_as_int(rays[2 * (size_t)(first + r)].w), P);
                    alive = true;
                    fresh = true;
                }
            }
        }
        if (__ballot(alive) == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (alive && !photon_bounce<W>(S, slots, P, ts, &ovf, &fl, nray, ntie, !fresh)) {
            photon_finish(slots, P);
            alive = false;
        }
        fresh = false;
    }
    if (ovf) fl |= 2u;
    if (fl) atomicOr(flags, fl);
    add_wave_count(counts, nray);
    add_wave_count(counts + 2, ntie);
}

// ---- host side -------------------------------------------------------------
template <int W, int L>
static int trace_blocks_wl() {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pw_trace<W, L>, WT_BLOCK, 0) != hipSuccess || nb < 1)
        nb = 4;
    return cus * nb;
}
int wave_lanes_ok(int width, int lanes) { return (lanes == 2 || lanes == 4 || lanes == 8) && lanes <= width; }
int wave_trace_blocks(int width, int lanes) {
    if (width == 2) return trace_blocks_wl<2, 2>();
    if (width == 4) return lanes == 2 ? trace_blocks_wl<4, 2>() : trace_blocks_wl<4, 4>();
    return lanes == 2 ? trace_blocks_wl<8, 2>() : lanes == 4 ? trace_blocks_wl<8, 4>() : trace_blocks_wl<8, 8>();
}
size_t wave_spill_entries(int lanes, int blocks, int cap) {
    const size_t groups = (size_t)blocks * WT_WAVES * (64 / lanes);
    return groups * (size_t)(cap > WT_LDS ? cap - WT_LDS : 1);
}

void launch_photon_wave(const SceneDev& S, const SlotsDev& slots, int N, int step, const int* step_dev,
                        const WaveBuffers& wb, int bounces, int lanes, int trace_blocks, const TraceSpill& tail,
                        const GuideEmit* guide, uint32_t* flags, unsigned long long* rays, hipStream_t st) {
    const int nctl = bounces + 1;
    const int P = wave_partition(N);
    // sizes of queue b at qctl + b * STRIPE_INTS, claim counters after all sizes
    int* const qsize = wb.qctl;
    int* const qclaim = wb.qctl + (size_t)nctl * STRIPE_INTS;
    const int ge = (N + 255) / 256;
    if (guide)
        hipLaunchKernelGGL(k_pw_emit<true>, dim3(ge), dim3(256), 0, st, S, slots, N, step, step_dev, *guide, wb.path,
                           wb.rays[0], wb.qctl, nctl, P, flags);
    else
        hipLaunchKernelGGL(k_pw_emit<false>, dim3(ge), dim3(256), 0, st, S, slots, N, step, step_dev, GuideEmit{},
                           wb.path, wb.rays[0], wb.qctl, nctl, P, flags);
    int kb = (P + 255) / 256;
    if (kb > 32) kb = 32;
    const int gs = STRIPES * kb;
    for (int b = 0; b < bounces; ++b) {
        float4* rin = wb.rays[b & 1];
        float4* rout = wb.rays[(b + 1) & 1];
        int* cin = qsize + (size_t)b * STRIPE_INTS;
        int* cout = qsize + (size_t)(b + 1) * STRIPE_INTS;
        int* clm = qclaim + (size_t)b * STRIPE_INTS;
#define SPPM_WTRACE(WW, LL)                                                                                          \
    hipLaunchKernelGGL((k_pw_trace<WW, LL>), dim3(trace_blocks), dim3(WT_BLOCK), 0, st, S, rin, cin, clm, P,        \
                       wb.hits, SPPM_RAY_OFFSET, wb.cap, wb.spill, flags, rays)
        if (S.width == 2) SPPM_WTRACE(2, 2);
        else if (S.width == 4) {
            if (lanes == 2) SPPM_WTRACE(4, 2);
            else SPPM_WTRACE(4, 4);
        } else {
            if (lanes == 2) SPPM_WTRACE(8, 2);
            else if (lanes == 4) SPPM_WTRACE(8, 4);
            else SPPM_WTRACE(8, 8);
        }
#undef SPPM_WTRACE
        hipLaunchKernelGGL(k_pw_shade, dim3(gs), dim3(256), 0, st, S, slots, wb.path, rin, wb.hits, cin, rout, cout,
                           P, flags);
    }
    float4* rlast = wb.rays[bounces & 1];
    int* clast = qsize + (size_t)bounces * STRIPE_INTS;
    int* mlast = qclaim + (size_t)bounces * STRIPE_INTS;
#define SPPM_WRESUME(WW)                                                                                             \
    hipLaunchKernelGGL(k_pw_resume<WW>, dim3(tail.blocks), dim3(PATH_BLOCK), 0, st, S, slots, wb.path, rlast,       \
                       clast, mlast, P, tail.cap, tail.buf, flags, rays)
    if (S.width == 2) SPPM_WRESUME(2);
    else if (S.width == 4) SPPM_WRESUME(4);
    else SPPM_WRESUME(8);
#undef SPPM_WRESUME
}
int wave_partition(int N) { return (N + STRIPES - 1) / STRIPES; }
size_t wave_queue_entries(int N) { return (size_t)STRIPES * (size_t)wave_partition(N); }
size_t wave_ctl_ints(int bounces) { return 2 * (size_t)(bounces + 1) * STRIPE_INTS; }

}  // namespace sppm
