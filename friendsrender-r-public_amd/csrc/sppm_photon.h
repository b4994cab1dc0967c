// sppm_photon.h -- one photon path of createPhotonMap (SPPM.cpp:495-680):
// emission (plain and guided), the loop's checks, shading after a closest
// hit, and the path's end.  Shared by the megakernel photon pass and the
// wavefront kernels, so both run the same arithmetic.
#pragma once
#include "sppm_guide.h"
#include "sppm_pass.h"
#include "sppm_stripes.h"
#include "sppm_trace.h"

namespace sppm {

// ---------------------------------------------------------------------------
// Alias-table draw: TableSampler::Sample(RandomMT&) (TableSampler.h:126-129)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int table_sample(const float* __restrict__ prob, const int* __restrict__ left,
                                            const int* __restrict__ right, int n, MTCursor& m, uint32_t* ovf) {
    double r2 = real1(m, ovf);
    int bin = (int)((double)n * r2);
    if (bin >= n) bin = n - 1;  // unreachable: r2 < 1
    double r1 = real1(m, ovf);
    int id = (r1 < (double)prob[bin]) ? left[bin] : right[bin];
    return id < 0 ? left[bin] : id;
}

__device__ __forceinline__ bool near1(float s) { return fabs_(s - 1.0f) < HI_1EM5; }  // SPPM.h:27-29, (double)|s-1| < 0.00001

// ---------------------------------------------------------------------------
// createPhotonMap (SPPM.cpp:518-672), one lane per photon path
// ---------------------------------------------------------------------------
// One photon path in flight (createPhotonMap loop state, SPPM.cpp:495-680).
struct PhotonPath {
    MTCursor m;
    f3 rad, T, cur_org, cur_dir, next_dir;
    float cur_index, next_index, PDF;
    int i, pidx, loopcount, depth;
};

// Adds a per-lane count to a striped 64-bit device counter, one atomic per
// wave: ctr points at the counter's word in the first of STRIPES lines of
// RAY_STRIPE words (256 bytes); a wave adds to line (wave index mod STRIPES),
// and readers sum the lines.
__device__ __forceinline__ void add_wave_count(unsigned long long* __restrict__ ctr, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(&ctr[(size_t)(wave % STRIPES) * RAY_STRIPE], x);
}

// LightSampler::SamplePointOnAreaLight (LightSampler.cpp:68-82): a point of
// light L, its radiance and its area density.
__device__ __forceinline__ void light_point(const SceneDev& S, const DevLight& L, MTCursor& m, uint32_t* ovf,
                                            f3& lpos, f3& rad, float& pdf_area) {
    f3 p0 = mk(L.p[0], L.p[1], L.p[2]), p1 = mk(L.p[3], L.p[4], L.p[5]), p2 = mk(L.p[6], L.p[7], L.p[8]);
    if (L.textured) {  // LightSampler.cpp:70-76
        const int off = L.tab_off;
        int k = table_sample(S.ltab_prob + off, S.ltab_left + off, S.ltab_right + off, 1000, m, ovf);
        float4 w = S.ltab_uvw[off + k];
        float4 Lk = S.ltab_L[off + k];
        lpos = (p0 * w.x + p1 * w.y) + p2 * (1.0f - w.x - w.y);
        rad = mk(Lk.x, Lk.y, Lk.z);
        float gp = w.w / L.tab_avg / 1000.0f;
        pdf_area = (gp * 1000.0f) / L.area;
    } else {  // PointOnTriangle (Sampling.cpp:339-349)
        float u = (float)real1(m, ovf);
        float v = (float)real1(m, ovf);
        if (v > u) { u *= 0.5f; v -= u; }
        else { v *= 0.5f; u -= v; }
        pdf_area = 2.0f / fabs_(norm(cross(p1 - p0, p2 - p0)));
        lpos = (p0 * u + p1 * v) + p2 * (1.0f - u - v);
        rad = mk(L.flux[0], L.flux[1], L.flux[2]) / (2.0f * SPPM_PI * L.area);
    }
}

// LightSampler::GetAreaLightPDF (LightSampler.h:87): TableSampler::GetPDF of
// the light pick table, f / average / N.
__device__ __forceinline__ float light_pick_pdf(const SceneDev& S, const DevLight& L) {
    return L.flux_norm / S.lpick_avg / (float)S.nlights;
}

// generateLightRay (SPPM.cpp:453-471) + the emission photon in slot 0
// (SPPM.cpp:530-540); the first stored surface photon overwrites that slot
// (SPPM.cpp:606), so writing it up front leaves the same slot contents.
__device__ __forceinline__ void photon_start(const SceneDev& S, const SlotsDev& slots, int N, int step, int i,
                                             PhotonPath& P, uint32_t* ovf) {
    P.i = i;
    mt_seed(P.m, (uint32_t)i + 135u + (uint32_t)step * (uint32_t)N);
    const int li = table_sample(S.lpick_prob, S.lpick_left, S.lpick_right, S.nlights, P.m, ovf);
    const DevLight& L = S.lights[li];
    f3 lpos, rad;
    float pdf_area;
    light_point(S, L, P.m, ovf, lpos, rad, pdf_area);
    pdf_area *= light_pick_pdf(S, L);
    const f3 lnorm = mk(L.normal[0], L.normal[1], L.normal[2]);
    float pdf_dir = 1.0f;
    f3 outdir = cos_sampling(P.m, ovf, lnorm, &pdf_dir);
    rad = rad * smin(1.0f, smax(0.0f, dot(lnorm, outdir)));
    P.rad = rad;
    P.cur_org = lpos + lnorm * SPPM_RAY_OFFSET;
    P.cur_dir = normalized(outdir);  // Ray ctor normalises
    P.cur_index = 1.0f;
    P.next_dir = P.cur_dir;
    P.next_index = 1.0f;
    P.PDF = pdf_area * pdf_dir;
    P.T = mk(1.0f, 1.0f, 1.0f);
    P.pidx = 0;
    P.loopcount = 0;
    P.depth = 0;
    const f3 e_flux = (rad * P.T) * SPPM_PI;
    slots.pos[i] = make_float4(P.cur_org.x, P.cur_org.y, P.cur_org.z, 1.0f);
    slots.flux[i] = make_float4(e_flux.x, e_flux.y, e_flux.z, P.PDF);
    slots.dir[i] = make_float4(P.cur_dir.x, P.cur_dir.y, P.cur_dir.z, 0.0f);
    slots.valid[i] = 1;
}

// Guided emission: draw from the mixture, evaluate the mixture density at the
// drawn (position, direction) and use it as the photon's PDF; the photon
// flux is L_e cos, as for plain emission, so flux / PDF stays an unbiased
// weight whichever component drew the sample.
//
// Textured light: positions are the table points.  Point k has plain
// probability pi_k and the plain PDF D_k * cos / pi, where D_k = pi_k * 1000 /
// area is the reference's area density of the point (LightSampler.cpp:70-76);
// the proposal picks a cell c, then a point of c's position cell with
// probability pi_k / mass(c).  The PDF is D_k * (w cos / pi + (1 - w) P(c)
// d(omega | c) / mass(c)), the mixture's probability of (k, omega) over pi_k
// times D_k, which keeps the plain emission's expectation.
__device__ __forceinline__ int cdf_search(const float* __restrict__ cdf, int lo, int hi, float x) {
    while (lo < hi) {  // first entry > x in [lo, hi]
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}
__device__ __forceinline__ void photon_start_guided(const SceneDev& S, const SlotsDev& slots, const GuideEmit& G,
                                                    int N, int step, int i, PhotonPath& P, uint32_t* ovf) {
    P.i = i;
    mt_seed(P.m, (uint32_t)i + 135u + (uint32_t)step * (uint32_t)N);
    const f3 P0 = mk(G.P0[0], G.P0[1], G.P0[2]), E1 = mk(G.e1[0], G.e1[1], G.e1[2]), E2 = mk(G.e2[0], G.e2[1], G.e2[2]);
    const f3 Nn = mk(G.n[0], G.n[1], G.n[2]), T1 = mk(G.t1[0], G.t1[1], G.t1[2]), T2 = mk(G.t2[0], G.t2[1], G.t2[2]);
    const float total = G.cdf[GUIDE_CELLS - 1];
    f3 lpos, outdir, rad0;
    int tj = -1, tl = 0;  // textured: table entry and light of the point
    if (real1(P.m, ovf) < (double)G.w) {
        // plain: light pick, then PointOnTriangle or a table point, cosine lobe (SPPM.cpp:453-468)
        const int li = table_sample(S.lpick_prob, S.lpick_left, S.lpick_right, S.nlights, P.m, ovf);
        const DevLight& L = S.lights[li];
        f3 p0 = mk(L.p[0], L.p[1], L.p[2]), p1 = mk(L.p[3], L.p[4], L.p[5]), p2 = mk(L.p[6], L.p[7], L.p[8]);
        if (G.textured) {
            const int off = L.tab_off;
            const int k = table_sample(S.ltab_prob + off, S.ltab_left + off, S.ltab_right + off, 1000, P.m, ovf);
            const float4 w = S.ltab_uvw[off + k];
            lpos = (p0 * w.x + p1 * w.y) + p2 * (1.0f - w.x - w.y);
            tj = off + k;
            tl = li;
        } else {
            float u = (float)real1(P.m, ovf);
            float v = (float)real1(P.m, ovf);
            if (v > u) { u *= 0.5f; v -= u; }
            else { v *= 0.5f; u -= v; }
            lpos = (p0 * u + p1 * v) + p2 * (1.0f - u - v);
        }
        float pdf_dir = 1.0f;
        outdir = normalized(cos_sampling(P.m, ovf, Nn, &pdf_dir));
    } else {
        // proposal: a cell by its CDF, then a position and a direction inside it
        // x < total strictly, so the cell drawn has probability > 0 (masked cells have none)
        const float x = fminf((float)real1(P.m, ovf) * total, __uint_as_float(__float_as_uint(total) - 1u));
        const int c = cdf_search(G.cdf, 0, GUIDE_CELLS - 1, x);
        const int it = c % GUIDE_GT, is = (c / GUIDE_GT) % GUIDE_GS, iv = (c / (GUIDE_GS * GUIDE_GT)) % GUIDE_GV,
                  iu = c / (GUIDE_GV * GUIDE_GS * GUIDE_GT);
        if (G.textured) {
            // a table point of the position cell, by the plain probabilities
            const int pc = iu * GUIDE_GV + iv;
            const int a = G.pc_start[pc], e = G.pc_start[pc + 1];
            const float x = (float)real1(P.m, ovf) * G.pcell_mass[pc];
            const int q = e > a ? cdf_search(G.pc_cum, a, e - 1, x) : a;
            const int2 pt = G.pc_pt[q];
            const DevLight& L = S.lights[pt.x];
            f3 p0 = mk(L.p[0], L.p[1], L.p[2]), p1 = mk(L.p[3], L.p[4], L.p[5]), p2 = mk(L.p[6], L.p[7], L.p[8]);
            const float4 w = S.ltab_uvw[pt.y];
            lpos = (p0 * w.x + p1 * w.y) + p2 * (1.0f - w.x - w.y);
            tj = pt.y;
            tl = pt.x;
        } else {
            const float u = ((float)iu + (float)real1(P.m, ovf)) / (float)GUIDE_GU;
            const float v = ((float)iv + (float)real1(P.m, ovf)) / (float)GUIDE_GV;
            lpos = (P0 + E1 * u) + E2 * v;
        }
        const float s = -1.0f + ((float)is + (float)real1(P.m, ovf)) * (2.0f / (float)GUIDE_GS);
        const float t = -1.0f + ((float)it + (float)real1(P.m, ovf)) * (2.0f / (float)GUIDE_GT);
        const float px = 0.5f * (s + t), py = 0.5f * (s - t);
        const float pz = fmaxf(0.0f, 1.0f - fabs_(px) - fabs_(py));
        outdir = normalized((T1 * px + T2 * py) + Nn * pz);
    }
    // mixture density at (lpos, outdir)
    const float lx = dot(outdir, T1), ly = dot(outdir, T2), lz = dot(outdir, Nn);
    const float cs = fmaxf(lz, 0.0f);
    const float l1 = fabs_(lx) + fabs_(ly) + cs;
    const float ox = lx / l1, oy = ly / l1, oz = cs / l1;  // octahedron point (|ox|+|oy|+oz = 1)
    const float s = ox + oy, t = ox - oy;
    const int is = min(max((int)((s + 1.0f) * (0.5f * (float)GUIDE_GS)), 0), GUIDE_GS - 1);
    const int it = min(max((int)((t + 1.0f) * (0.5f * (float)GUIDE_GT)), 0), GUIDE_GT - 1);
    int pcell;
    if (G.textured) {
        pcell = G.pt_cell[tj];
    } else {
        const f3 rel = lpos - P0;
        const float u = fminf(fmaxf(dot(rel, E1) * G.inv_e1, 0.0f), 0.99999994f);
        const float v = fminf(fmaxf(dot(rel, E2) * G.inv_e2, 0.0f), 0.99999994f);
        const int iu = min((int)(u * (float)GUIDE_GU), GUIDE_GU - 1), iv = min((int)(v * (float)GUIDE_GV), GUIDE_GV - 1);
        pcell = iu * GUIDE_GV + iv;
    }
    const int cell = (pcell * GUIDE_GS + is) * GUIDE_GT + it;
    const float plen = sqrtf(ox * ox + oy * oy + oz * oz);
    // direction density inside a cell: 1 / (cell area in (s, t)) * 2 |P|^3 (d omega = ds dt / (2 |P|^3))
    const float dd = (float)(GUIDE_GS * GUIDE_GT / 4) * (2.0f * plen * plen * plen);
    const float pc_prob = G.prob[cell] / total;
    float q, p_plain;
    if (G.textured) {
        const DevLight& L = S.lights[tl];
        const float4 w = S.ltab_uvw[tj];
        const float4 Lk = S.ltab_L[tj];
        // D_k as the plain emission computes it (photon_start)
        const float gp = w.w / L.tab_avg / 1000.0f;
        const float Dk = ((gp * 1000.0f) / L.area) * (L.flux_norm / S.lpick_avg / (float)S.nlights);
        const float mass = G.pcell_mass[pcell];
        const float prop = mass > 0.0f ? pc_prob * dd / mass : 0.0f;
        q = Dk * (G.w * cs / SPPM_PI + (1.0f - G.w) * prop);
        p_plain = Dk * cs / SPPM_PI;
        rad0 = mk(Lk.x, Lk.y, Lk.z);
    } else {
        // cell density in (u, v, s, t) -> area x solid angle: / A
        const float p_mlp = pc_prob * (float)(GUIDE_GU * GUIDE_GV) * G.inv_area * dd;
        p_plain = G.inv_area * cs / SPPM_PI;
        q = G.w * p_plain + (1.0f - G.w) * p_mlp;
        rad0 = mk(G.rad[0], G.rad[1], G.rad[2]);
    }
    G.path_cell[(size_t)(step & 1) * (size_t)N + i] = cell;
    if (G.path_ratio) G.path_ratio[(size_t)(step & 1) * (size_t)N + i] = p_plain > 0.0f ? q / p_plain : 0.0f;
    const f3 rad = rad0 * smin(1.0f, cs);
    P.rad = rad;
    P.cur_org = lpos + Nn * SPPM_RAY_OFFSET;
    P.cur_dir = outdir;
    P.cur_index = 1.0f;
    P.next_dir = P.cur_dir;
    P.next_index = 1.0f;
    P.PDF = q;
    P.T = mk(1.0f, 1.0f, 1.0f);
    P.pidx = 0;
    P.loopcount = 0;
    P.depth = 0;
    const f3 e_flux = (rad * P.T) * SPPM_PI;
    slots.pos[i] = make_float4(P.cur_org.x, P.cur_org.y, P.cur_org.z, 1.0f);
    slots.flux[i] = make_float4(e_flux.x, e_flux.y, e_flux.z, P.PDF);
    slots.dir[i] = make_float4(P.cur_dir.x, P.cur_dir.y, P.cur_dir.z, 0.0f);
    slots.valid[i] = 1;
}

// The createPhotonMap loop's checks before a path's next trace
// (SPPM.cpp:548-554); false ends the path.
__device__ __forceinline__ bool photon_pretrace(PhotonPath& P) {
    if (P.loopcount++ > 32 || P.pidx >= 31) return false;
    if (P.T.x < HI_1EM8 || P.T.y < HI_1EM8 || P.T.z < HI_1EM8) return false;  // (double)T < 1e-8
    return true;
}

// The rest of the loop body after the closest hit h of the path's ray
// (SPPM.cpp:556-671): geometry, transmission, Russian roulette, photon
// store, BSDF sample; false ends the path.
__device__ __forceinline__ bool photon_shade(const SceneDev& S, const SlotsDev& slots, PhotonPath& P,
                                             const HitRec& h, uint32_t* ovf, uint32_t* fl) {
    Geom g;
    calc_geometry(S, h, g);
    const f3 Nf = dot(g.N, P.cur_dir) > 0 ? -g.N : g.N;
    f3 next_org = g.pos;
    if (norm2(g.radiance) > LO_1EM5) return false;  // (double)norm2 > 0.00001
    if (g.transl > 0.5f) {
        P.T = P.T * g.transl;
        P.cur_org = g.pos + P.cur_dir * SPPM_RAY_OFFSET;
        return true;
    }
    if (P.depth > 16) return false;  // russianRulette_PhotonMapping (SPPM.h:33-52)
    float a0 = g.albedo.x, a1 = g.albedo.y, a2 = g.albedo.z;
    float pdf_rr = 0.98f * (a0 > a1 ? (a0 > a2 ? a0 : a2) : (a1 > a2 ? a1 : a2));
    if (real1(P.m, ovf) > (double)pdf_rr) return false;
    P.PDF *= pdf_rr;
    float Ii, Io = 1.0f;
    if (near1(P.cur_index)) { Ii = 1.0f; Io = g.index; }
    else { Ii = P.cur_index; Io = 1.0f; }
    FParam fp{P.cur_dir, P.next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
    if (g.is_sppm) {
        if (P.pidx >= NSLOT - 2) { *fl |= 8u; return false; }
        f3 ph_flux = (P.rad * P.T) * g.albedo;
        if (!(P.PDF > 0.0f) || !isfinite(ph_flux.x) || !isfinite(ph_flux.y) || !isfinite(ph_flux.z) ||
            !isfinite(P.PDF))
            *fl |= 16u;  // stored photon with pdf <= 0 or non-finite flux (propagated as the reference does)
        const size_t o = (size_t)P.pidx * (size_t)slots.npaths + P.i;
        slots.pos[o] = make_float4(g.pos.x, g.pos.y, g.pos.z, P.cur_index);
        slots.flux[o] = make_float4(ph_flux.x, ph_flux.y, ph_flux.z, P.PDF);
        slots.dir[o] = make_float4(P.cur_dir.x, P.cur_dir.y, P.cur_dir.z, 0.0f);
        slots.valid[o] = 1;
        ++P.pidx;
    }
    float pdf_s = 1.0f;
    P.next_dir = surf_sample(g.kind, fp, P.m, ovf, &pdf_s);
    if (!rutil_valid_dir(P.next_dir)) return false;
    if (g.opaque) {
        if (dot(P.next_dir, Nf) < HI_1EM6) return false;  // (double)dot < 1e-6
    } else {
        P.next_index = rutil_is_reflect(P.cur_dir, P.next_dir, Nf) ? P.cur_index : Io;
    }
    fp.out = P.next_dir;
    pdf_s = surf_pdf(g.kind, fp);
    f3 outN = dot(Nf, P.next_dir) > 0 ? Nf : -Nf;
    next_org = next_org + outN * SPPM_RAY_OFFSET;
    float brdf = surf_brdf(g.kind, fp);
    float cs = smax(0.0f, smin(1.0f, dot(P.next_dir, outN)));
    if (brdf < HI_1EM8) return false;  // (double)brdf < 1e-8
    if (pdf_s < 0) return false;
    P.T = P.T * ((g.albedo * cs) * brdf);
    P.PDF *= pdf_s;
    P.cur_org = next_org;
    P.cur_dir = P.next_dir;
    P.cur_index = P.next_index;
    ++P.depth;
    return true;
}

// One iteration of the createPhotonMap loop body; false when the path ends.
// pre = false skips the checks before the trace (already made by the caller).
template <int W>
__device__ __forceinline__ bool photon_bounce(const SceneDev& S, const SlotsDev& slots, PhotonPath& P,
                                              const TraceStack& ts, uint32_t* ovf, uint32_t* fl, uint32_t& nray,
                                              uint32_t& ntie, bool pre = true) {
    if (pre && !photon_pretrace(P)) return false;
    HitRec h;
    ++nray;
    if (!closest_hit<W>(S, P.cur_org, P.cur_dir, SPPM_RAY_OFFSET, ts, h, fl)) return false;
    ntie += h.tie ? 1u : 0u;
    return photon_shade(S, slots, P, h, ovf, fl);
}

// Idle lanes a wave of a persistent path kernel waits for before it takes
// new paths (one atomic per refill).
constexpr int PHOTON_REGEN = 16;

// End of a path: terminator at slot pidx + 1 (SPPM.cpp:672); slot pidx keeps
// its content from an earlier pass and is flattened if valid (SPPM.cpp:812-822).
__device__ __forceinline__ void photon_finish(const SlotsDev& slots, const PhotonPath& P) {
    const size_t NP = (size_t)slots.npaths;
    const int cnt = P.pidx == 0 ? 1 : P.pidx + (slots.valid[(size_t)P.pidx * NP + P.i] ? 1 : 0);
    slots.valid[(size_t)(P.pidx + 1) * NP + P.i] = 0;
    slots.count[P.i] = cnt;
}


}  // namespace sppm
