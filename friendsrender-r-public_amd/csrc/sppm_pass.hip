// sppm_pass.hip -- the per-pass kernels of the SPPM core on gfx950.
//
// One pass (SPPM.cpp:770-843):
//   k_camera_words   shared UniformSample stream words for this pass (1 wave)
//   k_photon<W>      createPhotonMap: one lane per photon path, writes slots
//   scan(count)      flatten offsets (SPPM.cpp:812-822 order)
//   k_grid_params    cell size from the largest radius (device-side)
//   k_photon_keys    per flattened photon: cell key + rank (atomic count)
//   scan(cells)      cell starts
//   k_photon_scatter cell-sorted photon records (kdtree::CreateTree replacement)
//   k_eye<W>         createHitpoints: one lane per pixel
//   k_gather         applyHitPointInfo + progressive update, one lane per
//                    hit point in 8x8 pixel tiles
//   k_reconstruct    reconstructImageFromSharedPoints
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_kernels.h"
#include "sppm_pass.h"
#include "sppm_trace.h"

namespace sppm {

// ---------------------------------------------------------------------------
// Shared camera sampler stream: UniformSample(ID) owns one mt19937 for the
// whole render (CameraSampler.h:19-31); pixel i of a pass draws 3 doubles in
// pixel order.  One wave keeps the 624-word state in LDS and twists it in
// 64-word chunks in index order (new x[k] needs new x[k-227], which an earlier
// chunk produced), emitting the tempered words for the pass.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t temper(uint32_t z) {
    z ^= (z >> 11);
    z ^= (z << 7) & 0x9d2c5680u;
    z ^= (z << 15) & 0xefc60000u;
    z ^= (z >> 18);
    return z;
}

__global__ void __launch_bounds__(64) k_camera_words(uint32_t* __restrict__ st_mt, int* __restrict__ st_pos,
                                                     uint32_t* __restrict__ words, long long need) {
    __shared__ uint32_t x[624];
    const int t = threadIdx.x;
    for (int k = t; k < 624; k += 64) x[k] = st_mt[k];
    int pos = *st_pos;
    __syncthreads();
    long long out = 0;
    while (out < need) {
        if (pos >= 624) {
            for (int base = 0; base < 624; base += 64) {
                const int k = base + t;
                uint32_t nv = 0;
                if (k < 624) {
                    uint32_t a = x[k];
                    uint32_t b = (k + 1 < 624) ? x[k + 1] : x[0];
                    uint32_t c = (k < 227) ? x[k + 397] : x[k - 227];
                    uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
                    nv = c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
                }
                __syncthreads();
                if (k < 624) x[k] = nv;
                __syncthreads();
            }
            pos = 0;
        }
        long long rem = need - out;
        int n = (624 - pos) < rem ? (624 - pos) : (int)rem;
        for (int j = t; j < n; j += 64) words[out + j] = temper(x[pos + j]);
        out += n;
        pos += n;
    }
    __syncthreads();
    for (int k = t; k < 624; k += 64) st_mt[k] = x[k];
    if (t == 0) *st_pos = pos;
}

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

__device__ __forceinline__ bool near1(float s) { return (double)fabs_(s - 1.0f) < 0.00001; }  // SPPM.h:27-29

// ---------------------------------------------------------------------------
// createPhotonMap (SPPM.cpp:518-672), one lane per photon path
// ---------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK) k_photon(SceneDev S, SlotsDev slots, int N, int step, int stack_cap,
                                                       uint32_t* __restrict__ flags) {
    extern __shared__ uint32_t lds[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint32_t* stk = lds + threadIdx.x;
    uint32_t ovf = 0, fl = 0;
    MTCursor m;
    mt_seed(m, (uint32_t)i + 135u + (uint32_t)step * (uint32_t)N);

    // generateLightRay (SPPM.cpp:453-471)
    const int li = table_sample(S.lpick_prob, S.lpick_left, S.lpick_right, S.nlights, m, &ovf);
    const DevLight& L = S.lights[li];
    f3 p0 = mk(L.p[0], L.p[1], L.p[2]), p1 = mk(L.p[3], L.p[4], L.p[5]), p2 = mk(L.p[6], L.p[7], L.p[8]);
    f3 lpos, rad;
    float pdf_area;
    if (L.textured) {  // LightSampler.cpp:70-76
        const int off = L.tab_off;
        int k = table_sample(S.ltab_prob + off, S.ltab_left + off, S.ltab_right + off, 1000, m, &ovf);
        float4 w = S.ltab_uvw[off + k];
        float4 Lk = S.ltab_L[off + k];
        lpos = (p0 * w.x + p1 * w.y) + p2 * (1.0f - w.x - w.y);
        rad = mk(Lk.x, Lk.y, Lk.z);
        float gp = w.w / L.tab_avg / 1000.0f;
        pdf_area = (gp * 1000.0f) / L.area;
    } else {  // PointOnTriangle (Sampling.cpp:339-349)
        float u = (float)real1(m, &ovf);
        float v = (float)real1(m, &ovf);
        if (v > u) { u *= 0.5f; v -= u; }
        else { v *= 0.5f; u -= v; }
        pdf_area = 2.0f / fabs_(norm(cross(p1 - p0, p2 - p0)));
        lpos = (p0 * u + p1 * v) + p2 * (1.0f - u - v);
        rad = mk(L.flux[0], L.flux[1], L.flux[2]) / (2.0f * SPPM_PI * L.area);
    }
    pdf_area *= L.flux_norm / S.lpick_avg / (float)S.nlights;
    const f3 lnorm = mk(L.normal[0], L.normal[1], L.normal[2]);
    float pdf_dir = 1.0f;
    f3 outdir = cos_sampling(m, &ovf, lnorm, &pdf_dir);
    rad = rad * smin(1.0f, smax(0.0f, dot(lnorm, outdir)));
    f3 cur_org = lpos + lnorm * SPPM_RAY_OFFSET;
    f3 cur_dir = normalized(outdir);  // Ray ctor normalises
    float cur_index = 1.0f;
    f3 next_org = cur_org, next_dir = cur_dir;
    float next_index = 1.0f;
    float PDF = pdf_area * pdf_dir;
    f3 T = mk(1.0f, 1.0f, 1.0f);
    // emission photon for slot 0 (kept only if no surface photon overwrites it)
    const f3 e_flux = (rad * T) * SPPM_PI;
    const float e_pdf = PDF;
    const f3 e_dir = cur_dir, e_pos = cur_org;

    const size_t NP = (size_t)slots.npaths;
    int pidx = 0, loopcount = 0, depth = 0;
    for (;;) {
        if (loopcount++ > 32 || pidx >= 31) break;
        if ((double)T.x < 1e-8 || (double)T.y < 1e-8 || (double)T.z < 1e-8) break;
        HitRec h;
        if (!closest_hit<W>(S, cur_org, cur_dir, SPPM_RAY_OFFSET, stk, PATH_BLOCK, stack_cap, h, &fl)) break;
        Geom g;
        calc_geometry(S, h, g);
        f3 Nf = dot(g.N, cur_dir) > 0 ? -g.N : g.N;
        next_org = g.pos;
        if ((double)norm2(g.radiance) > 0.00001) break;
        if (g.transl > 0.5f) {
            T = T * g.transl;
            cur_org = g.pos + cur_dir * SPPM_RAY_OFFSET;
            continue;
        }
        if (depth > 16) break;  // russianRulette_PhotonMapping (SPPM.h:33-52)
        float a0 = g.albedo.x, a1 = g.albedo.y, a2 = g.albedo.z;
        float pdf_rr = 0.98f * (a0 > a1 ? (a0 > a2 ? a0 : a2) : (a1 > a2 ? a1 : a2));
        if (real1(m, &ovf) > (double)pdf_rr) break;
        PDF *= pdf_rr;
        float Ii, Io = 1.0f;
        if (near1(cur_index)) { Ii = 1.0f; Io = g.index; }
        else { Ii = cur_index; Io = 1.0f; }
        FParam fp{cur_dir, next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
        if (g.is_sppm) {
            if (pidx >= NSLOT - 2) { fl |= 8u; break; }
            f3 ph_flux = (rad * T) * g.albedo;
            const size_t o = (size_t)pidx * NP + i;
            slots.pos[o] = make_float4(g.pos.x, g.pos.y, g.pos.z, cur_index);
            slots.flux[o] = make_float4(ph_flux.x, ph_flux.y, ph_flux.z, PDF);
            slots.dir[o] = make_float4(cur_dir.x, cur_dir.y, cur_dir.z, 0.0f);
            slots.valid[o] = 1;
            ++pidx;
        }
        float pdf_s = 1.0f;
        next_dir = surf_sample(g.kind, fp, m, &ovf, &pdf_s);
        if (!rutil_valid_dir(next_dir)) break;
        if (g.opaque) {
            if ((double)dot(next_dir, Nf) < 1e-6) break;
        } else {
            next_index = rutil_is_reflect(cur_dir, next_dir, Nf) ? cur_index : Io;
        }
        fp.out = next_dir;
        pdf_s = surf_pdf(g.kind, fp);
        f3 outN = dot(Nf, next_dir) > 0 ? Nf : -Nf;
        next_org = next_org + outN * SPPM_RAY_OFFSET;
        float brdf = surf_brdf(g.kind, fp);
        float cs = smax(0.0f, smin(1.0f, dot(next_dir, outN)));
        if ((double)brdf < 1e-8) break;
        if (pdf_s < 0) break;
        T = T * ((g.albedo * cs) * brdf);
        PDF *= pdf_s;
        cur_org = next_org;
        cur_dir = next_dir;
        cur_index = next_index;
        ++depth;
    }
    int cnt;
    if (pidx == 0) {
        const size_t o = (size_t)i;
        slots.pos[o] = make_float4(e_pos.x, e_pos.y, e_pos.z, 1.0f);
        slots.flux[o] = make_float4(e_flux.x, e_flux.y, e_flux.z, e_pdf);
        slots.dir[o] = make_float4(e_dir.x, e_dir.y, e_dir.z, 0.0f);
        slots.valid[o] = 1;
        cnt = 1;
    } else {
        // slot pidx keeps its content from an earlier pass (SPPM.cpp:672)
        cnt = pidx + (slots.valid[(size_t)pidx * NP + i] ? 1 : 0);
    }
    slots.valid[(size_t)(pidx + 1) * NP + i] = 0;
    slots.count[i] = cnt;
    if (ovf) fl |= 2u;
    if (fl) atomicOr(flags, fl);
}

// ---------------------------------------------------------------------------
// Grid over the stored photons.  Cell size = largest shared radius, so a
// query box [c - R, c + R] spans at most 3 cells per axis.  Cell coordinates
// are clamped to the scene box; the same monotone map is used for photons
// and query boxes, so no neighbour is ever missed.
// ---------------------------------------------------------------------------
__global__ void k_grid_params(GridParams* gp, const uint32_t* __restrict__ rmax_bits, float bmin_x, float bmin_y,
                              float bmin_z, float ext_x, float ext_y, float ext_z, int cell_cap,
                              const int* __restrict__ nstored, Counters* __restrict__ ctr) {
    float rmax = __uint_as_float(*rmax_bits);
    float cell = rmax > 0.0f ? rmax : 1e-3f;
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

// per path: key + rank of each flattened photon
__global__ void __launch_bounds__(256) k_photon_keys(SlotsDev slots, const int* __restrict__ off,
                                                     const GridParams* __restrict__ gp, int* __restrict__ cell_count,
                                                     int2* __restrict__ keyrank) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots.npaths) return;
    const GridParams g = *gp;
    const int c = slots.count[i];
    const int base = off[i];
    const size_t NP = (size_t)slots.npaths;
    for (int s = 0; s < c; ++s) {
        float4 p = slots.pos[(size_t)s * NP + i];
        int cx = cell_coord(p.x, g.bmin[0], g.inv_cell, g.nx);
        int cy = cell_coord(p.y, g.bmin[1], g.inv_cell, g.ny);
        int cz = cell_coord(p.z, g.bmin[2], g.inv_cell, g.nz);
        int key = (cz * g.ny + cy) * g.nx + cx;
        int rank = atomicAdd(&cell_count[key], 1);
        keyrank[base + s] = make_int2(key, rank);
    }
}

__global__ void __launch_bounds__(256) k_photon_scatter(SlotsDev slots, const int* __restrict__ off,
                                                        const int* __restrict__ cell_start,
                                                        const int2* __restrict__ keyrank, float4* __restrict__ gpos,
                                                        float4* __restrict__ gflux, float4* __restrict__ gdir) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots.npaths) return;
    const int c = slots.count[i];
    const int base = off[i];
    const size_t NP = (size_t)slots.npaths;
    for (int s = 0; s < c; ++s) {
        const size_t o = (size_t)s * NP + i;
        int2 kr = keyrank[base + s];
        int dst = cell_start[kr.x] + kr.y;
        float4 d = slots.dir[o];
        d.w = __int_as_float(base + s);
        gpos[dst] = slots.pos[o];
        gflux[dst] = slots.flux[o];
        gdir[dst] = d;
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
        float pdf_area = 1.0f / (c.LW * c.LH);
        f3 lens = (pos + (e1 * (smp.x - 0.5f)) * c.LW) + (e2 * (smp.y - 0.5f)) * c.LH;
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
// No environment light: after the hit point is made the rest of the eye path
// could only add IBL radiance (zero), so the path stops there.
// ---------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK) k_eye(SceneDev S, CamDev C, HitDev H, int step, int stack_cap,
                                                    const uint32_t* __restrict__ words, uint32_t* __restrict__ flags) {
    extern __shared__ uint32_t lds[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int NPIX = C.W * C.H;
    if (i >= NPIX) return;
    uint32_t* stk = lds + threadIdx.x;
    uint32_t ovf = 0, fl = 0;
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
    for (;;) {
        if (depth > 6) break;
        else if (depth > 4) {
            float rus = (float)exp((double)(4 - depth));
            if (real1(m, &ovf) > (double)rus) break;
            PDF *= rus;
        }
        HitRec h;
        if (!closest_hit<W>(S, cur_org, cur_dir, SPPM_RAY_OFFSET, stk, PATH_BLOCK, stack_cap, h, &fl)) {
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
        if (g.is_sppm) {  // makeHitPoint (SPPM.cpp:79-96)
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
            break;
        }
        float Ii, Io = 1.0f;
        if (near1(cur_index)) { Ii = 1.0f; Io = g.index; }
        else { Ii = cur_index; Io = 1.0f; }
        FParam fp{cur_dir, next_dir, Nf, Ii, Io, g.rough, g.absorb, g.refl};
        float pdf_s = 1.0f;
        next_dir = surf_sample(g.kind, fp, m, &ovf, &pdf_s);
        if (!rutil_valid_dir(next_dir)) break;
        if (g.opaque) {
            if ((double)dot(next_dir, Nf) < 1e-6) break;
        } else {
            next_index = !(dot(next_dir, Nf) < 0.0f) ? cur_index : Io;
        }
        fp.out = next_dir;
        pdf_s = surf_pdf(g.kind, fp);
        f3 outN = dot(Nf, next_dir) > 0 ? Nf : -Nf;
        next_org = next_org + outN * SPPM_RAY_OFFSET;
        float brdf = surf_brdf(g.kind, fp);
        float cs = smax(0.0f, smin(1.0f, dot(next_dir, outN)));
        if ((double)brdf < 1e-8) break;
        if (pdf_s < 0) break;
        T = T * ((g.albedo * cs) * brdf);
        PDF *= pdf_s;
        cur_org = next_org;
        cur_dir = next_dir;
        cur_index = next_index;
        ++depth;
    }
    hw = hw / campdf;  // createHitpoints: PixelWeight /= pdf (SPPM.cpp:346)
    H.pos[i] = make_float4(hpos.x, hpos.y, hpos.z, (float)type);
    H.nrm[i] = make_float4(hn.x, hn.y, hn.z, hidx);
    H.out[i] = make_float4(hout.x, hout.y, hout.z, hrough);
    H.weight[i] = make_float4(hw.x, hw.y, hw.z, hrefl);
    H.mat[i] = make_int2(kind, facet);
    H.absorb[i] = habs;
    if (ovf) fl |= 2u;
    if (fl) atomicOr(flags, fl);
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

__global__ void __launch_bounds__(GATHER_BLOCK) k_gather(HitDev H, SharedDev SH, GridDev G,
                                                         const GridParams* __restrict__ gp, int W, int Hh, float alpha,
                                                         int mode, int* __restrict__ counts,
                                                         const int* __restrict__ list_off,
                                                         int* __restrict__ lists, uint32_t* __restrict__ rmax_bits,
                                                         Counters* __restrict__ ctr, uint32_t* __restrict__ flags) {
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    const int tiles = ((W + 7) >> 3) * ((Hh + 7) >> 3) * 64;
    float rnew = 0.0f;
    int is_hit = 0, nb = 0;
    if (lin < tiles) {
        const int p = tiled_pixel(lin, W, Hh);
        if (p >= 0) {
            float4 hp = H.pos[p];
            float4 sr = SH.rad[p];
            rnew = sr.w;
            int type = (int)hp.w;
            int M = 0;
            if (type == 1) {
                const GridParams g = *gp;
                const f3 c = mk(hp.x, hp.y, hp.z);
                const float R = sr.w;
                const f3 bbmin = c - mk(1.0f, 1.0f, 1.0f) * R;
                const f3 bbmax = c + mk(1.0f, 1.0f, 1.0f) * R;
                const float R2 = R * R;
                const int x0 = cell_coord(bbmin.x, g.bmin[0], g.inv_cell, g.nx);
                const int x1 = cell_coord(bbmax.x, g.bmin[0], g.inv_cell, g.nx);
                const int y0 = cell_coord(bbmin.y, g.bmin[1], g.inv_cell, g.ny);
                const int y1 = cell_coord(bbmax.y, g.bmin[1], g.inv_cell, g.ny);
                const int z0 = cell_coord(bbmin.z, g.bmin[2], g.inv_cell, g.nz);
                const int z1 = cell_coord(bbmax.z, g.bmin[2], g.inv_cell, g.nz);
                float4 hn = H.nrm[p], ho = H.out[p], hw = H.weight[p];
                const int kind = H.mat[p].x;
                const float absorb = H.absorb[p];
                const f3 N = mk(hn.x, hn.y, hn.z);
                const f3 outd = mk(ho.x, ho.y, ho.z);
                const f3 wgt = mk(hw.x, hw.y, hw.z);
                f3 add = mk(0.0f, 0.0f, 0.0f);
                int lo = (mode == 2) ? list_off[p] : 0;
                for (int z = z0; z <= z1; ++z)
                    for (int y = y0; y <= y1; ++y) {
                        const int row = (z * g.ny + y) * g.nx;
                        const int qa = G.cell_start[row + x0];
                        const int qb = G.cell_start[row + x1 + 1];
                        for (int q = qa; q < qb; ++q) {
                            const float4 pp = G.pos[q];
                            // isIncludePoint (kdtree.h:78-88) then the strict radius test (SPPM.cpp:432)
                            if (bbmin.x > pp.x || bbmax.x < pp.x) continue;
                            if (bbmin.y > pp.y || bbmax.y < pp.y) continue;
                            if (bbmin.z > pp.z || bbmax.z < pp.z) continue;
                            const f3 d = mk(pp.x, pp.y, pp.z) - c;
                            if (!(norm2(d) < R2)) continue;
                            ++M;
                            if (mode == 1) continue;
                            const float4 pf = G.flux[q];
                            const float4 pd = G.dir[q];
                            const f3 indir = mk(pd.x, pd.y, pd.z);
                            const float Ii = pp.w;
                            const float Io = near1(pp.w) ? hn.w : 1.0f;
                            const f3 nrm = dot(indir, N) < 0 ? N : -N;
                            FParam fp{indir, outd, nrm, Ii, Io, ho.w, absorb, hw.w};
                            const float b = surf_brdf(kind, fp);
                            add = add + ((wgt * b) * mk(pf.x, pf.y, pf.z)) / pf.w;
                            if (mode == 2) lists[lo++] = __float_as_int(pd.w);
                        }
                    }
                if (mode != 1 && M > 0) {
                    int2 cn = SH.cnt[p];
                    const float num = (float)cn.x + alpha * (float)M;
                    const float den = (float)(cn.x + M);
                    f3 r = mk(sr.x, sr.y, sr.z);
                    r = ((r + add) * num) / den;
                    const float R1 = sr.w * sqrtf_(num / den);
                    SH.rad[p] = make_float4(r.x, r.y, r.z, R1);
                    SH.cnt[p] = make_int2((int)((float)cn.x + (float)M * alpha), cn.y + 1);
                    rnew = R1;
                    if (!(r.x == r.x) || !(r.y == r.y) || !(r.z == r.z) || __builtin_isinf(r.x) ||
                        __builtin_isinf(r.y) || __builtin_isinf(r.z))
                        atomicOr(flags, 1u);
                }
            }
            if (counts) counts[p] = M;
            is_hit = (type == 1);
            nb = M;
        }
    }
    // block max of the (new) radius -> next pass's cell size
    __shared__ float red[GATHER_BLOCK / 64];
    __shared__ int redh[GATHER_BLOCK / 64], redn[GATHER_BLOCK / 64];
    float v = rnew;
    for (int off = 32; off > 0; off >>= 1) {
        v = fmaxf(v, __shfl_xor(v, off, 64));
        is_hit += __shfl_xor(is_hit, off, 64);
        nb += __shfl_xor(nb, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
        redh[threadIdx.x >> 6] = is_hit;
        redn[threadIdx.x >> 6] = nb;
    }
    __syncthreads();
    if (threadIdx.x == 0 && mode != 1) {
        float mx = red[0];
        unsigned long long th = 0, tn = 0;
        for (int k = 0; k < GATHER_BLOCK / 64; ++k) {
            mx = fmaxf(mx, red[k]);
            th += (unsigned long long)redh[k];
            tn += (unsigned long long)redn[k];
        }
        atomicMax(rmax_bits, __float_as_uint(mx));  // radii are positive: bit order == value order
        atomicAdd(&ctr->hit_points, th);
        atomicAdd(&ctr->neighbours, tn);
    }
}

// reconstructImageFromSharedPoints (SPPM.cpp:473-493)
__global__ void k_reconstruct(SharedDev SH, int np, int total_photons, int passes, float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float4 sr = SH.rad[i];
    int2 cn = SH.cnt[i];
    // DirectLight is zero without an environment light: 0 / (0 + 1e-4) = 0
    const float dl = 0.0f / (0.0f + 0.0001f);
    float den = ((float)total_photons * ((float)cn.y + 0.0001f) / (0.0001f + (float)passes)) * SPPM_PI * sr.w * sr.w;
    f3 L = mk(dl, dl, dl) + mk(sr.x, sr.y, sr.z) / den;
    const float b = 1.2f;
    out[i] = make_float4(powf(L.x, 1.0f / b), powf(L.y, 1.0f / b), powf(L.z, 1.0f / b), 1.0f);
}

__global__ void k_init_shared(SharedDev SH, int np, float r0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    SH.rad[i] = make_float4(0.0f, 0.0f, 0.0f, r0);
    SH.cnt[i] = make_int2(0, 0);
}
__global__ void k_init_slots(SlotsDev s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)NSLOT * s.npaths) return;
    s.valid[i] = 0;
}
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// Batch closest hit (N_BVH::ClosestHit) for tests and tools.
template <int W>
__global__ void __launch_bounds__(PATH_BLOCK) k_trace(SceneDev S, int n, const float* __restrict__ org,
                                                      const float* __restrict__ dir, float tmin, int stack_cap,
                                                      int* __restrict__ facet, float* __restrict__ tuv,
                                                      uint32_t* __restrict__ flags) {
    extern __shared__ uint32_t lds[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t fl = 0;
    HitRec h;
    f3 o = mk(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
    f3 d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    bool hit = closest_hit<W>(S, o, d, tmin, lds + threadIdx.x, PATH_BLOCK, stack_cap, h, &fl);
    facet[i] = hit ? h.facet : -1;
    tuv[3 * i] = hit ? h.t : 3.40282347e38f;
    tuv[3 * i + 1] = hit ? h.u : 0.0f;
    tuv[3 * i + 2] = hit ? h.v : 0.0f;
    if (fl) atomicOr(flags, fl);
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static size_t stack_bytes(int cap) { return (size_t)cap * PATH_BLOCK * sizeof(uint32_t); }

void launch_camera_words(uint32_t* st_mt, int* st_pos, uint32_t* words, long long need, hipStream_t st) {
    hipLaunchKernelGGL(k_camera_words, dim3(1), dim3(64), 0, st, st_mt, st_pos, words, need);
}

void launch_photon(const SceneDev& S, const SlotsDev& slots, int N, int step, int cap, uint32_t* flags,
                   hipStream_t st) {
    dim3 g((N + PATH_BLOCK - 1) / PATH_BLOCK), b(PATH_BLOCK);
    if (S.width == 2) hipLaunchKernelGGL(k_photon<2>, g, b, stack_bytes(cap), st, S, slots, N, step, cap, flags);
    else if (S.width == 4) hipLaunchKernelGGL(k_photon<4>, g, b, stack_bytes(cap), st, S, slots, N, step, cap, flags);
    else hipLaunchKernelGGL(k_photon<8>, g, b, stack_bytes(cap), st, S, slots, N, step, cap, flags);
}

void launch_eye(const SceneDev& S, const CamDev& C, const HitDev& H, int step, int cap, const uint32_t* words,
                uint32_t* flags, hipStream_t st) {
    int np = C.W * C.H;
    dim3 g((np + PATH_BLOCK - 1) / PATH_BLOCK), b(PATH_BLOCK);
    if (S.width == 2) hipLaunchKernelGGL(k_eye<2>, g, b, stack_bytes(cap), st, S, C, H, step, cap, words, flags);
    else if (S.width == 4) hipLaunchKernelGGL(k_eye<4>, g, b, stack_bytes(cap), st, S, C, H, step, cap, words, flags);
    else hipLaunchKernelGGL(k_eye<8>, g, b, stack_bytes(cap), st, S, C, H, step, cap, words, flags);
}

void launch_trace(const SceneDev& S, int n, const float* org, const float* dir, float tmin, int cap, int* facet,
                  float* tuv, uint32_t* flags, hipStream_t st) {
    dim3 g((n + PATH_BLOCK - 1) / PATH_BLOCK), b(PATH_BLOCK);
    if (S.width == 2) hipLaunchKernelGGL(k_trace<2>, g, b, stack_bytes(cap), st, S, n, org, dir, tmin, cap, facet, tuv, flags);
    else if (S.width == 4) hipLaunchKernelGGL(k_trace<4>, g, b, stack_bytes(cap), st, S, n, org, dir, tmin, cap, facet, tuv, flags);
    else hipLaunchKernelGGL(k_trace<8>, g, b, stack_bytes(cap), st, S, n, org, dir, tmin, cap, facet, tuv, flags);
}

void launch_build_grid(const SlotsDev& slots, const int* off, const int* nstored, Counters* ctr, GridParams* gp,
                       const uint32_t* rmax_bits, const float* bmin, const float* ext, int cell_cap, int* cell_count,
                       int* cell_start, int* scan_tmp, int2* keyrank, float4* gpos, float4* gflux, float4* gdir,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_grid_params, dim3(1), dim3(1), 0, st, gp, rmax_bits, bmin[0], bmin[1], bmin[2], ext[0],
                       ext[1], ext[2], cell_cap, nstored, ctr);
    hipLaunchKernelGGL(k_zero_i32, dim3((cell_cap + 1 + 255) / 256), dim3(256), 0, st, cell_count, cell_cap + 1);
    int gP = (slots.npaths + 255) / 256;
    hipLaunchKernelGGL(k_photon_keys, dim3(gP), dim3(256), 0, st, slots, off, gp, cell_count, keyrank);
    // cell_start = exclusive scan over cell_cap+1 counts (unused cells are zero)
    exclusive_scan(cell_count, cell_start, cell_cap + 1, nullptr, scan_tmp, st);
    hipLaunchKernelGGL(k_photon_scatter, dim3(gP), dim3(256), 0, st, slots, off, cell_start, keyrank, gpos, gflux,
                       gdir);
}

void launch_photon_flat(const SlotsDev& slots, const int* off, float4* fpos, float4* fflux, float4* fdir,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_photon_flat, dim3((slots.npaths + 255) / 256), dim3(256), 0, st, slots, off, fpos, fflux,
                       fdir);
}

void launch_gather(const HitDev& H, const SharedDev& SH, const GridDev& G, const GridParams* gp, int W, int Hh,
                   float alpha, int mode, int* counts, const int* list_off, int* lists, uint32_t* rmax_bits,
                   Counters* ctr, uint32_t* flags, hipStream_t st) {
    int tiles = ((W + 7) >> 3) * ((Hh + 7) >> 3) * 64;
    hipLaunchKernelGGL(k_gather, dim3((tiles + GATHER_BLOCK - 1) / GATHER_BLOCK), dim3(GATHER_BLOCK), 0, st, H, SH, G,
                       gp, W, Hh, alpha, mode, counts, list_off, lists, rmax_bits, ctr, flags);
}

void launch_reconstruct(const SharedDev& SH, int np, int total, int passes, float4* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reconstruct, dim3((np + 255) / 256), dim3(256), 0, st, SH, np, total, passes, out);
}
void launch_init_shared(const SharedDev& SH, int np, float r0, hipStream_t st) {
    hipLaunchKernelGGL(k_init_shared, dim3((np + 255) / 256), dim3(256), 0, st, SH, np, r0);
}
void launch_init_slots(const SlotsDev& s, hipStream_t st) {
    size_t n = (size_t)NSLOT * s.npaths;
    hipLaunchKernelGGL(k_init_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s);
}
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, p, v);
}

}  // namespace sppm
