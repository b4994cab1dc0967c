// sppm_trace.h -- closest hit through the GPU-built wide BVH.
//
// Replaces N_BVH<Triangle_ID, Ray, 8>::ClosestHit (bvh8.h:568-641).  The
// triangle test is Isc::Intersect(Ray, Triangle_ID) (Intersect.cpp:67-89)
// with identical float operations; a hit is accepted when t > MinT and
// (t < best, or t == best and the facet id is lower): ties go to the lowest
// facet id, so the result is independent of the tree shape and width.
// Child boxes are padded by 0.001 like AABB::Expand (Intersect.cpp:556-575),
// and the slab test is inclusive, so no box containing an accepted hit is
// ever culled.  The slab distances are fma(bound, 1/d, -o/d): one FMA per
// plane, whose rounding differs from (bound - o)/d by far less than the
// padding (culling decides only which boxes are visited, never the hit).
//
// Traversal descends into the nearest intersected child and pushes the
// others far-to-near.  The first TRACE_LDS stack entries of a lane live in
// LDS (entry s of lane t at lds[s * 64 + t], bank-conflict free, with its
// box's entry distance at lds[(TRACE_LDS + s) * 64 + t], so a pop skips
// boxes that lie beyond the best hit found since the push); deeper entries
// spill to a global area (entry s at glob[(s - TRACE_LDS) * gstride]), so the
// LDS footprint, and with it the occupancy, does not depend on the tree
// depth.  The host sizes the total capacity from the measured depth.
#pragma once
#include "sppm_device.h"

namespace sppm {

struct TraceStack {
    uint32_t* lds;   // LDS base + lane; stride 64 (PATH_BLOCK)
    uint32_t* glob;  // spill area + the lane's slot; stride gstride
    int gstride;
    int cap;         // total entries
};

struct HitRec {
    float t, u, v;
    int facet;  // -1: miss
    int slot;   // leaf-triangle index (sorted order)
    bool tie;   // another facet is hit at exactly the same t: the reference keeps the
                // last-tested one (bvh8.h:591), which depends on its tree; here the lowest id
};

__device__ __forceinline__ bool tri_test(const float4* __restrict__ tri, int slot, f3 org, f3 dir, float& t,
                                         float& u, float& v) {
    float4 A = tri[3 * slot + 0];
    float4 B = tri[3 * slot + 1];
    float4 C = tri[3 * slot + 2];
    f3 p0 = mk(A.x, A.y, A.z);
    f3 e1 = mk(B.x, B.y, B.z);
    f3 e2 = mk(C.x, C.y, C.z);
    f3 ps = cross(dir, e2);
    float a = dot(e1, ps);
    if (a > -1.17549435e-38f && a < 1.17549435e-38f) return false;  // |a| < FLT_MIN
    float f = 1.0f / a;
    f3 s = org - p0;
    float uu = f * dot(s, ps);
    if (uu < 0.0f || uu > 1.0f) return false;
    f3 q = cross(s, e1);
    float vv = f * dot(dir, q);
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    t = f * dot(e2, q);
    u = uu;
    v = vv;
    return true;
}

__device__ __forceinline__ void cswap(float& da, uint32_t& ra, float& db, uint32_t& rb) {
    // keep ascending distance: (da, ra) <= (db, rb)
    if (db < da) {
        float td = da; da = db; db = td;
        uint32_t tr = ra; ra = rb; rb = tr;
    }
}
template <int W>
__device__ __forceinline__ void sort_children(float* d, uint32_t* r) {
    if constexpr (W == 2) {
        cswap(d[0], r[0], d[1], r[1]);
    } else if constexpr (W == 4) {
        cswap(d[0], r[0], d[1], r[1]); cswap(d[2], r[2], d[3], r[3]);
        cswap(d[0], r[0], d[2], r[2]); cswap(d[1], r[1], d[3], r[3]);
        cswap(d[1], r[1], d[2], r[2]);
    } else {
        // Batcher odd-even merge sort network for 8 inputs (19 comparators)
        cswap(d[0], r[0], d[1], r[1]); cswap(d[2], r[2], d[3], r[3]);
        cswap(d[4], r[4], d[5], r[5]); cswap(d[6], r[6], d[7], r[7]);
        cswap(d[0], r[0], d[2], r[2]); cswap(d[1], r[1], d[3], r[3]);
        cswap(d[4], r[4], d[6], r[6]); cswap(d[5], r[5], d[7], r[7]);
        cswap(d[1], r[1], d[2], r[2]); cswap(d[5], r[5], d[6], r[6]);
        cswap(d[0], r[0], d[4], r[4]); cswap(d[1], r[1], d[5], r[5]);
        cswap(d[2], r[2], d[6], r[6]); cswap(d[3], r[3], d[7], r[7]);
        cswap(d[2], r[2], d[4], r[4]); cswap(d[3], r[3], d[5], r[5]);
        cswap(d[1], r[1], d[2], r[2]); cswap(d[3], r[3], d[4], r[4]);
        cswap(d[5], r[5], d[6], r[6]);
    }
}

// Pops the next stack entry whose box may still hold an accepted hit: an
// LDS entry keeps its box's entry distance and is skipped once that is
// beyond best (the rule it was pushed under; a hit inside the box lies at or
// beyond it); spilled entries are always taken.  EMPTY_CHILD when none is left.
__device__ __forceinline__ uint32_t trace_pop(const TraceStack& ts, int& sp, float best) {
    while (sp > 0) {
        --sp;
        if (sp >= ts.cap) return EMPTY_CHILD;  // overflowed stack: stop (flagged)
        if (sp < TRACE_LDS) {
            if (__uint_as_float(ts.lds[(TRACE_LDS + sp) * 64]) > best) continue;
            return ts.lds[sp * 64];
        }
        return ts.glob[(size_t)(sp - TRACE_LDS) * ts.gstride];
    }
    return EMPTY_CHILD;
}
// Closest hit for a ray (origin, direction as stored in the Ray, MinT = tmin,
// MaxT = FLT_MAX).  Returns false on a miss.  *flags bit 2 is set if the
// stack capacity was exceeded (the host sizes it from the tree depth, so
// this never happens for a well-formed build).
template <int W>
__device__ bool closest_hit(const SceneDev& S, f3 org, f3 dir, float tmin, const TraceStack& ts, HitRec& h,
                            uint32_t* flags) {
    const float BIG = 1e16f;
    f3 inv;
    inv.x = fabs_(dir.x) < 1e-16f ? (dir.x > 0.0f ? BIG : -BIG) : 1.0f / dir.x;
    inv.y = fabs_(dir.y) < 1e-16f ? (dir.y > 0.0f ? BIG : -BIG) : 1.0f / dir.y;
    inv.z = fabs_(dir.z) < 1e-16f ? (dir.z > 0.0f ? BIG : -BIG) : 1.0f / dir.z;
    const f3 oi = mk(org.x * inv.x, org.y * inv.y, org.z * inv.z);
    float best = 3.40282347e38f;  // FLT_MAX = Ray::m_MaxT
    int bfacet = 0x7fffffff;
    int bslot = -1;
    float bu = 0.0f, bv = 0.0f;
    bool tie = false;
    const int NF = node_floats(W);
    int sp = 0;
    uint32_t ref = 0;  // root node
    for (;;) {
        if (ref & LEAF_BIT) {
            int start = (int)((ref & ~LEAF_BIT) >> 3);
            int count = (int)(ref & 7u) + 1;
            for (int k = 0; k < count; ++k) {
                int s = start + k;
                float t, u, v;
                if (tri_test(S.tri, s, org, dir, t, u, v)) {
                    int fid = __float_as_int(S.tri[3 * s].w);
                    if (t > tmin) {
                        tie = t == best ? (tie || fid != bfacet) : (t < best ? false : tie);
                        if (t < best || (t == best && fid < bfacet)) {
                            best = t; bu = u; bv = v; bfacet = fid; bslot = s;
                        }
                    }
                }
            }
        } else {
            const float* nd = S.nodes + (size_t)ref * NF;
            float d[W];
            uint32_t r[W];
            const uint32_t* cref = reinterpret_cast<const uint32_t*>(nd + 6 * W);
#pragma unroll
            for (int c = 0; c < W; ++c) {
                float tx0 = fmaf(nd[0 * W + c], inv.x, -oi.x), tx1 = fmaf(nd[3 * W + c], inv.x, -oi.x);
                float ty0 = fmaf(nd[1 * W + c], inv.y, -oi.y), ty1 = fmaf(nd[4 * W + c], inv.y, -oi.y);
                float tz0 = fmaf(nd[2 * W + c], inv.z, -oi.z), tz1 = fmaf(nd[5 * W + c], inv.z, -oi.z);
                float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
                float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
                uint32_t cr = cref[c];
                bool hit = (cr != EMPTY_CHILD) && (tf >= tn) && (tn <= best) && (tf >= 0.0f);
                d[c] = hit ? tn : 3.0e38f;
                r[c] = hit ? cr : EMPTY_CHILD;
            }
            sort_children<W>(d, r);
            if (r[0] != EMPTY_CHILD) {
                // push the farther children far-to-near, descend into the nearest
#pragma unroll
                for (int c = W - 1; c >= 1; --c) {
                    if (r[c] != EMPTY_CHILD) {
                        if (sp < TRACE_LDS) {
                            ts.lds[sp * 64] = r[c];
                            ts.lds[(TRACE_LDS + sp) * 64] = __float_as_uint(d[c]);
                        }
                        else if (sp < ts.cap) ts.glob[(size_t)(sp - TRACE_LDS) * ts.gstride] = r[c];
                        else *flags |= 4u;
                        ++sp;
                    }
                }
                ref = r[0];
                continue;
            }
        }
        ref = trace_pop(ts, sp, best);
        if (ref == EMPTY_CHILD) break;
    }
    if (bslot < 0 || !(best < 3.40282347e38f)) {
        h.facet = -1;
        h.slot = -1;
        h.t = 3.40282347e38f;
        h.tie = false;
        return false;
    }
    h.t = best; h.u = bu; h.v = bv; h.facet = bfacet; h.slot = bslot;
    h.tie = tie;
    return true;
}

// ---- calculateGeometry (SPPM.cpp:36-77) ------------------------------------
struct Geom {
    f3 pos, N;
    float uvx, uvy;
    int kind, opaque, is_sppm;
    f3 albedo, radiance;
    float rough, index, absorb, refl, transl;
};

__device__ inline void calc_geometry(const SceneDev& S, const HitRec& h, Geom& g) {
    const float4* tr = S.tri + 3 * h.slot;
    float4 A = tr[0], B = tr[1], C = tr[2];
    f3 p0 = mk(A.x, A.y, A.z), e1 = mk(B.x, B.y, B.z), e2 = mk(C.x, C.y, C.z);
    const float u = h.u, v = h.v;
    g.pos = (e1 * u + e2 * v) + p0;
    const float* nn = S.facet_nrm + 9 * (size_t)h.facet;
    f3 n0 = mk(nn[0], nn[1], nn[2]), n1 = mk(nn[3], nn[4], nn[5]), n2 = mk(nn[6], nn[7], nn[8]);
    g.N = normalized(((n1 - n0) * u + (n2 - n0) * v) + n0);
    const float2* tv = S.facet_uv + 3 * (size_t)h.facet;
    float2 t0 = tv[0], t1 = tv[1], t2 = tv[2];
    g.uvx = ((t1.x - t0.x) * u + (t2.x - t0.x) * v) + t0.x;
    g.uvy = ((t1.y - t0.y) * u + (t2.y - t0.y) * v) + t0.y;
    const DevAttr& a = S.attrs[S.facet_attr[h.facet]];
    g.kind = a.kind;
    g.opaque = a.opaque;
    g.is_sppm = a.is_sppm;
    g.albedo = mk(a.albedo[0], a.albedo[1], a.albedo[2]);
    g.radiance = mk(a.radiance[0], a.radiance[1], a.radiance[2]);
    g.rough = a.rough;
    g.index = a.index;
    g.absorb = a.absorb;
    g.refl = a.refl;
    g.transl = 0.0f;
    if ((double)a.radiance_norm2 > 1e-6) g.albedo = mk(1.0f, 1.0f, 1.0f);
    if (a.tex_rough >= 0) {
        TexDesc td = S.tex[a.tex_rough];
        float4 c = tex_bilinear(S.texels + td.offset, td.width, td.height, g.uvx, g.uvy);
        g.rough = smin(0.8f, smax(0.05f, 0.5f * c.x));
    }
    if (a.tex_albedo >= 0) {
        TexDesc td = S.tex[a.tex_albedo];
        float4 c = tex_bilinear(S.texels + td.offset, td.width, td.height, g.uvx, g.uvy);
        g.albedo = mk(c.x, c.y, c.z);
        g.radiance = g.radiance * g.albedo;
        g.transl = (float)(1.0 - (double)c.w);
    }
    if (a.tex_bump >= 0) {
        TexDesc td = S.tex[a.tex_bump];
        float4 c = tex_bilinear(S.texels + td.offset, td.width, td.height, g.uvx, g.uvy);
        f3 nm = normalized(mk(2.0f * c.x - 1.0f, 2.0f * c.y - 1.0f, 2.0f * c.z - 1.0f));
        nm = normalized(mk(nm.x, nm.z, nm.y));
        f3 E2 = g.N;
        f3 E1 = fabs_(E2.y) > LO_0P9 ? cross(E2, mk(1, 0, 0)) : cross(E2, mk(0, 1, 0));  // (double)|y| > 0.9
        f3 E3 = normalized(cross(E1, E2));
        // Matrix3 with columns E1,E2,E3; inverse via cofactors (Matrix.h:180-211)
        float m[9] = {E1.x, E2.x, E3.x, E1.y, E2.y, E3.y, E1.z, E2.z, E3.z};
        float C[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                int r0 = (i == 0) ? 1 : 0, r1 = (i == 2) ? 1 : 2;
                int c0 = (j == 0) ? 1 : 0, c1 = (j == 2) ? 1 : 2;
                float det2 = m[3 * r0 + c0] * m[3 * r1 + c1] - m[3 * r0 + c1] * m[3 * r1 + c0];
                C[3 * i + j] = det2 * (float)(((i + j) % 2) ? -1 : 1);
            }
        float det = m[0] * C[0] + m[3] * C[3] + m[6] * C[6];
        float ix = 1.0f / det;
        // inverse row i = ix * C[:, i] (adjugate = C^T)
        f3 r0 = mk(ix * C[0], ix * C[3], ix * C[6]);
        f3 r1 = mk(ix * C[1], ix * C[4], ix * C[7]);
        f3 r2 = mk(ix * C[2], ix * C[5], ix * C[8]);
        g.N = normalized(mk(dot(r0, nm), dot(r1, nm), dot(r2, nm)));
    }
}

}  // namespace sppm
