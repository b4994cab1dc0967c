// sppm_bvh.hip -- GPU construction of the closest-hit structure.
//
// Replaces N_BVH<Triangle_ID, Ray, N>::Build3 (bvh8.h:135-189), which the
// reference runs on the CPU with OpenMP.  Here the whole build is on the GPU:
//   1. padded triangle boxes (AABB::Expand, Intersect.cpp:556-575) + centroids
//   2. 30-bit Morton codes of the centroids, stable radix sort
//   3. binary radix tree (Karras 2012, ties broken by sorted index)
//   4. bottom-up box refit with one atomic flag per internal node
//   5. top-down collapse into W-wide nodes (W = 2, 4, 8): each wide node
//      repeatedly opens its child with the largest surface area; subtrees of
//      at most LEAF_MAX triangles become one leaf range
//   6. triangles reordered into leaf order as (v0, e1, e2) records
// Closest-hit results do not depend on the tree (ties go to the lowest facet
// id), so the build is free to optimise traversal cost.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <utility>
#include <vector>

#include "sppm_kernels.h"

namespace sppm {

__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_tri_prep(const float* __restrict__ pos, int n, float4* __restrict__ bmin, float4* __restrict__ bmax,
                           uint32_t* __restrict__ cbounds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = pos + 9 * (size_t)i;
    float mn[3], mx[3], c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = p[k], b = p[3 + k], d = p[6 + k];
        float hi = (a > b ? a : b) > d ? (a > b ? a : b) : d;
        float lo = (a < b ? a : b) < d ? (a < b ? a : b) : d;
        mx[k] = hi + 0.001f;
        mn[k] = lo - 0.001f;
        c[k] = (a + b + d) * (1.0f / 3.0f);
    }
    bmin[i] = make_float4(mn[0], mn[1], mn[2], 0.0f);
    bmax[i] = make_float4(mx[0], mx[1], mx[2], 0.0f);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicMin(&cbounds[k], f2ord(c[k]));
        atomicMax(&cbounds[3 + k], f2ord(c[k]));
    }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ void k_morton(const float* __restrict__ pos, int n, const uint32_t* __restrict__ cbounds,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = pos + 9 * (size_t)i;
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = ord2f(cbounds[k]), hi = ord2f(cbounds[3 + k]);
        float c = (p[k] + p[3 + k] + p[6 + k]) * (1.0f / 3.0f);
        float ext = hi - lo;
        float t = ext > 0.0f ? (c - lo) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        q[k] = (uint32_t)fminf(t * 1024.0f, 1023.0f);
    }
    keys[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    vals[i] = (uint32_t)i;
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// Internal node i (0..n-2); leaves are numbered n-1+k for sorted position k.
__global__ void k_karras(const uint32_t* __restrict__ keys, int n, int2* __restrict__ child,
                         int2* __restrict__ range, int* __restrict__ parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(lo, hi);
    parent[left] = i;
    parent[right] = i;
}

// Leaf k boxes copied to node n-1+k, then walk up; the second thread to reach
// a node computes its box (union of its two children); the walk ends at the
// root (node 0 of the radix tree, the root's split position for SAH).
__global__ void k_refit(const float4* __restrict__ tbmin, const float4* __restrict__ tbmax,
                        const uint32_t* __restrict__ order, int n, const int2* __restrict__ child,
                        const int* __restrict__ parent, float4* nbmin, float4* nbmax, int* __restrict__ flag,
                        int root) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int node = n - 1 + k;
    uint32_t tri = order[k];
    nbmin[node] = tbmin[tri];
    nbmax[node] = tbmax[tri];
    if (n == 1) return;
    for (;;) {
        if (node == root) return;
        __threadfence();
        int p = parent[node];
        if (p < 0 || p >= n - 1) return;  // not reached for a well-formed tree (host validation follows)
        if (atomicAdd(&flag[p], 1) == 0) return;
        __threadfence();
        int2 c = child[p];
        float4 a0 = nbmin[c.x];
        float4 a1 = nbmax[c.x];
        float4 b0 = nbmin[c.y];
        float4 b1 = nbmax[c.y];
        nbmin[p] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
        nbmax[p] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
        node = p;
    }
}

__device__ __forceinline__ float half_area(float4 a, float4 b) {
    float x = b.x - a.x, y = b.y - a.y, z = b.z - a.z;
    return x * y + y * z + z * x;
}

template <int W>
__global__ void k_collapse(const int2* __restrict__ qin, int nin, int2* __restrict__ qout, int* __restrict__ counters,
                           const int2* __restrict__ child, const int2* __restrict__ range,
                           const float4* __restrict__ nbmin, const float4* __restrict__ nbmax, int n,
                           float* __restrict__ nodes) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nin) return;
    int2 e = qin[q];
    int b = e.x, wnode = e.y;
    int c[W];
    int cnt = 2;
    c[0] = child[b].x;
    c[1] = child[b].y;
#pragma unroll
    for (int k = 2; k < W; ++k) c[k] = -1;
    auto size_of = [&](int node) -> int {
        if (node >= n - 1) return 1;
        int2 r = range[node];
        return r.y - r.x + 1;
    };
    while (cnt < W) {
        int best = -1;
        float ba = -1.0f;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (k < cnt && c[k] < n - 1 && size_of(c[k]) > LEAF_MAX) {
                float a = half_area(nbmin[c[k]], nbmax[c[k]]);
                if (a > ba) { ba = a; best = k; }
            }
        }
        if (best < 0) break;
        int open = c[best];
        int2 ch = child[open];
#pragma unroll
        for (int k = 0; k < W; ++k) if (k == best) c[k] = ch.x;
#pragma unroll
        for (int k = 0; k < W; ++k) if (k == cnt) c[k] = ch.y;
        ++cnt;
    }
    const int NF = node_floats(W);
    float* nd = nodes + (size_t)wnode * NF;
    uint32_t* cref = reinterpret_cast<uint32_t*>(nd + 6 * W);
#pragma unroll
    for (int k = 0; k < W; ++k) {
        if (k < cnt) {
            int node = c[k];
            float4 lo = nbmin[node], hi = nbmax[node];
            nd[0 * W + k] = lo.x; nd[1 * W + k] = lo.y; nd[2 * W + k] = lo.z;
            nd[3 * W + k] = hi.x; nd[4 * W + k] = hi.y; nd[5 * W + k] = hi.z;
            int sz = size_of(node);
            if (sz <= LEAF_MAX) {
                int first = node >= n - 1 ? node - (n - 1) : range[node].x;
                cref[k] = LEAF_BIT | ((uint32_t)first << 3) | (uint32_t)(sz - 1);
            } else {
                int w = atomicAdd(&counters[0], 1);
                int slot = atomicAdd(&counters[1], 1);
                qout[slot] = make_int2(node, w);
                cref[k] = (uint32_t)w;
            }
        } else {
            nd[0 * W + k] = 1e30f; nd[1 * W + k] = 1e30f; nd[2 * W + k] = 1e30f;
            nd[3 * W + k] = -1e30f; nd[4 * W + k] = -1e30f; nd[5 * W + k] = -1e30f;
            cref[k] = EMPTY_CHILD;
        }
    }
    for (int k = 7 * W; k < NF; ++k) nd[k] = 0.0f;
}

__global__ void k_tri_reorder(const float* __restrict__ pos, const uint32_t* __restrict__ order, int n,
                              float4* __restrict__ tri) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t f = order[k];
    const float* p = pos + 9 * (size_t)f;
    float v0x = p[0], v0y = p[1], v0z = p[2];
    tri[3 * k + 0] = make_float4(v0x, v0y, v0z, __int_as_float((int)f));
    tri[3 * k + 1] = make_float4(p[3] - v0x, p[4] - v0y, p[5] - v0z, 0.0f);
    tri[3 * k + 2] = make_float4(p[6] - v0x, p[7] - v0y, p[8] - v0z, 0.0f);
}

// ---- binned-SAH top-down builder ---------------------------------------------
// One workgroup per open segment [lo, hi] of the primitive order: centroid
// bounds, the axis of largest extent, 32 bins (counts and boxes), the split
// with the lowest A_L N_L + A_R N_R, then a stable in-place partition.  The
// segment's internal node id is its split position s (left = [lo, s]); every
// position boundary is split exactly once, so ids are unique and the result
// has the radix-tree format the refit and the wide collapse read: child
// (left, right) with leaves n-1+k, range (lo, hi), parent.
constexpr int SAH_BINS = 32;
constexpr int SAH_TB = 256;
struct SahSeg {
    int lo, hi, parent, side;  // side 0: left child of parent, 1: right
};
__device__ __forceinline__ int fkey(float f) {  // order-preserving float -> int
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float fkey_inv(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ void __launch_bounds__(SAH_TB) k_sah_split(const SahSeg* __restrict__ qin, int nin, SahSeg* __restrict__ qout,
                                                      int* __restrict__ qcount, uint32_t* __restrict__ ord,
                                                      uint32_t* __restrict__ tmp, const float4* __restrict__ tbmin,
                                                      const float4* __restrict__ tbmax, int n, int2* __restrict__ child,
                                                      int2* __restrict__ range, int* __restrict__ parent,
                                                      int* __restrict__ root) {
    __shared__ int cmin[3], cmax[3];
    __shared__ int bcnt[SAH_BINS];
    __shared__ int bmn[SAH_BINS][3], bmx[SAH_BINS][3];
    __shared__ int s_split, s_axis;
    __shared__ float s_lo, s_scale;
    __shared__ int wsum[SAH_TB / 64];
    const int t = threadIdx.x;
    const SahSeg sg = qin[blockIdx.x];
    const int lo = sg.lo, hi = sg.hi, cnt = hi - lo + 1;
    if (t < 3) { cmin[t] = 0x7fffffff; cmax[t] = (int)0x80000000; }
    if (t < SAH_BINS) {
        bcnt[t] = 0;
        for (int a = 0; a < 3; ++a) { bmn[t][a] = 0x7fffffff; bmx[t][a] = (int)0x80000000; }
    }
    __syncthreads();
    // centroid bounds
    for (int p = lo + t; p <= hi; p += SAH_TB) {
        const uint32_t f = ord[p];
        const float4 a = tbmin[f], b = tbmax[f];
        const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
        for (int k = 0; k < 3; ++k) {
            atomicMin(&cmin[k], fkey(c[k]));
            atomicMax(&cmax[k], fkey(c[k]));
        }
    }
    __syncthreads();
    if (t == 0) {
        int ax = 0;
        float ext = -1.0f;
        for (int k = 0; k < 3; ++k) {
            const float e = fkey_inv(cmax[k]) - fkey_inv(cmin[k]);
            if (e > ext) { ext = e; ax = k; }
        }
        s_axis = ax;
        s_lo = fkey_inv(cmin[ax]);
        s_scale = ext > 0.0f ? (float)SAH_BINS * (1.0f - 1e-6f) / ext : 0.0f;
    }
    __syncthreads();
    const int ax = s_axis;
    const float blo = s_lo, bs = s_scale;
    auto bin_of = [&](uint32_t f) {
        const float4 a = tbmin[f], b = tbmax[f];
        const float c = ax == 0 ? 0.5f * (a.x + b.x) : (ax == 1 ? 0.5f * (a.y + b.y) : 0.5f * (a.z + b.z));
        int k = (int)((c - blo) * bs);
        return k < 0 ? 0 : (k >= SAH_BINS ? SAH_BINS - 1 : k);
    };
    if (bs > 0.0f) {
        for (int p = lo + t; p <= hi; p += SAH_TB) {
            const uint32_t f = ord[p];
            const int k = bin_of(f);
            const float4 a = tbmin[f], b = tbmax[f];
            atomicAdd(&bcnt[k], 1);
            atomicMin(&bmn[k][0], fkey(a.x)); atomicMin(&bmn[k][1], fkey(a.y)); atomicMin(&bmn[k][2], fkey(a.z));
            atomicMax(&bmx[k][0], fkey(b.x)); atomicMax(&bmx[k][1], fkey(b.y)); atomicMax(&bmx[k][2], fkey(b.z));
        }
    }
    __syncthreads();
    if (t == 0) {
        int best = -1;
        if (bs > 0.0f) {
            // suffix areas / counts, then the sweep
            float ra[SAH_BINS];
            int rc[SAH_BINS];
            float mn[3] = {3e38f, 3e38f, 3e38f}, mx[3] = {-3e38f, -3e38f, -3e38f};
            int c = 0;
            for (int k = SAH_BINS - 1; k >= 0; --k) {
                if (bcnt[k]) {
                    for (int a = 0; a < 3; ++a) {
                        mn[a] = fminf(mn[a], fkey_inv(bmn[k][a]));
                        mx[a] = fmaxf(mx[a], fkey_inv(bmx[k][a]));
                    }
                }
                c += bcnt[k];
                const float x = mx[0] - mn[0], y = mx[1] - mn[1], z = mx[2] - mn[2];
                ra[k] = c ? x * y + y * z + z * x : 0.0f;
                rc[k] = c;
            }
            float lmn[3] = {3e38f, 3e38f, 3e38f}, lmx[3] = {-3e38f, -3e38f, -3e38f};
            int lc = 0;
            float bestc = 3.4e38f;
            for (int k = 0; k < SAH_BINS - 1; ++k) {
                if (bcnt[k]) {
                    for (int a = 0; a < 3; ++a) {
                        lmn[a] = fminf(lmn[a], fkey_inv(bmn[k][a]));
                        lmx[a] = fmaxf(lmx[a], fkey_inv(bmx[k][a]));
                    }
                }
                lc += bcnt[k];
                if (lc == 0 || rc[k + 1] == 0) continue;
                const float x = lmx[0] - lmn[0], y = lmx[1] - lmn[1], z = lmx[2] - lmn[2];
                const float cost = (x * y + y * z + z * x) * (float)lc + ra[k + 1] * (float)rc[k + 1];
                if (cost < bestc) { bestc = cost; best = k; }
            }
        }
        s_split = best;  // -1: split at the middle of the range
    }
    __syncthreads();
    const int split_bin = s_split;
    int nleft;
    if (split_bin < 0) {
        nleft = cnt / 2;
    } else {
        // stable partition through tmp: left keeps order, then right keeps order
        const int lane = t & 63, w = t >> 6;
        int base_l = 0, base_r = 0;
        int total_left = 0;
        for (int k = 0; k <= split_bin; ++k) total_left += bcnt[k];
        for (int c0 = lo; c0 <= hi; c0 += SAH_TB) {
            const int p = c0 + t;
            const bool valid = p <= hi;
            const uint32_t f = valid ? ord[p] : 0u;
            const bool left = valid && bin_of(f) <= split_bin;
            const unsigned long long bl = __ballot(left);
            const int rank_w = __popcll(bl & ((1ull << lane) - 1ull));
            if (lane == 0) wsum[w] = __popcll(bl);
            __syncthreads();
            int before = 0, tot = 0;
            for (int k = 0; k < SAH_TB / 64; ++k) {
                if (k < w) before += wsum[k];
                tot += wsum[k];
            }
            const int nchunk = min(SAH_TB, hi - c0 + 1);
            if (valid) {
                if (left) tmp[lo + base_l + before + rank_w] = f;
                else tmp[lo + total_left + base_r + (t - (before + rank_w))] = f;
            }
            base_l += tot;
            base_r += nchunk - tot;
            __syncthreads();
        }
        for (int p = lo + t; p <= hi; p += SAH_TB) ord[p] = tmp[p];
        nleft = total_left;
    }
    __syncthreads();
    if (t == 0) {
        const int s = lo + nleft - 1;  // this node's id
        const int lid = nleft == 1 ? (n - 1 + lo) : -1;
        const int rid = cnt - nleft == 1 ? (n - 1 + hi) : -1;
        child[s] = make_int2(lid, rid);
        range[s] = make_int2(lo, hi);
        if (sg.parent < 0) {
            *root = s;
        } else {
            if (sg.side == 0) child[sg.parent].x = s;
            else child[sg.parent].y = s;
            parent[s] = sg.parent;
        }
        if (lid >= 0) parent[lid] = s;
        if (rid >= 0) parent[rid] = s;
        if (nleft > 1) qout[atomicAdd(qcount, 1)] = SahSeg{lo, s, s, 0};
        if (cnt - nleft > 1) qout[atomicAdd(qcount, 1)] = SahSeg{s + 1, hi, s, 1};
    }
}

__global__ void k_iota_u32(uint32_t* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static void hfree(void* p) {
    if (p) (void)hipFree(p);
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            snprintf(err, errlen, "%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            goto fail;                                                                    \
        }                                                                                 \
    } while (0)

void free_bvh(BvhResult* r) {
    if (r->nodes) (void)hipFree(r->nodes);
    if (r->tri) (void)hipFree(r->tri);
    r->nodes = nullptr;
    r->tri = nullptr;
    r->nnodes = 0;
}

int build_bvh(const float* pos, int n, int width, bool sah, hipStream_t st, BvhResult* out, char* err,
              size_t errlen) {
    const int TB = 256;
    const int G = (n + TB - 1) / TB;
    const int NF = node_floats(width);
    float4 *tbmin = nullptr, *tbmax = nullptr, *nbmin = nullptr, *nbmax = nullptr;
    uint32_t *cb = nullptr, *keys = nullptr, *vals = nullptr, *tk = nullptr, *tv = nullptr, *hist = nullptr;
    int2 *child = nullptr, *range = nullptr, *qa = nullptr, *qb = nullptr;
    int *parent = nullptr, *flag = nullptr, *counters = nullptr;
    SahSeg *sa = nullptr, *sb = nullptr;
    int root_id = 0;
    int nint = n > 1 ? n - 1 : 1;
    int maxw = n > 1 ? n - 1 : 1;
    int hcnt[2];
    int nin = 1, depth = 0;
    free_bvh(out);
    out->width = width;
    if (n <= 0) {
        snprintf(err, errlen, "empty scene");
        return -1;
    }
    if (n > RADIX_SORT_MAX_N) {
        snprintf(err, errlen, "scene of %d triangles: at most %d", n, RADIX_SORT_MAX_N);
        return -1;
    }
    HIPCHK(hipMalloc(&tbmin, sizeof(float4) * n));
    HIPCHK(hipMalloc(&tbmax, sizeof(float4) * n));
    HIPCHK(hipMalloc(&cb, sizeof(uint32_t) * 6));
    HIPCHK(hipMalloc(&keys, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&vals, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&tk, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&tv, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&hist, sizeof(uint32_t) * radix_hist_entries(n)));
    HIPCHK(hipMalloc(&child, sizeof(int2) * nint));
    HIPCHK(hipMalloc(&range, sizeof(int2) * nint));
    HIPCHK(hipMalloc(&parent, sizeof(int) * (2 * n)));
    HIPCHK(hipMalloc(&flag, sizeof(int) * nint));
    HIPCHK(hipMalloc(&nbmin, sizeof(float4) * (2 * n)));
    HIPCHK(hipMalloc(&nbmax, sizeof(float4) * (2 * n)));
    HIPCHK(hipMalloc(&qa, sizeof(int2) * maxw));
    HIPCHK(hipMalloc(&qb, sizeof(int2) * maxw));
    HIPCHK(hipMalloc(&counters, sizeof(int) * 2));
    if (sah) {
        HIPCHK(hipMalloc(&sa, sizeof(SahSeg) * (size_t)n));
        HIPCHK(hipMalloc(&sb, sizeof(SahSeg) * (size_t)n));
    }
    HIPCHK(hipMalloc(&out->nodes, sizeof(float) * NF * (size_t)maxw));
    HIPCHK(hipMalloc(&out->tri, sizeof(float4) * 3 * (size_t)n));
    {
        uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
        HIPCHK(hipMemcpyAsync(cb, init, sizeof(init), hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_tri_prep, dim3(G), dim3(TB), 0, st, pos, n, tbmin, tbmax, cb);
    if (sah && n > 1) {
        // binned-SAH top-down, one launch per level (one workgroup per open segment)
        hipLaunchKernelGGL(k_iota_u32, dim3(G), dim3(TB), 0, st, vals, n);
        hipLaunchKernelGGL(k_fill_u32, dim3((2 * n + TB - 1) / TB), dim3(TB), 0, st, (uint32_t*)parent, 2 * n,
                           0xffffffffu);
        SahSeg r0{0, n - 1, -1, 0};
        HIPCHK(hipMemcpyAsync(sa, &r0, sizeof r0, hipMemcpyHostToDevice, st));
        int* rootp = counters;  // counters[0] = root id, counters[1] = next queue size
        int nseg = 1;
        while (nseg > 0) {
            int zero = 0;
            HIPCHK(hipMemcpyAsync(counters + 1, &zero, sizeof(int), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_sah_split, dim3(nseg), dim3(SAH_TB), 0, st, sa, nseg, sb, counters + 1, vals, tv,
                               tbmin, tbmax, n, child, range, parent, rootp);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hcnt, counters, sizeof(hcnt), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (hcnt[1] < 0 || hcnt[1] > n) {
                snprintf(err, errlen, "SAH build: bad segment count %d", hcnt[1]);
                goto fail;
            }
            nseg = hcnt[1];
            SahSeg* t = sa; sa = sb; sb = t;
        }
        root_id = hcnt[0];
        if (root_id < 0 || root_id >= n - 1) {
            snprintf(err, errlen, "SAH build: bad root %d", root_id);
            goto fail;
        }
        {
            // host check of the binary tree before the refit and the collapse
            // read it: every child reference in range, every node but the root
            // referenced exactly once
            std::vector<int2> hc((size_t)n - 1);
            HIPCHK(hipMemcpy(hc.data(), child, sizeof(int2) * hc.size(), hipMemcpyDeviceToHost));
            std::vector<unsigned char> refd((size_t)2 * n - 1, 0);
            bool ok = true;
            for (int i = 0; i < n - 1 && ok; ++i)
                for (int r : {hc[i].x, hc[i].y}) {
                    if (r < 0 || r > 2 * n - 2 || refd[r]) { ok = false; break; }
                    refd[r] = 1;
                }
            for (int i = 0; i < 2 * n - 1 && ok; ++i) ok = (i == root_id) ? !refd[i] : refd[i] != 0;
            if (!ok) {
                snprintf(err, errlen, "SAH build: binary tree validation failed");
                goto fail;
            }
        }
        hipLaunchKernelGGL(k_fill_u32, dim3((nint + TB - 1) / TB), dim3(TB), 0, st, (uint32_t*)flag, nint, 0u);
    } else {
        hipLaunchKernelGGL(k_morton, dim3(G), dim3(TB), 0, st, pos, n, cb, keys, vals);
        SortResult sr = radix_sort_pairs(keys, vals, tk, tv, hist, n, 30, st);
        if (sr.keys != keys) {  // the sorted pairs are in the second buffers
            std::swap(keys, tk);
            std::swap(vals, tv);
        }
        HIPCHK(hipGetLastError());
        if (n > 1) {
            hipLaunchKernelGGL(k_karras, dim3((n - 1 + TB - 1) / TB), dim3(TB), 0, st, keys, n, child, range, parent);
            hipLaunchKernelGGL(k_fill_u32, dim3((nint + TB - 1) / TB), dim3(TB), 0, st, (uint32_t*)flag, nint, 0u);
        }
    }
    hipLaunchKernelGGL(k_refit, dim3(G), dim3(TB), 0, st, tbmin, tbmax, vals, n, child, parent, nbmin, nbmax, flag,
                       root_id);
    hipLaunchKernelGGL(k_tri_reorder, dim3(G), dim3(TB), 0, st, pos, vals, n, out->tri);
    HIPCHK(hipGetLastError());
    if (n == 1) {
        // one leaf under a single wide node
        std::vector<float> nd(NF, 0.0f);
        float4 lo, hi;
        HIPCHK(hipMemcpyAsync(&lo, nbmin, sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(&hi, nbmax, sizeof(float4), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int k = 0; k < width; ++k) {
            bool real = k == 0;
            nd[0 * width + k] = real ? lo.x : 1e30f; nd[1 * width + k] = real ? lo.y : 1e30f;
            nd[2 * width + k] = real ? lo.z : 1e30f; nd[3 * width + k] = real ? hi.x : -1e30f;
            nd[4 * width + k] = real ? hi.y : -1e30f; nd[5 * width + k] = real ? hi.z : -1e30f;
            uint32_t r = real ? (LEAF_BIT | 0u) : EMPTY_CHILD;
            memcpy(&nd[6 * width + k], &r, 4);
        }
        HIPCHK(hipMemcpyAsync(out->nodes, nd.data(), sizeof(float) * NF, hipMemcpyHostToDevice, st));
        out->nnodes = 1;
        depth = 1;
    } else {
        int2 root = make_int2(root_id, 0);  // radix tree: node 0; SAH: the root's split position
        int init[2] = {1, 0};
        HIPCHK(hipMemcpyAsync(qa, &root, sizeof(int2), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, st));
        while (nin > 0) {
            int zero = 0;
            HIPCHK(hipMemcpyAsync(counters + 1, &zero, sizeof(int), hipMemcpyHostToDevice, st));
            int g = (nin + TB - 1) / TB;
            if (width == 2)
                hipLaunchKernelGGL(k_collapse<2>, dim3(g), dim3(TB), 0, st, qa, nin, qb, counters, child, range, nbmin,
                                   nbmax, n, out->nodes);
            else if (width == 4)
                hipLaunchKernelGGL(k_collapse<4>, dim3(g), dim3(TB), 0, st, qa, nin, qb, counters, child, range, nbmin,
                                   nbmax, n, out->nodes);
            else
                hipLaunchKernelGGL(k_collapse<8>, dim3(g), dim3(TB), 0, st, qa, nin, qb, counters, child, range, nbmin,
                                   nbmax, n, out->nodes);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hcnt, counters, sizeof(hcnt), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            ++depth;
            nin = hcnt[1];
            int2* t = qa; qa = qb; qb = t;
        }
        out->nnodes = hcnt[0];
    }
    out->depth = depth;
    HIPCHK(hipStreamSynchronize(st));
    // Host validation of the tree before any kernel traverses it: every
    // internal reference in range, every leaf range inside the triangle
    // array, every triangle in exactly one leaf, depth as measured.
    {
        std::vector<float> hn((size_t)NF * out->nnodes);
        HIPCHK(hipMemcpy(hn.data(), out->nodes, sizeof(float) * hn.size(), hipMemcpyDeviceToHost));
        std::vector<unsigned char> seen(n, 0);
        std::vector<int> stackv{0};
        std::vector<int> dep{1};
        std::vector<unsigned char> visited(out->nnodes, 0);
        int maxd = 0;
        bool ok = true;
        while (!stackv.empty() && ok) {
            int nd = stackv.back(), d = dep.back();
            stackv.pop_back();
            dep.pop_back();
            if (nd < 0 || nd >= out->nnodes || visited[nd]) { ok = false; break; }
            visited[nd] = 1;
            maxd = d > maxd ? d : maxd;
            const uint32_t* cref = reinterpret_cast<const uint32_t*>(hn.data() + (size_t)nd * NF + 6 * width);
            for (int k = 0; k < width; ++k) {
                uint32_t r = cref[k];
                if (r == EMPTY_CHILD) continue;
                if (r & LEAF_BIT) {
                    int start = (int)((r & ~LEAF_BIT) >> 3), cnt = (int)(r & 7u) + 1;
                    if (start < 0 || start + cnt > n) { ok = false; break; }
                    for (int t = start; t < start + cnt; ++t) {
                        if (seen[t]) ok = false;
                        seen[t] = 1;
                    }
                } else {
                    stackv.push_back((int)r);
                    dep.push_back(d + 1);
                }
            }
        }
        for (int t = 0; t < n && ok; ++t) ok = seen[t] != 0;
        if (!ok) {
            snprintf(err, errlen, "BVH validation failed (bad reference or triangle coverage)");
            goto fail;
        }
        out->depth = maxd;
    }
    hfree(tbmin); hfree(tbmax); hfree(cb); hfree(keys); hfree(vals); hfree(tk); hfree(tv);
    hfree(hist); hfree(child); hfree(range); hfree(parent); hfree(flag); hfree(nbmin);
    hfree(nbmax); hfree(qa); hfree(qb); hfree(counters); hfree(sa); hfree(sb);
    return 0;
fail:
    hfree(tbmin); hfree(tbmax); hfree(cb); hfree(keys);
    hfree(vals); hfree(tk); hfree(tv); hfree(hist);
    hfree(child); hfree(range); hfree(parent); hfree(flag);
    hfree(nbmin); hfree(nbmax); hfree(qa); hfree(qb);
    hfree(counters); hfree(sa); hfree(sb);
    free_bvh(out);
    return -1;
}

}  // namespace sppm
