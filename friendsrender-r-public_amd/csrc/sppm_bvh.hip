// sppm_bvh.hip -- GPU construction of the closest-hit structure.
//
// Replaces N_BVH<Triangle_ID, Ray, N>::Build3 (bvh8.h:135-189), which the
// reference runs on the CPU with OpenMP.  Here the binary tree is built on
// the GPU:
//   1. padded triangle boxes (AABB::Expand, Intersect.cpp:556-575) + centroids
//   2. binned SAH, top-down and level by level (default), or 30-bit Morton
//      codes, a stable radix sort and a binary radix tree (Karras 2012, ties
//      broken by sorted index)
//   3. bottom-up box refit with one atomic flag per internal node
//   4. triangles reordered into leaf order as (v0, e1, e2) records
// and collapsed on the host into W-wide nodes (W = 2, 4, 8) by a dynamic
// program that minimises the wide tree's SAH cost (collapse_dp below).
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
// bounds, 32 bins (counts and boxes) along each axis with a centroid extent,
// the split with the lowest A_L N_L + A_R N_R over the three axes, then a
// stable in-place partition.  The
// segment's internal node id is its split position s (left = [lo, s]); every
// position boundary is split exactly once, so ids are unique and the result
// has the radix-tree format the refit and the wide collapse read: child
// (left, right) with leaves n-1+k, range (lo, hi), parent.
constexpr int SAH_BINS = 32;  // 16: as fast, 64: slower traversal (full-material scene, BVH8)
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
    __shared__ int bcnt[3][SAH_BINS];
    __shared__ int bmn[3][SAH_BINS][3], bmx[3][SAH_BINS][3];
    __shared__ int s_split, s_axis;
    __shared__ float s_lo[3], s_scale[3];
    __shared__ float s_cost[3];
    __shared__ int s_best[3];
    __shared__ int wsum[SAH_TB / 64];
    const int t = threadIdx.x;
    const SahSeg sg = qin[blockIdx.x];
    const int lo = sg.lo, hi = sg.hi, cnt = hi - lo + 1;
    if (t < 3) { cmin[t] = 0x7fffffff; cmax[t] = (int)0x80000000; }
    if (t < 3 * SAH_BINS) {
        const int ax = t / SAH_BINS, k = t % SAH_BINS;
        bcnt[ax][k] = 0;
        for (int a = 0; a < 3; ++a) { bmn[ax][k][a] = 0x7fffffff; bmx[ax][k][a] = (int)0x80000000; }
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
    if (t < 3) {
        const float ext = fkey_inv(cmax[t]) - fkey_inv(cmin[t]);
        s_lo[t] = fkey_inv(cmin[t]);
        s_scale[t] = ext > 0.0f ? (float)SAH_BINS * (1.0f - 1e-6f) / ext : 0.0f;
    }
    __syncthreads();
    auto bin_of = [&](uint32_t f, int ax) {
        const float4 a = tbmin[f], b = tbmax[f];
        const float c = ax == 0 ? 0.5f * (a.x + b.x) : (ax == 1 ? 0.5f * (a.y + b.y) : 0.5f * (a.z + b.z));
        int k = (int)((c - s_lo[ax]) * s_scale[ax]);
        return k < 0 ? 0 : (k >= SAH_BINS ? SAH_BINS - 1 : k);
    };
    // bins along every axis with a centroid extent
    for (int p = lo + t; p <= hi; p += SAH_TB) {
        const uint32_t f = ord[p];
        const float4 a = tbmin[f], b = tbmax[f];
        for (int ax = 0; ax < 3; ++ax) {
            if (!(s_scale[ax] > 0.0f)) continue;
            const int k = bin_of(f, ax);
            atomicAdd(&bcnt[ax][k], 1);
            atomicMin(&bmn[ax][k][0], fkey(a.x)); atomicMin(&bmn[ax][k][1], fkey(a.y));
            atomicMin(&bmn[ax][k][2], fkey(a.z));
            atomicMax(&bmx[ax][k][0], fkey(b.x)); atomicMax(&bmx[ax][k][1], fkey(b.y));
            atomicMax(&bmx[ax][k][2], fkey(b.z));
        }
    }
    __syncthreads();
    if (t < 3) {
        // the SAH sweep of axis t: suffix areas / counts, then prefixes
        const int ax = t;
        int best = -1;
        float bestc = 3.4e38f;
        if (s_scale[ax] > 0.0f) {
            float ra[SAH_BINS];
            int rc[SAH_BINS];
            float mn[3] = {3e38f, 3e38f, 3e38f}, mx[3] = {-3e38f, -3e38f, -3e38f};
            int c = 0;
            for (int k = SAH_BINS - 1; k >= 0; --k) {
                if (bcnt[ax][k]) {
                    for (int a = 0; a < 3; ++a) {
                        mn[a] = fminf(mn[a], fkey_inv(bmn[ax][k][a]));
                        mx[a] = fmaxf(mx[a], fkey_inv(bmx[ax][k][a]));
                    }
                }
                c += bcnt[ax][k];
                const float x = mx[0] - mn[0], y = mx[1] - mn[1], z = mx[2] - mn[2];
                ra[k] = c ? x * y + y * z + z * x : 0.0f;
                rc[k] = c;
            }
            float lmn[3] = {3e38f, 3e38f, 3e38f}, lmx[3] = {-3e38f, -3e38f, -3e38f};
            int lc = 0;
            for (int k = 0; k < SAH_BINS - 1; ++k) {
                if (bcnt[ax][k]) {
                    for (int a = 0; a < 3; ++a) {
                        lmn[a] = fminf(lmn[a], fkey_inv(bmn[ax][k][a]));
                        lmx[a] = fmaxf(lmx[a], fkey_inv(bmx[ax][k][a]));
                    }
                }
                lc += bcnt[ax][k];
                if (lc == 0 || rc[k + 1] == 0) continue;
                const float x = lmx[0] - lmn[0], y = lmx[1] - lmn[1], z = lmx[2] - lmn[2];
                const float cost = (x * y + y * z + z * x) * (float)lc + ra[k + 1] * (float)rc[k + 1];
                if (cost < bestc) { bestc = cost; best = k; }
            }
        }
        s_cost[ax] = bestc;
        s_best[ax] = best;
    }
    __syncthreads();
    if (t == 0) {
        // the cheapest axis (ties: the lower axis); -1: split at the middle of the range
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (s_cost[k] < s_cost[ax]) ax = k;
        s_axis = ax;
        s_split = s_best[ax];
    }
    __syncthreads();
    const int split_bin = s_split, ax = s_axis;
    int nleft;
    if (split_bin < 0) {
        nleft = cnt / 2;
    } else {
        // stable partition through tmp: left keeps order, then right keeps order
        const int lane = t & 63, w = t >> 6;
        int base_l = 0, base_r = 0;
        int total_left = 0;
        for (int k = 0; k <= split_bin; ++k) total_left += bcnt[ax][k];
        for (int c0 = lo; c0 <= hi; c0 += SAH_TB) {
            const int p = c0 + t;
            const bool valid = p <= hi;
            const uint32_t f = valid ? ord[p] : 0u;
            const bool left = valid && bin_of(f, ax) <= split_bin;
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

// ---- SAH-optimal collapse of the binary tree to W-wide nodes (host) ----------
// A dynamic program over the binary tree (after Ylitie, Karras and Laine,
// "Efficient incoherent ray traversal on GPUs through compressed wide BVHs",
// HPG 2017): C(m, i) is the least SAH cost of covering subtree m with at most
// i entries of its parent's W child slots, where an entry is a leaf (a
// subtree of at most DP_LEAF_MAX triangles, cost SA * (C_LEAF + triangles *
// C_TRI)) or a wide node (cost SA * C_NODE plus the best split of its own W
// slots between its two binary children).  The wide tree then follows the
// decisions from the root, breadth first.  Node boxes are the binary nodes'
// refitted boxes, so every triangle stays inside the boxes above it.
// C_TRI = 0.15 with leaves of up to 8 triangles measured fastest on the
// full-material scene (BVH8, three-axis SAH: 4.06 ms per pass against 4.13
// at 0.1 and 4.27 at 0.07; with the longest-axis SAH 0.05 / 0.3 / 0.6 and
// 4-triangle leaves were slower, and the photon pass took 2.88 ms against
// 3.01 with the earlier greedy collapse, which opened the largest child
// until a node had W).
namespace {
constexpr float DP_C_NODE = 1.0f;
constexpr float DP_C_TRI = 0.15f;
constexpr float DP_C_LEAF = 0.5f;  // entering a leaf (pop, loop set-up): 4.03 ms per pass, 4.07 without
constexpr int DP_LEAF_MAX = 8;  // a leaf reference holds count - 1 in 3 bits
int collapse_dp(const std::vector<int2>& child, const std::vector<int2>& range, const std::vector<float4>& bmin,
                const std::vector<float4>& bmax, int n, int root, int W, std::vector<float>& nodes) {
    const int NN = 2 * n - 1;
    const int NF = node_floats(W);
    auto area = [&](int m) {
        const float x = bmax[m].x - bmin[m].x, y = bmax[m].y - bmin[m].y, z = bmax[m].z - bmin[m].z;
        return x * y + y * z + z * x;
    };
    auto tris = [&](int m) { return m >= n - 1 ? 1 : range[m].y - range[m].x + 1; };
    const float INF = 3.0e38f;
    // cost[m * 9 + i], i = 1..W; how[m * 9 + i]: 0 leaf, 1 wide node (i = 1),
    // 2 as with i - 1 entries, 10 + k split k / i - k between the children
    std::vector<float> cost((size_t)NN * 9, INF);
    std::vector<unsigned char> how((size_t)NN * 9, 0);
    std::vector<float> dist((size_t)(n - 1) * 9, INF);  // best split of i entries (internal nodes)
    std::vector<unsigned char> dk((size_t)(n - 1) * 9, 0);
    // post-order over the binary tree
    std::vector<int> order;
    order.reserve(NN);
    {
        std::vector<int> st{root};
        while (!st.empty()) {
            const int m = st.back();
            st.pop_back();
            order.push_back(m);
            if (m < n - 1) {
                st.push_back(child[m].x);
                st.push_back(child[m].y);
            }
        }
    }
    for (size_t q = order.size(); q-- > 0;) {
        const int m = order[q];
        const float a = area(m);
        if (m >= n - 1) {
            for (int i = 1; i <= W; ++i) {
                cost[(size_t)m * 9 + i] = a * (DP_C_LEAF + DP_C_TRI);
                how[(size_t)m * 9 + i] = i == 1 ? 0 : 2;
            }
            continue;
        }
        const int L = child[m].x, R = child[m].y;
        for (int j = 2; j <= W; ++j) {
            float b = INF;
            int bk = 1;
            for (int k = 1; k < j; ++k) {
                const float c = cost[(size_t)L * 9 + k] + cost[(size_t)R * 9 + (j - k)];
                if (c < b) { b = c; bk = k; }
            }
            dist[(size_t)m * 9 + j] = b;
            dk[(size_t)m * 9 + j] = (unsigned char)bk;
        }
        const int p = tris(m);
        const float c_leaf = p <= DP_LEAF_MAX ? a * (DP_C_LEAF + p * DP_C_TRI) : INF;
        const float c_node = dist[(size_t)m * 9 + W] + a * DP_C_NODE;
        cost[(size_t)m * 9 + 1] = c_leaf <= c_node ? c_leaf : c_node;
        how[(size_t)m * 9 + 1] = c_leaf <= c_node ? 0 : 1;
        for (int i = 2; i <= W; ++i) {
            const float d = dist[(size_t)m * 9 + i], prev = cost[(size_t)m * 9 + i - 1];
            if (d < prev) {
                cost[(size_t)m * 9 + i] = d;
                how[(size_t)m * 9 + i] = (unsigned char)(10 + dk[(size_t)m * 9 + i]);
            } else {
                cost[(size_t)m * 9 + i] = prev;
                how[(size_t)m * 9 + i] = 2;
            }
        }
    }
    // the entries of binary node m given i slots
    std::vector<int> ent;
    std::vector<std::pair<int, int>> work;
    auto collect = [&](int m, int i) {
        work.clear();
        work.push_back({m, i});
        while (!work.empty()) {
            auto [x, j] = work.back();
            work.pop_back();
            while (j > 1 && how[(size_t)x * 9 + j] == 2) --j;
            if (j == 1) {
                ent.push_back(x);
            } else {
                const int k = how[(size_t)x * 9 + j] - 10;
                work.push_back({child[x].y, j - k});
                work.push_back({child[x].x, k});
            }
        }
    };
    nodes.clear();
    std::vector<int> queue{root};
    size_t head = 0;
    while (head < queue.size()) {
        const int m = queue[head];
        const int w = (int)head;
        ++head;
        nodes.resize((size_t)(w + 1) * NF, 0.0f);
        float* nd = nodes.data() + (size_t)w * NF;
        uint32_t cref[8];
        ent.clear();
        // a wide node splits its W slots between its two binary children
        const int k = dk[(size_t)m * 9 + W];
        collect(child[m].x, k);
        collect(child[m].y, W - k);
        if ((int)ent.size() > W) return -1;
        for (int c = 0; c < W; ++c) {
            if (c < (int)ent.size()) {
                const int x = ent[c];
                nd[0 * W + c] = bmin[x].x; nd[1 * W + c] = bmin[x].y; nd[2 * W + c] = bmin[x].z;
                nd[3 * W + c] = bmax[x].x; nd[4 * W + c] = bmax[x].y; nd[5 * W + c] = bmax[x].z;
                const int p = tris(x);
                if (x >= n - 1 || how[(size_t)x * 9 + 1] == 0) {
                    if (p > DP_LEAF_MAX) return -1;
                    const int first = x >= n - 1 ? x - (n - 1) : range[x].x;
                    cref[c] = LEAF_BIT | ((uint32_t)first << 3) | (uint32_t)(p - 1);
                } else {
                    cref[c] = (uint32_t)queue.size();
                    queue.push_back(x);
                }
            } else {
                nd[0 * W + c] = 1e30f; nd[1 * W + c] = 1e30f; nd[2 * W + c] = 1e30f;
                nd[3 * W + c] = -1e30f; nd[4 * W + c] = -1e30f; nd[5 * W + c] = -1e30f;
                cref[c] = EMPTY_CHILD;
            }
        }
        memcpy(nd + 6 * W, cref, sizeof(uint32_t) * W);
    }
    return (int)queue.size();
}
}  // namespace

int build_bvh(const float* pos, int n, int width, bool sah, hipStream_t st, BvhResult* out, char* err,
              size_t errlen) {
    const int TB = 256;
    const int G = (n + TB - 1) / TB;
    const int NF = node_floats(width);
    float4 *tbmin = nullptr, *tbmax = nullptr, *nbmin = nullptr, *nbmax = nullptr;
    uint32_t *cb = nullptr, *keys = nullptr, *vals = nullptr, *tk = nullptr, *tv = nullptr, *hist = nullptr;
    int2 *child = nullptr, *range = nullptr;
    int *parent = nullptr, *flag = nullptr, *counters = nullptr;
    SahSeg *sa = nullptr, *sb = nullptr;
    int root_id = 0;
    int nint = n > 1 ? n - 1 : 1;
    int maxw = n > 1 ? n - 1 : 1;
    int hcnt[2];
    int depth = 0;
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
        std::vector<int2> hc((size_t)n - 1), hr((size_t)n - 1);
        std::vector<float4> hlo((size_t)2 * n - 1), hhi((size_t)2 * n - 1);
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpy(hc.data(), child, sizeof(int2) * hc.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hr.data(), range, sizeof(int2) * hr.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hlo.data(), nbmin, sizeof(float4) * hlo.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hhi.data(), nbmax, sizeof(float4) * hhi.size(), hipMemcpyDeviceToHost));
        std::vector<float> hn;
        const int nw = collapse_dp(hc, hr, hlo, hhi, n, root_id, width, hn);
        if (nw <= 0 || nw > maxw) {
            snprintf(err, errlen, "BVH collapse failed (%d nodes)", nw);
            goto fail;
        }
        HIPCHK(hipMemcpy(out->nodes, hn.data(), sizeof(float) * hn.size(), hipMemcpyHostToDevice));
        out->nnodes = nw;
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
    hfree(nbmax); hfree(counters); hfree(sa); hfree(sb);
    return 0;
fail:
    hfree(tbmin); hfree(tbmax); hfree(cb); hfree(keys);
    hfree(vals); hfree(tk); hfree(tv); hfree(hist);
    hfree(child); hfree(range); hfree(parent); hfree(flag);
    hfree(nbmin); hfree(nbmax);
    hfree(counters); hfree(sa); hfree(sb);
    free_bvh(out);
    return -1;
}

}  // namespace sppm
