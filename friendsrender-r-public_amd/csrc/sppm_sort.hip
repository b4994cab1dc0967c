// sppm_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs
// and a device-wide exclusive scan, both hand-written for 64-lane waves.
//
// Radix pass (4-bit digits): count per tile -> scan of the digit-major tile
// histogram -> stable scatter.  Within a tile, the rank of an element among
// equal digits is computed with 4 wave ballots (lane match mask) plus a
// per-wave prefix held in LDS, so equal keys keep their input order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_kernels.h"

namespace sppm {

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ITEMS = 16;                       // elements per thread per tile
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS; // 4096
constexpr int RADIX_BITS = 4;
constexpr int RADIX = 1 << RADIX_BITS;

// n_dev (optional): the element count on the device, at most n (the grid is
// sized for n; tiles past the count write empty histograms)
__device__ __forceinline__ int sort_count(int n, const int* n_dev) {
    if (!n_dev) return n;
    const int m = *n_dev;
    return m < n ? m : n;
}

__global__ void __launch_bounds__(SORT_THREADS) k_radix_count(const uint32_t* __restrict__ keys, int n,
                                                              const int* __restrict__ n_dev, int shift,
                                                              int nblocks, uint32_t* __restrict__ hist) {
    n = sort_count(n, n_dev);
    __shared__ uint32_t h[RADIX];
    if (threadIdx.x < RADIX) h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * SORT_TILE;
    for (int i = 0; i < SORT_ITEMS; ++i) {
        int idx = base + i * SORT_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < RADIX) hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Single-block exclusive scan of up to a few hundred thousand uint32 entries.
__global__ void __launch_bounds__(1024) k_scan_small(uint32_t* __restrict__ data, int n) {
    __shared__ uint32_t part[1024];
    int per = (n + 1023) / 1024;
    int b = threadIdx.x * per;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i)
        if (b + i < n) s += data[b + i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int i = 0; i < per; ++i)
        if (b + i < n) {
            uint32_t v = data[b + i];
            data[b + i] = run;
            run += v;
        }
}

__global__ void __launch_bounds__(SORT_THREADS) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, int n,
                                                                const int* __restrict__ n_dev, int shift,
                                                                int nblocks, const uint32_t* __restrict__ hist) {
    n = sort_count(n, n_dev);
    if ((int)blockIdx.x * SORT_TILE >= n) return;  // uniform: the whole tile is past the count
    __shared__ uint32_t base[RADIX];
    __shared__ uint32_t wcnt[SORT_WAVES][RADIX];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (threadIdx.x < RADIX) base[threadIdx.x] = hist[threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    const int tile = blockIdx.x * SORT_TILE;
    for (int it = 0; it < SORT_ITEMS; ++it) {
        int idx = tile + it * SORT_THREADS + threadIdx.x;
        bool valid = idx < n;
        uint32_t k = valid ? kin[idx] : 0u;
        uint32_t v = valid ? vin[idx] : 0u;
        uint32_t d = valid ? ((k >> shift) & (RADIX - 1)) : RADIX;  // RADIX: no digit
        // lanes of this wave with the same digit
        uint64_t act = __ballot(valid);
        uint64_t match = act;
#pragma unroll
        for (int b = 0; b < RADIX_BITS; ++b) {
            uint64_t m = __ballot(valid && ((d >> b) & 1u));
            match &= ((d >> b) & 1u) ? m : ~m;
        }
        uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        uint32_t rank_in_wave = (uint32_t)__popcll(match & below);
        // the lowest matching lane publishes the wave's count for its digit
        if (threadIdx.x < SORT_WAVES * RADIX) (&wcnt[0][0])[threadIdx.x] = 0;
        __syncthreads();
        if (valid && rank_in_wave == 0) wcnt[wave][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d];
            for (int w = 0; w < wave; ++w) off += wcnt[w][d];
            uint32_t dst = off + rank_in_wave;
            kout[dst] = k;
            vout[dst] = v;
        }
        __syncthreads();
        if (threadIdx.x < RADIX) {
            uint32_t t = 0;
            for (int w = 0; w < SORT_WAVES; ++w) t += wcnt[w][threadIdx.x];
            base[threadIdx.x] += t;
        }
        __syncthreads();
    }
}

// Sorts n pairs (or the first *n_dev of them, *n_dev <= n, when n_dev is
// given) by the low `bits` bits of the key, stably.  Buffers tmp_k/tmp_v have
// n entries; hist has RADIX * ceil(n / SORT_TILE) entries.  Result ends in
// keys/vals.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* tmp_k, uint32_t* tmp_v, uint32_t* hist, int n,
                      int bits, hipStream_t st, const int* n_dev) {
    if (n <= 1) return;
    int nblocks = (n + SORT_TILE - 1) / SORT_TILE;
    int passes = (bits + RADIX_BITS - 1) / RADIX_BITS;
    if (passes & 1) passes += 1;  // even number of passes: result lands back in keys/vals
    uint32_t *ka = keys, *va = vals, *kb = tmp_k, *vb = tmp_v;
    for (int p = 0; p < passes; ++p) {
        int shift = p * RADIX_BITS;
        if (shift >= 32) shift = 28;  // extra pass on a high digit keeps order (stable)
        hipLaunchKernelGGL(k_radix_count, dim3(nblocks), dim3(SORT_THREADS), 0, st, ka, n, n_dev, shift, nblocks,
                           hist);
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, st, hist, RADIX * nblocks);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(SORT_THREADS), 0, st, ka, va, kb, vb, n, n_dev,
                           shift, nblocks, hist);
        uint32_t* t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
}
size_t radix_hist_entries(int n) { return (size_t)RADIX * (size_t)((n + SORT_TILE - 1) / SORT_TILE + 1); }

// ---- device-wide exclusive scan of int32 (large n) -------------------------
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
    __shared__ int ws[SCAN_THREADS / 64];
    int inc = wave_incl_scan(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wave) pre += ws[w];
        tot += ws[w];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                             int n, int* __restrict__ tile_sums) {
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int loc[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        loc[i] = (base + i < n) ? in[base + i] : 0;
        s += loc[i];
    }
    int tot;
    int pre = block_excl_scan(s, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = pre;
        pre += loc[i];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(int* __restrict__ out, int n,
                                                           const int* __restrict__ tile_pre, int* __restrict__ total,
                                                           int ntiles) {
    int base = blockIdx.x * SCAN_TILE;
    int add = tile_pre[blockIdx.x];
    for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS)
        if (base + i < n) out[base + i] += add;
    if (blockIdx.x == ntiles - 1 && threadIdx.x == 0 && total) *total = tile_pre[ntiles];
}

// out[i] = sum_{j<i} in[j]; *total (device) = sum of all.  tmp has
// scan_tmp_entries(n) ints.  in and out may alias.
size_t scan_tmp_entries(int n) {
    int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)nt + 1 + 64;
}
void exclusive_scan(const int* in, int* out, int n, int* total, int* tmp, hipStream_t st) {
    if (n <= 0) return;
    int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SCAN_THREADS), 0, st, in, out, n, tmp);
    // tile sums -> exclusive prefix with total in tmp[nt] (nt tiles fit one small scan)
    hipLaunchKernelGGL(k_tile_total_prep, dim3(1), dim3(1), 0, st, tmp, nt);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, st, reinterpret_cast<uint32_t*>(tmp), nt + 1);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(SCAN_THREADS), 0, st, out, n, tmp, total, nt);
}
__global__ void k_tile_total_prep(int* tmp, int nt) { tmp[nt] = 0; }

}  // namespace sppm
