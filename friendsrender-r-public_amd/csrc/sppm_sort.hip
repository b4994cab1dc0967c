// sppm_sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs
// and a device-wide exclusive scan, both hand-written for 64-lane waves.
//
// Radix sort: 8-bit digits, one kernel per digit pass.  A prep kernel clears
// the work area and a histogram kernel counts the digits of every pass in one
// read of the keys; then each pass kernel claims tiles of 4096 pairs in order
// (atomic ticket), ranks the tile's pairs by digit, publishes the tile's digit
// counts and finds the counts of all earlier tiles by a look-back over their
// published states (aggregate / inclusive prefix), and writes the tile out
// through LDS so that each digit's run is stored contiguously.
//
// Stability: inside a tile wave w ranks pairs [w*1024, (w+1)*1024) in order,
// 64 at a time; equal digits within 64 lanes are ranked by lane through 8
// ballots, across iterations through a wave-private running count, across
// waves by the waves' exclusive prefix, across tiles by the look-back.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <utility>

#include "sppm_kernels.h"

namespace sppm {

constexpr int OS_THREADS = 256;
constexpr int OS_WAVES = OS_THREADS / 64;
constexpr int OS_ITEMS = 16;                         // pairs per lane per tile
constexpr int OS_WAVE_TILE = 64 * OS_ITEMS;          // 1024 consecutive pairs per wave
constexpr int OS_TILE = OS_WAVES * OS_WAVE_TILE;     // 4096
constexpr int OS_BITS = 8;
constexpr int OS_RADIX = 1 << OS_BITS;
constexpr int OS_MAX_PASSES = 4;
// work area (uint32 words): digit totals [4][256], tickets [64], then per
// pass one state word per (tile, digit)
constexpr size_t OS_TOT = 0;
constexpr size_t OS_TICKET = OS_MAX_PASSES * OS_RADIX;
constexpr size_t OS_STATUS = OS_TICKET + 64;
constexpr uint32_t ST_AGG = 1u << 30;   // the tile's own count is published
constexpr uint32_t ST_INC = 2u << 30;   // the count of all tiles up to and including it
constexpr uint32_t ST_VAL = ST_AGG - 1u;

// n_dev (optional): the element count on the device, at most n (the grid is
// sized for n; tiles past the count do nothing)
__device__ __forceinline__ int sort_count(int n, const int* n_dev) {
    if (!n_dev) return n;
    const int m = *n_dev;
    return m < n ? m : n;
}
static inline size_t os_tiles(int n) { return (size_t)((n + OS_TILE - 1) / OS_TILE); }

__global__ void __launch_bounds__(256) k_sort_prep(uint32_t* __restrict__ work, int n, const int* __restrict__ n_dev,
                                                   int passes, size_t stride) {
    n = sort_count(n, n_dev);
    const size_t used = (size_t)((n + OS_TILE - 1) / OS_TILE) * OS_RADIX;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = i0; i < OS_STATUS; i += step) work[i] = 0u;
    for (int p = 0; p < passes; ++p) {
        uint32_t* w = work + OS_STATUS + (size_t)p * stride;
        for (size_t i = i0; i < used; i += step) w[i] = 0u;
    }
}

__global__ void __launch_bounds__(256) k_sort_hist(const uint32_t* __restrict__ keys, int n,
                                                   const int* __restrict__ n_dev, int passes,
                                                   uint32_t* __restrict__ work) {
    n = sort_count(n, n_dev);
    __shared__ uint32_t h[OS_MAX_PASSES * OS_RADIX];
    for (int i = threadIdx.x; i < OS_MAX_PASSES * OS_RADIX; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p * OS_RADIX + ((k >> (p * OS_BITS)) & (OS_RADIX - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * OS_RADIX; i += blockDim.x)
        if (h[i]) atomicAdd(&work[OS_TOT + i], h[i]);
}

// exclusive scans over the 256 threads of a block of two values at once
__device__ __forceinline__ void block_excl_scan2(uint32_t a, uint32_t b, uint32_t* ea, uint32_t* eb) {
    __shared__ uint32_t wa[OS_WAVES], wb[OS_WAVES];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ta = __shfl_up(ia, off, 64), tb = __shfl_up(ib, off, 64);
        if (lane >= off) {
            ia += ta;
            ib += tb;
        }
    }
    if (lane == 63) {
        wa[wave] = ia;
        wb[wave] = ib;
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0;
    for (int w = 0; w < wave; ++w) {
        pa += wa[w];
        pb += wb[w];
    }
    *ea = pa + ia - a;
    *eb = pb + ib - b;
}

__device__ __forceinline__ void state_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t state_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One digit pass.  Every valid tile publishes its aggregate before it looks
// back, and tiles are claimed in ticket order, so each tile a look-back waits
// on is resident and has no wait of its own before it publishes: every wait
// ends.  Tiles past the count take a ticket and leave.
__global__ void __launch_bounds__(OS_THREADS) k_sort_pass(const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          int n, const int* __restrict__ n_dev, int shift,
                                                          const uint32_t* __restrict__ digit_tot,
                                                          uint32_t* __restrict__ ticket, uint32_t* __restrict__ state) {
    n = sort_count(n, n_dev);
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_cnt[OS_WAVES][OS_RADIX];
    __shared__ uint32_t s_lstart[OS_RADIX];
    __shared__ uint32_t s_gbase[OS_RADIX];
    __shared__ uint32_t s_k[OS_TILE];
    __shared__ uint32_t s_v[OS_TILE];
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < OS_WAVES * OS_RADIX; i += OS_THREADS) (&s_cnt[0][0])[i] = 0u;
    __syncthreads();
    const int tile = (int)s_tile;
    const long long t0 = (long long)tile * OS_TILE;
    if (t0 >= n) return;  // uniform over the block
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long wbase = t0 + (long long)wave * OS_WAVE_TILE + lane;
    uint32_t key[OS_ITEMS], val[OS_ITEMS], rank[OS_ITEMS];
#pragma unroll
    for (int it = 0; it < OS_ITEMS; ++it) {
        const long long idx = wbase + it * 64;
        const bool ok = idx < n;
        key[it] = ok ? kin[idx] : 0u;
        val[it] = ok ? vin[idx] : 0u;
    }
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int it = 0; it < OS_ITEMS; ++it) {
        const bool ok = wbase + it * 64 < n;
        const uint32_t d = (key[it] >> shift) & (OS_RADIX - 1);
        uint64_t match = __ballot(ok);
#pragma unroll
        for (int b = 0; b < OS_BITS; ++b) {
            const uint64_t m = __ballot(ok && ((d >> b) & 1u));
            match &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t r = (uint32_t)__popcll(match & below);
        uint32_t c = 0;
        if (ok) c = s_cnt[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (ok && r == 0) s_cnt[wave][d] = c + (uint32_t)__popcll(match);
        __builtin_amdgcn_wave_barrier();
        rank[it] = c + r;
    }
    __syncthreads();
    // thread d: the tile's count of digit d, the waves' offsets, publication,
    // look-back
    const int d = threadIdx.x;
    const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
    const uint32_t tot = c0 + c1 + c2 + c3;
    uint32_t* st = state + (size_t)tile * OS_RADIX + d;
    state_store(st, (tile == 0 ? ST_INC : ST_AGG) | tot);
    uint32_t lstart, gstart;
    block_excl_scan2(tot, digit_tot[d], &lstart, &gstart);
    uint32_t excl = 0;
    if (tile > 0) {
        const uint32_t* p = st - OS_RADIX;
        for (;;) {
            const uint32_t s = state_load(p);
            if ((s & ~ST_VAL) == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += s & ST_VAL;
            if (s & ST_INC) break;
            p -= OS_RADIX;
        }
        state_store(st, ST_INC | (excl + tot));
    }
    s_cnt[0][d] = 0u;
    s_cnt[1][d] = c0;
    s_cnt[2][d] = c0 + c1;
    s_cnt[3][d] = c0 + c1 + c2;
    s_lstart[d] = lstart;
    s_gbase[d] = gstart + excl - lstart;  // global position = s_gbase[d] + tile position (mod 2^32)
    __syncthreads();
#pragma unroll
    for (int it = 0; it < OS_ITEMS; ++it) {
        if (wbase + it * 64 < n) {
            const uint32_t dd = (key[it] >> shift) & (OS_RADIX - 1);
            const uint32_t pos = s_lstart[dd] + s_cnt[wave][dd] + rank[it];
            s_k[pos] = key[it];
            s_v[pos] = val[it];
        }
    }
    __syncthreads();
    const int tn = (int)((long long)n - t0 < OS_TILE ? (long long)n - t0 : OS_TILE);
    for (int i = threadIdx.x; i < tn; i += OS_THREADS) {
        const uint32_t k = s_k[i];
        const uint32_t dst = s_gbase[(k >> shift) & (OS_RADIX - 1)] + (uint32_t)i;
        kout[dst] = k;
        vout[dst] = s_v[i];
    }
}

// Sorts n pairs (or the first *n_dev of them, *n_dev <= n, when n_dev is
// given) by the low `bits` (<= 32) bits of the key, stably.  tmp_k/tmp_v have
// n entries; work has radix_hist_entries(n) words.  The result lands in
// keys/vals after an even number of 8-bit passes and in tmp_k/tmp_v after an
// odd number; the returned pair names where.
SortResult radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* tmp_k, uint32_t* tmp_v, uint32_t* work, int n,
                            int bits, hipStream_t st, const int* n_dev) {
    SortResult res{keys, vals};
    if (n <= 1) return res;
    const int passes = (bits + OS_BITS - 1) / OS_BITS;
    const size_t tiles = os_tiles(n), stride = tiles * OS_RADIX;
    const int gprep = (int)std::min<size_t>((OS_STATUS + passes * stride + 255) / 256, 2048);
    hipLaunchKernelGGL(k_sort_prep, dim3(gprep), dim3(256), 0, st, work, n, n_dev, passes, stride);
    const int ghist = (int)std::min<size_t>(tiles, 1024);
    hipLaunchKernelGGL(k_sort_hist, dim3(ghist), dim3(256), 0, st, keys, n, n_dev, passes, work);
    uint32_t *ka = keys, *va = vals, *kb = tmp_k, *vb = tmp_v;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_sort_pass, dim3((unsigned)tiles), dim3(OS_THREADS), 0, st, ka, va, kb, vb, n, n_dev,
                           p * OS_BITS, work + OS_TOT + (size_t)p * OS_RADIX, work + OS_TICKET + p,
                           work + OS_STATUS + (size_t)p * stride);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    res.keys = ka;
    res.vals = va;
    return res;
}
SortResult radix_sort_result(uint32_t* keys, uint32_t* vals, uint32_t* tmp_k, uint32_t* tmp_v, int n, int bits) {
    const int passes = (bits + OS_BITS - 1) / OS_BITS;
    if (n <= 1 || passes % 2 == 0) return SortResult{keys, vals};
    return SortResult{tmp_k, tmp_v};
}
size_t radix_hist_entries(int n) { return OS_STATUS + (size_t)OS_MAX_PASSES * os_tiles(n < 1 ? 1 : n) * OS_RADIX; }

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
