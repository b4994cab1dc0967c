// sppm_stripes.h -- work counters that many waves claim from at once.
//
// Atomics to one address from many waves serialise (measured on MI355X:
// about 11.5 ns per atomic on one address, 1.5 ns over 8 addresses, 0.3 ns
// over 64), so a single queue counter shared by every wave of a persistent
// kernel costs tens to hundreds of microseconds per launch.  Here the work
// is cut into STRIPES stripes with a counter each, 256 bytes apart.  A wave
// starts on stripe (global wave index mod STRIPES), claims items of that
// stripe, and moves on to the next stripe when it is exhausted; after
// STRIPES exhausted stripes it has no work left.  Claims never overlap and
// together cover every item exactly once, whichever wave makes them.
#pragma once
#include <hip/hip_runtime.h>

namespace sppm {

constexpr int STRIPES = 64;
constexpr int STRIPE_PAD = 64;                        // ints from one counter to the next (256 bytes)
constexpr int STRIPE_INTS = STRIPES * STRIPE_PAD;     // ints of one set of counters (zeroed before a launch)
constexpr int RAY_STRIPE = 32;                        // 64-bit words of one line of striped 64-bit counters

struct StripeCursor {
    int s;      // current stripe
    int tried;  // stripes found exhausted
};
__device__ __forceinline__ StripeCursor stripe_start() {
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    return StripeCursor{wave % STRIPES, 0};
}

// Items [0, n): stripe s holds [s * per, min((s + 1) * per, n)), per =
// ceil(n / STRIPES).  Claims up to `want` consecutive items; returns how many
// (0: no work left) and the first in *first.  Called by the whole wave
// (wave-uniform control flow); lane 0 makes the atomics.
__device__ __forceinline__ int stripe_claim(int* __restrict__ ctr, int n, StripeCursor& c, int want, int* first) {
    const int per = (n + STRIPES - 1) / STRIPES;
    while (c.tried < STRIPES) {
        const int s0 = c.s * per;
        const int len = n - s0 < per ? n - s0 : per;
        if (len > 0) {
            int old = 0;
            if ((threadIdx.x & 63) == 0) old = atomicAdd(&ctr[c.s * STRIPE_PAD], want);
            old = __shfl(old, 0, 64);
            if (old < len) {
                *first = s0 + old;
                return len - old < want ? len - old : want;
            }
        }
        c.s = c.s + 1 == STRIPES ? 0 : c.s + 1;
        ++c.tried;
    }
    return 0;
}

// Items [0, n) dealt round-robin: item k * STRIPES + s belongs to stripe s,
// claimed one at a time.  Waves on different stripes then claim neighbouring
// items at about the same time, so the claims still follow the items' order
// (work on items that share data stays together in the caches).  Returns
// false when no work is left.
__device__ __forceinline__ bool stripe_claim_rr(int* __restrict__ ctr, int n, StripeCursor& c, int* item) {
    while (c.tried < STRIPES) {
        int old = 0;
        if ((threadIdx.x & 63) == 0) old = atomicAdd(&ctr[c.s * STRIPE_PAD], 1);
        old = __shfl(old, 0, 64);
        const long long it = (long long)old * STRIPES + c.s;
        if (it < n) {
            *item = (int)it;
            return true;
        }
        c.s = c.s + 1 == STRIPES ? 0 : c.s + 1;
        ++c.tried;
    }
    return false;
}

// A queue kept as STRIPES partitions of capacity P: partition s holds
// entries [s * P, s * P + cnt[s * STRIPE_PAD]).  Claims up to `want`
// consecutive entries through the partitions' claim counters fct.
__device__ __forceinline__ int part_claim(const int* __restrict__ cnt, int* __restrict__ fct, int P, StripeCursor& c,
                                          int want, int* first) {
    while (c.tried < STRIPES) {
        const int len = cnt[c.s * STRIPE_PAD];
        if (len > 0) {
            int old = 0;
            if ((threadIdx.x & 63) == 0) old = atomicAdd(&fct[c.s * STRIPE_PAD], want);
            old = __shfl(old, 0, 64);
            if (old < len) {
                *first = c.s * P + old;
                return len - old < want ? len - old : want;
            }
        }
        c.s = c.s + 1 == STRIPES ? 0 : c.s + 1;
        ++c.tried;
    }
    return 0;
}

}  // namespace sppm
