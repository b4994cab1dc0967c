// sppm_kernels.h -- host-side launchers shared between the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_layout.h"

namespace sppm {

// ---- sort / scan (sppm_sort.hip) ------------------------------------------
// Stable sort by the low `bits` bits; n_dev (optional): sort only the first
// *n_dev <= n pairs.  n < 2^30.  The result is in keys/vals or in tmp_k/tmp_v
// (an odd number of 8-bit digit passes): SortResult names which.
struct SortResult {
    uint32_t* keys;
    uint32_t* vals;
};
// The buffers radix_sort_pairs leaves its output in for these arguments (no launch).
SortResult radix_sort_result(uint32_t* keys, uint32_t* vals, uint32_t* tmp_k, uint32_t* tmp_v, int n, int bits);
SortResult radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* tmp_k, uint32_t* tmp_v, uint32_t* work, int n,
                            int bits, hipStream_t st, const int* n_dev = nullptr);
constexpr int RADIX_SORT_MAX_N = (1 << 30) - 1;
size_t radix_hist_entries(int n);  // words of the sort's work area
__global__ void k_tile_total_prep(int* tmp, int nt);
__global__ void k_scan_small(uint32_t* data, int n);
void exclusive_scan(const int* in, int* out, int n, int* total, int* tmp, hipStream_t st);
size_t scan_tmp_entries(int n);

// ---- BVH build (sppm_bvh.hip) ---------------------------------------------
struct BvhResult {
    float* nodes = nullptr;  // device, nnodes * node_floats(width)
    float4* tri = nullptr;   // device, 3 * ntri
    int nnodes = 0;
    int depth = 0;           // wide levels
    int width = 0;
};
// pos: device array of 9 floats per triangle. Returns 0 on success.
// sah: binned-SAH top-down build; otherwise a radix tree over Morton codes.
int build_bvh(const float* pos, int ntri, int width, bool sah, hipStream_t st, BvhResult* out, char* err,
              size_t errlen);
void free_bvh(BvhResult* r);

// ---- pass kernels (sppm_pass.hip) -----------------------------------------
struct PassBuffers;  // defined in sppm_pass.hip

}  // namespace sppm
