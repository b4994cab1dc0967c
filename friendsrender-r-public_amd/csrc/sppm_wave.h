// sppm_wave.h -- buffers and launcher of the wavefront photon pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sppm_pass.h"

namespace sppm {

// State of every photon path between the wavefront kernels, one entry per
// path index (N entries each, 96 bytes per path).
struct PathStateDev {
    uint4* rng;     // MT cursor: a0, a1, b, k
    float4* org;    // cur_org, cur_index
    float4* dir;    // cur_dir, next_index
    float4* ndir;   // next_dir, PDF
    float4* rad;    // rad, pidx | loopcount << 8 | depth << 16 (as float bits)
    float4* thr;    // T
};
struct WaveBuffers {
    PathStateDev path;
    float4* rays[2];   // ray queues: (org, path index bits), (dir, 0) per entry, wave_queue_entries(N) each
    float4* hits;      // closest hit per queue entry: (t, u, v, slot | tie << 30, or -1 for a miss)
    int* qctl;         // per queue: partition sizes, then per queue the partitions' claim counters
                       // (wave_ctl_ints(bounces))
    uint32_t* spill;   // traversal stack entries past the LDS part (wave_spill_entries)
    int cap;           // traversal stack capacity per ray
};

constexpr int WAVE_BOUNCES_DEFAULT = 2;
constexpr int WAVE_LANES_DEFAULT = 4;
// Ray queues: STRIPES partitions of wave_partition(N) entries.
int wave_partition(int N);
size_t wave_queue_entries(int N);
size_t wave_ctl_ints(int bounces);
// Lanes per ray of the trace kernel: 2, 4 or 8, at most the BVH width.
int wave_lanes_ok(int width, int lanes);
// Blocks of the persistent trace kernel that fill the device, and the spill
// entries it needs for that grid.
int wave_trace_blocks(int width, int lanes);
size_t wave_spill_entries(int lanes, int blocks, int cap);

// The photon pass as wavefront kernels: emission, `bounces` rounds of closest
// hit (`lanes` lanes per ray) + shading over ray queues, then the one-lane-per-path loop for the
// paths still alive (tail: its grid and stack spill area).  step / step_dev /
// guide / flags / rays as launch_photon.
void launch_photon_wave(const SceneDev& S, const SlotsDev& slots, int N, int step, const int* step_dev,
                        const WaveBuffers& wb, int bounces, int lanes, int trace_blocks, const TraceSpill& tail,
                        const GuideEmit* guide, uint32_t* flags, unsigned long long* rays, hipStream_t st);

}  // namespace sppm
