// sppm_guide.h -- learned photon-emission proposal (no counterpart in the
// reference; it replaces the position x direction sampling of
// generateLightRay, SPPM/SPPM/SPPM.cpp:453-471, when guiding is on).
//
// An MLP (64 Fourier features -> `layers` hidden ReLU layers of `hidden`
// units -> 1) maps the cell centre of the 32x32 (light u, v) x 16x16
// (hemi-octahedral direction s, t) grid to a logit; the proposal is the
// softmax over the 262,144 cells, uniform inside a cell.  Training minimises
// -sum_c W_c log P_c over the histogram W of what the photons emitted from
// each cell deposited on visible hit points (hit points gathering a photon x
// its luminance / PDF), decayed over passes.  Forward and backward run on
// MFMA 16x16x32 in bf16 or fp16 with float32 accumulation; Adam keeps float32
// master weights.  Networks: hidden 64 or 128, layers 4 or 6.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sppm {

constexpr int GUIDE_GU = 32, GUIDE_GV = 32, GUIDE_GS = 16, GUIDE_GT = 16;
constexpr int GUIDE_CELLS = GUIDE_GU * GUIDE_GV * GUIDE_GS * GUIDE_GT;
constexpr int GUIDE_FEAT = 64;  // 4 coordinates x 8 frequencies x (sin, cos)

// Network shape.  Parameters (float32, in this order): for l = 1..layers
// W_l [in_l][hidden] (in_1 = GUIDE_FEAT, else hidden) and b_l [hidden]; then
// the output weights w_out [hidden] and bias b_out.
struct GuideNet {
    int hidden = 64, layers = 4, fp16 = 0;
};
constexpr int guide_in(const GuideNet& n, int l) { return l == 1 ? GUIDE_FEAT : n.hidden; }
constexpr int guide_w_off(const GuideNet& n, int l) {
    return l == 1 ? 0 : GUIDE_FEAT * n.hidden + n.hidden + (l - 2) * (n.hidden * n.hidden + n.hidden);
}
constexpr int guide_b_off(const GuideNet& n, int l) { return guide_w_off(n, l) + guide_in(n, l) * n.hidden; }
constexpr int guide_wout_off(const GuideNet& n) { return guide_b_off(n, n.layers) + n.hidden; }
constexpr int guide_nparams(const GuideNet& n) { return guide_wout_off(n) + n.hidden + 1; }
// low-precision weight copies: layer l starts at sum over l' < l of in_l' * hidden
constexpr int guide_lw_off(const GuideNet& n, int l) { return l == 1 ? 0 : GUIDE_FEAT * n.hidden + (l - 2) * n.hidden * n.hidden; }
constexpr int guide_lw_size(const GuideNet& n) { return guide_lw_off(n, n.layers + 1); }
constexpr bool guide_net_ok(const GuideNet& n) {
    return (n.hidden == 64 || n.hidden == 128) && (n.layers == 4 || n.layers == 6) && (n.fp16 == 0 || n.fp16 == 1);
}
// gradients of the hidden layers are carried scaled by this power of two (no
// fp16 subnormals for the ~1e-6 per-cell gradients; exact for bf16)
constexpr float GUIDE_GRAD_SCALE = 1024.0f;
constexpr int GUIDE_BLOCKS = 256;            // workgroups of the per-layer gradient kernels

// Emission from the mixture of the plain emission (probability w) and the
// proposal, for a scene whose only light is one rectangle split into two
// triangles: position P0 + u e1 + v e2, direction in the frame (t1, t2, n)
// through the hemi-octahedral map.  An untextured light emits from any point
// (the proposal is uniform inside a cell); a textured one emits from its
// 1000-point tables (LightSampler.cpp:30-66), so the proposal picks a table
// point inside the chosen position cell with the plain emission's own
// probabilities, and position cells without points get no proposal mass.
struct GuideEmit {
    float P0[3], e1[3], e2[3], n[3], t1[3], t2[3];
    float inv_e1, inv_e2;  // 1 / |e1|^2, 1 / |e2|^2
    float inv_area;        // plain emission's area density (uniform on the light)
    float rad[3];          // emitted radiance (LightSampler flux / (2 pi area))
    float w;               // probability of the plain emission
    const float* prob;     // proposal over cells (unnormalised: total = cdf[last])
    const float* cdf;
    int* path_cell;        // emission cell of each photon path (training record), 2 x N:
                           // the pass of step s writes block s & 1
    float* path_ratio;     // q / p_plain at each path's emission, laid out like path_cell
                           // (NULL: not recorded)
    // textured light only
    int textured;
    const int* pt_cell;        // position cell of each table point (indexed like the light tables)
    const float* pcell_mass;   // plain probability of the table points of each position cell
    const int* pc_start;       // points of position cell c: entries [pc_start[c], pc_start[c + 1])
    const int2* pc_pt;         // (light, table entry) of each point, in cell order
    const float* pc_cum;       // inclusive cumulative plain probability within the cell
};

// The guided training signal is summed in 64-bit fixed point (integer
// atomics), so its sums, and with them the trained proposal, do not depend on
// the order in which waves add: 2^24 per unit, each addition clamped to 1e9.
constexpr float GUIDE_FIX_SCALE = 16777216.0f;
constexpr float GUIDE_FIX_INV = 5.9604644775390625e-08f;
constexpr float GUIDE_FIX_MAX = 1.0e9f;
__host__ __device__ inline unsigned long long guide_fix(float v) {
    return (unsigned long long)((v < GUIDE_FIX_MAX ? v : GUIDE_FIX_MAX) * GUIDE_FIX_SCALE);
}

struct GuideBuffers {
    GuideNet net;
    int nparams = 0;
    float* pcell = nullptr;     // plain emission's probability of each cell
    float* tgt = nullptr;       // training target per cell (SPPM_GUIDE_TARGET_RMS)
    float* theta = nullptr;     // master weights
    float* m = nullptr;         // Adam moments
    float* v = nullptr;
    float* grad = nullptr;      // last gradient
    uint16_t* wfwd = nullptr;   // low-precision W_l in [hidden][in_l] layout
    uint16_t* wbwd = nullptr;   // low-precision W_l in [in_l][hidden] layout
    uint16_t* acts = nullptr;   // training forward: layers x cells x hidden activations
    uint16_t* dh = nullptr;     // backward: 2 x cells x hidden (scaled) activation gradients
    float* logits = nullptr;    // per cell
    float* prob = nullptr;      // softmax
    float* cdf = nullptr;       // inclusive prefix sum of prob
    // the tables the photon pass samples: prob / cdf as they were when the
    // build after the previous photon pass published them (guide_publish)
    float* eprob = nullptr;
    float* ecdf = nullptr;
    float* hist = nullptr;      // training histogram (decayed sum of the passes' deposits)
    unsigned long long* hpass = nullptr;  // deposits since the last fold (fixed point, guide_fix)
    float* part = nullptr;      // per-block gradient partials: GUIDE_BLOCKS * (nparams + 1)
    float* part2 = nullptr;     // per-block (max, sumexp), scan block sums, histogram partials
    float* scal = nullptr;      // [0] lse, [1] sum of hist, [2] loss, [3] Adam steps taken
    // textured light: table points by position cell (see GuideEmit)
    int* pt_cell = nullptr;
    float* pcell_mass = nullptr;  // also masks the proposal: cells of zero mass get probability 0
    int* pc_start = nullptr;
    int2* pc_pt = nullptr;
    float* pc_cum = nullptr;
};

// Allocates for the network g.net (set it first).
int guide_alloc(GuideBuffers& g);
void guide_free(GuideBuffers& g);
// theta (host, g.nparams floats) -> device, Adam state reset, low-precision copies
int guide_upload_params(GuideBuffers& g, const float* theta, hipStream_t st);
// Forward over all cells: logits, log-sum-exp, probabilities and their CDF.
void guide_infer(GuideBuffers& g, hipStream_t st);
// Sum of the training weights w (one per cell: g.hist, or g.tgt) into scal[1].
void guide_hist_sum(GuideBuffers& g, const float* w, hipStream_t st);
// Gradient of the loss for the weights w (needs the log-sum-exp of the same
// weights from guide_infer and the weights' sum); writes g.grad and the loss.
void guide_gradient(GuideBuffers& g, const float* w, hipStream_t st);
// g.tgt = sqrt(g.pcell * g.hist) (SPPM_GUIDE_TARGET_RMS)
void guide_rms_target(GuideBuffers& g, hipStream_t st);
// Adam step with g.grad (skipped on the device when the histogram is empty).
void guide_adam(GuideBuffers& g, float lr, hipStream_t st);
void guide_clear_hist(GuideBuffers& g, hipStream_t st);
// After theta, Adam state, histogram and scalars were copied in: the
// low-precision weight copies and the proposal.
void guide_restore(GuideBuffers& g, hipStream_t st);
// hist = decay * hist + hpass; hpass = 0
void guide_fold_hist(GuideBuffers& g, float decay, hipStream_t st);
// eprob / ecdf = prob / cdf (a copy kernel: the emission reads the copy while
// the next proposal is trained)
void guide_publish(GuideBuffers& g, hipStream_t st);

}  // namespace sppm
