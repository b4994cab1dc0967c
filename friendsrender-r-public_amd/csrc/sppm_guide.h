// sppm_guide.h -- learned photon-emission proposal (no counterpart in the
// reference; it replaces the position x direction sampling of
// generateLightRay, SPPM/SPPM/SPPM.cpp:453-471, when guiding is on).
//
// A 64 -> 64x4 (ReLU) -> 1 MLP maps the Fourier features of a cell centre of
// the 32x32 (light u, v) x 16x16 (hemi-octahedral direction s, t) grid to a
// logit; the proposal is the softmax over the 262,144 cells, uniform inside a
// cell.  Training minimises -sum_c W_c log P_c over the histogram W of what
// the photons emitted from each cell deposited on visible hit points (hit
// points gathering a photon x its luminance / PDF), decayed over passes.  Forward and backward run on bf16 MFMA (16x16x32) with float32
// accumulation; Adam keeps float32 master weights.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sppm {

constexpr int GUIDE_GU = 32, GUIDE_GV = 32, GUIDE_GS = 16, GUIDE_GT = 16;
constexpr int GUIDE_CELLS = GUIDE_GU * GUIDE_GV * GUIDE_GS * GUIDE_GT;
constexpr int GUIDE_FEAT = 64, GUIDE_HID = 64, GUIDE_LAYERS = 4;
// parameter layout (float32): W1 [in][out], b1, W2, b2, W3, b3, W4, b4, w5, b5
constexpr int GUIDE_WSZ = GUIDE_HID * GUIDE_HID;
constexpr int guide_w_off(int l) { return (l - 1) * (GUIDE_WSZ + GUIDE_HID); }   // l = 1..4
constexpr int guide_b_off(int l) { return guide_w_off(l) + GUIDE_WSZ; }
constexpr int GUIDE_W5 = 4 * (GUIDE_WSZ + GUIDE_HID);
constexpr int GUIDE_B5 = GUIDE_W5 + GUIDE_HID;
constexpr int GUIDE_NPARAMS = GUIDE_B5 + 1;  // 16,705
constexpr int GUIDE_BLOCKS = 256;            // training / inference grid (one wave per block)

// Emission from the mixture of the plain emission (probability w) and the
// proposal, for a scene whose only light is one untextured rectangle split
// into two triangles: position P0 + u e1 + v e2, direction in the frame
// (t1, t2, n) through the hemi-octahedral map.
struct GuideEmit {
    float P0[3], e1[3], e2[3], n[3], t1[3], t2[3];
    float inv_e1, inv_e2;  // 1 / |e1|^2, 1 / |e2|^2
    float inv_area;        // plain emission's area density (uniform on the light)
    float rad[3];          // emitted radiance (LightSampler flux / (2 pi area))
    float w;               // probability of the plain emission
    const float* prob;     // proposal over cells (unnormalised: total = cdf[last])
    const float* cdf;
    int* path_cell;        // emission cell of each photon path (training record)
};

struct GuideBuffers {
    float* theta = nullptr;     // master weights
    float* m = nullptr;         // Adam moments
    float* v = nullptr;
    float* grad = nullptr;      // last gradient (+ loss at index NPARAMS)
    uint16_t* wfwd = nullptr;   // bf16 W_l in [out][in] layout, 4 layers
    uint16_t* wbwd = nullptr;   // bf16 W_l in [in][out] layout, 4 layers
    float* logits = nullptr;    // per cell
    float* prob = nullptr;      // softmax
    float* cdf = nullptr;       // inclusive prefix sum of prob
    float* hist = nullptr;      // training histogram (decayed sum of the passes' deposits)
    float* hpass = nullptr;     // this pass's deposit histogram
    float* part = nullptr;      // per-block partials: GUIDE_BLOCKS * (GUIDE_NPARAMS + 1)
    float* part2 = nullptr;     // per-block (max, sumexp), scan block sums, histogram partials
    float* scal = nullptr;      // [0] lse, [1] sum of hist, [2] loss, [3] Adam steps taken
};

int guide_alloc(GuideBuffers& g);
void guide_free(GuideBuffers& g);
// theta (host, GUIDE_NPARAMS floats) -> device, Adam state reset, bf16 copies
int guide_upload_params(GuideBuffers& g, const float* theta, hipStream_t st);
// Forward over all cells: logits, log-sum-exp, probabilities and their CDF.
void guide_infer(GuideBuffers& g, hipStream_t st);
// Sum of the training histogram (scal[1]).
void guide_hist_sum(GuideBuffers& g, hipStream_t st);
// Gradient of the loss for the histogram in g.hist (needs the log-sum-exp of
// the same weights from guide_infer and the histogram sum); writes g.grad and
// the loss.
void guide_gradient(GuideBuffers& g, hipStream_t st);
// Adam step with g.grad (skipped on the device when the histogram is empty).
void guide_adam(GuideBuffers& g, float lr, hipStream_t st);
void guide_clear_hist(GuideBuffers& g, hipStream_t st);
// hist = decay * hist + hpass; hpass = 0
void guide_fold_hist(GuideBuffers& g, float decay, hipStream_t st);

}  // namespace sppm
