// sppm_guide.hip -- training and inference kernels of the learned emission
// proposal (MFMA 16x16x32 in bf16 or fp16, float32 accumulation).
//
// For v_mfma_f32_16x16x32_{bf16,f16} lane l supplies A[l%16][8*(l/16) .. +7]
// and B[8*(l/16) .. +7][l%16] and receives D[4*(l/16) + i][l%16], i = 0..3.
//
// Inference (and the forward of a training step) is one kernel: a wave takes
// 64 cells, runs every layer with the activations in LDS, and writes the
// logits (a training forward also writes every layer's activations).  The
// backward runs one kernel per layer over 64-cell tiles with 4 waves: the
// weight gradient H_{l-1}^T dH_l stays in the waves' accumulators across the
// block's tiles (split over the cells, summed over blocks afterwards), and
// dH_{l-1} = (dH_l W_l^T) * (H_{l-1} > 0) goes back to memory for the next
// layer.  LDS holds each tile in the two layouts the products read: [cell][k]
// and [k][cell].
#include "sppm_guide.h"

#include <math.h>

#include <vector>

namespace sppm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <bool F16>
struct Lowp;
template <>
struct Lowp<false> {  // bfloat16
    static __device__ __forceinline__ uint16_t from(float f) {  // round to nearest even
        uint32_t u = __float_as_uint(f);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
    static __device__ __forceinline__ float to(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
    static __device__ __forceinline__ f32x4 mma(const uint16_t* a, const uint16_t* b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                       *reinterpret_cast<const bf16x8*>(b), c, 0, 0, 0);
    }
};
template <>
struct Lowp<true> {  // IEEE half
    static __device__ __forceinline__ uint16_t from(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
    static __device__ __forceinline__ float to(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
    static __device__ __forceinline__ f32x4 mma(const uint16_t* a, const uint16_t* b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(*reinterpret_cast<const f16x8*>(a),
                                                      *reinterpret_cast<const f16x8*>(b), c, 0, 0, 0);
    }
};
// > 0 in either format: sign bit clear and not zero
__device__ __forceinline__ bool lowp_pos(uint16_t h) { return h != 0 && !(h & 0x8000u); }

// ---- input features of every cell (constant; computed once per device) ----
// coordinate j of cell c as the oracle defines it, then sin / cos at
// pi * 2^(k/2) in double, rounded to float, then to the network's format.
template <bool F16>
__global__ void k_guide_features(uint16_t* __restrict__ X, const float* __restrict__ freq) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= GUIDE_CELLS) return;
    const int it = c % GUIDE_GT, is = (c / GUIDE_GT) % GUIDE_GS, iv = (c / (GUIDE_GS * GUIDE_GT)) % GUIDE_GV,
              iu = c / (GUIDE_GV * GUIDE_GS * GUIDE_GT);
    float co[4];
    co[0] = 2.0f * (((float)iu + 0.5f) / (float)GUIDE_GU) - 1.0f;
    co[1] = 2.0f * (((float)iv + 0.5f) / (float)GUIDE_GV) - 1.0f;
    co[2] = -1.0f + ((float)is + 0.5f) * (2.0f / (float)GUIDE_GS);
    co[3] = -1.0f + ((float)it + 0.5f) * (2.0f / (float)GUIDE_GT);
    uint16_t* out = X + (size_t)c * GUIDE_FEAT;
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 8; ++k) {
            const double x = (double)co[j] * (double)freq[k];
            out[16 * j + 2 * k] = Lowp<F16>::from((float)sin(x));
            out[16 * j + 2 * k + 1] = Lowp<F16>::from((float)cos(x));
        }
}

// Copies `rows` x `cols` low-precision values between global memory (row
// stride cols) and LDS (row stride lds_stride), 16 bytes per lane-step.
template <int ROWS, int COLS, int LS, int NT>
__device__ __forceinline__ void tile_to_lds(uint16_t* __restrict__ lds, const uint16_t* __restrict__ g, int tid) {
    constexpr int CH = COLS / 8;
    static_assert((ROWS * CH) % NT == 0, "whole steps");
#pragma unroll
    for (int s = 0; s < ROWS * CH / NT; ++s) {
        const int c = tid + s * NT, r = c / CH, k = (c % CH) * 8;
        *reinterpret_cast<uint4*>(lds + r * LS + k) = *reinterpret_cast<const uint4*>(g + (size_t)r * COLS + k);
    }
}
template <int ROWS, int COLS, int LS, int NT>
__device__ __forceinline__ void lds_to_tile(uint16_t* __restrict__ g, const uint16_t* __restrict__ lds, int tid) {
    constexpr int CH = COLS / 8;
    static_assert((ROWS * CH) % NT == 0, "whole steps");
#pragma unroll
    for (int s = 0; s < ROWS * CH / NT; ++s) {
        const int c = tid + s * NT, r = c / CH, k = (c % CH) * 8;
        *reinterpret_cast<uint4*>(g + (size_t)r * COLS + k) = *reinterpret_cast<const uint4*>(lds + r * LS + k);
    }
}

// ---- forward: all layers of 64 cells per wave ---------------------------------
// STORE: write every layer's activations (training forward).
constexpr int FT = 64;

// One layer for the wave's 64 cells: out = relu(in * W^T + b) (low precision,
// LDS); the last layer instead adds h * w_out into lp (float32, unrounded h).
template <int H, int K, bool F16, bool LASTL, bool STORE>
__device__ __forceinline__ void fwd_layer(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                          const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                          const float* __restrict__ wout, float (&lp)[FT / 16][4], int lane) {
    constexpr int S = H + 8;
    using LP = Lowp<F16>;
    const int r16 = lane & 15, q = lane >> 4;
#pragma unroll 1
    for (int nt = 0; nt < H / 16; ++nt) {
        f32x4 acc[FT / 16];
#pragma unroll
        for (int mt = 0; mt < FT / 16; ++mt) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < K / 32; ++kb) {
            const uint4 bv = *reinterpret_cast<const uint4*>(W + (nt * 16 + r16) * K + kb * 32 + 8 * q);
#pragma unroll
            for (int mt = 0; mt < FT / 16; ++mt)
                acc[mt] = LP::mma(in + (mt * 16 + r16) * S + kb * 32 + 8 * q, reinterpret_cast<const uint16_t*>(&bv),
                                  acc[mt]);
        }
        const float bb = bias[nt * 16 + r16];
        const float wo = LASTL ? wout[nt * 16 + r16] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < FT / 16; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float h = fmaxf(acc[mt][i] + bb, 0.0f);
                if (!LASTL || STORE) out[(mt * 16 + 4 * q + i) * S + nt * 16 + r16] = LP::from(h);
                if (LASTL) lp[mt][i] += h * wo;
            }
    }
}

template <int H, int L, bool F16, bool STORE>
__global__ void __launch_bounds__(64) k_mlp_forward(const uint16_t* __restrict__ X, const uint16_t* __restrict__ wfwd,
                                                    const float* __restrict__ theta, uint16_t* __restrict__ acts,
                                                    float* __restrict__ logits, float* __restrict__ part2) {
    constexpr GuideNet NET{H, L, F16};
    constexpr int S = H + 8;  // LDS row stride (elements), padded against bank conflicts
    __shared__ __align__(16) uint16_t buf[2][FT * S];
    __shared__ float slog[FT];
    const int lane = threadIdx.x, r16 = lane & 15, q = lane >> 4;
    const float* wout = theta + guide_wout_off(NET);
    float m = -INFINITY, sum = 0.0f;
    for (int t = blockIdx.x; t < GUIDE_CELLS / FT; t += gridDim.x) {
        const int base = t * FT;
        tile_to_lds<FT, GUIDE_FEAT, S, 64>(buf[0], X + (size_t)base * GUIDE_FEAT, lane);
        __syncthreads();
        float lp[FT / 16][4];
#pragma unroll
        for (int mt = 0; mt < FT / 16; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) lp[mt][i] = 0.0f;
        // layer 1 (64 features in)
        fwd_layer<H, GUIDE_FEAT, F16, false, STORE>(buf[0], buf[1], wfwd, theta + guide_b_off(NET, 1), wout, lp, lane);
        __syncthreads();
        if (STORE) lds_to_tile<FT, H, S, 64>(acts + (size_t)base * H, buf[1], lane);
        int cur = 1;
#pragma unroll 1
        for (int l = 2; l <= L; ++l) {
            const uint16_t* W = wfwd + guide_lw_off(NET, l);
            if (l < L)
                fwd_layer<H, H, F16, false, STORE>(buf[cur], buf[cur ^ 1], W, theta + guide_b_off(NET, l), wout, lp,
                                                   lane);
            else
                fwd_layer<H, H, F16, true, STORE>(buf[cur], buf[cur ^ 1], W, theta + guide_b_off(NET, l), wout, lp,
                                                  lane);
            __syncthreads();
            if (STORE) lds_to_tile<FT, H, S, 64>(acts + ((size_t)(l - 1) * GUIDE_CELLS + base) * H, buf[cur ^ 1], lane);
            cur ^= 1;
            __syncthreads();
        }
        // logit = h_L . w_out + b_out: the 16 lanes of a row group hold its 16-column slices
#pragma unroll
        for (int mt = 0; mt < FT / 16; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float p = lp[mt][i];
                p += __shfl_xor(p, 1, 64);
                p += __shfl_xor(p, 2, 64);
                p += __shfl_xor(p, 4, 64);
                p += __shfl_xor(p, 8, 64);
                if (r16 == 0) slog[mt * 16 + 4 * q + i] = p + theta[guide_wout_off(NET) + H];
            }
        __syncthreads();
        const float lg = slog[lane];
        logits[base + lane] = lg;
        if (lg > m) {
            sum = sum * expf(m - lg) + 1.0f;
            m = lg;
        } else {
            sum += expf(lg - m);
        }
        __syncthreads();
    }
    // combine the lanes' running (max, sum)
    for (int off = 32; off > 0; off >>= 1) {
        const float m2 = __shfl_xor(m, off, 64), s2 = __shfl_xor(sum, off, 64);
        const float mm = fmaxf(m, m2);
        sum = (m == -INFINITY ? 0.0f : sum * expf(m - mm)) + (m2 == -INFINITY ? 0.0f : s2 * expf(m2 - mm));
        m = mm;
    }
    if (lane == 0) {
        part2[2 * blockIdx.x] = m;
        part2[2 * blockIdx.x + 1] = sum;
    }
}

// sum of the histogram: per-block partials (256 blocks x 1024 cells) ...
__global__ void __launch_bounds__(256) k_guide_hsum(const float* __restrict__ hist, double* __restrict__ hpart) {
    __shared__ double red[256];
    const int t = threadIdx.x, base = blockIdx.x * 1024 + 4 * t;
    red[t] = (double)hist[base] + (double)hist[base + 1] + (double)hist[base + 2] + (double)hist[base + 3];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) hpart[blockIdx.x] = red[0];
}
// ... and their total (fixed order) into scal[1]
__global__ void __launch_bounds__(256) k_guide_hsum_total(const double* __restrict__ hpart, float* __restrict__ scal) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    red[t] = hpart[t];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) scal[1] = (float)red[0];
}

// log-sum-exp over the block partials of k_guide_forward
__global__ void k_guide_lse(const float* __restrict__ part2, int nparts, float* __restrict__ scal) {
    __shared__ double red[256];
    __shared__ float mx[256];
    const int t = threadIdx.x;
    float m = -INFINITY;
    for (int b = t; b < nparts; b += 256) m = fmaxf(m, part2[2 * b]);
    mx[t] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) mx[t] = fmaxf(mx[t], mx[t + o]);
        __syncthreads();
    }
    const float M = mx[0];
    double s = 0.0;
    for (int b = t; b < nparts; b += 256) s += (double)part2[2 * b + 1] * exp((double)part2[2 * b] - (double)M);
    red[t] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) scal[0] = (float)((double)M + log(red[0]));
}

// probabilities and their inclusive prefix sum (per block of 1024, then offsets)
// mask (textured light): cells whose position cell holds no table point get 0
__global__ void __launch_bounds__(256) k_guide_prob(const float* __restrict__ logits, const float* __restrict__ scal,
                                                    float* __restrict__ prob, float* __restrict__ cdf,
                                                    float* __restrict__ bsum, const float* __restrict__ mask) {
    __shared__ float ws[4];
    const int t = threadIdx.x, base = blockIdx.x * 1024 + 4 * t;
    const float lse = scal[0];
    float v[4], run = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = expf(logits[base + k] - lse);
        if (mask && !(mask[(base + k) / (GUIDE_GS * GUIDE_GT)] > 0.0f)) v[k] = 0.0f;
        prob[base + k] = v[k];
        run += v[k];
        v[k] = run;  // thread-local inclusive
    }
    // inclusive scan of the thread totals within the block
    float x = run;
    const int lane = t & 63, w = t >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const float y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    float woff = 0.0f;
    for (int k = 0; k < w; ++k) woff += ws[k];
    const float excl = woff + x - run;
#pragma unroll
    for (int k = 0; k < 4; ++k) cdf[base + k] = excl + v[k];
    if (t == 255) bsum[blockIdx.x] = excl + run;
}
__global__ void __launch_bounds__(256) k_guide_cdf_offsets(float* __restrict__ cdf, const float* __restrict__ bsum) {
    __shared__ float off;
    if (threadIdx.x == 0) {
        float o = 0.0f;
        for (int b = 0; b < (int)blockIdx.x; ++b) o += bsum[b];
        off = o;
    }
    __syncthreads();
    const int base = blockIdx.x * 1024 + 4 * threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) cdf[base + k] += off;
}

// ---- backward of one layer l ---------------------------------------------------
// Tile: 64 cells, 4 waves.  In: H_{l-1} (the features for l = 1) and dH_l
// (FIRST, l = L: made here from g = softmax - W / sum W and H_L).  Out:
// dW_l (+ db_{l-1}, and for FIRST db_L, w_out, b_out, sum W * logit) as this
// block's partial sums, dH_{l-1} (not for l = 1) to memory.  Gradients are
// carried scaled by GUIDE_GRAD_SCALE.
constexpr int BT = 64;
template <int H, int KIN, bool F16, bool FIRST, bool LAST>
__global__ void __launch_bounds__(256) k_mlp_backward(const uint16_t* __restrict__ hin,
                                                      const uint16_t* __restrict__ dh_in,
                                                      const uint16_t* __restrict__ h_top,
                                                      const float* __restrict__ logits, const float* __restrict__ hist,
                                                      const float* __restrict__ scal, const float* __restrict__ theta,
                                                      const uint16_t* __restrict__ wbwd, uint16_t* __restrict__ dh_out,
                                                      float* __restrict__ part, int nparams, int w_off, int bprev_off,
                                                      int btop_off, int wout_off) {
    constexpr int HS = KIN + 8, TS = BT + 8, DS = H + 8;
    __shared__ __align__(16) uint16_t hT[KIN * TS];             // H_{l-1}^T [k][cell]
    __shared__ __align__(16) uint16_t hR[LAST ? 8 : BT * HS];   // H_{l-1} [cell][k]; then dH_{l-1}
    __shared__ __align__(16) uint16_t dR[LAST ? 8 : BT * DS];   // dH_l [cell][h]
    __shared__ __align__(16) uint16_t dT[H * TS];               // dH_l^T [h][cell]
    __shared__ float sg[BT];
    __shared__ float sred[4][KIN > H ? KIN : H];
    __shared__ float sred2[4][H];
    using LP = Lowp<F16>;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, r16 = lane & 15, q = lane >> 4;
    const float lse = scal[0];
    const float invW = scal[1] > 0.0f ? 1.0f / scal[1] : 0.0f;
    // accumulators: wave w owns input rows [w * KIN/4, (w+1) * KIN/4) of dW_l, all H columns
    constexpr int IT = KIN / 64, OT = H / 16;
    f32x4 acc[IT][OT];
#pragma unroll
    for (int a = 0; a < IT; ++a)
#pragma unroll
        for (int b = 0; b < OT; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // sred[w][k]: wave w's running column sums of dH_{l-1} (db_{l-1})
    for (int k = tid; k < 4 * (KIN > H ? KIN : H); k += 256) (&sred[0][0])[k] = 0.0f;
    __syncthreads();
    // FIRST: per-thread column sums of its 8 columns of dH_L (db_L) and g * H_L (w_out)
    constexpr int CH = H / 8;
    float dbt[8], dwo[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) dbt[k] = dwo[k] = 0.0f;
    float gsum = 0.0f, wl = 0.0f;
    for (int t = blockIdx.x; t < GUIDE_CELLS / BT; t += gridDim.x) {
        const int base = t * BT;
        // H_{l-1}: [cell][k] and [k][cell]
        for (int c = tid; c < BT * (KIN / 8); c += 256) {
            const int r = c / (KIN / 8), k = (c % (KIN / 8)) * 8;
            const uint4 v = *reinterpret_cast<const uint4*>(hin + (size_t)(base + r) * KIN + k);
            if (!LAST) *reinterpret_cast<uint4*>(hR + r * HS + k) = v;
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j) hT[(k + j) * TS + r] = e[j];
        }
        if (FIRST) {
            if (tid < BT) {
                const float lg = logits[base + tid], h = hist[base + tid];
                const float g = (expf(lg - lse) - h * invW) * GUIDE_GRAD_SCALE;
                sg[tid] = g;
                gsum += g;
                wl += h * lg;
            }
            __syncthreads();
        }
        // dH_l: [cell][h] and [h][cell]
        for (int c = tid; c < BT * CH; c += 256) {
            const int r = c / CH, k = (c % CH) * 8;
            uint4 v;
            uint16_t* e = reinterpret_cast<uint16_t*>(&v);
            if (FIRST) {
                const uint4 hv = *reinterpret_cast<const uint4*>(h_top + (size_t)(base + r) * H + k);
                const uint16_t* he = reinterpret_cast<const uint16_t*>(&hv);
                const float g = sg[r];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float hf = LP::to(he[j]);
                    const float d = lowp_pos(he[j]) ? g * theta[wout_off + k + j] : 0.0f;
                    dbt[j] += d;
                    dwo[j] += g * hf;
                    e[j] = LP::from(d);
                }
            } else {
                v = *reinterpret_cast<const uint4*>(dh_in + (size_t)(base + r) * H + k);
            }
            if (!LAST) *reinterpret_cast<uint4*>(dR + r * DS + k) = v;
#pragma unroll
            for (int j = 0; j < 8; ++j) dT[(k + j) * TS + r] = e[j];
        }
        __syncthreads();
        // dW_l[i][o] += sum over the tile's cells of H_{l-1}[c][i] dH_l[c][o]
#pragma unroll
        for (int kb = 0; kb < BT / 32; ++kb)
#pragma unroll
            for (int a = 0; a < IT; ++a) {
                const uint16_t* af = hT + (w * (KIN / 4) + a * 16 + r16) * TS + kb * 32 + 8 * q;
#pragma unroll
                for (int b = 0; b < OT; ++b)
                    acc[a][b] = LP::mma(af, dT + (b * 16 + r16) * TS + kb * 32 + 8 * q, acc[a][b]);
            }
        if (!LAST) {
            // dH_{l-1} for the wave's 16 cells, masked by H_{l-1} > 0, into hR
#pragma unroll 1
            for (int it = 0; it < KIN / 16; ++it) {
                f32x4 c4 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < H / 32; ++kb)
                    c4 = LP::mma(dR + (w * 16 + r16) * DS + kb * 32 + 8 * q,
                                 wbwd + (size_t)(it * 16 + r16) * H + kb * 32 + 8 * q, c4);
                float cs = 0.0f;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    uint16_t* hp = hR + (w * 16 + 4 * q + i) * HS + it * 16 + r16;
                    const float d = lowp_pos(*hp) ? c4[i] : 0.0f;
                    cs += d;
                    *hp = LP::from(d);
                }
                cs += __shfl_xor(cs, 16, 64);
                cs += __shfl_xor(cs, 32, 64);
                if (q == 0) sred[w][it * 16 + r16] += cs;
            }
            __syncthreads();
            lds_to_tile<BT, KIN, HS, 256>(dh_out + (size_t)base * KIN, hR, tid);
        }
        __syncthreads();
    }
    // block partials (the reduction over blocks divides by GUIDE_GRAD_SCALE)
    float* out = part + (size_t)blockIdx.x * (nparams + 1);
#pragma unroll
    for (int a = 0; a < IT; ++a)
#pragma unroll
        for (int b = 0; b < OT; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[w_off + (w * (KIN / 4) + a * 16 + 4 * q + i) * H + b * 16 + r16] = acc[a][b][i];
    if (!LAST) {
        __syncthreads();
        for (int k = tid; k < KIN; k += 256)
            out[bprev_off + k] = sred[0][k] + sred[1][k] + sred[2][k] + sred[3][k];
    }
    if (FIRST) {
        // column sums of the threads sharing a column slice, then the block
        __syncthreads();
        for (int k = tid; k < 4 * H; k += 256) {
            (&sred[0][0])[k] = 0.0f;
            (&sred2[0][0])[k] = 0.0f;
        }
        __syncthreads();
        const int k0 = (tid % CH) * 8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            atomicAdd(&sred[w][k0 + j], dbt[j]);
            atomicAdd(&sred2[w][k0 + j], dwo[j]);
        }
        for (int off = 32; off > 0; off >>= 1) {
            gsum += __shfl_xor(gsum, off, 64);
            wl += __shfl_xor(wl, off, 64);
        }
        __shared__ float s2[2][4];
        if (lane == 0) {
            s2[0][w] = gsum;
            s2[1][w] = wl;
        }
        __syncthreads();
        for (int k = tid; k < H; k += 256) {
            out[btop_off + k] = sred[0][k] + sred[1][k] + sred[2][k] + sred[3][k];
            out[wout_off + k] = sred2[0][k] + sred2[1][k] + sred2[2][k] + sred2[3][k];
        }
        if (tid == 0) {
            out[wout_off + H] = s2[0][0] + s2[0][1] + s2[0][2] + s2[0][3];
            out[nparams] = s2[1][0] + s2[1][1] + s2[1][2] + s2[1][3];
        }
    }
}

// sum of the block partials (fixed order): gradient = sum / GUIDE_GRAD_SCALE;
// loss = lse - sum(W * logit) / sum(W)
__global__ void k_guide_grad_reduce(const float* __restrict__ part, int nblocks, int nparams,
                                    const float* __restrict__ scal, float* __restrict__ grad,
                                    float* __restrict__ scal_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nparams) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += (double)part[(size_t)b * (nparams + 1) + k];
    if (k < nparams) {
        grad[k] = (float)(s / (double)GUIDE_GRAD_SCALE);
    } else {
        const double w = scal[1];
        scal_out[2] = w > 0.0 ? (float)((double)scal[0] - s / w) : 0.0f;
    }
}

// low-precision copies of the hidden-layer weights in both layouts
template <bool F16>
__global__ void k_guide_lowp(const float* __restrict__ theta, uint16_t* __restrict__ wfwd,
                             uint16_t* __restrict__ wbwd, GuideNet net) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= guide_lw_size(net)) return;
    int l = 1;
    while (l < net.layers && e >= guide_lw_off(net, l + 1)) ++l;
    const int K = guide_in(net, l), H = net.hidden;
    const int el = e - guide_lw_off(net, l), in = el / H, o = el % H;
    const uint16_t b = Lowp<F16>::from(theta[guide_w_off(net, l) + el]);
    wbwd[guide_lw_off(net, l) + in * H + o] = b;
    wfwd[guide_lw_off(net, l) + o * K + in] = b;
}

// Adam (float32 master weights); skipped when the histogram is empty.
// Adam step count t + 1 = scal[3] + 1 (kept on the device so graph replays count).
__global__ void k_guide_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ grad, const float* __restrict__ scal, float lr, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !(scal[1] > 0.0f)) return;
    const float t = scal[3] + 1.0f;
    const float c1 = 1.0f / (1.0f - powf(0.9f, t)), c2 = 1.0f / (1.0f - powf(0.999f, t));
    const float g = grad[k];
    const float mk = 0.9f * m[k] + 0.1f * g;
    const float vk = 0.999f * v[k] + 0.001f * g * g;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + 1e-8f);
}

__global__ void k_guide_fold(float* __restrict__ hist, unsigned long long* __restrict__ hpass, float decay) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GUIDE_CELLS) return;
    hist[k] = decay * hist[k] + (float)hpass[k] * GUIDE_FIX_INV;
    hpass[k] = 0ull;
}

__global__ void k_guide_tick(float* __restrict__ scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && scal[1] > 0.0f) scal[3] += 1.0f;
}

__global__ void k_guide_zero(float* __restrict__ p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0f;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static uint16_t* g_features[2] = {nullptr, nullptr};  // per-device constant input (bf16, fp16)
static int g_features_dev[2] = {-1, -1};

template <class T>
static int galloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1)) == hipSuccess ? 0 : -4;
}

static int ensure_features(int dev, int f16) {
    if (g_features_dev[f16] == dev) return 0;
    if (g_features[f16]) (void)hipFree(g_features[f16]);
    g_features[f16] = nullptr;
    if (galloc(&g_features[f16], (size_t)GUIDE_CELLS * GUIDE_FEAT)) return -4;
    float fr[8];
    for (int k = 0; k < 8; ++k) fr[k] = (float)(3.14159265358979323846 * pow(2.0, k / 2.0));
    float* dfr = nullptr;
    if (galloc(&dfr, 8)) return -4;
    (void)hipMemcpy(dfr, fr, sizeof fr, hipMemcpyHostToDevice);
    if (f16) hipLaunchKernelGGL(k_guide_features<true>, dim3(GUIDE_CELLS / 256), dim3(256), 0, 0, g_features[1], dfr);
    else hipLaunchKernelGGL(k_guide_features<false>, dim3(GUIDE_CELLS / 256), dim3(256), 0, 0, g_features[0], dfr);
    (void)hipDeviceSynchronize();
    (void)hipFree(dfr);
    g_features_dev[f16] = dev;
    return 0;
}

int guide_alloc(GuideBuffers& g) {
    int dev = 0;
    if (!guide_net_ok(g.net)) return -1;
    if (hipGetDevice(&dev) != hipSuccess) return -2;
    const GuideNet n = g.net;
    g.nparams = guide_nparams(n);
    const size_t acts = (size_t)n.layers * GUIDE_CELLS * n.hidden;
    if (galloc(&g.theta, g.nparams) || galloc(&g.m, g.nparams) || galloc(&g.v, g.nparams) ||
        galloc(&g.grad, g.nparams) || galloc(&g.wfwd, guide_lw_size(n)) || galloc(&g.wbwd, guide_lw_size(n)) ||
        galloc(&g.acts, acts) || galloc(&g.dh, 2 * (size_t)GUIDE_CELLS * n.hidden) ||
        galloc(&g.logits, GUIDE_CELLS) || galloc(&g.prob, GUIDE_CELLS) || galloc(&g.cdf, GUIDE_CELLS) ||
        galloc(&g.eprob, GUIDE_CELLS) || galloc(&g.ecdf, GUIDE_CELLS) ||
        galloc(&g.hist, GUIDE_CELLS) || galloc(&g.hpass, GUIDE_CELLS) ||
        galloc(&g.part, (size_t)GUIDE_BLOCKS * (g.nparams + 1)) || galloc(&g.part2, 4 * 1024) || galloc(&g.scal, 8)) {
        guide_free(g);
        return -4;
    }
    if (ensure_features(dev, n.fp16)) {
        guide_free(g);
        return -4;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

void guide_free(GuideBuffers& g) {
    for (void* p : {(void*)g.theta, (void*)g.m, (void*)g.v, (void*)g.grad, (void*)g.wfwd, (void*)g.wbwd,
                    (void*)g.acts, (void*)g.dh, (void*)g.logits, (void*)g.prob, (void*)g.cdf, (void*)g.eprob,
                    (void*)g.ecdf, (void*)g.hist,
                    (void*)g.hpass, (void*)g.part, (void*)g.part2, (void*)g.scal, (void*)g.pt_cell,
                    (void*)g.pcell_mass, (void*)g.pc_start, (void*)g.pc_pt, (void*)g.pc_cum, (void*)g.pcell,
                    (void*)g.tgt})
        if (p) (void)hipFree(p);
    const GuideNet n = g.net;
    g = GuideBuffers();
    g.net = n;
}

static void launch_lowp(GuideBuffers& g, hipStream_t st) {
    const int n = guide_lw_size(g.net);
    if (g.net.fp16)
        hipLaunchKernelGGL(k_guide_lowp<true>, dim3((n + 255) / 256), dim3(256), 0, st, g.theta, g.wfwd, g.wbwd, g.net);
    else
        hipLaunchKernelGGL(k_guide_lowp<false>, dim3((n + 255) / 256), dim3(256), 0, st, g.theta, g.wfwd, g.wbwd, g.net);
}

int guide_upload_params(GuideBuffers& g, const float* theta, hipStream_t st) {
    if (hipMemcpyAsync(g.theta, theta, sizeof(float) * g.nparams, hipMemcpyHostToDevice, st) != hipSuccess) return -2;
    hipLaunchKernelGGL(k_guide_zero, dim3((g.nparams + 255) / 256), dim3(256), 0, st, g.m, g.nparams);
    hipLaunchKernelGGL(k_guide_zero, dim3((g.nparams + 255) / 256), dim3(256), 0, st, g.v, g.nparams);
    launch_lowp(g, st);
    hipLaunchKernelGGL(k_guide_zero, dim3(1), dim3(64), 0, st, g.scal, 8);  // lse, sums, loss, Adam step count
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

constexpr int FWD_BLOCKS = 1024;

// Per-network launches.
template <int H, int L, bool F16>
struct NetLaunch {
    static constexpr GuideNet NET{H, L, F16};
    static void forward(GuideBuffers& g, bool store, hipStream_t st) {
        const uint16_t* X = g_features[F16 ? 1 : 0];
        if (store)
            hipLaunchKernelGGL((k_mlp_forward<H, L, F16, true>), dim3(FWD_BLOCKS), dim3(64), 0, st, X, g.wfwd, g.theta,
                               g.acts, g.logits, g.part2);
        else
            hipLaunchKernelGGL((k_mlp_forward<H, L, F16, false>), dim3(FWD_BLOCKS), dim3(64), 0, st, X, g.wfwd,
                               g.theta, g.acts, g.logits, g.part2);
    }
    static void backward(GuideBuffers& g, const float* w, hipStream_t st) {
        const size_t LS = (size_t)GUIDE_CELLS * H;
        const uint16_t* X = g_features[F16 ? 1 : 0];
        const int np = g.nparams, wo = guide_wout_off(NET);
        // layer L: dH_L from the logits; writes dH_{L-1} into dh[L & 1]
        hipLaunchKernelGGL((k_mlp_backward<H, H, F16, true, false>), dim3(GUIDE_BLOCKS), dim3(256), 0, st,
                           g.acts + (L - 2) * LS, (const uint16_t*)nullptr, g.acts + (L - 1) * LS, g.logits, w,
                           g.scal, g.theta, g.wbwd + guide_lw_off(NET, L), g.dh + (L & 1) * LS, g.part, np,
                           guide_w_off(NET, L), guide_b_off(NET, L - 1), guide_b_off(NET, L), wo);
        for (int l = L - 1; l >= 2; --l)
            hipLaunchKernelGGL((k_mlp_backward<H, H, F16, false, false>), dim3(GUIDE_BLOCKS), dim3(256), 0, st,
                               g.acts + (l - 2) * LS, g.dh + ((l + 1) & 1) * LS, (const uint16_t*)nullptr, g.logits,
                               w, g.scal, g.theta, g.wbwd + guide_lw_off(NET, l), g.dh + (l & 1) * LS, g.part, np,
                               guide_w_off(NET, l), guide_b_off(NET, l - 1), 0, wo);
        hipLaunchKernelGGL((k_mlp_backward<H, GUIDE_FEAT, F16, false, true>), dim3(GUIDE_BLOCKS), dim3(256), 0, st, X,
                           g.dh + (2 & 1) * LS, (const uint16_t*)nullptr, g.logits, w, g.scal, g.theta,
                           g.wbwd + guide_lw_off(NET, 1), (uint16_t*)nullptr, g.part, np, guide_w_off(NET, 1), 0, 0,
                           wo);
    }
};

template <class F>
static void with_net(const GuideNet& n, F&& f) {
    if (n.hidden == 64 && n.layers == 4) {
        if (n.fp16) f(NetLaunch<64, 4, true>{}); else f(NetLaunch<64, 4, false>{});
    } else if (n.hidden == 64 && n.layers == 6) {
        if (n.fp16) f(NetLaunch<64, 6, true>{}); else f(NetLaunch<64, 6, false>{});
    } else if (n.hidden == 128 && n.layers == 4) {
        if (n.fp16) f(NetLaunch<128, 4, true>{}); else f(NetLaunch<128, 4, false>{});
    } else {
        if (n.fp16) f(NetLaunch<128, 6, true>{}); else f(NetLaunch<128, 6, false>{});
    }
}

void guide_infer(GuideBuffers& g, hipStream_t st) {
    with_net(g.net, [&](auto op) { op.forward(g, false, st); });
    hipLaunchKernelGGL(k_guide_lse, dim3(1), dim3(256), 0, st, g.part2, FWD_BLOCKS, g.scal);
    hipLaunchKernelGGL(k_guide_prob, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, g.logits, g.scal, g.prob, g.cdf,
                       g.part2, (const float*)g.pcell_mass);
    hipLaunchKernelGGL(k_guide_cdf_offsets, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, g.cdf, g.part2);
}

void guide_hist_sum(GuideBuffers& g, const float* w, hipStream_t st) {
    double* hpart = reinterpret_cast<double*>(g.part2 + 2048);  // after the forward partials (2 * 1024 floats)
    hipLaunchKernelGGL(k_guide_hsum, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, w, hpart);
    hipLaunchKernelGGL(k_guide_hsum_total, dim3(1), dim3(256), 0, st, hpart, g.scal);
}

void guide_gradient(GuideBuffers& g, const float* w, hipStream_t st) {
    // training forward (activations kept), its log-sum-exp, then the layers backwards
    with_net(g.net, [&](auto op) { op.forward(g, true, st); });
    hipLaunchKernelGGL(k_guide_lse, dim3(1), dim3(256), 0, st, g.part2, FWD_BLOCKS, g.scal);
    with_net(g.net, [&](auto op) { op.backward(g, w, st); });
    hipLaunchKernelGGL(k_guide_grad_reduce, dim3((g.nparams + 1 + 255) / 256), dim3(256), 0, st, g.part,
                       GUIDE_BLOCKS, g.nparams, g.scal, g.grad, g.scal);
}

void guide_adam(GuideBuffers& g, float lr, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_adam, dim3((g.nparams + 255) / 256), dim3(256), 0, st, g.theta, g.m, g.v, g.grad,
                       g.scal, lr, g.nparams);
    hipLaunchKernelGGL(k_guide_tick, dim3(1), dim3(64), 0, st, g.scal);
    launch_lowp(g, st);
}

void guide_restore(GuideBuffers& g, hipStream_t st) {
    launch_lowp(g, st);
    guide_infer(g, st);
}

// tgt = sqrt(pcell * max(hist, 0)): the RMS target from the histogram of
// deposit^2 x q / p
__global__ void k_guide_rms_target(const float* __restrict__ hist, const float* __restrict__ pcell,
                                   float* __restrict__ tgt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GUIDE_CELLS) return;
    const float h = hist[k];
    tgt[k] = h > 0.0f ? sqrtf(pcell[k] * h) : 0.0f;
}
void guide_rms_target(GuideBuffers& g, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_rms_target, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hist, g.pcell, g.tgt);
}

void guide_clear_hist(GuideBuffers& g, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_zero, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hist, GUIDE_CELLS);
    hipLaunchKernelGGL(k_guide_zero, dim3(2 * GUIDE_CELLS / 256), dim3(256), 0, st, reinterpret_cast<float*>(g.hpass),
                       2 * GUIDE_CELLS);
}

void guide_fold_hist(GuideBuffers& g, float decay, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_fold, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hist, g.hpass, decay);
}

__global__ void k_guide_publish(const float* __restrict__ prob, const float* __restrict__ cdf,
                                float* __restrict__ eprob, float* __restrict__ ecdf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= GUIDE_CELLS) return;
    eprob[i] = prob[i];
    ecdf[i] = cdf[i];
}
void guide_publish(GuideBuffers& g, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_publish, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.prob, g.cdf, g.eprob, g.ecdf);
}

}  // namespace sppm
