// sppm_guide.hip -- training and inference kernels of the learned emission
// proposal (bf16 MFMA 16x16x32, float32 accumulation).
//
// Tiles of 32 cells; one wave per block.  For v_mfma_f32_16x16x32_bf16 lane l
// supplies A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16] and receives
// D[4*(l/16) + i][l%16], i = 0..3.  Activations live in LDS as bfloat16 in
// two layouts: [row][feature] (A operand of H * W and dH * W^T) and
// [feature][row] (A / B operands of the weight gradient H^T * dH, whose
// reduction runs over rows).  Weights are read from global memory in the
// layout each product needs ([out][in] forward, [in][out] backward).
#include "sppm_guide.h"

#include <math.h>

#include <vector>

namespace sppm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TR = 32;   // cells per tile
constexpr int RS = 72;   // stride (bf16) of [row][64] matrices (padded against bank conflicts)
constexpr int CS = 40;   // stride (bf16) of [64][row] matrices
constexpr int DS = 65;   // stride (float) of the [in][out] gradient accumulators

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---- input features of every cell (constant; computed once) ----------------
// coordinate j of cell c as the oracle defines it, then sin / cos at
// pi * 2^(k/2) in double, rounded to float, then to bf16.
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
            out[16 * j + 2 * k] = f2bf((float)sin(x));
            out[16 * j + 2 * k + 1] = f2bf((float)cos(x));
        }
}

struct __align__(16) GuideTile {
    uint16_t xr[TR * RS];
    uint16_t xt[64 * CS];
    uint16_t hr[3][TR * RS];
    uint16_t ht[3][64 * CS];
    uint16_t dr[TR * RS];
    uint16_t dt[64 * CS];
    float g[TR];
    float logit[TR];
};

// Loads the features of cells [base, base + 32) into xr (and xt when asked).
__device__ __forceinline__ void load_x(GuideTile& s, const uint16_t* __restrict__ X, int base, int lane, bool trans) {
    const int row = lane >> 1, half = lane & 1;  // two lanes per row, 32 features each
    const uint16_t* src = X + (size_t)(base + row) * GUIDE_FEAT + half * 32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = *reinterpret_cast<const uint4*>(src + 8 * q);
        *reinterpret_cast<uint4*>(&s.xr[row * RS + half * 32 + 8 * q]) = v;
        if (trans) {
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int t = 0; t < 8; ++t) s.xt[(half * 32 + 8 * q + t) * CS + row] = e[t];
        }
    }
}

// Y = relu(in * W_l + b_l) for the tile: acc[mt][nt] holds rows mt*16 + 4*(l/16) + i, column nt*16 + l%16.
__device__ __forceinline__ void layer_fwd(const uint16_t* in_r, const uint16_t* __restrict__ wT,
                                          const float* __restrict__ bias, f32x4 acc[2][4], int lane) {
    const int r16 = lane & 15, q = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
                c = mfma(ld8(in_r + (mt * 16 + r16) * RS + kb * 32 + 8 * q),
                         ld8(wT + (nt * 16 + r16) * 64 + kb * 32 + 8 * q), c);
            const float bb = bias[nt * 16 + r16];
#pragma unroll
            for (int i = 0; i < 4; ++i) c[i] = fmaxf(c[i] + bb, 0.0f);
            acc[mt][nt] = c;
        }
}

// Writes a D-layout tile set as bf16 into [row][64] (and [64][row]).
__device__ __forceinline__ void store_act(const f32x4 acc[2][4], uint16_t* out_r, uint16_t* out_t, int lane) {
    const int r16 = lane & 15, q = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int col = nt * 16 + r16, row = mt * 16 + 4 * q;
            uint16_t h[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                h[i] = f2bf(acc[mt][nt][i]);
                out_r[(row + i) * RS + col] = h[i];
            }
            if (out_t) {
                uint2 pk;
                pk.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
                pk.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
                *reinterpret_cast<uint2*>(&out_t[col * CS + row]) = pk;
            }
        }
}

// Full forward of a tile; activations of layers 1-3 stored (with the
// transposed copies when `trans`), layer 4 returned in acc4, logits in s.logit.
__device__ __forceinline__ void tile_forward(GuideTile& s, const uint16_t* __restrict__ X,
                                             const uint16_t* __restrict__ wfwd, const float* __restrict__ theta,
                                             int base, int lane, bool trans, f32x4 acc4[2][4]) {
    load_x(s, X, base, lane, trans);
    __syncthreads();
    f32x4 acc[2][4];
    const uint16_t* in_r = s.xr;
    for (int l = 1; l <= 3; ++l) {
        layer_fwd(in_r, wfwd + (l - 1) * GUIDE_WSZ, theta + guide_b_off(l), acc, lane);
        store_act(acc, s.hr[l - 1], trans ? s.ht[l - 1] : nullptr, lane);
        __syncthreads();
        in_r = s.hr[l - 1];
    }
    layer_fwd(in_r, wfwd + 3 * GUIDE_WSZ, theta + guide_b_off(4), acc4, lane);
    // logit = h4 . w5 + b5 (float32, h4 before rounding)
    const int r16 = lane & 15, q = lane >> 4;
    float w5[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) w5[nt] = theta[GUIDE_W5 + nt * 16 + r16];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float p = 0.0f;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) p += acc4[mt][nt][i] * w5[nt];
            p += __shfl_xor(p, 1, 64);
            p += __shfl_xor(p, 2, 64);
            p += __shfl_xor(p, 4, 64);
            p += __shfl_xor(p, 8, 64);
            if (r16 == 0) s.logit[mt * 16 + 4 * q + i] = p + theta[GUIDE_B5];
        }
    __syncthreads();
}

// ---- inference: logits of all cells + per-block (max, sum exp) -------------
__global__ void __launch_bounds__(64) k_guide_forward(const uint16_t* __restrict__ X, const uint16_t* __restrict__ wfwd,
                                                      const float* __restrict__ theta, float* __restrict__ logits,
                                                      float* __restrict__ part2) {
    __shared__ GuideTile s;
    const int lane = threadIdx.x;
    float m = -INFINITY, sum = 0.0f;
    for (int t = blockIdx.x; t < GUIDE_CELLS / TR; t += gridDim.x) {
        f32x4 acc4[2][4];
        tile_forward(s, X, wfwd, theta, t * TR, lane, false, acc4);
        if (lane < TR) {
            const float l = s.logit[lane];
            logits[t * TR + lane] = l;
            if (l > m) {
                sum = sum * expf(m - l) + 1.0f;
                m = l;
            } else {
                sum += expf(l - m);
            }
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
__global__ void __launch_bounds__(256) k_guide_prob(const float* __restrict__ logits, const float* __restrict__ scal,
                                                    float* __restrict__ prob, float* __restrict__ cdf,
                                                    float* __restrict__ bsum) {
    __shared__ float ws[4];
    const int t = threadIdx.x, base = blockIdx.x * 1024 + 4 * t;
    const float lse = scal[0];
    float v[4], run = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = expf(logits[base + k] - lse);
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

// ---- training: gradient partials per block ------------------------------------
__global__ void __launch_bounds__(64) k_guide_backward(const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ wfwd,
                                                       const uint16_t* __restrict__ wbwd,
                                                       const float* __restrict__ theta,
                                                       const float* __restrict__ hist,
                                                       const float* __restrict__ scal, float* __restrict__ part) {
    __shared__ GuideTile s;
    __shared__ float dW[4][64 * DS];
    __shared__ float db[5][64];  // b1..b4, w5
    const int lane = threadIdx.x, r16 = lane & 15, q = lane >> 4;
    for (int k = lane; k < 4 * 64 * DS; k += 64) (&dW[0][0])[k] = 0.0f;
    for (int k = lane; k < 5 * 64; k += 64) (&db[0][0])[k] = 0.0f;
    const float lse = scal[0];
    const float invW = scal[1] > 0.0f ? 1.0f / scal[1] : 0.0f;
    float db5 = 0.0f, wl = 0.0f;
    __syncthreads();
    for (int t = blockIdx.x; t < GUIDE_CELLS / TR; t += gridDim.x) {
        const int base = t * TR;
        f32x4 acc4[2][4];
        tile_forward(s, X, wfwd, theta, base, lane, true, acc4);
        if (lane < TR) {
            const float l = s.logit[lane], h = hist[base + lane];
            s.g[lane] = expf(l - lse) - h * invW;
            wl += h * l;
        }
        __syncthreads();
        if (lane == 0)
            for (int r = 0; r < TR; ++r) db5 += s.g[r];
        // dH4 = g * w5 * (h4 > 0); db4, dw5
        {
            float w5[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) w5[nt] = theta[GUIDE_W5 + nt * 16 + r16];
            f32x4 d4[2][4];
            float sb[4] = {0, 0, 0, 0}, sw[4] = {0, 0, 0, 0};
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float gr = s.g[mt * 16 + 4 * q + i];
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        const float h = acc4[mt][nt][i];
                        const float d = h > 0.0f ? gr * w5[nt] : 0.0f;
                        d4[mt][nt][i] = d;
                        sb[nt] += d;
                        sw[nt] += h * gr;
                    }
                }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                sb[nt] += __shfl_xor(sb[nt], 16, 64);
                sb[nt] += __shfl_xor(sb[nt], 32, 64);
                sw[nt] += __shfl_xor(sw[nt], 16, 64);
                sw[nt] += __shfl_xor(sw[nt], 32, 64);
                if (q == 0) {
                    db[3][nt * 16 + r16] += sb[nt];
                    db[4][nt * 16 + r16] += sw[nt];
                }
            }
            store_act(d4, s.dr, s.dt, lane);
        }
        __syncthreads();
        const uint16_t* cur_r = s.dr;
        const uint16_t* cur_t = s.dt;
        for (int l = 4; l >= 1; --l) {
            uint16_t* in_r = l == 1 ? s.xr : s.hr[l - 2];
            uint16_t* in_t = l == 1 ? s.xt : s.ht[l - 2];
            // dW_l += H_{l-1}^T dH_l  (reduction over the 32 rows of the tile)
            float* acc = dW[l - 1];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const int in0 = mt * 16 + 4 * q, out = nt * 16 + r16;
                    f32x4 c;
#pragma unroll
                    for (int i = 0; i < 4; ++i) c[i] = acc[(in0 + i) * DS + out];
                    c = mfma(ld8(in_t + (mt * 16 + r16) * CS + 8 * q), ld8(cur_t + (nt * 16 + r16) * CS + 8 * q), c);
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[(in0 + i) * DS + out] = c[i];
                }
            if (l == 1) break;
            // dH_{l-1} = (dH_l W_l^T) * (H_{l-1} > 0); db_{l-1}
            f32x4 d[2][4];
            float sb[4] = {0, 0, 0, 0};
            const uint16_t* wb = wbwd + (l - 1) * GUIDE_WSZ;
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb)
                        c = mfma(ld8(cur_r + (mt * 16 + r16) * RS + kb * 32 + 8 * q),
                                 ld8(wb + (nt * 16 + r16) * 64 + kb * 32 + 8 * q), c);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint16_t hb = in_r[(mt * 16 + 4 * q + i) * RS + nt * 16 + r16];
                        c[i] = hb != 0 && !(hb & 0x8000u) ? c[i] : 0.0f;  // bf16 > 0
                        sb[nt] += c[i];
                    }
                    d[mt][nt] = c;
                }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                sb[nt] += __shfl_xor(sb[nt], 16, 64);
                sb[nt] += __shfl_xor(sb[nt], 32, 64);
                if (q == 0) db[l - 2][nt * 16 + r16] += sb[nt];
            }
            __syncthreads();  // every lane has read in_t (dW) and its mask entries of in_r
            store_act(d, in_r, in_t, lane);
            __syncthreads();
            cur_r = in_r;
            cur_t = in_t;
        }
        __syncthreads();
    }
    // per-block partial gradient (+ sum of W * logit at NPARAMS)
    float* out = part + (size_t)blockIdx.x * (GUIDE_NPARAMS + 1);
    for (int l = 1; l <= 4; ++l)
        for (int k = lane; k < GUIDE_WSZ; k += 64) out[guide_w_off(l) + k] = dW[l - 1][(k >> 6) * DS + (k & 63)];
    for (int l = 1; l <= 4; ++l) out[guide_b_off(l) + lane] = db[l - 1][lane];
    out[GUIDE_W5 + lane] = db[4][lane];
    for (int off = 32; off > 0; off >>= 1) wl += __shfl_xor(wl, off, 64);
    if (lane == 0) {
        out[GUIDE_B5] = db5;
        out[GUIDE_NPARAMS] = wl;
    }
}

// sum of the block partials (fixed order); loss = lse - sum(W * logit) / sum(W)
__global__ void k_guide_grad_reduce(const float* __restrict__ part, int nblocks, const float* __restrict__ scal,
                                    float* __restrict__ grad, float* __restrict__ scal_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > GUIDE_NPARAMS) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += (double)part[(size_t)b * (GUIDE_NPARAMS + 1) + k];
    if (k < GUIDE_NPARAMS) {
        grad[k] = (float)s;
    } else {
        const double w = scal[1];
        scal_out[2] = w > 0.0 ? (float)((double)scal[0] - s / w) : 0.0f;
    }
}

// bf16 copies of the hidden-layer weights in both layouts
__global__ void k_guide_bf16(const float* __restrict__ theta, uint16_t* __restrict__ wfwd,
                             uint16_t* __restrict__ wbwd) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 4 * GUIDE_WSZ) return;
    const int l = k / GUIDE_WSZ, e = k % GUIDE_WSZ, in = e >> 6, out = e & 63;
    const uint16_t b = f2bf(theta[guide_w_off(l + 1) + e]);
    wbwd[l * GUIDE_WSZ + in * 64 + out] = b;
    wfwd[l * GUIDE_WSZ + out * 64 + in] = b;
}

// Adam (float32 master weights); skipped when the histogram is empty.
// Adam step count t + 1 = scal[3] + 1 (kept on the device so graph replays count).
__global__ void k_guide_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                             const float* __restrict__ grad, const float* __restrict__ scal, float lr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GUIDE_NPARAMS || !(scal[1] > 0.0f)) return;
    const float t = scal[3] + 1.0f;
    const float c1 = 1.0f / (1.0f - powf(0.9f, t)), c2 = 1.0f / (1.0f - powf(0.999f, t));
    const float g = grad[k];
    const float mk = 0.9f * m[k] + 0.1f * g;
    const float vk = 0.999f * v[k] + 0.001f * g * g;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk * c1) / (sqrtf(vk * c2) + 1e-8f);
}

__global__ void k_guide_fold(float* __restrict__ hist, float* __restrict__ hpass, float decay) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GUIDE_CELLS) return;
    hist[k] = decay * hist[k] + hpass[k];
    hpass[k] = 0.0f;
}

__global__ void k_guide_tick(float* __restrict__ scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && scal[1] > 0.0f) scal[3] += 1.0f;
}

__global__ void k_guide_zero(float* __restrict__ p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0f;
}

// ---------------------------------------------------------------------------
static uint16_t* g_features = nullptr;  // per-device constant input, allocated with the first GuideBuffers
static int g_features_dev = -1;

template <class T>
static int galloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1)) == hipSuccess ? 0 : -4;
}

int guide_alloc(GuideBuffers& g) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -2;
    if (galloc(&g.theta, GUIDE_NPARAMS) || galloc(&g.m, GUIDE_NPARAMS) || galloc(&g.v, GUIDE_NPARAMS) ||
        galloc(&g.grad, GUIDE_NPARAMS) || galloc(&g.wfwd, 4 * GUIDE_WSZ) || galloc(&g.wbwd, 4 * GUIDE_WSZ) ||
        galloc(&g.logits, GUIDE_CELLS) || galloc(&g.prob, GUIDE_CELLS) || galloc(&g.cdf, GUIDE_CELLS) ||
        galloc(&g.hist, GUIDE_CELLS) || galloc(&g.hpass, GUIDE_CELLS) || galloc(&g.part, (size_t)GUIDE_BLOCKS * (GUIDE_NPARAMS + 1)) ||
        galloc(&g.part2, 4 * 1024) || galloc(&g.scal, 8)) {
        guide_free(g);
        return -4;
    }
    if (g_features_dev != dev) {
        if (g_features) (void)hipFree(g_features);
        g_features = nullptr;
        if (galloc(&g_features, (size_t)GUIDE_CELLS * GUIDE_FEAT)) {
            guide_free(g);
            return -4;
        }
        float fr[8];
        for (int k = 0; k < 8; ++k) fr[k] = (float)(3.14159265358979323846 * pow(2.0, k / 2.0));
        float* dfr = nullptr;
        if (galloc(&dfr, 8)) return -4;
        (void)hipMemcpy(dfr, fr, sizeof fr, hipMemcpyHostToDevice);
        hipLaunchKernelGGL(k_guide_features, dim3(GUIDE_CELLS / 256), dim3(256), 0, 0, g_features, dfr);
        (void)hipDeviceSynchronize();
        (void)hipFree(dfr);
        g_features_dev = dev;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

void guide_free(GuideBuffers& g) {
    for (void* p : {(void*)g.theta, (void*)g.m, (void*)g.v, (void*)g.grad, (void*)g.wfwd, (void*)g.wbwd,
                    (void*)g.logits, (void*)g.prob, (void*)g.cdf, (void*)g.hist, (void*)g.hpass, (void*)g.part,
                    (void*)g.part2,
                    (void*)g.scal})
        if (p) (void)hipFree(p);
    g = GuideBuffers();
}

int guide_upload_params(GuideBuffers& g, const float* theta, hipStream_t st) {
    if (hipMemcpyAsync(g.theta, theta, sizeof(float) * GUIDE_NPARAMS, hipMemcpyHostToDevice, st) != hipSuccess)
        return -2;
    hipLaunchKernelGGL(k_guide_zero, dim3((GUIDE_NPARAMS + 255) / 256), dim3(256), 0, st, g.m, GUIDE_NPARAMS);
    hipLaunchKernelGGL(k_guide_zero, dim3((GUIDE_NPARAMS + 255) / 256), dim3(256), 0, st, g.v, GUIDE_NPARAMS);
    hipLaunchKernelGGL(k_guide_bf16, dim3(4 * GUIDE_WSZ / 256), dim3(256), 0, st, g.theta, g.wfwd, g.wbwd);
    hipLaunchKernelGGL(k_guide_zero, dim3(1), dim3(64), 0, st, g.scal, 8);  // lse, sums, loss, Adam step count
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

void guide_infer(GuideBuffers& g, hipStream_t st) {
    const int nb = 1024;
    hipLaunchKernelGGL(k_guide_forward, dim3(nb), dim3(64), 0, st, g_features, g.wfwd, g.theta, g.logits, g.part2);
    hipLaunchKernelGGL(k_guide_lse, dim3(1), dim3(256), 0, st, g.part2, nb, g.scal);
    hipLaunchKernelGGL(k_guide_prob, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, g.logits, g.scal, g.prob, g.cdf,
                       g.part2);
    hipLaunchKernelGGL(k_guide_cdf_offsets, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, g.cdf, g.part2);
}

void guide_hist_sum(GuideBuffers& g, hipStream_t st) {
    double* hpart = reinterpret_cast<double*>(g.part2 + 2048);  // after the forward partials (2 * 1024 floats)
    hipLaunchKernelGGL(k_guide_hsum, dim3(GUIDE_CELLS / 1024), dim3(256), 0, st, g.hist, hpart);
    hipLaunchKernelGGL(k_guide_hsum_total, dim3(1), dim3(256), 0, st, hpart, g.scal);
}

void guide_gradient(GuideBuffers& g, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_backward, dim3(GUIDE_BLOCKS), dim3(64), 0, st, g_features, g.wfwd, g.wbwd, g.theta,
                       g.hist, g.scal, g.part);
    hipLaunchKernelGGL(k_guide_grad_reduce, dim3((GUIDE_NPARAMS + 1 + 255) / 256), dim3(256), 0, st, g.part,
                       GUIDE_BLOCKS, g.scal, g.grad, g.scal);
}

void guide_adam(GuideBuffers& g, float lr, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_adam, dim3((GUIDE_NPARAMS + 255) / 256), dim3(256), 0, st, g.theta, g.m, g.v, g.grad,
                       g.scal, lr);
    hipLaunchKernelGGL(k_guide_tick, dim3(1), dim3(64), 0, st, g.scal);
    hipLaunchKernelGGL(k_guide_bf16, dim3(4 * GUIDE_WSZ / 256), dim3(256), 0, st, g.theta, g.wfwd, g.wbwd);
}

void guide_clear_hist(GuideBuffers& g, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_zero, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hist, GUIDE_CELLS);
    hipLaunchKernelGGL(k_guide_zero, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hpass, GUIDE_CELLS);
}

void guide_fold_hist(GuideBuffers& g, float decay, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_fold, dim3(GUIDE_CELLS / 256), dim3(256), 0, st, g.hist, g.hpass, decay);
}

}  // namespace sppm
