// Learned radiance cache (no counterpart in the reference): a 16-level hash-grid encoding
// (Mueller et al., Instant NGP 2022) of the position, beside per-query surface attributes (shading
// normal, albedo, roughness and a constant 1 that stands in for the first layer's bias), feeding a
// bias-free 64-64-64-64-64-3 ReLU MLP, trained online with a relative L2 loss (Mueller et al.,
// neural radiance caching 2021) and Adam (Kingma & Ba).
//
// Device layout:
//   table      FP16 [levels][2^log2_table][2]   (+ FP32 master, Adam m/v, BF16-pair gradient)
//   weights    FP32 master [out][in] per layer, concatenated (16,576 values for the default net)
//              BF16 forward copy  Wf_l [out_pad16][in]  and backward copy Wb_l = W_l^T [in][out_pad32]
// The MLP runs on CDNA4 matrix cores: v_mfma_f32_16x16x32_bf16 with neurons as rows (M), queries
// as columns (N) and FP32 accumulation. One wave owns a tile of 64 queries; activations live in LDS
// as BF16 rows [query][feature] so that every operand fetch is one 16-byte LDS read per lane.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/spk_pt.h"
#include "spk_cache.h"
#include "spk_context.h"

namespace spk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kQ = 64;        // queries per wave tile
constexpr int kEnc = 32;      // encoding width (levels * features)
constexpr int kAttr = 8;      // surface attributes per query (SPK_CACHE_ATTRS)
constexpr int kIn = 64;       // MLP input: encoding, attributes, zeros
constexpr int kH = 64;        // hidden width
constexpr int kOutPad = 16;   // output rows padded to one MFMA tile
constexpr int kStrideH = kH + 8;   // LDS row stride (bf16), padded against bank conflicts
constexpr int kStrideG = kStrideH;

struct CacheDev {
    const __half2 *table;      // [L*T]
    const __bf16 *wf[5];       // forward weights
    const __bf16 *wb[5];       // backward (transposed) weights
    uint32_t res[16];          // grid resolution per level
    uint32_t dense[16];        // 1 when the level is indexed densely
    uint32_t log2_table;
    float lo[3], inv_ext[3];   // world -> [0,1]^3
};

__device__ __forceinline__ uint32_t grid_index(const CacheDev &c, uint32_t l, uint32_t x, uint32_t y, uint32_t z) {
    if (c.dense[l]) {
        const uint32_t side = c.res[l] + 1u;
        return x + side * (y + side * z);
    }
    return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & ((1u << c.log2_table) - 1u);
}

__device__ __forceinline__ float half_bits(uint32_t b) { return __half2float(__ushort_as_half((unsigned short)b)); }

// trilinear hash-grid features of one point in [0,1]^3 (16 levels x 2 features), written as BF16 to
// row[0..31] (zeros when !valid)
__device__ __forceinline__ void encode_point(const CacheDev &c, float px, float py, float pz, bool valid,
                                             __bf16 *row) {
    const uint32_t T = 1u << c.log2_table;
#pragma unroll 2
    for (uint32_t l = 0; l < 16; ++l) {
        const float res = (float)c.res[l];
        float pos[3] = {px * res, py * res, pz * res};
        uint32_t cell[3];
        float fr[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float cf = floorf(pos[a]);
            cf = fminf(fmaxf(cf, 0.f), res - 1.f);
            cell[a] = (uint32_t)cf;
            fr[a] = pos[a] - cf;
        }
        float f0 = 0.f, f1 = 0.f;
        const uint32_t *lv = reinterpret_cast<const uint32_t *>(c.table) + (size_t)l * T;
#pragma unroll
        for (int yz = 0; yz < 4; ++yz) {
            const uint32_t cy = cell[1] + (yz & 1), cz = cell[2] + (yz >> 1);
            const uint32_t i0 = grid_index(c, l, cell[0], cy, cz), i1 = grid_index(c, l, cell[0] + 1u, cy, cz);
            // the two x-neighbours fall into one aligned quad of entries when i0 ^ i1 < 4: on a dense level
            // unless i0 % 4 == 3, on a hashed one (x has hash prime 1, so i1 = i0 ^ (x ^ (x + 1))) unless
            // x % 4 == 3. One 16-byte load brings the quad, a second load the other neighbour otherwise
            // (1.25 loads per pair on average). The gathers are bound by L2 requests, not bytes. The
            // quad stays inside the level: T is a multiple of 4.
            const uint4 quad = *reinterpret_cast<const uint4 *>(lv + (i0 & ~3u));
            const uint32_t e0 = (i0 & 2u) ? ((i0 & 1u) ? quad.w : quad.z) : ((i0 & 1u) ? quad.y : quad.x);
            const uint32_t q1 = (i1 & 2u) ? ((i1 & 1u) ? quad.w : quad.z) : ((i1 & 1u) ? quad.y : quad.x);
            const uint32_t e1 = (i0 ^ i1) < 4u ? q1 : lv[i1];
            const float wyz = ((yz & 1) ? fr[1] : 1.f - fr[1]) * ((yz >> 1) ? fr[2] : 1.f - fr[2]);
            const float w0 = (1.f - fr[0]) * wyz, w1 = fr[0] * wyz;
            f0 += w0 * half_bits(e0 & 0xFFFFu);
            f1 += w0 * half_bits(e0 >> 16);
            f0 += w1 * half_bits(e1 & 0xFFFFu);
            f1 += w1 * half_bits(e1 >> 16);
        }
        row[2 * l] = (__bf16)(valid ? f0 : 0.f);
        row[2 * l + 1] = (__bf16)(valid ? f1 : 0.f);
    }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// scatter d(loss)/d(features) of one point (BF16 row g[0..31]) into the table gradient: BF16 pairs,
// one packed atomic add per corner for both features (global_atomic_pk_add_bf16), half the atomics
// of two FP32 adds; Adam normalises the step by the gradient's running RMS, so 8 mantissa bits of
// accumulated gradient are enough
__device__ __forceinline__ void encode_backward(const CacheDev &c, float px, float py, float pz, const __bf16 *g,
                                                bf16x2 *gtable) {
    const uint32_t T = 1u << c.log2_table;
#pragma unroll 1
    for (uint32_t l = 0; l < 16; ++l) {
        const float res = (float)c.res[l];
        float pos[3] = {px * res, py * res, pz * res};
        uint32_t cell[3];
        float fr[3];
        for (int a = 0; a < 3; ++a) {
            float cf = floorf(pos[a]);
            cf = fminf(fmaxf(cf, 0.f), res - 1.f);
            cell[a] = (uint32_t)cf;
            fr[a] = pos[a] - cf;
        }
        const float g0 = (float)g[2 * l], g1 = (float)g[2 * l + 1];
        if (g0 == 0.f && g1 == 0.f) continue;
        for (int k = 0; k < 8; ++k) {
            const uint32_t cx = cell[0] + (k & 1), cy = cell[1] + ((k >> 1) & 1), cz = cell[2] + ((k >> 2) & 1);
            const float w = ((k & 1) ? fr[0] : 1.f - fr[0]) * (((k >> 1) & 1) ? fr[1] : 1.f - fr[1]) *
                            (((k >> 2) & 1) ? fr[2] : 1.f - fr[2]);
            const size_t e = (size_t)l * T + grid_index(c, l, cx, cy, cz);
            const bf16x2 v = {(__bf16)(w * g0), (__bf16)(w * g1)};
            __builtin_amdgcn_global_atomic_fadd_v2bf16(gtable + e, v);
        }
    }
}

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// D[M_TILES*16 x 64] = W[M_TILES*16 x K_STEPS*32] * Hin^T, Hin rows [query][k] in LDS.
// The result is written to Hout rows [query][neuron] as BF16, after ReLU when RELU (or as FP32 to
// out_f32 [query][m_valid]). RECORD sets bit n of mask_out[query] where the ReLU output is positive;
// MASK multiplies by bit n of mask_in[query] (the ReLU derivative of the forward pass).
template <int M_TILES, int K_STEPS, bool RELU, bool MASK, bool RECORD>
__device__ __forceinline__ void mlp_layer(const __bf16 *__restrict__ W, const __bf16 *Hin, int sin, __bf16 *Hout,
                                          int sout, const uint64_t *mask_in, uint64_t *mask_out, float *out_f32,
                                          int m_valid) {
    const int lane = __lane_id();
    const int r = lane & 15, g = lane >> 4;
    bf16x8 a[M_TILES][K_STEPS];
#pragma unroll
    for (int mt = 0; mt < M_TILES; ++mt)
#pragma unroll
        for (int ks = 0; ks < K_STEPS; ++ks) a[mt][ks] = ld8(W + (size_t)(16 * mt + r) * (32 * K_STEPS) + 32 * ks + 8 * g);
#pragma unroll
    for (int nt = 0; nt < kQ / 16; ++nt) {
        const int q = 16 * nt + r;
        bf16x8 b[K_STEPS];
#pragma unroll
        for (int ks = 0; ks < K_STEPS; ++ks) b[ks] = ld8(Hin + q * sin + 32 * ks + 8 * g);
        const uint64_t min_bits = MASK ? mask_in[q] : 0ull;
        uint64_t bits = 0ull;
#pragma unroll
        for (int mt = 0; mt < M_TILES; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < K_STEPS; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][ks], b[ks], acc, 0, 0, 0);
            // lane holds rows 4g..4g+3 of this tile for query q
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int n = 16 * mt + 4 * g + i;
                float v = acc[i];
                if (RELU) v = fmaxf(v, 0.f);
                if (MASK) v = ((min_bits >> n) & 1ull) ? v : 0.f;
                if (RECORD && v > 0.f) bits |= 1ull << n;
                if (out_f32) {
                    if (n < m_valid) out_f32[q * m_valid + n] = v;
                } else {
                    Hout[q * sout + n] = (__bf16)v;
                }
            }
        }
        if (RECORD) atomicOr((unsigned long long *)&mask_out[q], (unsigned long long)bits);
    }
}

struct TrainDev {
    const float *targets;  // [n][3]
    float *loss;           // accumulated (sum of per-element losses / (n*3))
    bf16x2 *gtable;        // table gradient, BF16 feature pairs
    __bf16 *act_t;         // transposed activations: 5 blocks [feature][n] (X is 32 rows, H 64 rows)
    __bf16 *grad_t;        // transposed output gradients: 5 blocks [neuron][n] (dH0..dH3 64 rows, dY 16 rows)
    uint32_t n_pad;        // column stride of act_t / grad_t
};

// Fused encode + MLP forward (and, when TRAIN, loss and backward to the table) for 64 queries per wave.
// LDS: one activation tile [64][72] BF16 updated in place (each 16-query slice of a layer reads its
// inputs before it writes its outputs, and a wave's LDS operations complete in order), the outputs,
// and (training) the ReLU masks of the four hidden layers as one 64-bit word per query; activations
// and gradients also go to the transposed global copies for the weight-gradient GEMM as soon as they
// are produced. About 10 KB of LDS per wave.
template <bool TRAIN>
__device__ __forceinline__ void cache_tile(const CacheDev &c, const float *__restrict__ points,
                                           const float *__restrict__ attrs, uint32_t n, int points_are_world,
                                           float *__restrict__ out, const TrainDev &td, uint32_t base) {
    __shared__ __bf16 P[kQ * kStrideH];
    __shared__ float Y[kQ * 3];
    __shared__ uint64_t MB[TRAIN ? 4 : 1][kQ];
    const int lane = __lane_id();
    const uint32_t q = base + lane;
    const bool valid = q < n;
    float p[3] = {0.f, 0.f, 0.f};
    if (valid) {
        for (int a = 0; a < 3; ++a) {
            float v = points[3 * (size_t)q + a];
            if (points_are_world) v = (v - c.lo[a]) * c.inv_ext[a];
            p[a] = fminf(fmaxf(v, 0.f), 1.f);
        }
    }
    __bf16 *p0 = P, *p1 = P;  // the same tile: every layer runs in place
    encode_point(c, p[0], p[1], p[2], valid, p0 + lane * kStrideH);
    {
        // the query's surface attributes after the encoding, then zeros up to the MLP's input width
        __bf16 *row = p0 + lane * kStrideH;
#pragma unroll
        for (int k = 0; k < kAttr; ++k) row[kEnc + k] = (__bf16)(valid && attrs ? attrs[(size_t)kAttr * q + k] : 0.f);
#pragma unroll
        for (int k = kEnc + kAttr; k < kIn; ++k) row[k] = (__bf16)0.f;
    }
    if (TRAIN)
        for (int l = 0; l < 4; ++l) MB[l][lane] = 0ull;
    __syncthreads();
    // transposed stores for the weight-gradient GEMMs
    const size_t blk = (size_t)64 * td.n_pad;
    auto store_t = [&](const __bf16 *src, int stride, int rows, __bf16 *dst) {
#pragma unroll 1
        for (int k = 0; k < rows; k += 8) {
            const bf16x8 v = ld8(src + lane * stride + k);
#pragma unroll
            for (int i = 0; i < 8; ++i) dst[(size_t)(k + i) * td.n_pad + base + lane] = v[i];
        }
    };
    // forward: X (p0) -> H0 (p1) -> H1 (p0) -> H2 (p1) -> H3 (p0) -> Y
    if (TRAIN) store_t(p0, kStrideH, kIn, td.act_t);
    mlp_layer<4, kIn / 32, true, false, TRAIN>(c.wf[0], p0, kStrideH, p1, kStrideH, nullptr, MB[0], nullptr, 0);
    __syncthreads();
    if (TRAIN) store_t(p1, kStrideH, 64, td.act_t + blk);
    mlp_layer<4, 2, true, false, TRAIN>(c.wf[1], p1, kStrideH, p0, kStrideH, nullptr, MB[TRAIN ? 1 : 0], nullptr, 0);
    __syncthreads();
    if (TRAIN) store_t(p0, kStrideH, 64, td.act_t + 2 * blk);
    mlp_layer<4, 2, true, false, TRAIN>(c.wf[2], p0, kStrideH, p1, kStrideH, nullptr, MB[TRAIN ? 2 : 0], nullptr, 0);
    __syncthreads();
    if (TRAIN) store_t(p1, kStrideH, 64, td.act_t + 3 * blk);
    mlp_layer<4, 2, true, false, TRAIN>(c.wf[3], p1, kStrideH, p0, kStrideH, nullptr, MB[TRAIN ? 3 : 0], nullptr, 0);
    __syncthreads();
    if (TRAIN) store_t(p0, kStrideH, 64, td.act_t + 4 * blk);
    mlp_layer<1, 2, false, false, false>(c.wf[4], p0, kStrideH, nullptr, 0, nullptr, nullptr, Y, 3);
    __syncthreads();
    if (!TRAIN) {
        if (valid)
            for (int o = 0; o < 3; ++o) out[3 * (size_t)q + o] = Y[lane * 3 + o];
        return;
    }
    // relative L2 loss, d/dy = 2 (y - t) / (y^2 + 0.01) / (n * 3) with the denominator held constant
    const float inv = 1.f / ((float)n * 3.f);
    float lsum = 0.f;
    {
        __bf16 *dy = p1;
        for (int o = 0; o < 32; ++o) dy[lane * kStrideG + o] = (__bf16)0.f;
        if (valid) {
            for (int o = 0; o < 3; ++o) {
                const float y = Y[lane * 3 + o], t = td.targets[3 * (size_t)q + o];
                const float den = y * y + 0.01f;
                lsum += (y - t) * (y - t) / den * inv;
                dy[lane * kStrideG + o] = (__bf16)(2.f * (y - t) / den * inv);
                if (out) out[3 * (size_t)q + o] = y;
            }
        }
    }
    // wave reduction of the loss, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) atomicAdd(td.loss, lsum);
    __syncthreads();
    // backward: dH3 = (W5^T dY) * relu'(H3), dH2, dH1, dH0, dX = W1^T dH0
    store_t(p1, kStrideG, 16, td.grad_t + 4 * blk);  // dY (3 rows used, 16 stored)
    mlp_layer<4, 1, false, true, false>(c.wb[4], p1, kStrideG, p0, kStrideG, MB[3], nullptr, nullptr, 0);  // dH3
    __syncthreads();
    store_t(p0, kStrideG, 64, td.grad_t + 3 * blk);
    mlp_layer<4, 2, false, true, false>(c.wb[3], p0, kStrideG, p1, kStrideG, MB[2], nullptr, nullptr, 0);  // dH2
    __syncthreads();
    store_t(p1, kStrideG, 64, td.grad_t + 2 * blk);
    mlp_layer<4, 2, false, true, false>(c.wb[2], p1, kStrideG, p0, kStrideG, MB[1], nullptr, nullptr, 0);  // dH1
    __syncthreads();
    store_t(p0, kStrideG, 64, td.grad_t + 1 * blk);
    mlp_layer<4, 2, false, true, false>(c.wb[1], p0, kStrideG, p1, kStrideG, MB[0], nullptr, nullptr, 0);  // dH0
    __syncthreads();
    store_t(p1, kStrideG, 64, td.grad_t);
    // dX of the encoding only (the first kEnc inputs; the attributes take no gradient)
    mlp_layer<kEnc / 16, 2, false, false, false>(c.wb[0], p1, kStrideG, p0, kStrideG, nullptr, nullptr, nullptr, 0);
    __syncthreads();
    if (valid && td.gtable) encode_backward(c, p[0], p[1], p[2], p0 + lane * kStrideG, td.gtable);
}

// Persistent over 64-query tiles; the query count comes from n_dev (device) when given, capped at n.
template <bool TRAIN>
__global__ __launch_bounds__(64) void k_cache_mlp(CacheDev c, const float *__restrict__ points,
                                                  const float *__restrict__ attrs, uint32_t n,
                                                  const uint32_t *__restrict__ n_dev, int points_are_world,
                                                  float *__restrict__ out, TrainDev td) {
    const uint32_t count = n_dev ? min(*n_dev, n) : n;
    if (TRAIN && n_dev && count < kMinTrainSamples) return;  // the frame takes no training step
    for (uint32_t tile = blockIdx.x; tile * (uint32_t)kQ < count; tile += gridDim.x) {
        cache_tile<TRAIN>(c, points, attrs, count, points_are_world, out, td, tile * (uint32_t)kQ);
        __syncthreads();
    }
}

// Weight gradients dW_l[out][in] = sum_q dH_l[out][q] * A_{l-1}[in][q] as a K = n GEMM on MFMA.
// grid.x = K chunks of 256 queries, grid.y = layer; one wave per (16-row, 16-col) output tile pair set.
__global__ __launch_bounds__(256) void k_cache_dw(const __bf16 *__restrict__ act_t, const __bf16 *__restrict__ grad_t,
                                                  uint32_t n_pad, uint32_t n, const uint32_t *__restrict__ n_dev,
                                                  float *__restrict__ dw) {
    const uint32_t count = n_dev ? min(*n_dev, n) : n;
    if (n_dev && count < kMinTrainSamples) return;
    // columns up to the last 64-sample tile were written by the tile kernel (zeros past count)
    const uint32_t kmax = (count + (uint32_t)kQ - 1u) / (uint32_t)kQ * (uint32_t)kQ;
    const uint32_t k0 = blockIdx.x * 256u;
    if (k0 >= kmax) return;
    const int layer = blockIdx.y;
    const int rows = layer == 4 ? 3 : 64;
    const int cols = layer == 0 ? kIn : 64;
    const size_t blk = (size_t)64 * n_pad;
    const __bf16 *A = grad_t + layer * blk;  // [rows][n]
    const __bf16 *B = act_t + layer * blk;   // [cols][n]
    size_t woff = 0;
    for (int l = 0; l < layer; ++l) woff += (size_t)(l == 0 ? 64 * kIn : 64 * 64);
    const int wave = threadIdx.x >> 6;
    const int lane = __lane_id();
    const int r = lane & 15, g = lane >> 4;
    const int mtiles = (rows + 15) / 16, ntiles = cols / 16;
    const bf16x8 zero = {};
    for (int tile = wave; tile < mtiles * ntiles; tile += 4) {
        const int mt = tile / ntiles, nt = tile % ntiles;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t ks = 0; ks < 256; ks += 32) {
            const uint32_t kk = k0 + ks + 8 * g;
            const bf16x8 a = kk < kmax ? ld8(A + (size_t)(16 * mt + r) * n_pad + kk) : zero;
            const bf16x8 b = kk < kmax ? ld8(B + (size_t)(16 * nt + r) * n_pad + kk) : zero;
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        // lane holds rows 4g..4g+3 (out) of column r (in) of this tile
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mt + 4 * g + i, in_ = 16 * nt + r;
            if (o < rows) atomicAdd(&dw[woff + (size_t)o * cols + in_], acc[i]);
        }
    }
}

// Adam (Kingma & Ba 2015) over FP32 master parameters, with bias correction; writes the low-precision
// copies used by the kernels.
// One thread: decides whether this step runs (enough samples) and advances the step counter.
__global__ void k_adam_begin(AdamStep *st, const uint32_t *n_dev, uint32_t cap, float b1, float b2) {
    const uint32_t count = n_dev ? min(*n_dev, cap) : cap;
    const bool active = !n_dev || count >= kMinTrainSamples;
    st->active = active ? 1u : 0u;
    if (active) {
        const uint32_t step = st->step + 1u;
        st->step = step;
        st->c1 = 1.f - powf(b1, (float)step);
        st->c2 = 1.f - powf(b2, (float)step);
    }
}

// Adam over the MLP weights (dense: every weight has a gradient every step)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, uint64_t count, float lr, float b1, float b2, float eps,
                       const AdamStep *__restrict__ st) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || !st->active) return;
    const float c1 = st->c1, c2 = st->c2;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
}

// Adam over the hash table, lazily (the rule of torch.optim.SparseAdam): an entry (two features,
// one BF16 gradient pair) whose gradient is zero this step keeps its parameters and moments; a
// touched one is updated and its gradient word cleared for the next step, so the table gradient is
// all zero again after the step without a separate clear. A step touches a fraction of the 16 x 2^19
// entries (few on the coarse levels, where the samples share corners), and an untouched entry costs
// a 4-byte read instead of 28 bytes of parameters, moments and copies.
__global__ void k_adam_table(float2 *__restrict__ p, uint32_t *__restrict__ grad, float2 *__restrict__ m,
                             float2 *__restrict__ v, uint64_t entries, float lr, float b1, float b2, float eps,
                             const AdamStep *__restrict__ st, __half2 *__restrict__ half_copy) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= entries || !st->active) return;
    const uint32_t gw = grad[i];
    if (gw == 0u) return;
    grad[i] = 0u;
    const float c1 = st->c1, c2 = st->c2;
    const float g0 = __uint_as_float(gw << 16), g1 = __uint_as_float(gw & 0xFFFF0000u);
    const float2 mi = m[i], vi = v[i], pi = p[i];
    const float m0 = b1 * mi.x + (1.f - b1) * g0, m1 = b1 * mi.y + (1.f - b1) * g1;
    const float v0 = b2 * vi.x + (1.f - b2) * g0 * g0, v1 = b2 * vi.y + (1.f - b2) * g1 * g1;
    const float p0 = pi.x - lr * (m0 / c1) / (sqrtf(v0 / c2) + eps);
    const float p1 = pi.y - lr * (m1 / c1) / (sqrtf(v1 / c2) + eps);
    m[i] = make_float2(m0, m1);
    v[i] = make_float2(v0, v1);
    p[i] = make_float2(p0, p1);
    half_copy[i] = __floats2half2_rn(p0, p1);
}

// FP32 master weights -> BF16 forward [out_pad][in] and backward [in][out_pad32] copies
__global__ void k_pack_weights(const float *__restrict__ w, __bf16 *__restrict__ wf, __bf16 *__restrict__ wb) {
    // layer geometry: in, out, forward out pad, backward out pad, offsets
    const int ins[5] = {kIn, 64, 64, 64, 64}, outs[5] = {64, 64, 64, 64, 3};
    const int fpad[5] = {64, 64, 64, 64, 16}, bpad[5] = {64, 64, 64, 64, 32};
    int off = 0, foff = 0, boff = 0;
    for (int l = 0; l < 5; ++l) {
        const int total_f = fpad[l] * ins[l];
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total_f; e += gridDim.x * blockDim.x) {
            const int o = e / ins[l], i = e % ins[l];
            wf[foff + e] = (__bf16)(o < outs[l] ? w[off + o * ins[l] + i] : 0.f);
        }
        const int total_b = ins[l] * bpad[l];
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total_b; e += gridDim.x * blockDim.x) {
            const int i = e / bpad[l], o = e % bpad[l];
            wb[boff + e] = (__bf16)(o < outs[l] ? w[off + o * ins[l] + i] : 0.f);
        }
        off += outs[l] * ins[l];
        foff += total_f;
        boff += total_b;
    }
}

// clears n 32-bit words (a kernel rather than hipMemsetAsync, so the training step can sit inside a
// captured frame graph)
__global__ void k_zero_words(uint32_t *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_half_to_f32(const __half *__restrict__ h, float *__restrict__ f, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = __half2float(h[i]);
}
__global__ void k_f32_to_half(const float *__restrict__ f, __half *__restrict__ h, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = __float2half(f[i]);
}

// ---------------------------------------------------------------------------------------------
static const int kLayerIn[5] = {kIn, 64, 64, 64, 64}, kLayerOut[5] = {64, 64, 64, 64, 3};
static const int kFPad[5] = {64, 64, 64, 64, 16}, kBPad[5] = {64, 64, 64, 64, 32};

size_t cache_num_weights() {
    size_t s = 0;
    for (int l = 0; l < 5; ++l) s += (size_t)kLayerIn[l] * kLayerOut[l];
    return s;
}

// BF16 values in each of the packed forward and backward weight copies (the larger of the two)
static size_t cache_packed_bf16() {
    size_t f = 0, b = 0;
    for (int l = 0; l < 5; ++l) {
        f += (size_t)kFPad[l] * kLayerIn[l];
        b += (size_t)kLayerIn[l] * kBPad[l];
    }
    return f > b ? f : b;
}

static CacheDev make_dev(const CacheState &cs) {
    CacheDev d;
    std::memset(&d, 0, sizeof(d));
    d.table = (const __half2 *)cs.table_h.ptr;
    size_t fo = 0, bo = 0;
    for (int l = 0; l < 5; ++l) {
        d.wf[l] = (const __bf16 *)cs.wf.ptr + fo;
        d.wb[l] = (const __bf16 *)cs.wb.ptr + bo;
        fo += (size_t)kFPad[l] * kLayerIn[l];
        bo += (size_t)kLayerIn[l] * kBPad[l];
    }
    for (int l = 0; l < 16; ++l) {
        d.res[l] = cs.res[l];
        d.dense[l] = cs.dense[l];
    }
    d.log2_table = cs.cfg.log2_table;
    for (int a = 0; a < 3; ++a) {
        d.lo[a] = cs.lo[a];
        d.inv_ext[a] = cs.inv_ext[a];
    }
    return d;
}

hipError_t cache_pack(CacheState &cs, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_weights, dim3(16), dim3(256), 0, s, (const float *)cs.w.ptr, (__bf16 *)cs.wf.ptr,
                       (__bf16 *)cs.wb.ptr);
    const uint64_t n = cs.table_params();
    hipLaunchKernelGGL(k_f32_to_half, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                       (const float *)cs.table_f.ptr, (__half *)cs.table_h.ptr, n);
    return hipGetLastError();
}

hipError_t cache_infer(CacheState &cs, const float *points, const float *attrs, uint32_t n, int world, float *out,
                       hipStream_t s) {
    if (n == 0) return hipSuccess;
    TrainDev td;
    std::memset(&td, 0, sizeof(td));
    hipLaunchKernelGGL(k_cache_mlp<false>, dim3((n + kQ - 1) / kQ), dim3(64), 0, s, make_dev(cs), points, attrs, n,
                       (const uint32_t *)nullptr, world, out, td);
    return hipGetLastError();
}

hipError_t cache_infer_queue(CacheState &cs, const float *points, const float *attrs, const uint32_t *count_dev,
                             uint32_t capacity, float *out, int grid, hipStream_t s) {
    TrainDev td;
    std::memset(&td, 0, sizeof(td));
    hipLaunchKernelGGL(k_cache_mlp<false>, dim3(grid), dim3(64), 0, s, make_dev(cs), points, attrs, capacity,
                       count_dev, 1, out, td);
    return hipGetLastError();
}

hipError_t cache_prepare(CacheState &cs, uint32_t n) {
    const uint32_t n_pad = ((n + 255u) / 256u) * 256u;
    hipError_t e = cs.act_t.ensure((size_t)5 * 64 * n_pad * 2);
    if (e == hipSuccess) e = cs.grad_t.ensure((size_t)5 * 64 * n_pad * 2);
    return e;
}

hipError_t cache_gradients(CacheState &cs, const float *points, const float *attrs, const float *targets, uint32_t n,
                           int world, float *out, hipStream_t s, const uint32_t *n_dev) {
    hipError_t e;
    const uint32_t n_pad = ((n + 255u) / 256u) * 256u;
    if ((e = cache_prepare(cs, n)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_zero_words, dim3(64), dim3(256), 0, s, (uint32_t *)cs.gw.ptr, (uint64_t)cache_num_weights());
    if (!cs.gtable_clean)  // the table Adam clears what it consumes; otherwise clear it here
        hipLaunchKernelGGL(k_zero_words, dim3(1024), dim3(256), 0, s, (uint32_t *)cs.gtable.ptr, cs.table_params() / 2);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, (uint32_t *)&cache_state(cs)->loss, (uint64_t)1);
    cs.gtable_clean = n == 0;
    if (n == 0) return hipGetLastError();
    TrainDev td;
    td.targets = targets;
    td.loss = &cache_state(cs)->loss;
    td.gtable = (bf16x2 *)cs.gtable.ptr;
    td.act_t = (__bf16 *)cs.act_t.ptr;
    td.grad_t = (__bf16 *)cs.grad_t.ptr;
    td.n_pad = n_pad;
    hipLaunchKernelGGL(k_cache_mlp<true>, dim3((n + kQ - 1) / kQ), dim3(64), 0, s, make_dev(cs), points, attrs, n,
                       n_dev, world, out, td);
    hipLaunchKernelGGL(k_cache_dw, dim3(n_pad / 256, 5), dim3(256), 0, s, (const __bf16 *)cs.act_t.ptr,
                       (const __bf16 *)cs.grad_t.ptr, n_pad, n, n_dev, (float *)cs.gw.ptr);
    return hipGetLastError();
}

hipError_t cache_adam(CacheState &cs, hipStream_t s, const uint32_t *n_dev, uint32_t cap) {
    const float b1 = cs.cfg_beta1, b2 = cs.cfg_beta2, eps = cs.cfg_eps, lr = cs.cfg.learning_rate;
    AdamStep *st = cache_state(cs);
    hipLaunchKernelGGL(k_adam_begin, dim3(1), dim3(1), 0, s, st, n_dev, cap, b1, b2);
    const uint64_t nw = cache_num_weights();
    hipLaunchKernelGGL(k_adam, dim3((uint32_t)((nw + 255) / 256)), dim3(256), 0, s, (float *)cs.w.ptr,
                       (const float *)cs.gw.ptr, (float *)cs.mw.ptr, (float *)cs.vw.ptr, nw, lr, b1, b2, eps,
                       (const AdamStep *)st);
    const uint64_t entries = cs.table_params() / 2;
    hipLaunchKernelGGL(k_adam_table, dim3((uint32_t)((entries + 255) / 256)), dim3(256), 0, s,
                       (float2 *)cs.table_f.ptr, (uint32_t *)cs.gtable.ptr, (float2 *)cs.mt.ptr, (float2 *)cs.vt.ptr,
                       entries, lr, b1, b2, eps, (const AdamStep *)st, (__half2 *)cs.table_h.ptr);
    // an active step cleared every gradient word it read; an inactive one (fewer than
    // kMinTrainSamples samples through n_dev) followed a gradient pass that wrote nothing
    cs.gtable_clean = true;
    hipLaunchKernelGGL(k_pack_weights, dim3(16), dim3(256), 0, s, (const float *)cs.w.ptr, (__bf16 *)cs.wf.ptr,
                       (__bf16 *)cs.wb.ptr);
    return hipGetLastError();
}

}  // namespace spk

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------
using namespace spk;

namespace {
struct HostRng {
    uint64_t s;
    explicit HostRng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * (1.0 / 16777216.0)); }
};

int cache_check(spk_context *ctx) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->cache) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "no radiance cache (spk_cache_create)");
    return SPK_OK;
}
}  // namespace

extern "C" {

// Initial parameters from the configured seed: table U(-1e-4, 1e-4) (Instant NGP), weights
// Xavier-uniform, Adam moments and step count zero.
static int cache_init_params(spk_context *ctx, CacheState &cs) {
    const uint64_t ntab = cs.table_params();
    const size_t nw = cache_num_weights();
    hipError_t e;
    HostRng rng(cs.cfg.seed);
    std::vector<float> tab(ntab), w(nw);
    for (auto &x : tab) x = rng.uniform(-1e-4f, 1e-4f);
    size_t off = 0;
    for (int l = 0; l < 5; ++l) {
        const int in = kLayerIn[l], out = kLayerOut[l];
        const float lim = std::sqrt(6.f / (float)(in + out));
        for (int k = 0; k < in * out; ++k) w[off + k] = rng.uniform(-lim, lim);
        off += (size_t)in * out;
    }
    if ((e = hipMemcpy(cs.table_f.ptr, tab.data(), ntab * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(cs.w.ptr, w.data(), nw * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(cs.mt.ptr, 0, ntab * 4)) != hipSuccess || (e = hipMemset(cs.vt.ptr, 0, ntab * 4)) != hipSuccess ||
        (e = hipMemset(cs.mw.ptr, 0, nw * 4)) != hipSuccess || (e = hipMemset(cs.vw.ptr, 0, nw * 4)) != hipSuccess ||
        (e = hipMemset(cs.wf.ptr, 0, cache_packed_bf16() * 2)) != hipSuccess ||
        (e = hipMemset(cs.wb.ptr, 0, cache_packed_bf16() * 2)) != hipSuccess ||
        (e = hipMemset(cs.state.ptr, 0, sizeof(AdamStep))) != hipSuccess ||
        (e = hipMemset(cs.gtable.ptr, 0, ntab * 2)) != hipSuccess ||
        (e = hipMemset(cs.st_count.ptr, 0, 4)) != hipSuccess)  // no staged samples to train on
        return hip_fail(ctx, e, "cache initialisation");
    if ((e = cache_pack(cs, ctx->stream)) != hipSuccess || (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return hip_fail(ctx, e, "cache pack");
    cs.gtable_clean = true;
    return SPK_OK;
}

int spk_cache_create(spk_context *ctx, const spk_cache_config *config) {
    if (!ctx || !config) return SPK_ERR_INVALID_ARGUMENT;
    if (!ctx->scene.valid) return set_error(ctx, SPK_ERR_NO_SCENE, "the cache is sized from the uploaded scene box");
    if (config->levels != 16 || config->features != 2)
        return set_error(ctx, SPK_ERR_UNSUPPORTED, "the cache kernels are built for 16 levels x 2 features");
    if (config->log2_table < 10 || config->log2_table > 24)
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "log2_table must be in [10, 24]");
    if (config->base_resolution < 2 || !(config->per_level_scale >= 1.f) || !(config->learning_rate > 0.f))
        return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "invalid cache resolution or learning rate");
    delete ctx->cache;
    ctx->cache = new CacheState();
    ctx->cache_last_loss = 0.f;
    CacheState &cs = *ctx->cache;
    cs.cfg = *config;
    if (cs.cfg.batch == 0) cs.cfg.batch = 1u << 16;
    const uint64_t T = 1ull << cs.cfg.log2_table;
    for (uint32_t l = 0; l < 16; ++l) {
        const double r = std::floor((double)cs.cfg.base_resolution * std::pow((double)cs.cfg.per_level_scale, (double)l));
        cs.res[l] = (uint32_t)std::max(1.0, std::min(r, 65535.0));
        const uint64_t side = (uint64_t)cs.res[l] + 1;
        cs.dense[l] = side * side * side <= T ? 1u : 0u;
    }
    // scene box with a small margin
    const std::vector<float> &v = ctx->host_scene.verts;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < v.size(); i += 3)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], v[i + a]);
            hi[a] = std::max(hi[a], v[i + a]);
        }
    float ext = 0.f;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    for (int a = 0; a < 3; ++a) {
        // a cube around the scene keeps the grid cells isotropic
        const float c = 0.5f * (lo[a] + hi[a]);
        cs.lo[a] = c - 0.51f * ext;
        cs.inv_ext[a] = 1.f / (1.02f * ext);
    }
    const uint64_t ntab = cs.table_params();
    const size_t nw = cache_num_weights();
    hipError_t e;
    if ((e = cs.table_h.ensure(ntab * 2)) != hipSuccess || (e = cs.table_f.ensure(ntab * 4)) != hipSuccess ||
        (e = cs.mt.ensure(ntab * 4)) != hipSuccess || (e = cs.vt.ensure(ntab * 4)) != hipSuccess ||
        (e = cs.gtable.ensure(ntab * 2)) != hipSuccess || (e = cs.w.ensure(nw * 4)) != hipSuccess ||
        (e = cs.mw.ensure(nw * 4)) != hipSuccess || (e = cs.vw.ensure(nw * 4)) != hipSuccess ||
        (e = cs.gw.ensure(nw * 4)) != hipSuccess || (e = cs.wf.ensure(cache_packed_bf16() * 2)) != hipSuccess ||
        (e = cs.wb.ensure(cache_packed_bf16() * 2)) != hipSuccess || (e = cs.state.ensure(sizeof(AdamStep))) != hipSuccess ||
        (e = cs.st_count.ensure(4)) != hipSuccess)
        return hip_fail(ctx, e, "cache allocation");
    int rc = cache_init_params(ctx, cs);
    if (rc != SPK_OK) return rc;
    ctx->on_scene_changed();
    return SPK_OK;
}

int spk_cache_reset(spk_context *ctx) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cache reset");
    ctx->cache_last_loss = 0.f;
    return cache_init_params(ctx, *ctx->cache);
}

int spk_cache_destroy(spk_context *ctx) {
    if (!ctx) return SPK_ERR_INVALID_ARGUMENT;
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx->cache;
    ctx->cache = nullptr;
    ctx->on_scene_changed();
    return SPK_OK;
}

int spk_cache_infer(spk_context *ctx, const float *points_dev, const float *attrs_dev, uint32_t n, int world,
                    float *out_dev) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    if (n && (!points_dev || !out_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null array");
    hipError_t e = cache_infer(*ctx->cache, points_dev, attrs_dev, n, world, out_dev, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cache inference");
    return SPK_OK;
}

int spk_cache_gradients(spk_context *ctx, const float *points_dev, const float *attrs_dev, const float *targets_dev,
                        uint32_t n, int world, float *loss, float *dweights, float *dtable) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    if (n && (!points_dev || !targets_dev)) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null array");
    CacheState &cs = *ctx->cache;
    hipError_t e = cache_gradients(cs, points_dev, attrs_dev, targets_dev, n, world, nullptr, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && loss) e = hipMemcpy(loss, &cache_state(cs)->loss, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && dweights) e = hipMemcpy(dweights, cs.gw.ptr, cache_num_weights() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && dtable) {
        // the table gradient is kept as BF16; widen it for the caller
        std::vector<uint16_t> g(cs.table_params());
        e = hipMemcpy(g.data(), cs.gtable.ptr, g.size() * 2, hipMemcpyDeviceToHost);
        for (size_t k = 0; e == hipSuccess && k < g.size(); ++k) {
            const uint32_t bits = (uint32_t)g[k] << 16;
            std::memcpy(&dtable[k], &bits, 4);
        }
    }
    if (e == hipSuccess) {  // no Adam step consumes this gradient: clear it for the next pass
        hipLaunchKernelGGL(k_zero_words, dim3(1024), dim3(256), 0, ctx->stream, (uint32_t *)cs.gtable.ptr,
                           cs.table_params() / 2);
        e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) cs.gtable_clean = true;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "cache gradients");
    return SPK_OK;
}

int spk_cache_train(spk_context *ctx, const float *points_dev, const float *attrs_dev, const float *targets_dev,
                    uint32_t n, int world, float *loss) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    if (n == 0) return SPK_OK;
    if (!points_dev || !targets_dev) return set_error(ctx, SPK_ERR_INVALID_ARGUMENT, "null array");
    CacheState &cs = *ctx->cache;
    hipError_t e = cache_gradients(cs, points_dev, attrs_dev, targets_dev, n, world, nullptr, ctx->stream);
    if (e == hipSuccess) e = cache_adam(cs, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float l = 0.f;
    if (e == hipSuccess) e = hipMemcpy(&l, &cache_state(*ctx->cache)->loss, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "cache training step");
    ctx->cache_last_loss = l;
    if (loss) *loss = l;
    return SPK_OK;
}

int spk_cache_get_params(spk_context *ctx, float *weights, float *table) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    CacheState &cs = *ctx->cache;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && weights) e = hipMemcpy(weights, cs.w.ptr, cache_num_weights() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && table) e = hipMemcpy(table, cs.table_f.ptr, cs.table_params() * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "cache parameters");
    return SPK_OK;
}

int spk_cache_set_params(spk_context *ctx, const float *weights, const float *table) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    CacheState &cs = *ctx->cache;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && weights) e = hipMemcpy(cs.w.ptr, weights, cache_num_weights() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && table) e = hipMemcpy(cs.table_f.ptr, table, cs.table_params() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = cache_pack(cs, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "cache parameters");
    return SPK_OK;
}

int spk_cache_info(spk_context *ctx, uint64_t *num_weights, uint64_t *num_table_params, uint32_t *steps,
                   float *last_loss) {
    int rc = cache_check(ctx);
    if (rc != SPK_OK) return rc;
    if (num_weights) *num_weights = cache_num_weights();
    if (num_table_params) *num_table_params = ctx->cache->table_params();
    if (steps) {
        AdamStep st{};
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(&st, ctx->cache->state.ptr, sizeof(st), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(ctx, e, "cache state");
        *steps = st.step;
    }
    if (last_loss) *last_loss = ctx->cache_last_loss;
    return SPK_OK;
}

}  // extern "C"
