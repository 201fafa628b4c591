// Guide network kernels: BF16 MFMA forward, backward and Adam for the 4-layer
// 64-wide MLP of guide.hpp, FP32 accumulation.
//
// One wave owns a tile of 64 samples.  Activations live in LDS as bf16
// [sample][feature] rows (stride 68 so the 16 rows a fragment touches fall in
// different banks).  A layer is three loops of v_mfma_f32_16x16x16_bf16:
//   forward   Y[s][o]   = sum_i X[s][i] W[o][i]      (A = X rows, B = W rows)
//   backward  D'[s][i]  = sum_o D[s][o] W[o][i]      (A = D rows, B = W^T rows)
//   gradient  dW[o][i] += sum_s D[s][o] X[s][i]      (A, B = LDS columns)
// Weights are kept as fp32 masters (with biases and Adam moments) and as bf16
// copies in both [out][in] and [in][out] order, so every B fragment is one
// 8-byte load.
#include <algorithm>

#include "guide.hpp"
#include "guide_net.hpp"
#include "launch.hpp"

namespace pt {

typedef short bf16x4 __attribute__((ext_vector_type(4)));

constexpr int kLd = 68;                 // bf16 row stride of an LDS tile
constexpr int kLdOut = 65;              // fp32 row stride of the output tile

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);     // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 row4(const uint16_t* t, int row, int col) {
    return *reinterpret_cast<const bf16x4*>(t + row * kLd + col);
}
__device__ __forceinline__ bf16x4 col4(const uint16_t* t, int row0, int col) {
    bf16x4 v;
    v[0] = (short)t[(row0 + 0) * kLd + col];
    v[1] = (short)t[(row0 + 1) * kLd + col];
    v[2] = (short)t[(row0 + 2) * kLd + col];
    v[3] = (short)t[(row0 + 3) * kLd + col];
    return v;
}
__device__ __forceinline__ bf16x4 glb4(const uint16_t* w, int row, int col) {
    return *reinterpret_cast<const bf16x4*>(w + row * kGuideDim + col);
}

// The training kernel splits each 64-sample tile over four waves: wave w
// computes the 16 result columns 16 w .. 16 w + 15 of every product.

// acc[mt] <- columns 16 nt .. 16 nt + 15 of A . B^T, A a [64][64] LDS tile and
// B a [64][64] global matrix stored row-major (B row n = output n).  Lane l
// holds element (16 mt + 4 (l >> 4) + j, 16 nt + (l & 15)) in acc[mt][j].
__device__ __forceinline__ void mm_abt(const uint16_t* a, const uint16_t* b, int nt, f32x4 acc[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        const bf16x4 bf = glb4(b, 16 * nt + r, 16 * kt + 4 * kq);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 af = row4(a, 16 * mt + r, 16 * kt + 4 * kq);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf, acc[mt], 0, 0, 0);
        }
    }
}

// acc[mt] <- columns 16 nt .. of D^T . X over the 64 samples (LDS tiles).
__device__ __forceinline__ void mm_dtx(const uint16_t* d, const uint16_t* x, int nt, f32x4 acc[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        const bf16x4 xf = col4(x, 16 * kt + 4 * kq, 16 * nt + r);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 df = col4(d, 16 * kt + 4 * kq, 16 * mt + r);
            acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(df, xf, acc[mt], 0, 0, 0);
        }
    }
}

// Hidden layer, columns 16 nt ..: out = relu(in W^T + b) as bf16.
__device__ __forceinline__ void layer_hidden(const uint16_t* in, uint16_t* out, const uint16_t* w, const float* bias,
                                             int nt) {
    f32x4 acc[4];
    mm_abt(in, w, nt, acc);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    const int o = 16 * nt + r;
    const float b = bias[o];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) out[(16 * mt + 4 * kq + j) * kLd + o] = to_bf16(fmaxf(acc[mt][j] + b, 0.0f));
}

// Output layer, columns 16 nt ..: fp32 [sample][kLdOut].
__device__ __forceinline__ void layer_output(const uint16_t* in, float* out, const uint16_t* w, const float* bias,
                                             int nt) {
    f32x4 acc[4];
    mm_abt(in, w, nt, acc);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    const int o = 16 * nt + r;
    const float b = bias[o];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) out[(16 * mt + 4 * kq + j) * kLdOut + o] = acc[mt][j] + b;
}

// Writes the features of this lane's sample as one bf16 row (zeros past 45).
__device__ __forceinline__ void store_features(uint16_t* x, int s, const float f[kGuideFeatures], bool valid) {
    for (int i = 0; i < kGuideDim; i++) x[s * kLd + i] = to_bf16(valid && i < kGuideFeatures ? f[i] : 0.0f);
}

struct GuideNet {
    const uint16_t* w;      // bf16 [layer][out][in]
    const uint16_t* wt;     // bf16 [layer][in][out]
    const float* w32;       // fp32 masters: [layer] { W[out][in], b[out] }
};

__device__ __forceinline__ const float* bias_of(const GuideNet& n, int l) {
    return n.w32 + l * kGuideLayerParams + kGuideDim * kGuideDim;
}

// Raw outputs of given inputs (hook for tests), in the network's output order.
__global__ __launch_bounds__(64 * kForwardWaves) void k_guide_forward(const uint16_t* wf, const float* bf,
                                                                      const float* features, uint32_t n, float* out) {
    __shared__ InferLds<kForwardWaves, 4> lds;
    bf16x8* wsh = lds.w;
    float* bsh = lds.b;
    bf16x8 (*xs)[2 * 4 * 64] = lds.x;
    stage_weights(wsh, bsh, wf, bf, InferLds<kForwardWaves, 4>::kWeights);
    const int wave = threadIdx.x >> 6, s = threadIdx.x & 63, r = s & 15, kq = s >> 4;
    const uint32_t tiles = (n + 63) / 64, stride = gridDim.x * kForwardWaves;
    for (uint32_t t = blockIdx.x * kForwardWaves + wave; t < tiles; t += stride) {
        const uint32_t idx = t * 64u + (uint32_t)s;
        float f[kGuideDim];
#pragma unroll
        for (int i = 0; i < kGuideDim; i++) f[i] = idx < n ? features[(size_t)idx * kGuideDim + i] : 0.0f;
        store_input_frags(xs[wave], s, f);
        wave_lds_sync();
        f32x4 o[4][4];
        tile_forward<4>(wsh, bsh, xs[wave], o);
        wave_lds_sync();
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const uint32_t e = t * 64u + 16u * nt + (uint32_t)r;
                if (e < n)
#pragma unroll
                    for (int j = 0; j < 4; j++) out[(size_t)e * kGuideDim + guide_out_row(16 * mt + 4 * kq + j)] = o[mt][nt][j];
            }
    }
}

// Inputs given as rows of 64 floats (hook for tests).
struct InferFeatureSrc {
    const float* x;
    uint32_t n;
    __device__ __forceinline__ uint32_t count() const { return n; }
    __device__ __forceinline__ void features(uint32_t e, float f[kGuideDim]) const {
#pragma unroll
        for (int k = 0; k < kGuideDim; k++) f[k] = x[(size_t)e * kGuideDim + k];
    }
};

// Packed mixtures (guide.hpp) of every input of `src`, word k of entry e at
// mix[k * ms + e].
template <class Src>
__global__ __launch_bounds__(64 * kInferWaves) void k_guide_infer(Src src, const uint16_t* wf, const float* bf,
                                                                 uint32_t* mix, uint32_t ms) {
    __shared__ InferLds<kInferWaves, 3> lds;
    bf16x8* wsh = lds.w;
    float* bsh = lds.b;
    bf16x8 (*xs)[2 * 4 * 64] = lds.x;
    stage_weights(wsh, bsh, wf, bf, InferLds<kInferWaves, 3>::kWeights);
    const uint32_t total = src.count();
    const int wave = threadIdx.x >> 6, s = threadIdx.x & 63, r = s & 15, kq = s >> 4;
    const uint32_t tiles = (total + 63) / 64, stride = gridDim.x * kInferWaves;
    for (uint32_t t = blockIdx.x * kInferWaves + wave; t < tiles; t += stride) {
        const uint32_t e = t * 64u + (uint32_t)s;
        float f[kGuideDim];
#pragma unroll
        for (int i = 0; i < kGuideDim; i++) f[i] = 0.0f;
        if (e < total) src.features(e, f);
        store_input_frags(xs[wave], s, f);
        wave_lds_sync();
        f32x4 o[4][4];
        tile_forward<3>(wsh, bsh, xs[wave], o);
        wave_lds_sync();
        mixture_words(o, [&](int k, uint32_t sidx, uint32_t v) {
            const uint32_t ee = t * 64u + sidx;
            if (ee < total) mix[(size_t)k * ms + ee] = v;
        });
    }
}

// ---------------------------------------------------------------- training

// Index of batch sample b among `count` records: a 32-bit integer hash
// (Wellons' lowbias32) of the step seed and b.
__device__ __forceinline__ uint32_t pick_record(uint32_t seed, uint32_t b, uint32_t count) {
    uint32_t x = seed * 0x9E3779B9u ^ (b + 0x7F4A7C15u);
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x % count;
}

PT_HD float guide_weight(const GuideRecord& r, float w_max) {
    const float lum = 0.2126f * r.li[0] + 0.7152f * r.li[1] + 0.0722f * r.li[2];
    const float w = r.pdf > 0.0f ? lum * r.cos_i / r.pdf : 0.0f;
    return fminf(fmaxf(w, 0.0f), w_max);
}

// Gradient of sum_b -w_b log q(wi_b) over `batch` records (picked by hash
// when `seed` != 0, else records [0, batch)): block k writes the gradient of
// its 64 records to part[k][kGuideParams] (no atomics: the sum over blocks is
// taken by the Adam step); the loss is summed into *loss.
constexpr int kTrainWaves = 4;       // waves sharing one 64-record tile

__global__ __launch_bounds__(64 * kTrainWaves) void k_guide_train(GuideNet net, const GuideRecord* recs, uint32_t count,
                                                                 uint32_t batch, uint32_t seed, GuideConfig g,
                                                                 float w_max, float* part, float* loss) {
    __shared__ uint16_t x0[64 * kLd], h1[64 * kLd], h2[64 * kLd], h3[64 * kLd], d4[64 * kLd];
    __shared__ float o32[64 * kLdOut];
    const int wave = threadIdx.x >> 6, s = threadIdx.x & 63;
    const uint32_t b = blockIdx.x * 64u + (uint32_t)s;
    const bool valid = b < batch && count > 0;
    // wave 0 owns the per-record work (features, loss), every wave a quarter of
    // each product
    GuideRecord rec{};
    if (wave == 0) {
        float f[kGuideFeatures];
        if (valid) {
            rec = recs[seed ? pick_record(seed, b, count) : b];
            guide_features(g, v3(rec.p[0], rec.p[1], rec.p[2]), v3(rec.n[0], rec.n[1], rec.n[2]),
                           v3(rec.wo[0], rec.wo[1], rec.wo[2]), f);
        }
        store_features(x0, s, f, valid);
    }
    __syncthreads();
    layer_hidden(x0, h1, net.w, bias_of(net, 0), wave);
    __syncthreads();
    layer_hidden(h1, h2, net.w + 1 * kGuideDim * kGuideDim, bias_of(net, 1), wave);
    __syncthreads();
    layer_hidden(h2, h3, net.w + 2 * kGuideDim * kGuideDim, bias_of(net, 2), wave);
    __syncthreads();
    layer_output(h3, o32, net.w + 3 * kGuideDim * kGuideDim, bias_of(net, 3), wave);
    __syncthreads();

    if (wave == 0) {
        // d loss / d outputs of this lane's record
        float gout[kGuideOut];
        float lsum = 0.0f;
        if (valid) {
            float o[kGuideOut];
            for (int k = 0; k < kGuideOut; k++) o[k] = o32[s * kLdOut + k];
            const float w = guide_weight(rec, w_max);
            const float logq = guide_loss_grad(o, v3(rec.wi[0], rec.wi[1], rec.wi[2]), w, gout);
            lsum = -w * logq;
        } else {
            for (int k = 0; k < kGuideOut; k++) gout[k] = 0.0f;
        }
        for (int k = 0; k < kGuideDim; k++) d4[s * kLd + k] = to_bf16(k < kGuideOut ? gout[k] : 0.0f);
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (s == 0) atomicAdd(loss, lsum);
    }
    __syncthreads();

    const int r = s & 15, kq = s >> 4, col = 16 * wave + r;
    uint16_t* dl = d4;                         // delta of the current layer's outputs
#pragma unroll
    for (int l = kGuideLayers - 1; l >= 0; l--) {
        uint16_t* act = l == 0 ? x0 : l == 1 ? h1 : l == 2 ? h2 : h3;   // input of layer l
        float* gw = part + (size_t)blockIdx.x * kGuideParams + l * kGuideLayerParams;
        f32x4 acc[4];
        mm_dtx(dl, act, wave, acc);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int j = 0; j < 4; j++) gw[(16 * mt + 4 * kq + j) * kGuideDim + col] = acc[mt][j];
        // bias gradient of output `col`: the four lane groups sum 16 records each
        float db = 0.0f;
#pragma unroll
        for (int t = 0; t < 16; t++) db += from_bf16(dl[(16 * kq + t) * kLd + col]);
        db += __shfl_xor(db, 16);
        db += __shfl_xor(db, 32);
        if (kq == 0) gw[kGuideDim * kGuideDim + col] = db;
        __syncthreads();
        if (l == 0) break;
        // delta of layer l-1 outputs = (dl W_l) * relu'(act), written over act
        mm_abt(dl, net.wt + l * kGuideDim * kGuideDim, wave, acc);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int at = (16 * mt + 4 * kq + j) * kLd + col;
                act[at] = to_bf16(from_bf16(act[at]) > 0.0f ? acc[mt][j] : 0.0f);
            }
        __syncthreads();
        dl = act;
    }
}

// Adam (Kingma and Ba 2015) on the fp32 masters, then the bf16 copies.
// Writes parameter k of value x into the bf16 copies (weights) or the
// inference bias table.
__device__ __forceinline__ void guide_copies(uint32_t k, float x, uint16_t* w, uint16_t* wt, uint16_t* wf, float* bf) {
    const uint32_t l = k / kGuideLayerParams, off = k % kGuideLayerParams;
    const bool last = l + 1 == (uint32_t)kGuideLayers;
    if (off < (uint32_t)(kGuideDim * kGuideDim)) {
        const uint32_t o = off / kGuideDim, i = off % kGuideDim;
        const uint16_t h = to_bf16(x);
        w[l * kGuideDim * kGuideDim + o * kGuideDim + i] = h;
        wt[l * kGuideDim * kGuideDim + i * kGuideDim + o] = h;
        wf[guide_frag_index((int)l, last ? guide_out_slot((int)o) : (int)o, (int)i)] = h;
    } else {
        const uint32_t o = off - kGuideDim * kGuideDim;
        bf[l * kGuideDim + (last ? (uint32_t)guide_out_slot((int)o) : o)] = x;
    }
}

// Sum of the partial gradients of parameters [64 blockIdx.x, +64): the 16
// waves of the block sum every 16th partial (coalesced 256-byte rows, eight
// loads in flight per lane), then combine through LDS in a fixed order.
// Returns the sum in the first wave (other waves: 0).
constexpr int kSumBlock = 1024;
__device__ __forceinline__ float sum_parts(const float* part, uint32_t parts, uint32_t k) {
    __shared__ float red[kSumBlock];
    const uint32_t wave = threadIdx.x >> 6;
    constexpr uint32_t kStep = kSumBlock / 64, kUnroll = 8;
    float acc[kUnroll] = {};
    if (k < (uint32_t)kGuideParams) {
        uint32_t b = wave;
        for (; b + (kUnroll - 1) * kStep < parts; b += kUnroll * kStep)
#pragma unroll
            for (uint32_t u = 0; u < kUnroll; u++) acc[u] += part[(size_t)(b + u * kStep) * kGuideParams + k];
        for (; b < parts; b += kStep) acc[0] += part[(size_t)b * kGuideParams + k];
    }
    float s = 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < kUnroll; u++) s += acc[u];
    red[threadIdx.x] = s;
    __syncthreads();
    float t = 0.0f;
    if (wave == 0)
#pragma unroll
        for (uint32_t w = 0; w < kStep; w++) t += red[threadIdx.x + 64 * w];
    return t;
}

// Gradient = sum of the per-block partial gradients (hook for tests).
__global__ __launch_bounds__(kSumBlock) void k_guide_sum(const float* part, uint32_t parts, float* grad) {
    const uint32_t k = blockIdx.x * 64u + (threadIdx.x & 63u);
    const float g = sum_parts(part, parts, k);
    if (threadIdx.x < 64u && k < (uint32_t)kGuideParams) grad[k] = g;
}

__global__ __launch_bounds__(kSumBlock) void k_guide_adam(float* w32, const float* part, uint32_t parts, float* m,
                                                          float* v, uint16_t* w, uint16_t* wt, uint16_t* wf, float* bf,
                                                          float lr, float beta1, float beta2, float eps, float c1,
                                                          float c2, float grad_scale) {
    const uint32_t k = blockIdx.x * 64u + (threadIdx.x & 63u);
    const float gsum = sum_parts(part, parts, k);
    if (threadIdx.x >= 64u || k >= (uint32_t)kGuideParams) return;
    const float gk = gsum * grad_scale;
    const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
    const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w32[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w32[k] = wk;
    guide_copies(k, wk, w, wt, wf, bf);
}

// bf16 copies of the masters (after an upload).
__global__ void k_guide_sync(const float* w32, uint16_t* w, uint16_t* wt, uint16_t* wf, float* bf) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint32_t)kGuideParams) return;
    guide_copies(k, w32[k], w, wt, wf, bf);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_guide_sync(hipStream_t s, const GuideDev& g) {
    hipLaunchKernelGGL(k_guide_sync, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g.w32, g.w, g.wt,
                       g.wf, g.bf);
    return hipGetLastError();
}

hipError_t launch_guide_forward(hipStream_t s, const GuideDev& g, const float* features, uint32_t n, float* out) {
    if (n == 0) return hipSuccess;
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t grid = std::max(1u, std::min((tiles + kForwardWaves - 1) / kForwardWaves, 1024u));
    hipLaunchKernelGGL(k_guide_forward, dim3(grid), dim3(64 * kForwardWaves), 0, s, g.wf, g.bf, features, n, out);
    return hipGetLastError();
}

hipError_t launch_guide_mixture(hipStream_t s, const GuideDev& g, const float* features, uint32_t n, uint32_t* words) {
    if (n == 0) return hipSuccess;
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t grid = std::max(1u, std::min((tiles + kInferWaves - 1) / kInferWaves, 1024u));
    hipLaunchKernelGGL(k_guide_infer<InferFeatureSrc>, dim3(grid), dim3(64 * kInferWaves), 0, s, InferFeatureSrc{features, n},
                       g.wf, g.bf, words, n);
    return hipGetLastError();
}

hipError_t launch_guide_train(hipStream_t s, const GuideDev& g, const GuideRecord* recs, uint32_t count,
                              uint32_t batch, uint32_t seed, float w_max, float* part, float* loss) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_train, dim3(guide_parts(batch)), dim3(64 * kTrainWaves), 0, s, GuideNet{g.w, g.wt, g.w32}, recs,
                       count, batch, seed, g.cfg, w_max, part, loss);
    return hipGetLastError();
}

hipError_t launch_guide_sum(hipStream_t s, const float* part, uint32_t parts, float* grad) {
    hipLaunchKernelGGL(k_guide_sum, dim3((kGuideParams + 63) / 64), dim3(kSumBlock), 0, s, part, parts, grad);
    return hipGetLastError();
}

hipError_t launch_guide_adam(hipStream_t s, const GuideDev& g, const float* part, uint32_t parts, float lr,
                             uint32_t step, float grad_scale) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 63) / 64), dim3(kSumBlock), 0, s, g.w32, part, parts, g.m, g.v,
                       g.w, g.wt, g.wf, g.bf, lr, beta1, beta2, eps, c1, c2, grad_scale);
    return hipGetLastError();
}

}  // namespace pt
