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
#include "guide.hpp"
#include "launch.hpp"

namespace pt {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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

// acc[mt][nt] <- A . B^T with A a [64][64] LDS tile and B a [64][64] global
// matrix stored row-major (B row n = output n).  Lane l holds element
// (16 mt + 4 (l >> 4) + j, 16 nt + (l & 15)) of the result in acc[mt][nt][j].
__device__ __forceinline__ void mm_abt(const uint16_t* a, const uint16_t* b, f32x4 acc[4][4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        bf16x4 bf[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) bf[nt] = glb4(b, 16 * nt + r, 16 * kt + 4 * kq);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 af = row4(a, 16 * mt + r, 16 * kt + 4 * kq);
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf[nt], acc[mt][nt], 0, 0, 0);
        }
    }
}

// acc[mt][nt] <- D^T . X over the 64 samples: D and X are LDS tiles.
__device__ __forceinline__ void mm_dtx(const uint16_t* d, const uint16_t* x, f32x4 acc[4][4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        bf16x4 xf[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) xf[nt] = col4(x, 16 * kt + 4 * kq, 16 * nt + r);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 df = col4(d, 16 * kt + 4 * kq, 16 * mt + r);
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(df, xf[nt], acc[mt][nt], 0, 0, 0);
        }
    }
}

// Hidden layer: out = relu(in W^T + b) as bf16.
__device__ __forceinline__ void layer_hidden(const uint16_t* in, uint16_t* out, const uint16_t* w, const float* bias) {
    f32x4 acc[4][4];
    mm_abt(in, w, acc);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int o = 16 * nt + r;
            const float b = bias[o];
#pragma unroll
            for (int j = 0; j < 4; j++) out[(16 * mt + 4 * kq + j) * kLd + o] = to_bf16(fmaxf(acc[mt][nt][j] + b, 0.0f));
        }
}

// Output layer: fp32 [sample][kLdOut].
__device__ __forceinline__ void layer_output(const uint16_t* in, float* out, const uint16_t* w, const float* bias) {
    f32x4 acc[4][4];
    mm_abt(in, w, acc);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int o = 16 * nt + r;
            const float b = bias[o];
#pragma unroll
            for (int j = 0; j < 4; j++) out[(16 * mt + 4 * kq + j) * kLdOut + o] = acc[mt][nt][j] + b;
        }
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

// ---------------------------------------------------------------- inference forward
//
// Inference computes each layer transposed, D = W . X^T (A = W rows from
// global memory, B = X rows from LDS), so that lane l holds outputs
// (16 mt + 4 (l >> 4) + j) of sample 16 nt + (l & 15): four consecutive
// features of one sample, stored with one 8-byte LDS write.  The two bf16
// ping-pong tiles double as the fp32 output tile (17 KB of LDS per wave).

typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void mm_wxt(const uint16_t* x, const uint16_t* w, f32x4 acc[4][4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
        bf16x4 wf[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) wf[mt] = glb4(w, 16 * mt + r, 16 * kt + 4 * kq);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const bf16x4 xf = row4(x, 16 * nt + r, 16 * kt + 4 * kq);
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf[mt], xf, acc[mt][nt], 0, 0, 0);
        }
    }
}

__device__ __forceinline__ void infer_hidden(const uint16_t* in, uint16_t* out, const uint16_t* w, const float* bias) {
    f32x4 acc[4][4];
    mm_wxt(in, w, acc);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const int o = 16 * mt + 4 * kq;
        const float4 b = *reinterpret_cast<const float4*>(bias + o);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            u16x4 v;
            v[0] = to_bf16(fmaxf(acc[mt][nt][0] + b.x, 0.0f));
            v[1] = to_bf16(fmaxf(acc[mt][nt][1] + b.y, 0.0f));
            v[2] = to_bf16(fmaxf(acc[mt][nt][2] + b.z, 0.0f));
            v[3] = to_bf16(fmaxf(acc[mt][nt][3] + b.w, 0.0f));
            *reinterpret_cast<u16x4*>(out + (16 * nt + r) * kLd + o) = v;
        }
    }
}

constexpr int kLdInfer = 68;            // fp32 row stride of the inference output tile

// Four layers on the tile in buf (bf16 [64][kLd] features in the first half);
// leaves the fp32 outputs as [sample][kLdInfer] over the whole buffer.
__device__ __forceinline__ void infer_forward(uint16_t* buf, const GuideNet& net) {
    uint16_t* xa = buf;
    uint16_t* xb = buf + 64 * kLd;
    __syncthreads();
    infer_hidden(xa, xb, net.w, bias_of(net, 0));
    __syncthreads();
    infer_hidden(xb, xa, net.w + 1 * kGuideDim * kGuideDim, bias_of(net, 1));
    __syncthreads();
    infer_hidden(xa, xb, net.w + 2 * kGuideDim * kGuideDim, bias_of(net, 2));
    __syncthreads();
    f32x4 acc[4][4];
    mm_wxt(xb, net.w + 3 * kGuideDim * kGuideDim, acc);
    __syncthreads();   // every read of xb is done: the outputs overwrite both tiles
    float* o32 = reinterpret_cast<float*>(buf);
    const float* bias = bias_of(net, 3);
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const int o = 16 * mt + 4 * kq;
        const float4 b = *reinterpret_cast<const float4*>(bias + o);
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
            *reinterpret_cast<float4*>(o32 + (16 * nt + r) * kLdInfer + o) =
                make_float4(acc[mt][nt][0] + b.x, acc[mt][nt][1] + b.y, acc[mt][nt][2] + b.z, acc[mt][nt][3] + b.w);
    }
    __syncthreads();
}

// One bf16 feature row (zeros past the features) with 8-byte stores.
__device__ __forceinline__ void store_row(uint16_t* x, int s, const float f[kGuideFeatures], bool valid) {
#pragma unroll
    for (int i = 0; i < kGuideDim; i += 4) {
        u16x4 v;
#pragma unroll
        for (int c = 0; c < 4; c++) v[c] = to_bf16(valid && i + c < kGuideFeatures ? f[i + c] : 0.0f);
        *reinterpret_cast<u16x4*>(x + s * kLd + i) = v;
    }
}

constexpr int kInferLds = 2 * 64 * kLd;   // uint16 elements: two bf16 tiles = one fp32 [64][68] tile
static_assert(kInferLds * 2 == 64 * kLdInfer * 4, "the output tile fits the two feature tiles");

// Raw outputs of given inputs (hook for tests).
__global__ __launch_bounds__(64) void k_guide_forward(GuideNet net, const float* features, uint32_t n, float* out) {
    __shared__ __attribute__((aligned(16))) uint16_t buf[kInferLds];
    const int s = threadIdx.x;
    const uint32_t idx = blockIdx.x * 64u + (uint32_t)s;
    const bool valid = idx < n;
    float f[kGuideFeatures];
    for (int i = 0; i < kGuideFeatures; i++) f[i] = valid ? features[(size_t)idx * kGuideDim + i] : 0.0f;
    store_row(buf, s, f, valid);
    // inputs past the features are part of the hook's contract too
    for (int i = kGuideFeatures; i < kGuideDim; i++) buf[s * kLd + i] = to_bf16(valid ? features[(size_t)idx * kGuideDim + i] : 0.0f);
    infer_forward(buf, net);
    const float* o32 = reinterpret_cast<const float*>(buf);
    if (valid)
        for (int o = 0; o < kGuideDim; o++) out[(size_t)idx * kGuideDim + o] = o32[s * kLdInfer + o];
}

// Mixture of every diffuse and GGX queue entry of the current bounce, in the
// shade kernel's virtual order [diffuse | ggx].
__global__ __launch_bounds__(64) void k_guide_infer(PathState p, SceneDev sc, GuideNet net, int cur) {
    __shared__ __attribute__((aligned(16))) uint16_t buf[kInferLds];
    const uint32_t c0 = p.ctr->queue[0], c1 = p.ctr->queue[1];
    const uint32_t total = c0 + c1;
    const GuideConfig g = p.guide.cfg;
    const int s = threadIdx.x;
    for (uint32_t base = blockIdx.x * 64u; base < total; base += gridDim.x * 64u) {
        const uint32_t e = base + (uint32_t)s;
        const bool valid = e < total;
        uint16_t* row = buf + s * kLd;
#pragma unroll
        for (int k = 0; k < kGuideDim; k += 4) *reinterpret_cast<u16x4*>(row + k) = u16x4{0, 0, 0, 0};
        if (valid) {
            const uint32_t i = e < c0 ? p.q[0][e] : p.q[1][e - c0];
            const float4 o4 = (cur ? p.ray_o[1] : p.ray_o[0])[i];
            const float4 d4 = (cur ? p.ray_d[1] : p.ray_d[0])[i];
            const float4 hh = p.hit[i];
            const V3 d = xyz(d4);
            const V3 x = xyz(o4) + d * hh.x;
            const float* gn = sc.shade[__float_as_uint(hh.w)].gn;
            V3 n = v3(gn[0], gn[1], gn[2]);
            if (dot(n, d) > 0.0f) n = -n;
            guide_features_each(g, x, n, -d, [&](int k, float v) { row[k] = to_bf16(v); });
        }
        // re-read the weights from cache every tile: hoisting them out of the
        // loop takes 128 registers and leaves one wave per SIMD
        uint32_t opaque = 0;
        asm volatile("" : "+s"(opaque));
        GuideNet tile_net = net;
        tile_net.w += opaque;
        tile_net.w32 += opaque;
        infer_forward(buf, tile_net);
        if (valid) {
            // lobe by lobe from the output row straight to the mixture record
            const float* o = reinterpret_cast<const float*>(buf) + s * kLdInfer;
            float mx, inv_sum;
            guide_softmax(o, mx, inv_sum);
            float* m = p.guide.mix + e;
            const size_t st = p.guide.mix_stride;
            for (int j = 0; j < kGuideLobes; j++) {
                float mu[3], kappa, pi, c;
                guide_lobe(o + 5 * j, mx, inv_sum, mu, kappa, pi, c);
                m[(3 * j + 0) * st] = mu[0];
                m[(3 * j + 1) * st] = mu[1];
                m[(3 * j + 2) * st] = mu[2];
                m[(24 + j) * st] = kappa;
                m[(32 + j) * st] = pi;
                m[(40 + j) * st] = c;
            }
        }
        __syncthreads();
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
// when `seed` != 0, else records [0, batch)), added to grad; loss summed into
// *loss.
__global__ __launch_bounds__(64) void k_guide_train(GuideNet net, const GuideRecord* recs, uint32_t count,
                                                   uint32_t batch, uint32_t seed, GuideConfig g, float w_max,
                                                   float* grad, float* loss) {
    __shared__ uint16_t x0[64 * kLd], h1[64 * kLd], h2[64 * kLd], h3[64 * kLd], d4[64 * kLd];
    __shared__ float o32[64 * kLdOut];
    const int s = threadIdx.x;
    const uint32_t b = blockIdx.x * 64u + (uint32_t)s;
    const bool valid = b < batch && count > 0;
    GuideRecord rec{};
    float f[kGuideFeatures];
    if (valid) {
        rec = recs[seed ? pick_record(seed, b, count) : b];
        guide_features(g, v3(rec.p[0], rec.p[1], rec.p[2]), v3(rec.n[0], rec.n[1], rec.n[2]),
                       v3(rec.wo[0], rec.wo[1], rec.wo[2]), f);
    }
    store_features(x0, s, f, valid);
    __syncthreads();
    layer_hidden(x0, h1, net.w, bias_of(net, 0));
    __syncthreads();
    layer_hidden(h1, h2, net.w + 1 * kGuideDim * kGuideDim, bias_of(net, 1));
    __syncthreads();
    layer_hidden(h2, h3, net.w + 2 * kGuideDim * kGuideDim, bias_of(net, 2));
    __syncthreads();
    layer_output(h3, o32, net.w + 3 * kGuideDim * kGuideDim, bias_of(net, 3));
    __syncthreads();

    // d loss / d outputs of this lane's sample
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
    __syncthreads();

    const int lane = s, r = lane & 15, kq = lane >> 4;
    uint16_t* dl = d4;                         // delta of the current layer's outputs
#pragma unroll
    for (int l = kGuideLayers - 1; l >= 0; l--) {
        uint16_t* act = l == 0 ? x0 : l == 1 ? h1 : l == 2 ? h2 : h3;   // input of layer l
        float* gw = grad + l * kGuideLayerParams;
        f32x4 acc[4][4];
        mm_dtx(dl, act, acc);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    atomicAdd(gw + (16 * mt + 4 * kq + j) * kGuideDim + 16 * nt + r, acc[mt][nt][j]);
        float db = 0.0f;
        for (int t = 0; t < 64; t++) db += from_bf16(dl[t * kLd + lane]);
        atomicAdd(gw + kGuideDim * kGuideDim + lane, db);
        __syncthreads();
        if (l == 0) break;
        // delta of layer l-1 outputs = (dl W_l) * relu'(act), written over act
        mm_abt(dl, net.wt + l * kGuideDim * kGuideDim, acc);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int at = (16 * mt + 4 * kq + j) * kLd + 16 * nt + r;
                    act[at] = to_bf16(from_bf16(act[at]) > 0.0f ? acc[mt][nt][j] : 0.0f);
                }
        __syncthreads();
        dl = act;
    }
}

// Adam (Kingma and Ba 2015) on the fp32 masters, then the bf16 copies.
__global__ void k_guide_adam(float* w32, float* grad, float* m, float* v, uint16_t* w, uint16_t* wt, float lr,
                             float beta1, float beta2, float eps, float c1, float c2, float grad_scale) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint32_t)kGuideParams) return;
    const float gk = grad[k] * grad_scale;
    const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
    const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w32[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w32[k] = wk;
    grad[k] = 0.0f;
    const uint32_t l = k / kGuideLayerParams, off = k % kGuideLayerParams;
    if (off < (uint32_t)(kGuideDim * kGuideDim)) {
        const uint32_t o = off / kGuideDim, i = off % kGuideDim;
        const uint16_t h = to_bf16(wk);
        w[l * kGuideDim * kGuideDim + o * kGuideDim + i] = h;
        wt[l * kGuideDim * kGuideDim + i * kGuideDim + o] = h;
    }
}

// bf16 copies of the masters (after an upload).
__global__ void k_guide_sync(const float* w32, uint16_t* w, uint16_t* wt) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint32_t)kGuideParams) return;
    const uint32_t l = k / kGuideLayerParams, off = k % kGuideLayerParams;
    if (off < (uint32_t)(kGuideDim * kGuideDim)) {
        const uint32_t o = off / kGuideDim, i = off % kGuideDim;
        const uint16_t h = to_bf16(w32[k]);
        w[l * kGuideDim * kGuideDim + o * kGuideDim + i] = h;
        wt[l * kGuideDim * kGuideDim + i * kGuideDim + o] = h;
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_guide_sync(hipStream_t s, const GuideDev& g) {
    hipLaunchKernelGGL(k_guide_sync, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g.w32, g.w, g.wt);
    return hipGetLastError();
}

hipError_t launch_guide_forward(hipStream_t s, const GuideDev& g, const float* features, uint32_t n, float* out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_forward, dim3((n + 63) / 64), dim3(64), 0, s, GuideNet{g.w, g.wt, g.w32}, features, n,
                       out);
    return hipGetLastError();
}

hipError_t launch_guide_infer(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, uint32_t grid) {
    hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(64), 0, s, p, sc,
                       GuideNet{p.guide.w, p.guide.wt, p.guide.w32}, cur);
    return hipGetLastError();
}

hipError_t launch_guide_train(hipStream_t s, const GuideDev& g, const GuideRecord* recs, uint32_t count,
                              uint32_t batch, uint32_t seed, float w_max, float* loss) {
    if (batch == 0) return hipSuccess;
    hipLaunchKernelGGL(k_guide_train, dim3((batch + 63) / 64), dim3(64), 0, s, GuideNet{g.w, g.wt, g.w32}, recs,
                       count, batch, seed, g.cfg, w_max, g.grad, loss);
    return hipGetLastError();
}

hipError_t launch_guide_adam(hipStream_t s, const GuideDev& g, float lr, uint32_t step, float grad_scale) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g.w32, g.grad, g.m, g.v, g.w,
                       g.wt, lr, beta1, beta2, eps, c1, c2, grad_scale);
    return hipGetLastError();
}

}  // namespace pt
