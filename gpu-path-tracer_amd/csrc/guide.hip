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
// Inference keeps the activations in registers.  Each layer is computed
// transposed, D = W . X^T, with v_mfma_f32_16x16x32_bf16: lane l = r + 16 kq
// of a 16x16 output tile holds outputs 4 kq + j (j < 4) of sample r, which is
// exactly what the lane needs as B operand of the next layer if the 32 inputs
// of an MFMA are ordered (4 kq + j, 16 + 4 kq + j) per lane -- the weights are
// stored in that order.  Only the first layer's input goes through LDS.  The
// bf16 weights of all layers (32 KB) are staged in LDS once per block.
//
// The output layer's rows are reordered so that lane kq of output tile 0 (1)
// holds lobe kq's (4 + kq's) mu xyz and log kappa, and of tile 2 the logits of
// lobes kq and 4 + kq (then two unused outputs): inference output o' maps to
// network output guide_out_row(o').

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

PT_HD int guide_out_row(int o) {
    if (o < 32) return 5 * (o >> 2) + (o & 3);
    if (o < 48) {
        const int kq = (o - 32) >> 2, c = (o - 32) & 3;
        return c == 0 ? 5 * kq + 4 : c == 1 ? 5 * (4 + kq) + 4 : 40 + 2 * kq + (c - 2);
    }
    return o;
}
PT_HD int guide_out_slot(int row) {     // inverse of guide_out_row
    if (row >= 48) return row;
    if (row >= 40) return 32 + 4 * ((row - 40) >> 1) + 2 + ((row - 40) & 1);
    const int j = row / 5, c = row % 5;
    if (c < 4) return 4 * j + c;
    return j < 4 ? 32 + 4 * j : 32 + 4 * (j - 4) + 1;
}

// Position of weight W_l[o'][i] (o' in inference order) in the fragment-ordered
// copy: [layer][mt][kt][lane][8].
PT_HD uint32_t guide_frag_index(int l, int o, int i) {
    const int mt = o >> 4, r = o & 15, kt = i >> 5, w = i & 31;
    const int half = w >> 4, kq = (w & 15) >> 2, j = (w & 3) + 4 * half;
    return ((((uint32_t)l * 4u + (uint32_t)mt) * 2u + (uint32_t)kt) * 64u + (uint32_t)(r + 16 * kq)) * 8u + (uint32_t)j;
}

constexpr int kFragWeights = kGuideLayers * 4 * 2 * 64;   // bf16x8 fragments of all layers

// LDS of an inference block: the weight fragments of layers 0-2 and of the
// first MT 16-row tiles of the output layer, the biases, and one input tile
// per wave.  One block per CU: 16 waves when the unused output rows are left
// out (159 KB), 12 with every output row.
template <int WAVES, int MT>
struct InferLds {
    static constexpr int kWeights = (3 * 4 + MT) * 2 * 64;
    bf16x8 w[kWeights];
    float b[kGuideLayers * kGuideDim];
    bf16x8 x[WAVES][2 * 4 * 64];
};
constexpr int kInferWaves = 16;                       // waves of an inference block
constexpr int kForwardWaves = 12;                     // waves of a (hook) forward block
static_assert(sizeof(InferLds<kInferWaves, 3>) <= 160 * 1024, "inference LDS");
static_assert(sizeof(InferLds<kForwardWaves, 4>) <= 160 * 1024, "forward LDS");

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two floats to packed bf16 (round to nearest even, as to_bf16): one
// v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float lo, float hi) {
    const bf16x2 h = __builtin_convertvector(f32x2{lo, hi}, bf16x2);
    uint32_t r;
    __builtin_memcpy(&r, &h, sizeof(r));
    return r;
}

__device__ __forceinline__ bf16x8 to_bf16x8(const f32x8 v) {
    const uint4 u = make_uint4(cvt_pk_bf16(v[0], v[1]), cvt_pk_bf16(v[2], v[3]), cvt_pk_bf16(v[4], v[5]),
                               cvt_pk_bf16(v[6], v[7]));
    bf16x8 r;
    __builtin_memcpy(&r, &u, sizeof(r));
    return r;
}

// Writes sample s's 64 inputs into the wave's fragment tile
// xf[kt][nt][lane] (bf16x8), nt = s / 16 (round to nearest even, as to_bf16).
__device__ __forceinline__ void store_input_frags(bf16x8* xf, int s, const float f[kGuideDim]) {
    const int nt = s >> 4, r = s & 15;
#pragma unroll
    for (int kt = 0; kt < 2; kt++)
#pragma unroll
        for (int kq = 0; kq < 4; kq++) {
            f32x8 v;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                v[j] = f[32 * kt + 4 * kq + j];
                v[4 + j] = f[32 * kt + 16 + 4 * kq + j];
            }
            xf[(kt * 4 + nt) * 64 + r + 16 * kq] = to_bf16x8(v);
        }
}

// Forward pass of the wave's 64-sample tile: hidden layers in registers, the
// output layer's first MT 16-row tiles into out[mt][nt].  The accumulators
// start from the biases.
template <int MT>
__device__ __forceinline__ void tile_forward(const bf16x8* wsh, const float* bsh, const bf16x8* xf, f32x4 out[4][4]) {
    const int lane = threadIdx.x & 63, kq = lane >> 4;
    bf16x8 b[2][4];
#pragma unroll
    for (int kt = 0; kt < 2; kt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) b[kt][nt] = xf[(kt * 4 + nt) * 64 + lane];
#pragma unroll
    for (int l = 0; l < kGuideLayers; l++) {
        const int mts = l + 1 < kGuideLayers ? 4 : MT;
        const float* bias = bsh + l * kGuideDim;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const float4 bb = *reinterpret_cast<const float4*>(bias + 16 * mt + 4 * kq);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) out[mt][nt] = f32x4{bb.x, bb.y, bb.z, bb.w};
        }
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            if (mt >= mts) break;
#pragma unroll
            for (int kt = 0; kt < 2; kt++) {
                const bf16x8 a = wsh[((l * 4 + mt) * 2 + kt) * 64 + lane];
#pragma unroll
                for (int nt = 0; nt < 4; nt++)
                    out[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[kt][nt], out[mt][nt], 0, 0, 0);
            }
        }
        if (l + 1 == kGuideLayers) break;
        // ReLU -> the next layer's B operands: round to bf16, then clear the
        // negative values as 16-bit integers (a set sign bit is a negative
        // int16), which equals rounding the ReLU output
#pragma unroll
        for (int kt = 0; kt < 2; kt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                f32x8 v;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    v[j] = out[2 * kt][nt][j];
                    v[4 + j] = out[2 * kt + 1][nt][j];
                }
                const bf16x8 h = to_bf16x8(v);
                i16x8 iv;
                __builtin_memcpy(&iv, &h, sizeof(iv));
                iv = __builtin_elementwise_max(iv, i16x8{0, 0, 0, 0, 0, 0, 0, 0});
                __builtin_memcpy(&b[kt][nt], &iv, sizeof(iv));
            }
    }
}

// Stages the first `frags` weight fragments and the biases in LDS (whole block).
__device__ __forceinline__ void stage_weights(bf16x8* wsh, float* bsh, const uint16_t* wf, const float* bf, int frags) {
    const bf16x8* src = reinterpret_cast<const bf16x8*>(wf);
    for (int k = threadIdx.x; k < frags; k += blockDim.x) wsh[k] = src[k];
    for (int k = threadIdx.x; k < kGuideLayers * kGuideDim; k += blockDim.x) bsh[k] = bf[k];
    __syncthreads();
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

// Inputs of the inference kernel: the diffuse and GGX queue entries of the
// current bounce, in the shade kernel's virtual order [diffuse | ggx].
struct InferPathSrc {
    PathState p;
    SceneDev sc;
    int cur;
    __device__ __forceinline__ uint32_t count() const { return p.ctr->queue[0] + p.ctr->queue[1]; }
    __device__ __forceinline__ void features(uint32_t e, float f[kGuideDim]) const {
        const uint32_t c0 = p.ctr->queue[0];
        const uint32_t i = e < c0 ? p.q[0][e] : p.q[1][e - c0];
        const float4 o4 = (cur ? p.ray_o[1] : p.ray_o[0])[i];
        const float4 d4 = (cur ? p.ray_d[1] : p.ray_d[0])[i];
        const float4 hh = p.hit[i];
        const V3 d = xyz(d4);
        const V3 x = xyz(o4) + d * hh.x;
        const float* gn = sc.shade[__float_as_uint(hh.w)].gn;
        V3 n = v3(gn[0], gn[1], gn[2]);
        if (dot(n, d) > 0.0f) n = -n;
        guide_features_each(p.guide.cfg, x, n, -d, [&](int k, float v) { f[k] = v; });
    }
};

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
        // lane (r, kq) holds, for samples 16 nt + r: lobe kq in o[0], lobe 4 + kq
        // in o[1], their logits in o[2][.][0] and [1]; the softmax spans the
        // four lanes r + 16 kq.  The stored fp16 mixture is renormalised where
        // it is read, so fast exp and reciprocals suffice here.
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const float la = o[2][nt][0], lb = o[2][nt][1];
            float mx = fmaxf(la, lb);
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float ea = __expf(la - mx), eb = __expf(lb - mx);
            float sum = ea + eb;
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            const float inv = __frcp_rn(sum);
            const float pa = ea * inv, pb = eb * inv;
            float mu[2][3], kap[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const f32x4 v = o[h][nt];
                const float len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                const bool ok = len2 > 1e-24f;
                const float il = rsqrtf(fmaxf(len2, 1e-24f));
                mu[h][0] = ok ? v[0] * il : 0.0f;
                mu[h][1] = ok ? v[1] * il : 0.0f;
                mu[h][2] = ok ? v[2] * il : 1.0f;
                kap[h] = __expf(clampf(v[3], kGuideLogKappaMin, kGuideLogKappaMax));
            }
            const uint32_t ee = t * 64u + 16u * nt + (uint32_t)r;
            if (ee < total) {
                uint32_t* w = mix + ee;
                w[(size_t)(0 + kq) * ms] = guide_pack_half2(mu[0][0], mu[1][0]);
                w[(size_t)(4 + kq) * ms] = guide_pack_half2(mu[0][1], mu[1][1]);
                w[(size_t)(8 + kq) * ms] = guide_pack_half2(mu[0][2], mu[1][2]);
                w[(size_t)(12 + kq) * ms] = guide_pack_half2(kap[0], kap[1]);
                w[(size_t)(16 + kq) * ms] = guide_pack_half2(pa, pb);
            }
        }
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
__global__ __launch_bounds__(64) void k_guide_train(GuideNet net, const GuideRecord* recs, uint32_t count,
                                                   uint32_t batch, uint32_t seed, GuideConfig g, float w_max,
                                                   float* part, float* loss) {
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
        float* gw = part + (size_t)blockIdx.x * kGuideParams + l * kGuideLayerParams;
        f32x4 acc[4][4];
        mm_dtx(dl, act, acc);
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int nt = 0; nt < 4; nt++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    gw[(16 * mt + 4 * kq + j) * kGuideDim + 16 * nt + r] = acc[mt][nt][j];
        float db = 0.0f;
        for (int t = 0; t < 64; t++) db += from_bf16(dl[t * kLd + lane]);
        gw[kGuideDim * kGuideDim + lane] = db;
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

// Sum of the partial gradients of parameters [64 blockIdx.x, +64): the four
// waves of the block sum every fourth partial (coalesced 256-byte rows), then
// combine through LDS.  Returns the sum in the first wave (other waves: 0).
constexpr int kSumBlock = 256;
__device__ __forceinline__ float sum_parts(const float* part, uint32_t parts, uint32_t k) {
    __shared__ float red[kSumBlock];
    const uint32_t wave = threadIdx.x >> 6;
    float s = 0.0f;
    if (k < (uint32_t)kGuideParams)
        for (uint32_t b = wave; b < parts; b += kSumBlock / 64) s += part[(size_t)b * kGuideParams + k];
    red[threadIdx.x] = s;
    __syncthreads();
    float t = 0.0f;
    if (wave == 0) t = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
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

hipError_t launch_guide_infer(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, uint32_t grid) {
    hipLaunchKernelGGL(k_guide_infer<InferPathSrc>, dim3(grid), dim3(64 * kInferWaves), 0, s, InferPathSrc{p, sc, cur},
                       p.guide.wf, p.guide.bf, p.guide.mix, p.guide.mix_stride);
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
    hipLaunchKernelGGL(k_guide_train, dim3(guide_parts(batch)), dim3(64), 0, s, GuideNet{g.w, g.wt, g.w32}, recs,
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
