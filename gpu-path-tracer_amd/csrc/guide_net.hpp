// Guide network inference shared by the inference kernel and the guided shade
// kernel: weight staging, the register-resident MFMA forward pass of a
// 64-sample tile, and the packed mixture output.
#pragma once

#include <hip/hip_runtime.h>

#include "guide.hpp"
#include "launch.hpp"

namespace pt {

typedef float f32x4 __attribute__((ext_vector_type(4)));

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

// Two floats to packed bf16 (round to nearest even): one
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
// xf[kt][nt][lane] (bf16x8), nt = s / 16, rounded to nearest even.
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
            if (mt >= mts) break;
            // the first product of each tile takes the bias as its accumulator input
            const float4 bb = *reinterpret_cast<const float4*>(bias + 16 * mt + 4 * kq);
            const f32x4 b4 = f32x4{bb.x, bb.y, bb.z, bb.w};
            const bf16x8 a0 = wsh[((l * 4 + mt) * 2 + 0) * 64 + lane];
            const bf16x8 a1 = wsh[((l * 4 + mt) * 2 + 1) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                out[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b[0][nt], b4, 0, 0, 0);
                out[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b[1][nt], out[mt][nt], 0, 0, 0);
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

// Packed mixture words (guide.hpp) of the wave's tile from the output-layer
// tiles o[0..2]: store(word k, sample 0..63 of the tile, value) is called by
// the lane that computes the word.
template <class Store>
__device__ __forceinline__ void mixture_words(const f32x4 o[4][4], Store store) {
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
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
        const uint32_t sidx = 16u * nt + (uint32_t)r;
        store(0 + kq, sidx, guide_pack_half2(mu[0][0], mu[1][0]));
        store(4 + kq, sidx, guide_pack_half2(mu[0][1], mu[1][1]));
        store(8 + kq, sidx, guide_pack_half2(mu[0][2], mu[1][2]));
        store(12 + kq, sidx, guide_pack_half2(kap[0], kap[1]));
        store(16 + kq, sidx, guide_pack_half2(pa, pb));
    }
}

}  // namespace pt
