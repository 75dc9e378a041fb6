// rz_mlp.h — the learned shading surrogate: a 12 -> 64 -> 64 -> 3 MLP (ReLU, ReLU,
// linear) evaluated per wavefront on gfx950 matrix cores, BF16 operands with FP32
// accumulation.
//
// Inputs per fragment (12): the interpolated camera-space normal N (3), the unit view
// vector from the surface to the eye (3), the albedo (3), and the decoded normal-map
// sample t = 2s-1 ((0,0,1) without a map) (3).  Output: the colour / 255 (3).
//
// Every layer is computed transposed, Y^T = W X^T.  Layer 1 (K = 12, padded to 16)
// uses v_mfma_f32_16x16x16_bf16, layers 2 and 3 (K = 64) v_mfma_f32_16x16x32_bf16:
//   x16 A (16x16): lane l holds A[l%16][4*(l/16) + j], j = 0..3
//   x16 B (16x16): lane l holds B[4*(l/16) + j][l%16]
//   x32 A (16x32): lane l holds A[l%16][8*(l/16) + j], j = 0..7
//   x32 B (32x16): lane l holds B[8*(l/16) + j][l%16]
//   D (16x16, both): lane l holds D[4*(l/16) + j][l%16]
// A layer's accumulator tile h leaves hidden unit 16h + 4*(l/16) + j of fragment l%16
// in lane l.  The x32 B operand of k-tile t is the lane's own tiles 2t and 2t+1
// concatenated, so its row 8g + j (g = l/16) is hidden unit
//   u(t, g, j) = 16*(2t + j/4) + 4g + j%4,
// and the A fragments are packed with the same permutation of K (k_mlp_pack): no data
// moves between layers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rz {

constexpr int kMlpIn = 12, kMlpHidden = 64, kMlpOut = 3;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpOut * kMlpHidden + kMlpOut;   // 5187
// offsets of the FP32 parameter vector: W1[64][12], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + 768, kOffW2 = kOffB1 + 64, kOffB2 = kOffW2 + 4096,
              kOffW3 = kOffB2 + 64, kOffB3 = kOffW3 + 192;
constexpr int kMlpSampleFloats = kMlpIn + kMlpOut;   // training sample: inputs, target colour/255

// Packed BF16 fragments: layer 1 [4 h][64 lanes][4], layer 2 [4 h][2 t][64][8],
// layer 3 [2 t][64][8] (uint16 each), then biases b1[64], b2[64], b3[16] as floats.
constexpr int kFragL1 = 0, kFragL2 = kFragL1 + 4 * 64 * 4, kFragL3 = kFragL2 + 16 * 64 * 4,
              kFragEnd = kFragL3 + 4 * 64 * 4;   // uint16 count: 6144
constexpr int kPackedBytes = kFragEnd * 2 + (64 + 64 + 16) * 4;

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef short mlp_s8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_b8 __attribute__((ext_vector_type(8)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// Hidden unit held in row 8g + j of the x32 B operand of k-tile t.
__host__ __device__ __forceinline__ int mlp_k32_unit(int t, int g, int j) {
    return 16 * (2 * t + j / 4) + 4 * g + j % 4;
}

__device__ __forceinline__ mlp_f4 mfma_k32(mlp_s8 a, mlp_s8 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_b8, a), __builtin_bit_cast(mlp_b8, b), c,
                                                   0, 0, 0);
}

// float -> bf16, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}
__host__ __device__ __forceinline__ float round_bf16(float f) { return from_bf16(to_bf16(f)); }

// Two floats to packed bf16 (a in the low half) with gfx950's conversion instruction:
// round to nearest even, so equal to to_bf16 on finite values.
__device__ __forceinline__ uint32_t bf16x2(float a, float b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)a) | ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)b) << 16);
}

__device__ __forceinline__ mlp_s4 relu_bf16x4(mlp_f4 acc, const float* bias) {
    const float v0 = fmaxf(acc[0] + bias[0], 0.f), v1 = fmaxf(acc[1] + bias[1], 0.f);
    const float v2 = fmaxf(acc[2] + bias[2], 0.f), v3 = fmaxf(acc[3] + bias[3], 0.f);
    return __builtin_bit_cast(mlp_s4, make_uint2(bf16x2(v0, v1), bf16x2(v2, v3)));
}

// Copies the packed network (kPackedBytes) into LDS; the whole workgroup calls it and
// a barrier must follow before mlp_wave reads it.
__device__ __forceinline__ void stage_net(const uint16_t* __restrict__ net, uint16_t* s_net, int tid,
                                          int nthreads) {
    const uint4* src = reinterpret_cast<const uint4*>(net);
    uint4* dst = reinterpret_cast<uint4*>(s_net);
    for (int e = tid; e < kPackedBytes / 16; e += nthreads) dst[e] = src[e];
}

// Evaluates the MLP for the 64 fragments of one wave.  net: the packed network (in LDS,
// staged by stage_net); s_in: this wave's LDS staging of [64][16] bf16 inputs (row =
// lane that owns the fragment, inputs 12..15 zero); s_out: [64][4] floats receiving the
// three outputs.  Every lane of the wave calls it.  Weight fragments are read where
// each product needs them, so the network does not occupy registers.
__device__ __forceinline__ void mlp_wave(const uint16_t* net, const uint16_t* s_in, float* s_out, int lane) {
    const mlp_s4* fr = reinterpret_cast<const mlp_s4*>(net);
    const mlp_s8* fr8 = reinterpret_cast<const mlp_s8*>(net);
    const float* bias = reinterpret_cast<const float*>(net + kFragEnd);
    const int row = 4 * (lane >> 4), col = lane & 15;
#pragma unroll 1
    for (int pt = 0; pt < 4; ++pt) {
        const mlp_s4 xb = *reinterpret_cast<const mlp_s4*>(s_in + (16 * pt + col) * 16 + row);
        mlp_s4 h1[4], h2[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr[kFragL1 / 4 + h * 64 + lane], xb, acc, 0, 0, 0);
            h1[h] = relu_bf16x4(acc, bias + 16 * h + row);
        }
        const mlp_s8 hb0 = __builtin_shufflevector(h1[0], h1[1], 0, 1, 2, 3, 4, 5, 6, 7);
        const mlp_s8 hb1 = __builtin_shufflevector(h1[2], h1[3], 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma_k32(fr8[kFragL2 / 8 + (h * 2 + 0) * 64 + lane], hb0, acc);
            acc = mfma_k32(fr8[kFragL2 / 8 + (h * 2 + 1) * 64 + lane], hb1, acc);
            h2[h] = relu_bf16x4(acc, bias + 64 + 16 * h + row);
        }
        mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_k32(fr8[kFragL3 / 8 + lane], __builtin_shufflevector(h2[0], h2[1], 0, 1, 2, 3, 4, 5, 6, 7), acc);
        acc = mfma_k32(fr8[kFragL3 / 8 + 64 + lane], __builtin_shufflevector(h2[2], h2[3], 0, 1, 2, 3, 4, 5, 6, 7),
                       acc);
        if (lane < 16) {
#pragma unroll
            for (int j = 0; j < 3; ++j) s_out[(16 * pt + col) * 4 + j] = acc[j] + bias[128 + row + j];
        }
    }
}

}  // namespace rz
