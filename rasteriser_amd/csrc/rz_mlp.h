// rz_mlp.h — the learned shading surrogate: a 12 -> 64 -> 64 -> 3 MLP (ReLU, ReLU,
// linear) evaluated per wavefront on gfx950 matrix cores, BF16 operands with FP32
// accumulation.
//
// Inputs per fragment (12): the interpolated camera-space normal N (3), the unit view
// vector from the surface to the eye (3), the albedo (3), and the decoded normal-map
// sample t = 2s-1 ((0,0,1) without a map) (3).  Output: the colour / 255 (3).
//
// Every layer is computed transposed, Y^T = W X^T, with v_mfma_f32_16x16x16_bf16:
//   A (16x16, rows m, cols k): lane l holds A[l%16][4*(l/16) + j], j = 0..3
//   B (16x16, rows k, cols n): lane l holds B[4*(l/16) + j][l%16]
//   D (16x16):                 lane l holds D[4*(l/16) + j][l%16]
// so a layer's accumulator tile (hidden unit 4*(l/16)+j of tile h, fragment l%16) is
// already the next layer's B operand for k-tile h: no data moves between layers.
// Weights are packed once into per-lane A fragments (k_mlp_pack).
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

// Packed BF16 fragments: layer 1 [4 h][64 lanes][4], layer 2 [4 h][4 kt][64][4],
// layer 3 [4 kt][64][4] (uint16 each), then biases b1[64], b2[64], b3[16] as floats.
constexpr int kFragL1 = 0, kFragL2 = kFragL1 + 4 * 64 * 4, kFragL3 = kFragL2 + 16 * 64 * 4,
              kFragEnd = kFragL3 + 4 * 64 * 4;   // uint16 count: 6144
constexpr int kPackedBytes = kFragEnd * 2 + (64 + 64 + 16) * 4;

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

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
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = acc[j] + bias[16 * h + row + j];
                h1[h][j] = (short)to_bf16(v > 0.f ? v : 0.f);
            }
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr[kFragL2 / 4 + (h * 4 + kt) * 64 + lane], h1[kt],
                                                               acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = acc[j] + bias[64 + 16 * h + row + j];
                h2[h][j] = (short)to_bf16(v > 0.f ? v : 0.f);
            }
        }
        mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fr[kFragL3 / 4 + kt * 64 + lane], h2[kt], acc, 0, 0, 0);
        if (lane < 16) {
#pragma unroll
            for (int j = 0; j < 3; ++j) s_out[(16 * pt + col) * 4 + j] = acc[j] + bias[128 + row + j];
        }
    }
}

}  // namespace rz
