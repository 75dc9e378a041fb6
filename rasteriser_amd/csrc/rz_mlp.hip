// rz_mlp.hip — packing, inference and training kernels of the learned shader (rz_mlp.h).
//
// Training is minibatch Adam on FP32 master weights.  The forward pass rounds weights
// and activations to BF16 exactly where inference does (inputs, hidden activations),
// so the trained network is the one the matrix cores evaluate; gradients are FP32 and
// pass straight through the rounding.  Loss: mean over the batch of the squared error
// summed over the three channels.
#include "rz_device.h"
#include "rz_mlp.h"

namespace rz {

// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ W, uint16_t* __restrict__ packed) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kFragL2) {
        const int h = i / 256, lane = (i / 4) % 64, j = i % 4;
        const int o = 16 * h + lane % 16, k = 4 * (lane / 16) + j;
        packed[i] = to_bf16(k < kMlpIn ? W[kOffW1 + o * kMlpIn + k] : 0.f);
    } else if (i < kFragL3) {
        // [h][t][lane][8]: A[m][8g + j] = W2[16h + m][u(t, g, j)]
        const int q = i - kFragL2, ht = q / 512, lane = (q / 8) % 64, j = q % 8;
        const int h = ht / 2, t = ht % 2;
        const int o = 16 * h + lane % 16, k = mlp_k32_unit(t, lane / 16, j);
        packed[i] = to_bf16(W[kOffW2 + o * kMlpHidden + k]);
    } else if (i < kFragEnd) {
        // [t][lane][8]: A[m][8g + j] = W3[m][u(t, g, j)] (rows 3..15 zero)
        const int q = i - kFragL3, t = q / 512, lane = (q / 8) % 64, j = q % 8;
        const int o = lane % 16, k = mlp_k32_unit(t, lane / 16, j);
        packed[i] = to_bf16(o < kMlpOut ? W[kOffW3 + o * kMlpHidden + k] : 0.f);
    } else if (i < kFragEnd + 144) {
        float* bias = reinterpret_cast<float*>(packed + kFragEnd);
        const int b = i - kFragEnd;
        bias[b] = b < 64 ? W[kOffB1 + b] : b < 128 ? W[kOffB2 + b - 64] : (b - 128 < kMlpOut ? W[kOffB3 + b - 128] : 0.f);
    }
}

// Inference over a feature array [n][12] -> [n][3] (the shading pass runs the same
// mlp_wave on fragments it computes itself).
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ x, int64_t n,
                                                   const uint16_t* __restrict__ packed, float* __restrict__ y) {
    __shared__ __align__(16) uint16_t s_net[kPackedBytes / 2];
    __shared__ __align__(16) uint16_t s_in[4][64 * 16];
    __shared__ __align__(16) float s_out[4][64 * 4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t i = (int64_t)blockIdx.x * 256 + tid;
    stage_net(packed, s_net, tid, 256);
#pragma unroll
    for (int k = 0; k < 16; ++k)
        s_in[wave][lane * 16 + k] = to_bf16((i < n && k < kMlpIn) ? x[i * kMlpIn + k] : 0.f);
    __syncthreads();
    mlp_wave(s_net, s_in[wave], s_out[wave], lane);
    __syncthreads();
    if (i < n)
        for (int c = 0; c < kMlpOut; ++c) y[i * kMlpOut + c] = s_out[wave][lane * 4 + c];
}

// ---------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One workgroup = 64 samples drawn at random from the training set; writes the
// workgroup's summed parameter gradient and loss.
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ samples, int64_t n,
                                                 const float* __restrict__ W, float* __restrict__ gpart,
                                                 float* __restrict__ lpart, uint64_t seed, uint32_t step,
                                                 float inv_batch) {
    __shared__ float sW[kMlpParams];     // weights rounded to bf16, biases as they are
    __shared__ float sX[64][kMlpIn + 1];
    __shared__ float sH1[64][65], sH2[64][65], sD1[64][65], sD2[64][65];
    __shared__ float sDY[64][4];
    __shared__ float sLoss[64];
    const int t = threadIdx.x;
    for (int e = t; e < kMlpParams; e += 64) {
        const bool bias = (e >= kOffB1 && e < kOffW2) || (e >= kOffB2 && e < kOffW3) || e >= kOffB3;
        sW[e] = bias ? W[e] : round_bf16(W[e]);
    }
    const uint64_t gid = (uint64_t)blockIdx.x * 64 + t;
    const int64_t idx = (int64_t)(mix64(seed + 0x9E3779B97F4A7C15ull * ((uint64_t)step * 1048576u + gid + 1)) %
                                  (uint64_t)n);
    const float* smp = samples + idx * kMlpSampleFloats;
    for (int k = 0; k < kMlpIn; ++k) sX[t][k] = round_bf16(smp[k]);
    __syncthreads();
    // forward
    for (int o = 0; o < kMlpHidden; ++o) {
        float a = 0.f;
        for (int k = 0; k < kMlpIn; ++k) a += sW[kOffW1 + o * kMlpIn + k] * sX[t][k];
        a += sW[kOffB1 + o];
        sD1[t][o] = a;                                   // pre-activation, kept for relu'
        sH1[t][o] = round_bf16(a > 0.f ? a : 0.f);
    }
    for (int o = 0; o < kMlpHidden; ++o) {
        float a = 0.f;
        for (int k = 0; k < kMlpHidden; ++k) a += sW[kOffW2 + o * kMlpHidden + k] * sH1[t][k];
        a += sW[kOffB2 + o];
        sD2[t][o] = a;
        sH2[t][o] = round_bf16(a > 0.f ? a : 0.f);
    }
    float loss = 0.f;
    for (int c = 0; c < kMlpOut; ++c) {
        float a = 0.f;
        for (int k = 0; k < kMlpHidden; ++k) a += sW[kOffW3 + c * kMlpHidden + k] * sH2[t][k];
        a += sW[kOffB3 + c];
        const float r = a - smp[kMlpIn + c];
        loss += r * r;
        sDY[t][c] = 2.f * r * inv_batch;
    }
    sLoss[t] = loss;
    // backward through the activations
    for (int k = 0; k < kMlpHidden; ++k) {
        float g = 0.f;
        for (int c = 0; c < kMlpOut; ++c) g += sW[kOffW3 + c * kMlpHidden + k] * sDY[t][c];
        sD2[t][k] = sD2[t][k] > 0.f ? g : 0.f;
    }
    for (int k = 0; k < kMlpHidden; ++k) {
        float g = 0.f;
        for (int o = 0; o < kMlpHidden; ++o) g += sW[kOffW2 + o * kMlpHidden + k] * sD2[t][o];
        sD1[t][k] = sD1[t][k] > 0.f ? g : 0.f;
    }
    __syncthreads();
    // parameter gradients summed over the workgroup's 64 samples
    float* out = gpart + (size_t)blockIdx.x * kMlpParams;
    for (int e = t; e < kMlpParams; e += 64) {
        float g = 0.f;
        if (e < kOffB1) {
            const int o = e / kMlpIn, k = e % kMlpIn;
            for (int s = 0; s < 64; ++s) g += sD1[s][o] * sX[s][k];
        } else if (e < kOffW2) {
            for (int s = 0; s < 64; ++s) g += sD1[s][e - kOffB1];
        } else if (e < kOffB2) {
            const int o = (e - kOffW2) / kMlpHidden, k = (e - kOffW2) % kMlpHidden;
            for (int s = 0; s < 64; ++s) g += sD2[s][o] * sH1[s][k];
        } else if (e < kOffW3) {
            for (int s = 0; s < 64; ++s) g += sD2[s][e - kOffB2];
        } else if (e < kOffB3) {
            const int c = (e - kOffW3) / kMlpHidden, k = (e - kOffW3) % kMlpHidden;
            for (int s = 0; s < 64; ++s) g += sDY[s][c] * sH2[s][k];
        } else {
            for (int s = 0; s < 64; ++s) g += sDY[s][e - kOffB3];
        }
        out[e] = g;
    }
    if (t == 0) {
        float l = 0.f;
        for (int s = 0; s < 64; ++s) l += sLoss[s];
        lpart[blockIdx.x] = l;
    }
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ W, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ gpart, int nparts,
                                                  const float* __restrict__ lpart, float* __restrict__ loss_out,
                                                  float lr, float c1, float c2) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < kMlpParams) {
        float g = 0.f;
        for (int b = 0; b < nparts; ++b) g += gpart[(size_t)b * kMlpParams + e];
        const float mm = 0.9f * m[e] + 0.1f * g;
        const float vv = 0.999f * v[e] + 0.001f * g * g;
        m[e] = mm;
        v[e] = vv;
        W[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + 1e-8f);
    }
    if (e == 0 && loss_out) {
        float l = 0.f;
        for (int b = 0; b < nparts; ++b) l += lpart[b];
        *loss_out = l;
    }
}

hipError_t launch_mlp_pack(const float* W, uint16_t* packed, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kFragEnd + 144 + 255) / 256), dim3(256), 0, s, W, packed);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* x, int64_t n, const uint16_t* packed, float* y, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, packed, y);
    return hipGetLastError();
}

hipError_t launch_mlp_step(const float* samples, int64_t n, float* W, float* m, float* v, float* gpart,
                           float* lpart, float* loss_out, int nblocks, uint64_t seed, uint32_t step, float lr,
                           float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(64), 0, s, samples, n, (const float*)W, gpart, lpart, seed,
                       step, 1.0f / (float)(nblocks * 64));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, W, m, v, (const float*)gpart,
                       nblocks, (const float*)lpart, loss_out, lr, c1, c2);
    return hipGetLastError();
}

}  // namespace rz
