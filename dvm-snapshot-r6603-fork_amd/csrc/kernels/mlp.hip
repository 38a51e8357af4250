// Kernels of the tuner's learned cost model: a 24 -> 64 -> 64 -> 1 ReLU
// perceptron in FP32 that predicts log(seconds per iteration point) of a
// loop launch configuration.
//
// Mapping for gfx950: one wave64 walks a contiguous stream of samples, lane j
// owning hidden unit j of both hidden layers.  Weights are staged in LDS once
// per workgroup; W1 and W2 are kept transposed for the forward products (lane j
// reads consecutive words) and W2 also in its natural layout for the backward
// product (lane k reads column k: consecutive words again), so no LDS access
// of the inner loops conflicts.  The output unit is a butterfly sum across the
// wave.  Gradients accumulate in each lane's registers over its wave's samples
// and are written as one partial record per wave; a second kernel folds the
// partials in wave order, so gradients are deterministic.  Adam updates the
// parameters elementwise.
//
// Parameter vector layout (NPARAM floats): W1[H][F] | b1[H] | W2[H][H] | b2[H] |
// W3[H] | b3.  Loss: mean over samples of (out - y)^2.
#include <hip/hip_runtime.h>

#include "launch.h"

namespace {

constexpr int F = DVMH_MLP_F, H = DVMH_MLP_H;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * F, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
static_assert(NPARAM == DVMH_MLP_NPARAM, "parameter count");
static_assert(H == 64, "one lane per hidden unit of a wave64");
constexpr int WAVES = 4, THREADS = 64 * WAVES;

__device__ inline float waveSum(float v) {
    // butterfly: every lane ends with the same value (addition is commutative)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct Stage {
    float w1t[F * H];   // w1t[f*H + j] = W1[j][f]
    float w2t[H * H];   // w2t[k*H + j] = W2[j][k]
    float w2[H * H];    // w2[j*H + k]  = W2[j][k]
    float b1[H], b2[H], w3[H];
    float b3;
};

__device__ void stageWeights(const float *__restrict__ P, Stage &s, bool natural) {
    for (int i = threadIdx.x; i < H * F; i += THREADS) {
        int j = i / F, f = i % F;
        s.w1t[f * H + j] = P[OFF_W1 + i];
    }
    for (int i = threadIdx.x; i < H * H; i += THREADS) {
        int j = i / H, k = i % H;
        float w = P[OFF_W2 + i];
        s.w2t[k * H + j] = w;
        if (natural) s.w2[i] = w;
    }
    for (int i = threadIdx.x; i < H; i += THREADS) {
        s.b1[i] = P[OFF_B1 + i];
        s.b2[i] = P[OFF_B2 + i];
        s.w3[i] = P[OFF_W3 + i];
    }
    if (threadIdx.x == 0) s.b3 = P[OFF_B3];
}

// Per-wave sample range [s0, s0 + spw) clipped to n.  Every wave of the grid
// iterates spw times (tail iterations are masked) so workgroup barriers stay
// uniform.
__global__ __launch_bounds__(THREADS) void mlpGradKernel(const float *__restrict__ P, const float *__restrict__ X,
        const float *__restrict__ Y, long n, long spw, float scale, float *__restrict__ part, float *__restrict__ lossPart) {
    __shared__ Stage s;
    __shared__ float sx[WAVES][F];
    __shared__ float sh1[WAVES][H];
    __shared__ float sd2[WAVES][H];
    stageWeights(P, s, true);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long gw = (long)blockIdx.x * WAVES + wv;
    const long s0 = gw * spw;
    float gW1[F], gW2[H];
#pragma unroll
    for (int f = 0; f < F; f++) gW1[f] = 0.f;
#pragma unroll
    for (int k = 0; k < H; k++) gW2[k] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gW3 = 0.f, gb3 = 0.f, loss = 0.f;
    __syncthreads();
    for (long it = 0; it < spw; it++) {
        const long smp = s0 + it;
        const bool valid = smp < n;
        if (lane < F) sx[wv][lane] = valid ? X[smp * F + lane] : 0.f;
        __syncthreads();
        float a1 = s.b1[lane];
#pragma unroll
        for (int f = 0; f < F; f++) a1 += s.w1t[f * H + lane] * sx[wv][f];
        const float h1 = a1 > 0.f ? a1 : 0.f;
        sh1[wv][lane] = h1;
        __syncthreads();
        float a2 = s.b2[lane];
#pragma unroll 16
        for (int k = 0; k < H; k++) a2 += s.w2t[k * H + lane] * sh1[wv][k];
        const float h2 = a2 > 0.f ? a2 : 0.f;
        const float out = waveSum(s.w3[lane] * h2) + s.b3;
        const float e = valid ? out - Y[smp] : 0.f;
        loss += e * e;
        const float dout = scale * e;
        gb3 += dout;
        gW3 += dout * h2;
        const float d2 = a2 > 0.f ? dout * s.w3[lane] : 0.f;
        gb2 += d2;
#pragma unroll
        for (int k = 0; k < H; k++) gW2[k] += d2 * sh1[wv][k];
        sd2[wv][lane] = d2;
        __syncthreads();
        float d1 = 0.f;
#pragma unroll 16
        for (int j = 0; j < H; j++) d1 += s.w2[j * H + lane] * sd2[wv][j];
        d1 = a1 > 0.f ? d1 : 0.f;
        gb1 += d1;
#pragma unroll
        for (int f = 0; f < F; f++) gW1[f] += d1 * sx[wv][f];
        __syncthreads();
    }
    float *pp = part + gw * NPARAM;
#pragma unroll
    for (int f = 0; f < F; f++) pp[OFF_W1 + lane * F + f] = gW1[f];
#pragma unroll
    for (int k = 0; k < H; k++) pp[OFF_W2 + lane * H + k] = gW2[k];
    pp[OFF_B1 + lane] = gb1;
    pp[OFF_B2 + lane] = gb2;
    pp[OFF_W3 + lane] = gW3;
    if (lane == 0) {
        pp[OFF_B3] = gb3;
        lossPart[gw] = loss;
    }
}

// grad[p] = sum over waves w (in order) of part[w][p]; loss = sum lossPart / n.
__global__ void mlpFoldKernel(const float *__restrict__ part, const float *__restrict__ lossPart, long waves, float invN,
        float *__restrict__ grad, float *__restrict__ loss) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) {
        float g = 0.f;
        for (long w = 0; w < waves; w++) g += part[w * NPARAM + p];
        grad[p] = g;
    } else if (p == NPARAM && loss) {
        float l = 0.f;
        for (long w = 0; w < waves; w++) l += lossPart[w];
        *loss = l * invN;
    }
}

// Adam (Kingma & Ba 2015): c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void mlpAdamKernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M, float *__restrict__ V,
        float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = G[p];
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * (g * g);
    M[p] = m;
    V[p] = v;
    const float mh = m / c1, vh = v / c2;
    P[p] = P[p] - lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(THREADS) void mlpForwardKernel(const float *__restrict__ P, const float *__restrict__ X, long n,
        long spw, float *__restrict__ out) {
    __shared__ Stage s;
    __shared__ float sx[WAVES][F];
    __shared__ float sh1[WAVES][H];
    stageWeights(P, s, false);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long s0 = ((long)blockIdx.x * WAVES + wv) * spw;
    __syncthreads();
    for (long it = 0; it < spw; it++) {
        const long smp = s0 + it;
        const bool valid = smp < n;
        if (lane < F) sx[wv][lane] = valid ? X[smp * F + lane] : 0.f;
        __syncthreads();
        float a1 = s.b1[lane];
#pragma unroll
        for (int f = 0; f < F; f++) a1 += s.w1t[f * H + lane] * sx[wv][f];
        sh1[wv][lane] = a1 > 0.f ? a1 : 0.f;
        __syncthreads();
        float a2 = s.b2[lane];
#pragma unroll 16
        for (int k = 0; k < H; k++) a2 += s.w2t[k * H + lane] * sh1[wv][k];
        const float h2 = a2 > 0.f ? a2 : 0.f;
        const float o = waveSum(s.w3[lane] * h2) + s.b3;
        if (valid && lane == 0) out[smp] = o;
        __syncthreads();
    }
}

// Waves for n samples: enough to fill the chip, at least 8 samples per wave.
void shape(long n, long &blocks, long &spw) {
    long waves = (n + 7) / 8;
    if (waves > 2048) waves = 2048;
    if (waves < 1) waves = 1;
    blocks = (waves + WAVES - 1) / WAVES;
    spw = (n + blocks * WAVES - 1) / (blocks * WAVES);
    if (spw < 1) spw = 1;
}

}  // namespace

extern "C" long dvmh_hip_mlp_grad_scratch_bytes(long n) {
    long blocks, spw;
    shape(n, blocks, spw);
    long waves = blocks * WAVES;
    return waves * (long)NPARAM * (long)sizeof(float) + waves * (long)sizeof(float) + 256;
}

extern "C" void dvmh_hip_mlp_grad(const float *P, const float *X, const float *Y, long n, float *grad, float *loss, void *scratch,
        hipStream_t st) {
    long blocks, spw;
    shape(n, blocks, spw);
    long waves = blocks * WAVES;
    float *part = (float *)scratch;
    float *lossPart = part + waves * NPARAM;
    float scale = n > 0 ? 2.0f / (float)n : 0.f;
    hipLaunchKernelGGL(mlpGradKernel, dim3((unsigned)blocks), dim3(THREADS), 0, st, P, X, Y, n, spw, scale, part, lossPart);
    hipLaunchKernelGGL(mlpFoldKernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, (const float *)part, (const float *)lossPart,
            waves, n > 0 ? 1.0f / (float)n : 0.f, grad, loss);
}

extern "C" void dvmh_hip_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float beta1, float beta2, float eps,
        float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(mlpAdamKernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, P, G, M, V, lr, beta1, beta2, eps, c1, c2);
}

extern "C" void dvmh_hip_mlp_forward(const float *P, const float *X, long n, float *out, hipStream_t st) {
    if (n <= 0) return;
    long blocks, spw;
    shape(n, blocks, spw);
    hipLaunchKernelGGL(mlpForwardKernel, dim3((unsigned)blocks), dim3(THREADS), 0, st, P, X, n, spw, out);
}
