// Streaming kernels of known traffic (gfx950): a device copy that reads and
// writes every byte once, and a read-only pass that reads every byte once.
// They give the practical HBM ceiling of a read-once/write-once loop kernel
// on the part in use, and the scale of the FETCH_SIZE / WRITE_SIZE counters
// (their bytes per dispatch are known exactly).
//
// Each lane moves 16-byte vectors, four in flight per loop trip, grid-stride
// over the buffer; the tail that is not a multiple of 16 bytes is handled by
// lane-sized scalar steps of the first workgroup.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.h"

namespace {

struct V16 {
    uint64_t a, b;
};

constexpr int UNROLL = 4;

__global__ __launch_bounds__(256) void streamCopyKernel(const V16 *__restrict__ src, V16 *__restrict__ dst, long n16) {
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (UNROLL - 1) * stride < n16; i += UNROLL * stride) {
        V16 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) dst[i + u * stride] = v[u];
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void streamReadKernel(const V16 *__restrict__ src, long n16, unsigned long long *sink) {
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc = 0;
    for (; i + (UNROLL - 1) * stride < n16; i += UNROLL * stride) {
        V16 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) acc ^= v[u].a ^ v[u].b;
    }
    for (; i < n16; i += stride) acc ^= src[i].a ^ src[i].b;
    // the value only keeps the loads alive; one atomic per wave
    for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) atomicXor(sink, (unsigned long long)acc);
}

__global__ void tailCopyKernel(const unsigned char *src, unsigned char *dst, long n) {
    for (long i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

long streamGrid(long n16) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    long want = (long)cus * 8;
    long need = (n16 + 255) / 256;
    return need < want ? (need > 0 ? need : 1) : want;
}

}  // namespace

extern "C" void dvmh_hip_stream_copy(void *dst, const void *src, long bytes, hipStream_t s) {
    const long n16 = bytes / 16;
    if (n16 > 0)
        hipLaunchKernelGGL(streamCopyKernel, dim3((unsigned)streamGrid(n16)), dim3(256), 0, s, (const V16 *)src, (V16 *)dst, n16);
    if (bytes > n16 * 16)
        hipLaunchKernelGGL(tailCopyKernel, dim3(1), dim3(64), 0, s, (const unsigned char *)src + n16 * 16,
                (unsigned char *)dst + n16 * 16, bytes - n16 * 16);
}

extern "C" void dvmh_hip_stream_read(const void *src, long bytes, unsigned long long *sink, hipStream_t s) {
    const long n16 = bytes / 16;
    if (n16 > 0)
        hipLaunchKernelGGL(streamReadKernel, dim3((unsigned)streamGrid(n16)), dim3(256), 0, s, (const V16 *)src, n16, sink);
}
