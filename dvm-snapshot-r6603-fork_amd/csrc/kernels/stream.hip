// Streaming kernels of known traffic (gfx950): a device copy that reads and
// writes every byte once, and a read-only pass that reads every byte once.
// They give the practical HBM ceiling of a read-once/write-once loop kernel
// on the part in use, and the scale of the FETCH_SIZE / WRITE_SIZE counters
// (their bytes per dispatch are known exactly).
//
// A workgroup of 256 lanes moves chunks of 256 x 4 16-byte vectors (16 KB),
// block-cyclic over the buffer: four loads in flight per lane, consecutive
// lanes on consecutive vectors.  The copy loads and stores with the
// nontemporal hint (the data is touched once); the read-only pass keeps a
// per-wave checksum of what it read so the loads cannot be dropped, stored by
// lane 0 of every wave into its own slot of `sink` (no atomics).  Tails that
// are not a whole chunk (or not a multiple of 16 bytes) take single steps.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.h"

namespace {

struct V16 {
    uint64_t a, b;
};

constexpr int UNROLL = 4;
constexpr int THREADS = 256;
constexpr long CHUNK = (long)THREADS * UNROLL;   // vectors per chunk

__device__ inline V16 loadNT(const V16 *p) {
    V16 v;
    v.a = __builtin_nontemporal_load(&p->a);
    v.b = __builtin_nontemporal_load(&p->b);
    return v;
}

__device__ inline void storeNT(V16 *p, const V16 &v) {
    __builtin_nontemporal_store(v.a, &p->a);
    __builtin_nontemporal_store(v.b, &p->b);
}

__global__ __launch_bounds__(THREADS) void streamCopyKernel(const V16 *__restrict__ src, V16 *__restrict__ dst, long n16) {
    const long chunks = n16 / CHUNK;
    for (long c = blockIdx.x; c < chunks; c += gridDim.x) {
        const long b = c * CHUNK + threadIdx.x;
        V16 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) v[u] = loadNT(src + b + u * THREADS);
#pragma unroll
        for (int u = 0; u < UNROLL; u++) storeNT(dst + b + u * THREADS, v[u]);
    }
    for (long i = chunks * CHUNK + (long)blockIdx.x * THREADS + threadIdx.x; i < n16; i += (long)gridDim.x * THREADS) dst[i] = src[i];
}

__global__ __launch_bounds__(THREADS) void streamReadKernel(const V16 *__restrict__ src, long n16, uint64_t *sink) {
    const long chunks = n16 / CHUNK;
    uint64_t acc = 0;
    for (long c = blockIdx.x; c < chunks; c += gridDim.x) {
        const long b = c * CHUNK + threadIdx.x;
        V16 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) v[u] = src[b + u * THREADS];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) acc ^= v[u].a ^ v[u].b;
    }
    for (long i = chunks * CHUNK + (long)blockIdx.x * THREADS + threadIdx.x; i < n16; i += (long)gridDim.x * THREADS)
        acc ^= src[i].a ^ src[i].b;
    for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sink[(long)blockIdx.x * (THREADS / 64) + (threadIdx.x >> 6)] = acc;
}

__global__ void tailCopyKernel(const unsigned char *src, unsigned char *dst, long n) {
    for (long i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

long streamGrid(long n16, int perCu) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const long want = (long)cus * perCu;
    const long need = (n16 + CHUNK - 1) / CHUNK;
    return need < want ? (need > 0 ? need : 1) : want;
}

}  // namespace

extern "C" long dvmh_hip_stream_sink_bytes(long bytes) { return streamGrid(bytes / 16, 16) * (THREADS / 64) * 8; }

extern "C" void dvmh_hip_stream_copy(void *dst, const void *src, long bytes, hipStream_t s) {
    const long n16 = bytes / 16;
    if (n16 > 0)
        hipLaunchKernelGGL(streamCopyKernel, dim3((unsigned)streamGrid(n16, 32)), dim3(THREADS), 0, s, (const V16 *)src, (V16 *)dst,
                n16);
    if (bytes > n16 * 16)
        hipLaunchKernelGGL(tailCopyKernel, dim3(1), dim3(64), 0, s, (const unsigned char *)src + n16 * 16,
                (unsigned char *)dst + n16 * 16, bytes - n16 * 16);
}

extern "C" void dvmh_hip_stream_read(const void *src, long bytes, uint64_t *sink, hipStream_t s) {
    const long n16 = bytes / 16;
    if (n16 > 0)
        hipLaunchKernelGGL(streamReadKernel, dim3((unsigned)streamGrid(n16, 16)), dim3(THREADS), 0, s, (const V16 *)src, n16, sink);
}
