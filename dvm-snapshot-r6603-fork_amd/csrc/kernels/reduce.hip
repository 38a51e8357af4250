// Final reductions and record fill for the runtime (gfx950, wave64).
//
// The reference folds one partial per GPU thread with a grid-stride kernel,
// copies one result per warp to the host and folds those on the CPU
// (rts/dvmh/cuda_reduction.cu:168-312).  Here the emitted loop kernels already
// reduce inside each workgroup, so the runtime folds one partial per workgroup
// (a few thousand values) in a single 1024-thread workgroup and only the final
// record goes back to the host.  Order is fixed (thread t folds t, t+1024, ...
// sequentially, then a fixed LDS tree), so the result is deterministic; for
// MAX/MIN with loc the lowest partial index wins ties.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "dvmhlib_const.h"
#include "launch.h"

namespace {

constexpr int FIN_THREADS = 1024;

template <typename T>
__device__ inline T identityOf(int f) {
    if (f == rf_PROD) return (T)1;
    if constexpr (std::is_integral<T>::value) {
        if (f == rf_AND) return (T)~(T)0;
    }
    return (T)0;
}

// acc <- acc (op) v.  Returns true when v replaced acc (MAX/MIN family).
template <typename T>
__device__ inline bool combine(int f, T &acc, long &accIdx, T v, long vIdx) {
    switch (f) {
        case rf_SUM: acc = acc + v; return false;
        case rf_PROD: acc = acc * v; return false;
        case rf_MAX: case rf_MAXLOC:
            if (acc < v || (!(v < acc) && vIdx < accIdx)) { acc = v; accIdx = vIdx; return true; }
            return false;
        case rf_MIN: case rf_MINLOC:
            if (acc > v || (!(v > acc) && vIdx < accIdx)) { acc = v; accIdx = vIdx; return true; }
            return false;
        default: return false;
    }
}

template <typename T>
__device__ inline bool combineBits(int f, T &acc, long &accIdx, T v, long vIdx) {
    switch (f) {
        case rf_AND: acc = acc & v; return false;
        case rf_OR: acc = acc | v; return false;
        case rf_XOR: case rf_NE: acc = acc ^ v; return false;
        case rf_EQU: case rf_EQ: acc = (T)(acc == v); return false;
        default: return combine<T>(f, acc, accIdx, v, vIdx);
    }
}

template <typename T, bool BITS>
__device__ inline void comb(int f, T &acc, long &accIdx, T v, long vIdx) {
    if constexpr (BITS) combineBits<T>(f, acc, accIdx, v, vIdx);
    else combine<T>(f, acc, accIdx, v, vIdx);
}

// One workgroup folds vals[begin, end) -> outVal (+ the winner's loc record).
template <typename T, bool BITS>
__device__ void foldRange(int f, long begin, long end, const T *__restrict__ vals, const char *__restrict__ locs,
        long locSize, T *outVal, char *outLoc) {
    __shared__ T sVal[FIN_THREADS];
    __shared__ long sIdx[FIN_THREADS];
    __shared__ int sHas[FIN_THREADS];
    const int t = threadIdx.x;
    T acc = identityOf<T>(f);
    long accIdx = -1;
    int has = 0;
    for (long i = begin + t; i < end; i += FIN_THREADS) {
        T v = vals[i];
        if (!has) { acc = v; accIdx = i; has = 1; }
        else comb<T, BITS>(f, acc, accIdx, v, i);
    }
    sVal[t] = acc;
    sIdx[t] = accIdx;
    sHas[t] = has;
    __syncthreads();
    for (int w = FIN_THREADS / 2; w > 0; w >>= 1) {
        if (t < w && sHas[t + w]) {
            if (!sHas[t]) { sVal[t] = sVal[t + w]; sIdx[t] = sIdx[t + w]; sHas[t] = 1; }
            else {
                T a = sVal[t];
                long ai = sIdx[t];
                comb<T, BITS>(f, a, ai, sVal[t + w], sIdx[t + w]);
                sVal[t] = a;
                sIdx[t] = ai;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        T r = sHas[0] ? sVal[0] : identityOf<T>(f);
        *outVal = r;
    }
    if (outLoc && locSize > 0 && sHas[0]) {
        long w = sIdx[0];
        for (long b = t; b < locSize; b += FIN_THREADS) outLoc[b] = locs[w * locSize + b];
    }
}

template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void finishKernel(int f, long n, const T *vals, const char *locs, long locSize,
        T *outVal, char *outLoc) {
    foldRange<T, BITS>(f, 0, n, vals, locs, locSize, outVal, outLoc);
}

template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void segmentsKernel(int f, const long *offsets, const T *vals, const char *locs,
        long locSize, T *outVals, char *outLocs) {
    const long g = blockIdx.x;
    foldRange<T, BITS>(f, offsets[g], offsets[g + 1], vals, locs, locSize, outVals + g,
            outLocs ? outLocs + g * locSize : nullptr);
}

__global__ void fillBytesKernel(char *dst, const char *rec, long bytes, long count) {
    long total = bytes * count;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        dst[i] = rec[i % bytes];
}

struct Rec64 {
    uint64_t w[8];
};
__global__ void fillSmallKernel(char *dst, Rec64 rec, long bytes, long count) {
    long total = bytes * count;
    const char *r = (const char *)rec.w;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        dst[i] = r[i % bytes];
}

template <typename T, bool BITS>
void launchFinish(int f, long n, const void *vals, const void *locs, long locSize, void *outVal, void *outLoc,
        hipStream_t s) {
    hipLaunchKernelGGL((finishKernel<T, BITS>), dim3(1), dim3(FIN_THREADS), 0, s, f, n, (const T *)vals,
            (const char *)locs, locSize, (T *)outVal, (char *)outLoc);
}

template <typename T, bool BITS>
void launchSegments(int f, long groups, const long *offsets, const void *vals, const void *locs, long locSize,
        void *outVals, void *outLocs, hipStream_t s) {
    if (groups <= 0) return;
    hipLaunchKernelGGL((segmentsKernel<T, BITS>), dim3((unsigned)groups), dim3(FIN_THREADS), 0, s, f, offsets,
            (const T *)vals, (const char *)locs, locSize, (T *)outVals, (char *)outLocs);
}

}  // namespace

#define DISPATCH_TYPES(LAUNCH, ...)                                                        \
    switch (rtType) {                                                                      \
        case rt_CHAR: case rt_UCHAR: LAUNCH<int8_t, true>(__VA_ARGS__); break;             \
        case rt_INT: case rt_UINT: case rt_LOGICAL: LAUNCH<int32_t, true>(__VA_ARGS__); break; \
        case rt_LONG: case rt_ULONG: case rt_LLONG: case rt_ULLONG: LAUNCH<int64_t, true>(__VA_ARGS__); break; \
        case rt_FLOAT: LAUNCH<float, false>(__VA_ARGS__); break;                           \
        case rt_DOUBLE: LAUNCH<double, false>(__VA_ARGS__); break;                         \
        default: break;                                                                    \
    }

extern "C" void dvmh_hip_reduce_finish(int func, int rtType, long n, const void *vals, const void *locs, long locSize,
        void *outVal, void *outLoc, hipStream_t s) {
    DISPATCH_TYPES(launchFinish, func, n, vals, locs, locSize, outVal, outLoc, s)
}

extern "C" void dvmh_hip_reduce_segments(int func, int rtType, long groups, const long *offsets, const void *vals,
        const void *locs, long locSize, void *outVals, void *outLocs, hipStream_t s) {
    DISPATCH_TYPES(launchSegments, func, groups, offsets, vals, locs, locSize, outVals, outLocs, s)
}

extern "C" void dvmh_hip_fill(void *dst, const void *value, long valueBytes, long count, hipStream_t s) {
    if (count <= 0 || valueBytes <= 0) return;
    long total = valueBytes * count;
    int threads = 256;
    long blocks = (total + threads - 1) / threads;
    if (blocks > 4096) blocks = 4096;
    if (valueBytes <= (long)sizeof(Rec64)) {
        Rec64 rec;
        std::memset(&rec, 0, sizeof(rec));
        std::memcpy(rec.w, value, (size_t)valueBytes);
        hipLaunchKernelGGL(fillSmallKernel, dim3((unsigned)blocks), dim3(threads), 0, s, (char *)dst, rec, valueBytes,
                count);
    } else {
        // Large records: value must already be device-accessible (the runtime stages it).
        hipLaunchKernelGGL(fillBytesKernel, dim3((unsigned)blocks), dim3(threads), 0, s, (char *)dst,
                (const char *)value, valueBytes, count);
    }
}
