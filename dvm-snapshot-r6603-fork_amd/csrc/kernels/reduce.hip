// Final reductions and record fill for the runtime (gfx950, wave64).
//
// The reference folds one partial per GPU thread with a grid-stride kernel,
// copies one result per warp to the host and folds those on the CPU
// (rts/dvmh/cuda_reduction.cu:168-312).  Here the emitted loop kernels already
// reduce inside each workgroup, so the runtime folds one partial per workgroup
// (a few thousand values) in a single 1024-thread workgroup and only the final
// record goes back to the host.  Order is fixed (thread t folds t, t+1024, ...
// sequentially, then fixed wave64 shuffle trees), so the result is
// deterministic.  MAXLOC/MINLOC ties keep the lower loc when the loc record is
// a scalar integer (1, 2, 4 or 8 bytes, read as signed), as inside the loop
// kernels, so the winner is the sequential program's whenever the loc grows
// with the iteration; other records keep the lower partial index.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
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

// Tie-break key of partial e for MAXLOC/MINLOC: its loc record read as a
// signed integer of `bytes` bytes (1, 2, 4 or 8), or 0 when bytes is 0 or e
// is no position (-1, the winner of an empty range).
struct LocKey {
    const char *locs;
    long size;
    int bytes;
    __device__ inline long operator()(long e) const {
        if (e < 0) return 0;
        const char *p = locs + e * size;
        switch (bytes) {
            case 8: { long long v; __builtin_memcpy(&v, p, 8); return (long)v; }
            case 4: { int v; __builtin_memcpy(&v, p, 4); return v; }
            case 2: { short v; __builtin_memcpy(&v, p, 2); return v; }
            case 1: return (long)(signed char)p[0];
            default: return 0;
        }
    }
};

__host__ __device__ inline int locKeyBytes(int f, const void *locs, long locSize) {
    if ((f != rf_MAXLOC && f != rf_MINLOC) || !locs) return 0;
    return (locSize == 1 || locSize == 2 || locSize == 4 || locSize == 8) ? (int)locSize : 0;
}

// Partial being folded: value, tie-break key, position in the partial buffer.
template <typename T>
struct Acc {
    T v;
    long key, pos;
};

// a <- a (op) b for the arithmetic and MAX/MIN families.  Equal (unordered)
// values keep the lower key, then the lower position.
template <typename T>
__device__ inline void combine(int f, Acc<T> &a, const Acc<T> &b) {
    const bool lower = b.key < a.key || (b.key == a.key && b.pos < a.pos);
    switch (f) {
        case rf_SUM: a.v = a.v + b.v; return;
        case rf_PROD: a.v = a.v * b.v; return;
        case rf_MAX: case rf_MAXLOC:
            if (a.v < b.v || (!(b.v < a.v) && lower)) a = b;
            return;
        case rf_MIN: case rf_MINLOC:
            if (a.v > b.v || (!(b.v > a.v) && lower)) a = b;
            return;
        default: return;
    }
}

template <typename T>
__device__ inline void combineBits(int f, Acc<T> &a, const Acc<T> &b) {
    switch (f) {
        case rf_AND: a.v = a.v & b.v; return;
        case rf_OR: a.v = a.v | b.v; return;
        case rf_XOR: case rf_NE: a.v = a.v ^ b.v; return;
        case rf_EQU: case rf_EQ: a.v = (T)(a.v == b.v); return;
        default: combine<T>(f, a, b); return;
    }
}

template <typename T, bool BITS>
__device__ inline void comb(int f, Acc<T> &a, const Acc<T> &b) {
    if constexpr (BITS) combineBits<T>(f, a, b);
    else combine<T>(f, a, b);
}

// Bit-exact lane shuffle of any element type used here (1, 4 or 8 bytes).
template <typename T>
__device__ inline T shflDownAny(T v, int o) {
    if constexpr (sizeof(T) == 8) {
        long long x;
        __builtin_memcpy(&x, &v, 8);
        x = __shfl_down(x, o, 64);
        __builtin_memcpy(&v, &x, 8);
        return v;
    } else if constexpr (sizeof(T) == 4) {
        int x;
        __builtin_memcpy(&x, &v, 4);
        x = __shfl_down(x, o, 64);
        __builtin_memcpy(&v, &x, 4);
        return v;
    } else {
        return (T)__shfl_down((int)v, o, 64);
    }
}

// Tree over the 64 lanes of a wavefront; lane 0 ends with the fold of all.
// KEYS: the partials carry loc keys (MAXLOC/MINLOC), shuffled with them.
template <typename T, bool BITS, bool KEYS>
__device__ inline void waveFold(int f, Acc<T> &acc, int &has) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Acc<T> b;
        b.v = shflDownAny(acc.v, o);
        b.pos = shflDownAny(acc.pos, o);
        b.key = KEYS ? shflDownAny(acc.key, o) : 0;
        const int vh = __shfl_down(has, o, 64);
        if (vh && lane + o < 64) {
            if (!has) { acc = b; has = 1; }
            else comb<T, BITS>(f, acc, b);
        }
    }
}

// One workgroup (FIN_THREADS threads) folds vals[begin, end).  Element e of
// the range sits at position idx ? idx[e] : e of the partial buffer, and its
// tie-break key is lk(position).  The result and the winner's position go to
// outVal / outIdx (position -1 for an empty range).  Thread t folds t,
// t + FIN_THREADS, ... in order (8 loads in flight), then a wave64 shuffle
// tree per wavefront and one more over the wavefront results: a fixed order,
// so the result is deterministic.
template <typename T, bool BITS, bool KEYS>
__device__ void foldRangeK(int f, long begin, long end, const T *__restrict__ vals, const long *__restrict__ idx, LocKey lk,
        T *outVal, long *outIdx) {
    constexpr int NW = FIN_THREADS / 64;
    __shared__ T sVal[NW];
    __shared__ long sPos[NW];
    __shared__ long sKey[NW];
    __shared__ int sHas[NW];
    const int t = threadIdx.x;
    Acc<T> acc{identityOf<T>(f), 0, -1};
    int has = 0;
    long i = begin + t;
    for (; i + 7 * FIN_THREADS < end; i += 8 * FIN_THREADS) {
        Acc<T> v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            v[q].v = vals[i + q * FIN_THREADS];
            v[q].pos = idx ? idx[i + q * FIN_THREADS] : i + q * FIN_THREADS;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) v[q].key = KEYS ? lk(v[q].pos) : 0;
        if (!has) { acc = v[0]; has = 1; } else comb<T, BITS>(f, acc, v[0]);
#pragma unroll
        for (int q = 1; q < 8; q++) comb<T, BITS>(f, acc, v[q]);
    }
    for (; i < end; i += FIN_THREADS) {
        Acc<T> v;
        v.v = vals[i];
        v.pos = idx ? idx[i] : i;
        v.key = KEYS ? lk(v.pos) : 0;
        if (!has) { acc = v; has = 1; }
        else comb<T, BITS>(f, acc, v);
    }
    waveFold<T, BITS, KEYS>(f, acc, has);
    const int w = t >> 6, lane = t & 63;
    if (lane == 0) {
        sVal[w] = acc.v;
        sPos[w] = has ? acc.pos : -1;
        sKey[w] = acc.key;
        sHas[w] = has;
    }
    __syncthreads();
    if (w == 0) {
        acc.v = lane < NW ? sVal[lane] : identityOf<T>(f);
        acc.pos = lane < NW ? sPos[lane] : -1;
        acc.key = lane < NW ? sKey[lane] : 0;
        has = lane < NW ? sHas[lane] : 0;
        waveFold<T, BITS, KEYS>(f, acc, has);
        if (lane == 0) {
            *outVal = has ? acc.v : identityOf<T>(f);
            if (outIdx) *outIdx = has ? acc.pos : -1;
        }
    }
}

template <typename T, bool BITS>
__device__ void foldRange(int f, long begin, long end, const T *__restrict__ vals, const long *__restrict__ idx, LocKey lk,
        T *outVal, long *outIdx) {
    if (lk.bytes) foldRangeK<T, BITS, true>(f, begin, end, vals, idx, lk, outVal, outIdx);
    else foldRangeK<T, BITS, false>(f, begin, end, vals, idx, lk, outVal, outIdx);
}

// Copies the winner's loc record; *winShared was written by thread 0.
template <typename T>
__device__ void copyLoc(const char *locs, long locSize, const long *winShared, char *outLoc) {
    __syncthreads();
    const long w = *winShared;
    if (outLoc && locSize > 0 && w >= 0)
        for (long b = threadIdx.x; b < locSize; b += blockDim.x) outLoc[b] = locs[w * locSize + b];
}

// pass 1: block b folds [b * chunk, (b + 1) * chunk)
template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void chunkKernel(int f, long n, long chunk, const T *vals, const char *locs,
        long locSize, T *outVals, long *outIdx) {
    long b = blockIdx.x;
    long e = (b + 1) * chunk < n ? (b + 1) * chunk : n;
    const LocKey lk{locs, locSize, locKeyBytes(f, locs, locSize)};
    foldRange<T, BITS>(f, b * chunk, e, vals, nullptr, lk, outVals + b, outIdx + b);
}

// final pass over n values (with their original indices if idx != 0)
template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void finishKernel(int f, long n, const T *vals, const long *idx, const char *locs,
        long locSize, T *outVal, char *outLoc) {
    __shared__ long win;
    const LocKey lk{locs, locSize, locKeyBytes(f, locs, locSize)};
    foldRange<T, BITS>(f, 0, n, vals, idx, lk, outVal, &win);
    copyLoc<T>(locs, locSize, &win, outLoc);
}

template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void segmentsKernel(int f, const long *offsets, const T *vals, const char *locs,
        long locSize, T *outVals, char *outLocs) {
    __shared__ long win;
    const long g = blockIdx.x;
    const LocKey lk{locs, locSize, locKeyBytes(f, locs, locSize)};
    foldRange<T, BITS>(f, offsets[g], offsets[g + 1], vals, nullptr, lk, outVals + g, &win);
    copyLoc<T>(locs, locSize, &win, outLocs ? outLocs + g * locSize : nullptr);
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

constexpr long FIN_CHUNK = 16 * FIN_THREADS;

template <typename T, bool BITS>
void launchFinish(int f, long n, const void *vals, const void *locs, long locSize, void *outVal, void *outLoc, void *scratch,
        hipStream_t s) {
    if (n > FIN_CHUNK && scratch) {
        long g = (n + FIN_CHUNK - 1) / FIN_CHUNK;
        T *pv = (T *)scratch;
        long *pi = (long *)((char *)scratch + ((g * sizeof(T) + 15) & ~15L));
        hipLaunchKernelGGL((chunkKernel<T, BITS>), dim3((unsigned)g), dim3(FIN_THREADS), 0, s, f, n, FIN_CHUNK, (const T *)vals,
                (const char *)locs, locSize, pv, pi);
        hipLaunchKernelGGL((finishKernel<T, BITS>), dim3(1), dim3(FIN_THREADS), 0, s, f, g, (const T *)pv, (const long *)pi,
                (const char *)locs, locSize, (T *)outVal, (char *)outLoc);
        return;
    }
    hipLaunchKernelGGL((finishKernel<T, BITS>), dim3(1), dim3(FIN_THREADS), 0, s, f, n, (const T *)vals, (const long *)nullptr,
            (const char *)locs, locSize, (T *)outVal, (char *)outLoc);
}

// ---- several reductions of one element type in one launch
// Block b belongs to reduction r with first[r] <= b < first[r + 1].  A
// reduction of at most MULTI_SINGLE partials gets one block, which folds them
// all and writes the result.  A larger one gets a block per MULTI_CHUNK
// partials: block c = b - first[r] folds [c * MULTI_CHUNK, (c + 1) * MULTI_CHUNK)
// into slot c, and the block that arrives last at r's counter folds r's slots
// in slot order (ties as in foldRange), writes the result and the
// winner's loc record, and resets the counter for the next launch.
constexpr long MULTI_CHUNK = 8 * FIN_THREADS;
constexpr long MULTI_SINGLE = 16 * FIN_THREADS;

__host__ __device__ inline long multiBlocks(long n) { return n <= MULTI_SINGLE ? 1 : (n + MULTI_CHUNK - 1) / MULTI_CHUNK; }

struct MultiArgs {
    int count;
    int func[DVMH_MULTI_MAX];
    long n[DVMH_MULTI_MAX];
    const void *vals[DVMH_MULTI_MAX];
    const char *locs[DVMH_MULTI_MAX];
    long locSize[DVMH_MULTI_MAX];
    void *outVal[DVMH_MULTI_MAX];
    char *outLoc[DVMH_MULTI_MAX];
    void *slots[DVMH_MULTI_MAX];
    long first[DVMH_MULTI_MAX + 1];
    unsigned *counters;
    unsigned long long *doneFlag;   // 0, or a word the launch sets to doneSeq once every result is written
    unsigned long long doneSeq;
};

// Called by every thread of a block that has just written reduction r's
// result.  The block that completes the last of the launch's results
// publishes doneSeq: every result (host memory) is visible to the host
// before the word is.
__device__ inline void multiResultWritten(const MultiArgs &a) {
    if (!a.doneFlag) return;
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned before = atomicAdd(&a.counters[DVMH_MULTI_MAX], 1u);
        if (before == (unsigned)(a.count - 1)) {
            a.counters[DVMH_MULTI_MAX] = 0u;
            __threadfence_system();
            __hip_atomic_store(a.doneFlag, a.doneSeq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

template <typename T, bool BITS>
__global__ __launch_bounds__(FIN_THREADS) void multiFinishKernel(MultiArgs a) {
    __shared__ long win;
    __shared__ int isLast;
    const long b = blockIdx.x;
    int r = 0;
    while (r + 1 < a.count && b >= a.first[r + 1]) r++;
    const long c = b - a.first[r];
    const long nb = a.first[r + 1] - a.first[r];
    const LocKey lk{a.locs[r], a.locSize[r], locKeyBytes(a.func[r], a.locs[r], a.locSize[r])};
    if (nb == 1) {
        foldRange<T, BITS>(a.func[r], 0, a.n[r], (const T *)a.vals[r], nullptr, lk, (T *)a.outVal[r], &win);
        copyLoc<T>(a.locs[r], a.locSize[r], &win, a.outLoc[r]);
        multiResultWritten(a);
        return;
    }
    T *slotVal = (T *)a.slots[r];
    long *slotIdx = (long *)((char *)a.slots[r] + ((nb * (long)sizeof(T) + 15) & ~15L));
    const long lo = c * MULTI_CHUNK;
    const long hi = lo + MULTI_CHUNK < a.n[r] ? lo + MULTI_CHUNK : a.n[r];
    foldRange<T, BITS>(a.func[r], lo, hi, (const T *)a.vals[r], nullptr, lk, slotVal + c, slotIdx + c);
    if (threadIdx.x == 0) {
        __threadfence();   // the slot is visible before the arrival is counted
        const unsigned before = atomicAdd(&a.counters[r], 1u);
        isLast = before == (unsigned)(nb - 1);
    }
    __syncthreads();
    if (!isLast) return;
    __threadfence();
    foldRange<T, BITS>(a.func[r], 0, nb, slotVal, slotIdx, lk, (T *)a.outVal[r], &win);
    copyLoc<T>(a.locs[r], a.locSize[r], &win, a.outLoc[r]);
    if (threadIdx.x == 0) a.counters[r] = 0u;
    multiResultWritten(a);
}

template <typename T, bool BITS>
void launchMulti(int count, const DvmhMultiRed *reds, unsigned *counters, unsigned long long *doneFlag,
        unsigned long long doneSeq, hipStream_t s) {
    MultiArgs a;
    a.count = count;
    a.counters = counters;
    a.doneFlag = doneFlag;
    a.doneSeq = doneSeq;
    long blocks = 0;
    for (int r = 0; r < count; r++) {
        a.func[r] = reds[r].func;
        a.n[r] = reds[r].n;
        a.vals[r] = reds[r].vals;
        a.locs[r] = (const char *)reds[r].locs;
        a.locSize[r] = reds[r].locSize;
        a.outVal[r] = reds[r].outVal;
        a.outLoc[r] = (char *)reds[r].outLoc;
        a.slots[r] = reds[r].slots;
        a.first[r] = blocks;
        blocks += multiBlocks(reds[r].n);
    }
    a.first[count] = blocks;
    hipLaunchKernelGGL((multiFinishKernel<T, BITS>), dim3((unsigned)blocks), dim3(FIN_THREADS), 0, s, a);
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
        case rt_CHAR: LAUNCH<int8_t, true>(__VA_ARGS__); break;                            \
        case rt_UCHAR: LAUNCH<uint8_t, true>(__VA_ARGS__); break;                          \
        case rt_INT: case rt_LOGICAL: LAUNCH<int32_t, true>(__VA_ARGS__); break;           \
        case rt_UINT: LAUNCH<uint32_t, true>(__VA_ARGS__); break;                          \
        case rt_LONG: case rt_LLONG: LAUNCH<int64_t, true>(__VA_ARGS__); break;            \
        case rt_ULONG: case rt_ULLONG: LAUNCH<uint64_t, true>(__VA_ARGS__); break;         \
        case rt_FLOAT: LAUNCH<float, false>(__VA_ARGS__); break;                           \
        case rt_DOUBLE: LAUNCH<double, false>(__VA_ARGS__); break;                         \
        default:                                                                           \
            fprintf(stderr, "DVMH internal error: reduction element type %d has no GPU fold\n", rtType); \
            abort();                                                                       \
    }

extern "C" long dvmh_hip_reduce_scratch_bytes(long n) {
    long g = (n + FIN_CHUNK - 1) / FIN_CHUNK;
    return n > FIN_CHUNK ? g * 24 + 64 : 0;
}

extern "C" void dvmh_hip_reduce_finish(int func, int rtType, long n, const void *vals, const void *locs, long locSize,
        void *outVal, void *outLoc, void *scratch, hipStream_t s) {
    DISPATCH_TYPES(launchFinish, func, n, vals, locs, locSize, outVal, outLoc, scratch, s)
}

extern "C" void dvmh_hip_reduce_segments(int func, int rtType, long groups, const long *offsets, const void *vals,
        const void *locs, long locSize, void *outVals, void *outLocs, hipStream_t s) {
    DISPATCH_TYPES(launchSegments, func, groups, offsets, vals, locs, locSize, outVals, outLocs, s)
}

extern "C" long dvmh_hip_multi_slot_bytes(long n) {
    const long nb = multiBlocks(n);
    return ((nb * 8 + 15) & ~15L) + nb * 8 + 64;
}

extern "C" void dvmh_hip_reduce_finish_multi(int rtType, int count, const DvmhMultiRed *reds, unsigned *counters,
        unsigned long long *doneFlag, unsigned long long doneSeq, hipStream_t s) {
    if (count <= 0 || count > DVMH_MULTI_MAX) return;
    DISPATCH_TYPES(launchMulti, count, reds, counters, doneFlag, doneSeq, s)
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
