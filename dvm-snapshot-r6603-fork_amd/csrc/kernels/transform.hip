// Layout transforms of device arrays (gfx950): axis permutation and the
// natural <-> diagonal plane layout used by ACROSS loops with two or more
// dependent axes.  Semantics follow rts/dvmh/cuda_transform.cu (what is moved
// where); the kernels are new: a 64x64 LDS tile transpose for permutations that
// move the fastest axis, a gather kernel for the rest, and a row-walking
// diagonal kernel.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "launch.h"
#include "dvmh_hip_device.h"

namespace {

constexpr int MAXR = 8;

struct PermDesc {
    int rank;
    long dstSizes[MAXR];
    long srcStrideOfDst[MAXR];   // src element stride of the axis at each dst position
    long total;
};

template <typename T>
__global__ void permuteGather(const T *__restrict__ src, T *__restrict__ dst, PermDesc d) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < d.total; i += (long)gridDim.x * blockDim.x) {
        long rest = i, off = 0;
        for (int p = d.rank - 1; p >= 0; p--) {
            long c = rest % d.dstSizes[p];
            rest /= d.dstSizes[p];
            off += c * d.srcStrideOfDst[p];
        }
        dst[i] = src[off];
    }
}

// Transpose of the two fastest-moving axes of a batch: src viewed as
// [outer][A][mid][B] with B contiguous, dst as [outer][B][mid][A] with A contiguous.
constexpr int TT = 64;
template <typename T>
__global__ __launch_bounds__(256) void transposeTiles(const T *__restrict__ src, T *__restrict__ dst, long A, long mid,
        long B, long outerCount) {
    __shared__ T tile[TT][TT + 1];
    const long tilesA = (A + TT - 1) / TT, tilesB = (B + TT - 1) / TT;
    const long perOuter = tilesA * tilesB * mid;
    for (long job = blockIdx.x; job < perOuter * outerCount; job += gridDim.x) {
        long o = job / perOuter, r = job % perOuter;
        long m = r / (tilesA * tilesB);
        r %= tilesA * tilesB;
        long ta = r / tilesB, tb = r % tilesB;
        const T *s = src + o * A * mid * B + m * B;
        T *dd = dst + o * B * mid * A + m * A;
        // load: rows a (stride mid*B), columns b contiguous
        for (int e = threadIdx.x; e < TT * TT; e += 256) {
            int ra = e / TT, cb = e % TT;
            long a = ta * TT + ra, b = tb * TT + cb;
            if (a < A && b < B) tile[ra][cb] = s[a * mid * B + b];
        }
        __syncthreads();
        // store: rows b (stride mid*A), columns a contiguous
        for (int e = threadIdx.x; e < TT * TT; e += 256) {
            int rb = e / TT, ca = e % TT;
            long a = ta * TT + ca, b = tb * TT + rb;
            if (a < A && b < B) dd[b * mid * A + a] = tile[ca][rb];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void diagonalKernel(const T *__restrict__ src, T *__restrict__ dst, long Rx, long Ry, long Rz, int back,
        int slash) {
    const long plane = Rx * Ry;
    const long total = plane * Rz;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        long z = i / plane, r = i % plane;
        long y = r / Rx, x = r % Rx;
        long d = z * plane + dvmh_dev::convertXY<long>(x, y, Rx, Ry, slash);
        if (!back) dst[d] = src[i];
        else dst[i] = src[d];
    }
}

struct Bytes16 { uint64_t a, b; };

long gridFor(long total) {
    long b = (total + 255) / 256;
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

}  // namespace

// The diagonal index map the kernels use, callable from the host (tests, tools).
extern "C" long dvmh_amd_convert_xy_C(long x, long y, long Rx, long Ry, long slash) {
    return dvmh_dev::convertXY<long>(x, y, Rx, Ry, (int)slash);
}

extern "C" void dvmh_hip_permute(const void *src, void *dst, long typeSize, int rank, const long *srcSizes,
        const int *dstFromSrc, hipStream_t s) {
    // the gather stays inside the buffer only for a true permutation
    unsigned seen = 0;
    for (int p = 0; p < rank; p++) {
        const int q = dstFromSrc[p];
        if (rank > MAXR || q < 0 || q >= rank || ((seen >> q) & 1u)) {
            fprintf(stderr, "dvmh_hip_permute: not a permutation of %d axes\n", rank);
            abort();
        }
        seen |= 1u << q;
    }
    long srcStride[MAXR];
    long coll = 1;
    for (int q = rank - 1; q >= 0; q--) { srcStride[q] = coll; coll *= srcSizes[q]; }
    const long total = coll;
    if (total == 0) return;
    // Fast path: only the last two positions exchange their axes with some
    // axes in between kept in place -> batched tile transpose.
    // Pattern: dst = [outer..][B][mid..][A] where src = [outer..][A][mid..][B].
    int pa = -1;
    for (int q = 0; q < rank; q++) if (dstFromSrc[rank - 1] == q) pa = q;
    bool tileable = pa >= 0 && pa < rank - 1 && dstFromSrc[pa] == rank - 1;
    for (int q = 0; q < rank && tileable; q++)
        if (q != pa && q != rank - 1 && dstFromSrc[q] != q) tileable = false;
    if (tileable && (typeSize == 8 || typeSize == 4)) {
        long outer = 1, mid = 1;
        for (int q = 0; q < pa; q++) outer *= srcSizes[q];
        for (int q = pa + 1; q < rank - 1; q++) mid *= srcSizes[q];
        long A = srcSizes[pa], B = srcSizes[rank - 1];
        long jobs = outer * mid * ((A + TT - 1) / TT) * ((B + TT - 1) / TT);
        unsigned blocks = (unsigned)(jobs < 65536 ? jobs : 65536);
        if (typeSize == 8)
            hipLaunchKernelGGL(transposeTiles<double>, dim3(blocks), dim3(256), 0, s, (const double *)src, (double *)dst,
                    A, mid, B, outer);
        else
            hipLaunchKernelGGL(transposeTiles<float>, dim3(blocks), dim3(256), 0, s, (const float *)src, (float *)dst, A,
                    mid, B, outer);
        return;
    }
    PermDesc d;
    d.rank = rank;
    d.total = total;
    for (int p = 0; p < rank; p++) {
        d.dstSizes[p] = srcSizes[dstFromSrc[p]];
        d.srcStrideOfDst[p] = srcStride[dstFromSrc[p]];
    }
    unsigned blocks = (unsigned)gridFor(total);
    switch (typeSize) {
        case 1: hipLaunchKernelGGL(permuteGather<uint8_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst, d); break;
        case 2: hipLaunchKernelGGL(permuteGather<uint16_t>, dim3(blocks), dim3(256), 0, s, (const uint16_t *)src, (uint16_t *)dst, d); break;
        case 4: hipLaunchKernelGGL(permuteGather<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t *)src, (uint32_t *)dst, d); break;
        case 8: hipLaunchKernelGGL(permuteGather<uint64_t>, dim3(blocks), dim3(256), 0, s, (const uint64_t *)src, (uint64_t *)dst, d); break;
        case 16: hipLaunchKernelGGL(permuteGather<Bytes16>, dim3(blocks), dim3(256), 0, s, (const Bytes16 *)src, (Bytes16 *)dst, d); break;
        default: {
            // arbitrary record size: permute bytes with one extra fastest axis
            PermDesc b;
            b.rank = rank + 1;
            b.total = total * typeSize;
            for (int p = 0; p < rank; p++) {
                b.dstSizes[p] = d.dstSizes[p];
                b.srcStrideOfDst[p] = d.srcStrideOfDst[p] * typeSize;
            }
            b.dstSizes[rank] = typeSize;
            b.srcStrideOfDst[rank] = 1;
            hipLaunchKernelGGL(permuteGather<uint8_t>, dim3((unsigned)gridFor(b.total)), dim3(256), 0, s,
                    (const uint8_t *)src, (uint8_t *)dst, b);
        }
    }
}

extern "C" void dvmh_hip_diagonal(const void *src, void *dst, long typeSize, long Rx, long Ry, long Rz, int back,
        int slash, hipStream_t s) {
    long total = Rx * Ry * Rz;
    if (total == 0) return;
    unsigned blocks = (unsigned)gridFor(total);
    switch (typeSize) {
        case 1: hipLaunchKernelGGL(diagonalKernel<uint8_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 2: hipLaunchKernelGGL(diagonalKernel<uint16_t>, dim3(blocks), dim3(256), 0, s, (const uint16_t *)src, (uint16_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 4: hipLaunchKernelGGL(diagonalKernel<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t *)src, (uint32_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 8: hipLaunchKernelGGL(diagonalKernel<uint64_t>, dim3(blocks), dim3(256), 0, s, (const uint64_t *)src, (uint64_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 16: hipLaunchKernelGGL(diagonalKernel<Bytes16>, dim3(blocks), dim3(256), 0, s, (const Bytes16 *)src, (Bytes16 *)dst, Rx, Ry, Rz, back, slash); break;
        default:
            fprintf(stderr, "dvmh_hip_diagonal: element size %ld has no diagonal kernel\n", typeSize);
            abort();
    }
}
