// Strided sub-block copy between two natural-layout arrays (gfx950).
// Used for sub-array actualisation and shadow-edge pack/unpack: the block
// [lo, hi] of each axis moves from one array to another; a dense destination
// is a pack into a staging buffer, a dense source an unpack from one.
// Semantics of rts/dvmh/cuda_copying.cu:46-303 (copyCutting0d..7d); here one
// kernel covers every rank: the innermost axis maps to consecutive threads
// (coalesced on both sides) and outer axes are decoded once per row.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.h"

namespace {

constexpr int MAXR = 8;

struct BlockDesc {
    int rank;
    long ext[MAXR];          // block extents
    long srcStride[MAXR], dstStride[MAXR];
    long srcBase, dstBase;   // element offsets of the block's first element
    long rows, rowLen;
};

template <typename T>
__global__ __launch_bounds__(256) void copyBlockKernel(const T *__restrict__ src, T *__restrict__ dst, BlockDesc d) {
    // one (row, column-chunk) per workgroup iteration; a row is the innermost axis
    const long chunks = (d.rowLen + 255) / 256;
    for (long job = blockIdx.x; job < d.rows * chunks; job += gridDim.x) {
        long row = job / chunks, c0 = (job % chunks) * 256;
        long rest = row, so = d.srcBase, doff = d.dstBase;
        for (int a = d.rank - 2; a >= 0; a--) {
            long c = rest % d.ext[a];
            rest /= d.ext[a];
            so += c * d.srcStride[a];
            doff += c * d.dstStride[a];
        }
        long x = c0 + threadIdx.x;
        if (x < d.rowLen) dst[doff + x * d.dstStride[d.rank - 1]] = src[so + x * d.srcStride[d.rank - 1]];
    }
}

struct B16 { uint64_t a, b; };

}  // namespace

extern "C" void dvmh_hip_copy_block(const void *src, const long *srcSizes, const long *srcOrigin, void *dst,
        const long *dstSizes, const long *dstOrigin, const long *lo, const long *hi, int rank, long typeSize,
        hipStream_t s) {
    BlockDesc d;
    d.rank = rank;
    long sc = 1, dc = 1;
    d.srcBase = 0;
    d.dstBase = 0;
    for (int a = rank - 1; a >= 0; a--) {
        d.ext[a] = hi[a] - lo[a] + 1;
        if (d.ext[a] <= 0) return;
        d.srcStride[a] = sc;
        d.dstStride[a] = dc;
        d.srcBase += (lo[a] - srcOrigin[a]) * sc;
        d.dstBase += (lo[a] - dstOrigin[a]) * dc;
        sc *= srcSizes[a];
        dc *= dstSizes[a];
    }
    d.rowLen = d.ext[rank - 1];
    d.rows = 1;
    for (int a = 0; a < rank - 1; a++) d.rows *= d.ext[a];
    long jobs = d.rows * ((d.rowLen + 255) / 256);
    unsigned blocks = (unsigned)(jobs < 65536 ? jobs : 65536);
    switch (typeSize) {
        case 1: hipLaunchKernelGGL(copyBlockKernel<uint8_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst, d); break;
        case 2: hipLaunchKernelGGL(copyBlockKernel<uint16_t>, dim3(blocks), dim3(256), 0, s, (const uint16_t *)src, (uint16_t *)dst, d); break;
        case 4: hipLaunchKernelGGL(copyBlockKernel<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t *)src, (uint32_t *)dst, d); break;
        case 8: hipLaunchKernelGGL(copyBlockKernel<uint64_t>, dim3(blocks), dim3(256), 0, s, (const uint64_t *)src, (uint64_t *)dst, d); break;
        case 16: hipLaunchKernelGGL(copyBlockKernel<B16>, dim3(blocks), dim3(256), 0, s, (const B16 *)src, (B16 *)dst, d); break;
        default: {
            // other record sizes: copy as bytes with one more (innermost) axis
            BlockDesc b = d;
            b.rank = rank + 1;
            for (int a = 0; a < rank; a++) {
                b.srcStride[a] = d.srcStride[a] * typeSize;
                b.dstStride[a] = d.dstStride[a] * typeSize;
            }
            b.ext[rank] = typeSize;
            b.srcStride[rank] = 1;
            b.dstStride[rank] = 1;
            b.srcBase = d.srcBase * typeSize;
            b.dstBase = d.dstBase * typeSize;
            b.rowLen = typeSize;
            b.rows = d.rows * d.rowLen;
            long j2 = b.rows * ((b.rowLen + 255) / 256);
            hipLaunchKernelGGL(copyBlockKernel<uint8_t>, dim3((unsigned)(j2 < 65536 ? j2 : 65536)), dim3(256), 0, s,
                    (const uint8_t *)src, (uint8_t *)dst, b);
        }
    }
}
