// Layout transforms of device arrays (gfx950): axis permutation and the
// natural <-> diagonal plane layout used by ACROSS loops with two or more
// dependent axes.  Semantics follow rts/dvmh/cuda_transform.cu (what is moved
// where); the kernels are new: a 64x64 LDS tile transpose for permutations that
// move the fastest axis, a gather kernel for the rest, and an LDS tile kernel
// for the diagonal layout.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "launch.h"
#include "dvmh_hip_device.h"

namespace {

constexpr int MAXR = 9;   // array rank + 1 (the byte axis of odd record sizes)

struct PermDesc {
    int rank;
    int rowThreads;              // lanes per destination row (a power of two up to 256)
    long dstSizes[MAXR];
    long srcStrideOfDst[MAXR];   // src element stride of the axis at each dst position
    long total;
};

// General permutation as a gather over destination rows (the innermost
// destination axis): the row's source offset is worked out once per row, then
// the row's lanes walk its elements with the source stride of that axis, so
// stores are contiguous and the index arithmetic is per row, not per element.
template <typename T>
__global__ __launch_bounds__(256) void permuteGather(const T *__restrict__ src, T *__restrict__ dst, PermDesc d) {
    const long n = d.dstSizes[d.rank - 1], sn = d.srcStrideOfDst[d.rank - 1];
    const long rows = d.total / n;
    const int perBlock = 256 / d.rowThreads;
    const int lr = threadIdx.x / d.rowThreads, lc = threadIdx.x % d.rowThreads;
    for (long row = (long)blockIdx.x * perBlock + lr; row < rows; row += (long)gridDim.x * perBlock) {
        long rest = row, off = 0;
        for (int p = d.rank - 2; p >= 0; p--) {
            const long c = rest % d.dstSizes[p];
            rest /= d.dstSizes[p];
            off += c * d.srcStrideOfDst[p];
        }
        const T *s = src + off;
        T *o = dst + row * n;
        for (long c = lc; c < n; c += d.rowThreads) o[c] = s[c * sn];
    }
}

static int rowThreadsFor(long n) {
    int t = 1;
    while (t < 256 && t < n) t <<= 1;
    return t;
}

static unsigned gatherGrid(const PermDesc &d) {
    const long rows = d.total / d.dstSizes[d.rank - 1], perBlock = 256 / d.rowThreads;
    const long b = (rows + perBlock - 1) / perBlock;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
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

// Natural <-> diagonal plane layout through DT x DT tiles of a plane staged in
// LDS.  The natural side moves whole tile rows (x contiguous); on the diagonal
// side each wave moves one diagonal of the tile at a time (x + y constant, or
// x - y for `slash`), whose elements convertXY places at consecutive positions,
// so both the loads and the stores of a wave cover contiguous runs.  The row
// pitch DT + 2 keeps the diagonal walks through LDS free of bank conflicts
// (element strides DT + 1 and DT + 3, both odd).
template <typename T, int DT>
__global__ __launch_bounds__(256) void diagonalTiles(const T *__restrict__ src, T *__restrict__ dst, long Rx, long Ry, long Rz,
        int back, int slash) {
    __shared__ T tile[DT][DT + 2];
    const long tilesX = (Rx + DT - 1) / DT, tilesY = (Ry + DT - 1) / DT;
    const long perPlane = tilesX * tilesY, plane = Rx * Ry;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long job = blockIdx.x; job < perPlane * Rz; job += gridDim.x) {
        const long z = job / perPlane, r = job % perPlane;
        const long y0 = (r / tilesX) * DT, x0 = (r % tilesX) * DT;
        const int w = (int)(Rx - x0 < DT ? Rx - x0 : DT), h = (int)(Ry - y0 < DT ? Ry - y0 : DT);
        const T *s = src + z * plane;
        T *d = dst + z * plane;
        for (int phase = 0; phase < 2; phase++) {
            // phase 0 reads the source side into LDS, phase 1 writes the other side
            const bool natural = (phase == 0) == !back;
            if (natural) {
                for (int e = threadIdx.x; e < DT * DT; e += 256) {
                    const int ly = e / DT, lx = e % DT;
                    if (lx < w && ly < h) {
                        const long g = (y0 + ly) * Rx + x0 + lx;
                        if (phase == 0) tile[ly][lx] = s[g];
                        else d[g] = tile[ly][lx];
                    }
                }
            } else {
                for (int t = wave; t < w + h - 1; t += 4) {
                    int lyLo, lyHi;
                    if (!slash) { lyLo = t - (w - 1) > 0 ? t - (w - 1) : 0; lyHi = t < h - 1 ? t : h - 1; }
                    else { lyLo = (h - 1) - t > 0 ? (h - 1) - t : 0; lyHi = (h - 1) - t + (w - 1) < h - 1 ? (h - 1) - t + (w - 1) : h - 1; }
                    const int ly = lyLo + lane;
                    if (ly > lyHi) continue;
                    const int lx = slash ? t - (h - 1) + ly : t - ly;
                    const long g = dvmh_dev::convertXY<long>(x0 + lx, y0 + ly, Rx, Ry, slash);
                    if (phase == 0) tile[ly][lx] = s[g];
                    else d[g] = tile[ly][lx];
                }
            }
            __syncthreads();
        }
    }
}

struct Bytes16 { uint64_t a, b; };


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
        if (rank >= MAXR || q < 0 || q >= rank || ((seen >> q) & 1u)) {
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
    d.rowThreads = rowThreadsFor(d.dstSizes[rank - 1]);
    unsigned blocks = gatherGrid(d);
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
            b.rowThreads = rowThreadsFor(typeSize);
            hipLaunchKernelGGL(permuteGather<uint8_t>, dim3(gatherGrid(b)), dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst, b);
        }
    }
}

extern "C" void dvmh_hip_diagonal(const void *src, void *dst, long typeSize, long Rx, long Ry, long Rz, int back,
        int slash, hipStream_t s) {
    long total = Rx * Ry * Rz;
    if (total == 0) return;
    const long jobs64 = ((Rx + 63) / 64) * ((Ry + 63) / 64) * Rz, jobs32 = ((Rx + 31) / 32) * ((Ry + 31) / 32) * Rz;
    const dim3 g64((unsigned)(jobs64 < 65536 ? jobs64 : 65536)), g32((unsigned)(jobs32 < 65536 ? jobs32 : 65536));
    switch (typeSize) {
        case 1: hipLaunchKernelGGL((diagonalTiles<uint8_t, 64>), g64, dim3(256), 0, s, (const uint8_t *)src, (uint8_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 2: hipLaunchKernelGGL((diagonalTiles<uint16_t, 64>), g64, dim3(256), 0, s, (const uint16_t *)src, (uint16_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 4: hipLaunchKernelGGL((diagonalTiles<uint32_t, 64>), g64, dim3(256), 0, s, (const uint32_t *)src, (uint32_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 8: hipLaunchKernelGGL((diagonalTiles<uint64_t, 64>), g64, dim3(256), 0, s, (const uint64_t *)src, (uint64_t *)dst, Rx, Ry, Rz, back, slash); break;
        case 16: hipLaunchKernelGGL((diagonalTiles<Bytes16, 32>), g32, dim3(256), 0, s, (const Bytes16 *)src, (Bytes16 *)dst, Rx, Ry, Rz, back, slash); break;
        default:
            fprintf(stderr, "dvmh_hip_diagonal: element size %ld has no diagonal kernel\n", typeSize);
            abort();
    }
}
