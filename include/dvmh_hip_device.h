/*
 * Device-side helpers for HIP loop handlers emitted for libdvmh_amd (gfx950).
 *
 *   View       array access A[i][j][k] through the natural base and the
 *              per-axis element strides of dvmh_fill_header_C; any axis
 *              permutation of the device layout is just another stride set.
 *   DiagView   the same for a layout whose two fastest storage axes are
 *              diagonalised (planes stored anti-diagonal by anti-diagonal), the
 *              index map of rts/dvmh/include/dvmhlib_device.h:56-98.
 *   MarchView2/3, TileView3, LineView
 *              the same syntax over copies staged in LDS: a rolling window of
 *              three planes of a read-only stencil array (march variant of
 *              PARALLEL loops), a tile with halo of an ACROSS array (tiled
 *              wavefront), and a chunk of one line of an ACROSS array (tiled
 *              line sweep along the innermost axis).
 *   reductions wave64 shuffle trees plus one LDS step across the waves of a
 *              workgroup; every thread of the workgroup must call them.
 *   xcdTile    XCD-aware renumbering of workgroups: the hardware deals
 *              consecutive workgroup ids round-robin over the 8 XCDs, so
 *              renumbering gives each XCD (and its L2) a contiguous run of tiles.
 *
 * The role of the reference's cdvmh/include/cdvmh_helpers.h (DvmhArrayHelper,
 * DvmhDiagonalizedArrayHelper) and rts/dvmh/include/dvmhlib_{warp,block}_red.h,
 * written for 64-wide wavefronts.
 */
#ifndef DVMH_HIP_DEVICE_H
#define DVMH_HIP_DEVICE_H

#include <hip/hip_runtime.h>

#include "dvmhlib_types.h"
#include "dvmhlib_const.h"

namespace dvmh_dev {

// ------------------------------------------------------------------ resources
// Registers per thread of a compiled kernel (0 if the runtime cannot say): the
// figure handlers pass to dvmh_loop_cuda_get_config_C so the block is fitted to
// the register file, as the reference's handlers pass the counts ptxas reports
// (driver/src/dvm_comp:400-432 -> rts/dvmh/loop.cpp:2177-2194).
static inline int kernelRegs(const void *kernel) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, kernel) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return a.numRegs > 0 ? a.numRegs : 0;
}

// ------------------------------------------------------------------ views
template <typename T, typename I, int R, bool UNIT>
struct View {
    T *base;
    I s[R];
    __device__ __forceinline__ View<T, I, R - 1, UNIT> operator[](I i) const {
        View<T, I, R - 1, UNIT> v;
        v.base = base + i * s[0];
#pragma unroll
        for (int k = 0; k < R - 1; k++) v.s[k] = s[k + 1];
        return v;
    }
};

template <typename T, typename I, bool UNIT>
struct View<T, I, 1, UNIT> {
    T *base;
    I s[1];
    __device__ __forceinline__ T &operator[](I i) const { return UNIT ? base[i] : base[i * s[0]]; }
};

// A device address held as an integer (descriptor tables of batch kernels) as
// a pointer the compiler knows to be global memory: the cast through address
// space 1 lets its loads and stores compile to global_* rather than flat_*
// instructions (no LDS/scratch aperture check, no LGKM counter wait).
template <typename T>
__device__ __forceinline__ T *globalPtr(long a) {
    return (T *)(__attribute__((address_space(1))) T *)a;
}

template <typename T, typename I, int R, bool UNIT>
__device__ __forceinline__ View<T, I, R, UNIT> makeView(T *base, const I *strides) {
    View<T, I, R, UNIT> v;
    v.base = base;
#pragma unroll
    for (int k = 0; k < R; k++) v.s[k] = strides[k];
    return v;
}

// ------------------------------------------------------------------ LDS march views
// A read-only stencil array staged in LDS as a rolling window of three planes
// of the outermost loop axis (planes i-1, i, i+1 of the point being computed).
// A plane holds `rows` x `cols` elements: the workgroup's tile of the inner
// axes plus a halo of one on each side; row 0 / column 0 hold array indices
// j0 / k0.  `A[i'][j'][k']` with |i' - i| <= 1 reads the LDS copy, so a loop
// body written against the global array runs unchanged.
template <typename T, typename I>
struct MarchRow {
    const T *p;
    I k0;
    __device__ __forceinline__ const T &operator[](I k) const { return p[k - k0]; }
};

template <typename T, typename I>
struct MarchPlane {
    const T *p;
    I cols, j0, k0;
    __device__ __forceinline__ MarchRow<T, I> operator[](I j) const { return MarchRow<T, I>{p + (j - j0) * cols, k0}; }
};

__device__ __forceinline__ int marchSlot(int cur, int delta) {
    int s = cur + delta;
    return s < 0 ? s + 3 : (s > 2 ? s - 3 : s);
}

// rank 3: planes of rows x cols
template <typename T, typename I>
struct MarchView3 {
    const T *lds;
    I rows, cols, j0, k0;
    I iCur;
    int sCur;
    __device__ __forceinline__ MarchPlane<T, I> operator[](I i) const {
        return MarchPlane<T, I>{lds + (I)marchSlot(sCur, (int)(i - iCur)) * rows * cols, cols, j0, k0};
    }
};

// rank 2: planes are single rows of cols elements
template <typename T, typename I>
struct MarchView2 {
    const T *lds;
    I cols, k0;
    I iCur;
    int sCur;
    __device__ __forceinline__ MarchRow<T, I> operator[](I i) const {
        return MarchRow<T, I>{lds + (I)marchSlot(sCur, (int)(i - iCur)) * cols, k0};
    }
};

// A tile of a rank-3 array staged in LDS (read and written): element (0,0,0)
// holds array index (i0, j0, k0), planes of rows x cols elements.
template <typename T, typename I>
struct TileRow {
    T *p;
    I k0;
    __device__ __forceinline__ T &operator[](I k) const { return p[k - k0]; }
};

template <typename T, typename I>
struct TilePlane {
    T *p;
    I cols, j0, k0;
    __device__ __forceinline__ TileRow<T, I> operator[](I j) const { return TileRow<T, I>{p + (j - j0) * cols, k0}; }
};

template <typename T, typename I>
struct TileView3 {
    T *p;
    I rows, cols, i0, j0, k0;
    __device__ __forceinline__ TilePlane<T, I> operator[](I i) const {
        return TilePlane<T, I>{p + (i - i0) * rows * cols, cols, j0, k0};
    }
};

// One line of an ACROSS array staged in LDS, for bodies that index it as
// X[i][j][k + d] with the loop's own i, j and |d| <= 1 along the line: the
// leading subscripts select nothing (they are the line the thread owns), the
// last one maps array index k to slot (k - kb) * dir of the line's chunk.
template <typename T, typename I, int D>
struct LineView {
    T *p;
    I kb, dir;
    __device__ __forceinline__ LineView<T, I, D - 1> operator[](I) const { return LineView<T, I, D - 1>{p, kb, dir}; }
};

template <typename T, typename I>
struct LineView<T, I, 1> {
    T *p;
    I kb, dir;
    __device__ __forceinline__ T &operator[](I k) const { return p[(k - kb) * dir]; }
};

// A thread's chunk of its own line of an ACROSS array held in a local buffer
// (the line sweep with register staging, lines along a non-innermost axis):
// for X[i][j][k] with the loop's own indices, the subscript on axis AX selects
// slot (x - x0) * DIR + 1 of the buffer, the others select nothing (they are
// the thread's line).  With the chunk's steps unrolled every slot index is a
// compile-time constant, so the buffer is promoted to registers.
template <typename T, typename I, int D, int AX, int DIR>
struct RegLineView {
    T *buf;
    I x0, slot;
    __device__ __forceinline__ RegLineView<T, I, D - 1, AX - 1, DIR> operator[](I x) const {
        return RegLineView<T, I, D - 1, AX - 1, DIR>{buf, x0, AX == 0 ? (x - x0) * DIR + 1 : slot};
    }
};

template <typename T, typename I, int AX, int DIR>
struct RegLineView<T, I, 1, AX, DIR> {
    T *buf;
    I x0, slot;
    __device__ __forceinline__ T &operator[](I x) const { return buf[AX == 0 ? (x - x0) * DIR + 1 : slot]; }
};

// Diagonal position inside an Rx x Ry plane (x fastest).  Also compiled for
// the host, so the layout-transform kernels, the handlers and host-side checks
// share one definition.
template <typename I>
__host__ __device__ __forceinline__ I convertXY(I x, I y, I Rx, I Ry, int slash) {
    I idx;
    if (!slash) {
        I d = x + y;
        if (Rx == Ry) {
            if (d < Rx) idx = y + (1 + d) * d / 2;
            else idx = Rx * (Rx - 1) + x - (2 * Rx - d - 1) * (2 * Rx - d - 2) / 2;
        } else if (Rx < Ry) {
            if (d < Rx) idx = y + (1 + d) * d / 2;
            else if (d < Ry) idx = (1 + Rx) * Rx / 2 + Rx - x - 1 + Rx * (d - Rx);
            else idx = Rx * Ry - Ry + y - (Rx + Ry - d - 1) * (Rx + Ry - d - 2) / 2;
        } else {
            if (d < Ry) idx = x + (1 + d) * d / 2;
            else if (d < Rx) idx = (1 + Ry) * Ry / 2 + (Ry - y - 1) + Ry * (d - Ry);
            else idx = Rx * Ry - Rx + x - (Rx + Ry - d - 1) * (Rx + Ry - d - 2) / 2;
        }
    } else {
        if (Rx == Ry) {
            I e = x + Rx - 1 - y;
            if (e < Rx) idx = Rx - 1 - y + (e + 1) * e / 2;
            else idx = Rx * (Rx - 1) + x - (Rx - x + y) * (Rx - x + y - 1) / 2;
        } else if (Rx < Ry) {
            I e = x + Ry - 1 - y;
            if (e < Rx) idx = Ry - 1 - y + (e + 1) * e / 2;
            else if (e < Ry) idx = (1 + Rx) * Rx / 2 + Rx - x - 1 + Rx * (e - Rx);
            else idx = Rx * Ry - 1 - y - (Rx + y - x) * (Rx + y - x - 1) / 2;
        } else {
            I e = x + Ry - 1 - y;
            if (e < Ry) idx = x + (1 + e) * e / 2;
            else if (e < Rx) idx = (1 + Ry) * Ry / 2 + y + Ry * (x - y - 1);
            else idx = Rx * Ry - Rx + x - (Rx + y - x) * (Rx + y - x - 1) / 2;
        }
    }
    return idx;
}

// Header extension of a diagonalised array (dvmh_fill_header_C's extendedParams).
template <typename I>
struct DiagInfo {
    int xAxis, yAxis, slash;   // 1-based original axes
    I xFirst, xLen, yFirst, yLen;
};

template <typename T, typename I, int R, int D>
struct DiagView {
    T *base;
    I s[R];
    I off, x, y;
    DiagInfo<I> di;
    __device__ __forceinline__ DiagView<T, I, R, D + 1> operator[](I i) const {
        DiagView<T, I, R, D + 1> v;
        v.base = base;
#pragma unroll
        for (int k = 0; k < R; k++) v.s[k] = s[k];
        v.di = di;
        v.off = off + i * s[D];
        v.x = (di.xAxis == D + 1) ? i : x;
        v.y = (di.yAxis == D + 1) ? i : y;
        return v;
    }
};

template <typename T, typename I, int R>
struct DiagView<T, I, R, R - 1> {
    T *base;
    I s[R];
    I off, x, y;
    DiagInfo<I> di;
    __device__ __forceinline__ T &operator[](I i) const {
        I o = off + i * s[R - 1];
        I xx = (di.xAxis == R) ? i : x, yy = (di.yAxis == R) ? i : y;
        return base[o + convertXY<I>(xx - di.xFirst, yy - di.yFirst, di.xLen, di.yLen, di.slash)];
    }
};

template <typename T, typename I, int R>
__device__ __forceinline__ DiagView<T, I, R, 0> makeDiagView(T *base, const I *strides, const DiagInfo<I> &di) {
    DiagView<T, I, R, 0> v;
    v.base = base;
#pragma unroll
    for (int k = 0; k < R; k++) v.s[k] = strides[k];
    v.off = 0;
    v.x = 0;
    v.y = 0;
    v.di = di;
    return v;
}

// ------------------------------------------------------------------ reductions
// Lane shuffle of any reduction element type (1, 2, 4 or 8 bytes, signed or
// unsigned, integer or floating), bit for bit.
template <typename T>
__device__ __forceinline__ T shflDown(T v, int o) {
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "reduction element size");
    if constexpr (sizeof(T) == 8) {
        long long x;
        __builtin_memcpy(&x, &v, 8);
        x = __shfl_down(x, o, 64);
        __builtin_memcpy(&v, &x, 8);
    } else {
        int x = 0;
        __builtin_memcpy(&x, &v, sizeof(T));
        x = __shfl_down(x, o, 64);
        __builtin_memcpy(&v, &x, sizeof(T));
    }
    return v;
}

struct OpSum {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpProd {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a * b; }
};
struct OpMax {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a < b ? b : a; }
};
struct OpMin {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? b : a; }
};
struct OpAnd {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a & b; }
};
struct OpOr {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a | b; }
};
struct OpXor {
    template <typename T> __device__ __forceinline__ T operator()(T a, T b) const { return a ^ b; }
};

// Initial values of a reduction array (N elements), passed to a kernel by value.
template <typename T, int N>
struct RedVec {
    T v[N];
};
template <typename T, int N>
__host__ __device__ inline RedVec<T, N> redVec(const T (&a)[N]) {
    RedVec<T, N> r;
    for (int i = 0; i < N; i++) r.v[i] = a[i];
    return r;
}

// Reduce over the workgroup; the result is valid in thread 0.  Fixed order:
// shuffle tree inside each wave, then wave results in wave order.
template <typename T, typename Op>
__device__ __forceinline__ T blockReduce(T v, Op op) {
    __shared__ T part[16];
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T other = shflDown(v, o);
        if (lane + o < 64 && t + o < nthreads) v = op(v, other);
    }
    __syncthreads();
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (t == 0) {
        const int nw = (nthreads + 63) >> 6;
        for (int w = 1; w < nw; w++) v = op(v, part[w]);
    }
    __syncthreads();
    return v;
}

// MAXLOC / MINLOC: larger (smaller) value wins, equal values keep the lower loc.
template <typename T, typename L, bool MAXF>
__device__ __forceinline__ void blockReduceLoc(T &v, L &loc) {
    __shared__ T pv[16];
    __shared__ L pl[16];
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T ov = shflDown(v, o);
        L ol = shflDown(loc, o);
        if (lane + o < 64 && t + o < nthreads) {
            bool take = MAXF ? (v < ov || (!(ov < v) && ol < loc)) : (v > ov || (!(ov > v) && ol < loc));
            if (take) { v = ov; loc = ol; }
        }
    }
    __syncthreads();
    if (lane == 0) { pv[wave] = v; pl[wave] = loc; }
    __syncthreads();
    if (t == 0) {
        const int nw = (nthreads + 63) >> 6;
        for (int w = 1; w < nw; w++) {
            bool take = MAXF ? (v < pv[w] || (!(pv[w] < v) && pl[w] < loc)) : (v > pv[w] || (!(pv[w] > v) && pl[w] < loc));
            if (take) { v = pv[w]; loc = pl[w]; }
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------ scheduling
// Workgroup id -> logical tile id such that each XCD gets a contiguous range.
__device__ __forceinline__ long xcdTile(long wg, long n, int enable) {
    if (!enable) return wg;
    const long X = 8;
    long per = n / X, rem = n % X;
    long x = wg % X, q = wg / X;
    return x < rem ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

}  // namespace dvmh_dev

#endif
