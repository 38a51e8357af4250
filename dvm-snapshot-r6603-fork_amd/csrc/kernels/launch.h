// Launchers of the runtime's own HIP kernels (csrc/kernels/*.hip).  Host code
// of the runtime calls only these; every launcher enqueues on the given stream
// and returns without synchronising.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Fill count records of valueBytes bytes each with the same record
// (rts/dvmh/cuda_reduction.cu:21-82, performReplicate*).
void dvmh_hip_fill(void *dst, const void *value, long valueBytes, long count, hipStream_t s);

// Fold n partial values (and their loc records) into one result.
// func: rf_* code; rtType: rt_* element type.  Deterministic: fixed partition,
// fixed per-thread order and fixed tree; on equal MAX/MIN values the lower
// partial index wins.
// (replaces rts/dvmh/cuda_reduction.cu:168-312: no per-warp D2H and host fold)
// Large n is folded in two passes (blocks of 16384 partials, then one
// workgroup) through `scratch` of dvmh_hip_reduce_scratch_bytes(n) bytes.
long dvmh_hip_reduce_scratch_bytes(long n);
void dvmh_hip_reduce_finish(int func, int rtType, long n, const void *vals, const void *locs, long locSize,
        void *outVal, void *outLoc, void *scratch, hipStream_t s);

// Same for many independent segments: segment g is [offsets[g], offsets[g+1])
// (offsets in device memory), result g goes to outVals[g] / outLocs[g].
void dvmh_hip_reduce_segments(int func, int rtType, long groups, const long *offsets, const void *vals,
        const void *locs, long locSize, void *outVals, void *outLocs, hipStream_t s);

// Finish of up to DVMH_MULTI_MAX reductions of one element type (rtType) in
// one launch, deterministic like dvmh_hip_reduce_finish.  Partials are folded
// in blocks of 2048 into per-reduction slots (`slots`:
// dvmh_hip_multi_slot_bytes(n) bytes of device memory); the last block of a
// reduction to finish, found with counters[r] (zero before the launch, zero
// again after it), folds the slots and writes the value and the winner's loc
// record to outVal / outLoc, which may point into mapped pinned host memory.
#define DVMH_MULTI_MAX 8
struct DvmhMultiRed {
    int func;
    long n;
    const void *vals;
    const void *locs;
    long locSize;
    void *outVal;
    void *outLoc;
    void *slots;
};
long dvmh_hip_multi_slot_bytes(long n);
// counters: DVMH_MULTI_MAX + 1 words, zero between launches.  doneFlag (mapped
// host memory, may be 0) is set to doneSeq once every result of the launch is
// written, so the host can wait for the results without a stream sync.
void dvmh_hip_reduce_finish_multi(int rtType, int count, const struct DvmhMultiRed *reds, unsigned *counters,
        unsigned long long *doneFlag, unsigned long long doneSeq, hipStream_t s);

// Axis permutation copy.  src is stored with storage-position sizes srcSizes[0..rank-1]
// (position 0 outermost); dst storage position p holds src storage position
// dstFromSrc[p].  (rts/dvmh/cuda_transform.cu:273-652 semantics)
void dvmh_hip_permute(const void *src, void *dst, long typeSize, int rank, const long *srcSizes,
        const int *dstFromSrc, hipStream_t s);

// Natural <-> diagonal layout of Rz planes of Ry x Rx elements (x fastest).
// back = 0: natural -> diagonal; back = 1: diagonal -> natural.
// (rts/dvmh/cuda_transform.cu:26-73 semantics; index map of dvmhlib_device.h:56-98)
void dvmh_hip_diagonal(const void *src, void *dst, long typeSize, long Rx, long Ry, long Rz, int back, int slash,
        hipStream_t s);

// Strided sub-block copy (copyCutting semantics, rts/dvmh/cuda_copying.cu:46-303):
// copies the block [lo[a], hi[a]] of every axis a from an array with sizes
// srcSizes / origin srcOrigin to one with sizes dstSizes / origin dstOrigin
// (both row-major, natural layout).  A contiguous dst (dstSizes == block
// extents, dstOrigin == lo) is a pack; a contiguous src is an unpack.
void dvmh_hip_copy_block(const void *src, const long *srcSizes, const long *srcOrigin, void *dst, const long *dstSizes,
        const long *dstOrigin, const long *lo, const long *hi, int rank, long typeSize, hipStream_t s);

// Cost-model perceptron (mlp.hip): F inputs, two ReLU layers of H units, one
// output; parameters W1[H][F] b1[H] W2[H][H] b2[H] W3[H] b3 as one FP32 vector.
#define DVMH_MLP_F 24
#define DVMH_MLP_H 64
#define DVMH_MLP_NPARAM (DVMH_MLP_H * DVMH_MLP_F + DVMH_MLP_H + DVMH_MLP_H * DVMH_MLP_H + DVMH_MLP_H + DVMH_MLP_H + 1)
// Gradient of mean((out - y)^2) over n samples (X row-major n x F) into grad
// (NPARAM floats) and the loss into *loss; scratch of
// dvmh_hip_mlp_grad_scratch_bytes(n) bytes.  Deterministic.
long dvmh_hip_mlp_grad_scratch_bytes(long n);
void dvmh_hip_mlp_grad(const float *P, const float *X, const float *Y, long n, float *grad, float *loss, void *scratch,
        hipStream_t s);
// One Adam step; c1 = 1 - beta1^t, c2 = 1 - beta2^t for step t (from 1).
void dvmh_hip_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float beta1, float beta2, float eps, float c1,
        float c2, hipStream_t s);
// out[i] = network(X[i]) for n samples.
void dvmh_hip_mlp_forward(const float *P, const float *X, long n, float *out, hipStream_t s);

// Streaming copy (every byte read once and written once) and read-only pass
// (every byte read once; per-wave checksums of the data stored into `sink`,
// dvmh_hip_stream_sink_bytes(bytes) bytes of device memory) of `bytes`.
void dvmh_hip_stream_copy(void *dst, const void *src, long bytes, hipStream_t s);
long dvmh_hip_stream_sink_bytes(long bytes);
void dvmh_hip_stream_read(const void *src, long bytes, uint64_t *sink, hipStream_t s);

#ifdef __cplusplus
}
#endif
