// Reduction fold without any allocation: the variant of
// dvmh_amd_reduce_partials_C for callers that may not allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its scratch is a fixed array in the code object's device
// memory: every GPU the library runs on holds one copy until the process
// ends, and every call on that GPU uses it, so calls that overlap on one GPU
// race.  The fold needs no cleared memory: the two-pass fold writes every
// scratch record it reads.
#include <hip/hip_runtime.h>

#include "rt.h"
#include "launch.h"

namespace dvmh {

constexpr long NOALLOC_SCRATCH_BYTES = 1L << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` per device ordinal (each GPU holds its own copy
// at its own address); the array's name in host code is not that address.
static void *scratchPtr[NOALLOC_MAX_DEVICES];

// Called once per device with that device current (by the runtime's GPU
// set-up, and again by the entry point for a device set up elsewhere).
void noallocInit(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES || scratchPtr[device]) return;
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratchPtr[device] = p;
}

// Most partials the fixed scratch folds: the two-pass fold needs 24 bytes per
// block of 16384 partials plus 64.
static long noallocMaxPartials() {
    long n = ((NOALLOC_SCRATCH_BYTES - 64) / 24) * 16384;
    while (dvmh_hip_reduce_scratch_bytes(n) > NOALLOC_SCRATCH_BYTES) n -= 16384;
    return n;
}

void checkFoldArgs(const char *where, long func, long rtType, long n, const void *vals, const void *locs, long locSize,
        const void *outVal, const void *outLoc);

}  // namespace dvmh

using namespace dvmh;

extern "C" DvmType dvmh_amd_reduce_partials_noalloc_C(DvmType func, DvmType rtType, DvmType n, const void *vals, const void *locs,
        DvmType locSize, void *outVal, void *outLoc) {
    runtimeInit();
    if (!gpu.present) return -1;
    checkFoldArgs("dvmh_amd_reduce_partials_noalloc_C", (long)func, (long)rtType, (long)n, vals, locs, (long)locSize, outVal, outLoc);
    const long sb = dvmh_hip_reduce_scratch_bytes((long)n);
    if (sb > NOALLOC_SCRATCH_BYTES) {
        logMsg(LOG_ERROR, "dvmh_amd_reduce_partials_noalloc_C: %ld partials need %ld bytes of scratch, more than its fixed %ld "
               "(at most %ld partials); nothing was folded", (long)n, sb, NOALLOC_SCRATCH_BYTES, noallocMaxPartials());
        return -2;
    }
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    DVMH_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "dvmh_amd_reduce_partials_noalloc_C: device %d", dev);
    if (!scratchPtr[dev]) noallocInit(dev);
    const bool loc = func == rf_MAXLOC || func == rf_MINLOC;
    dvmh_hip_reduce_finish((int)func, (int)rtType, (long)n, vals, loc ? locs : nullptr, loc ? (long)locSize : 0, outVal,
            loc ? outLoc : nullptr, sb > 0 ? scratchPtr[dev] : nullptr, gpu.stream);
    HIP_CHECK(hipStreamSynchronize(gpu.stream));
    return 0;
}
