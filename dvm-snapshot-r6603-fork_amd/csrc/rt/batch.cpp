// Batching of many small loops into one fused launch per loop site, and the
// host slot count used by host handlers.
#include <algorithm>

#include "rt.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace dvmh {

long graphReplays = 0, batchGroups = 0, batchLoops = 0;
static bool batching = false;
static std::vector<Loop *> queued;

bool batchQueue(Loop *loop) {
    (void)loop;
    return false;
}

bool loopIsQueued(Loop *loop) { return std::find(queued.begin(), queued.end(), loop) != queued.end(); }

int hostSlotCount() {
    if (settings.hostThreads > 0) return settings.hostThreads;
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // namespace dvmh

using namespace dvmh;

extern "C" void dvmh_amd_batch_begin_C(void) { batching = true; }
extern "C" void dvmh_amd_batch_end_C(void) { batching = false; }
