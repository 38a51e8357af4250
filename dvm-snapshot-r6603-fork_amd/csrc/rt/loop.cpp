// Loop execution: handler trampoline, dvmh_loop_perform_C's sequence, GPU
// reductions and their (optionally deferred) posting to host variables.
#include <algorithm>
#include <cstring>
#include <memory>
#include <utility>

#include "rt.h"
#include "../kernels/launch.h"

namespace dvmh {

// ------------------------------------------------------------ trampoline
// The runtime calls a registered handler as f(pLoopRef, params...)
// (rts/dvmh/util.cpp:659-720).  Every parameter is pointer sized, so the call
// goes through a non-variadic pointer type of the right arity.
namespace {
template <size_t... I>
void callWith(void *f, DvmType *ref, void *const *a, std::index_sequence<I...>) {
    using Fn = void (*)(DvmType *, decltype((void)I, (void *)nullptr)...);
    ((Fn)f)(ref, a[I]...);
}
template <size_t N>
void callN(void *f, DvmType *ref, void *const *a) { callWith(f, ref, a, std::make_index_sequence<N>()); }

template <size_t... N>
constexpr auto makeTable(std::index_sequence<N...>) {
    return std::array<void (*)(void *, DvmType *, void *const *), sizeof...(N)>{&callN<N>...};
}
}  // namespace
}  // namespace dvmh

#include <array>

namespace dvmh {

void callHandler(const HandlerDesc &h, DvmType *pLoopRef, const std::vector<void *> *override) {
    static const auto table = makeTable(std::make_index_sequence<MAX_PARAMS + 1>());
    const std::vector<void *> &p = override ? *override : h.params;
    DVMH_CHECK(p.size() <= (size_t)MAX_PARAMS, "Handler with %zu parameters is not supported (max %d)", p.size(), MAX_PARAMS);
    table[p.size()]((void *)h.func, pLoopRef, p.data());
}

// ------------------------------------------------------------ loop helpers
Loop::~Loop() { magic = 0; }

long Loop::iterations() const {
    long n = 1;
    for (int i = 0; i < rank; i++) {
        long c = step[i] > 0 ? (hi[i] - lo[i]) / step[i] + 1 : (lo[i] - hi[i]) / (-step[i]) + 1;
        if (c <= 0) return 0;
        n *= c;
    }
    return n;
}

// Alignment of an array to the object a loop is mapped on: follow Data::align
// links (see abi.cpp) to compose array axis -> template axis.
extern bool alignToRoot(const Data *d, const Object *root, long tAxisOfArrayAxis[], long sign[]);

// Array axis (0-based index) -> signed 1-based loop axis, 0 if none.
std::vector<long> Loop::correspondence(const Data *d) const {
    std::vector<long> res(d->rank, 0);
    for (auto &c : correspondences)
        if (c.first == d) return c.second;
    const Object *root = mapData ? (const Object *)mapData : (const Object *)mapTempl;
    if (root) {
        long tAxis[MAX_RANK], sgn[MAX_RANK];
        if (alignToRoot(d, root, tAxis, sgn)) {
            for (int a = 0; a < d->rank; a++) {
                long t = tAxis[a];
                if (t <= 0) continue;
                long la = mapAxis[t - 1];
                if (la > 0) res[a] = (mapMult[t - 1] * sgn[a] < 0) ? -la : la;
            }
            return res;
        }
    }
    // unmapped loop: identity on the common leading axes
    for (int a = 0; a < d->rank && a < rank; a++) res[a] = a + 1;
    return res;
}

// DvmhLoop::renewDependencyMask (rts/dvmh/loop.cpp:1117-1150): bit (rank - loopAxis)
// for every loop axis tied to an array axis with a non-zero ACROSS width.
void Loop::computeDependencyMask() {
    depMask = 0;
    long firstDep = -1;
    for (auto &ac : acrosses) {
        std::vector<long> corr = correspondence(ac.data);
        for (int j = 0; j < ac.data->rank; j++) {
            if (ac.low[j] > 0 || ac.high[j] > 0) {
                long la = corr[j] < 0 ? -corr[j] : corr[j];
                DVMH_CHECK(la > 0, "ACROSS array has dependency on dimension %d, which can not be tied with the parallel loop", j + 1);
                if (firstDep == -1 || la < firstDep) firstDep = la;
                long cnt = step[la - 1] > 0 ? (hi[la - 1] - lo[la - 1]) / step[la - 1] + 1 : (lo[la - 1] - hi[la - 1]) / (-step[la - 1]) + 1;
                if (cnt > 1) depMask |= 1L << (rank - la);
            }
        }
    }
    if (firstDep > 0 && depMask == 0) depMask |= 1L << (rank - firstDep);
}

// ------------------------------------------------------------ deferred posts
// A GPU reduction result lands in pinned host memory; it is merged into the
// program's variable at the next synchronisation point: right after the loop
// by default, or when results are needed with DVMH_AMD_SYNC=1.
struct PendingPost {
    std::shared_ptr<Reduction> red;
    char *val;
    char *loc;
    bool addInitial;
};
static std::vector<PendingPost> pending;

struct PinnedPool {
    std::vector<char *> blocks, devBlocks;   // host addresses and their device-side aliases
    size_t used = 0, blockSize = 1 << 20;
    char *take(size_t bytes) {
        bytes = (bytes + 63) & ~(size_t)63;
        size_t b = used / blockSize, off = used % blockSize;
        if (off + bytes > blockSize) { b++; off = 0; used = b * blockSize; }
        while (b >= blocks.size()) {
            void *p = nullptr, *d = nullptr;
            HIP_CHECK(hipHostMalloc(&p, blockSize, hipHostMallocDefault));
            HIP_CHECK(hipHostGetDevicePointer(&d, p, 0));
            blocks.push_back((char *)p);
            devBlocks.push_back((char *)d);
        }
        used += bytes;
        return blocks[b] + off;
    }
    char *device(char *h) const {
        for (size_t b = 0; b < blocks.size(); b++)
            if (h >= blocks[b] && h < blocks[b] + blockSize) return devBlocks[b] + (h - blocks[b]);
        return nullptr;
    }
    void reset() { used = 0; }
};
static PinnedPool pinnedPool;

char *pinnedTake(size_t bytes) { return pinnedPool.take(bytes); }

// Device-side address of mapped pinned host memory (pool blocks or graph slots).
char *mappedDevicePtr(char *h) {
    if (char *d = pinnedPool.device(h)) return d;
    void *d = nullptr;
    HIP_CHECK(hipHostGetDevicePointer(&d, h, 0));
    return (char *)d;
}

// Completion word value that covers every pending post, 0 when one of them
// came from a path without it (batches, graph replays): then the stream is
// synchronised.
static unsigned long long pendingDone = 0;

void flushPostsFor(const void *addr) {
    for (auto &p : pending)
        if (p.red->arrayAddr == addr || p.red->locAddr == addr) {
            flushDeferredPosts();
            return;
        }
}

long pendingPostCount() { return (long)pending.size(); }

void flushDeferredPosts() {
    if (pending.empty()) return;
    if (pendingDone && settings.pollDone) gpu.waitDone(pendingDone);
    else gpu.sync();
    pendingDone = 0;
    for (auto &p : pending) {
        p.red->postValues(p.val, p.loc);
        if (p.addInitial) p.red->postValues(p.red->initArr.data(), p.red->isLoc() ? p.red->initLoc.data() : nullptr, false);
    }
    pending.clear();
    pinnedPool.reset();
}

void addPendingPost(const std::shared_ptr<Reduction> &r, char *val, char *loc, bool addInitial) {
    pending.push_back(PendingPost{r, val, loc, addInitial});
    pendingDone = 0;   // until a completion word covers it
}

// ------------------------------------------------------------ GPU reductions
// DvmhReductionCuda::prepare (rts/dvmh/loop.cpp:274-297): `count` partial slots
// per reduction element, optionally pre-filled with the handler-local initial
// value.  A reduction array of L elements keeps its partials element-major:
// element e owns slots [e * count, (e + 1) * count) (the layout the reference's
// array fold walks, rts/dvmh/cuda_reduction.cu:182-199).
void gpuRedPrepare(SpecLoop *sl, Reduction *r, long count, bool fill) {
    DVMH_CHECK(count > 0, "Reduction partial count must be positive");
    DVMH_CHECK(r->gpuArr == nullptr, "dvmh_loop_cuda_red_prepare_C called twice for one reduction");
    r->items = count;
    r->gpuArr = gpu.alloc((size_t)(r->elemCount * r->elemSize * count));
    if (r->gpuArrPtr) *r->gpuArrPtr = r->gpuArr;
    if (r->isLoc()) {
        r->gpuLoc = gpu.alloc((size_t)(r->elemCount * r->locSize * count));
        if (r->gpuLocPtr) *r->gpuLocPtr = r->gpuLoc;
    }
    if (fill) {
        for (long e = 0; e < r->elemCount; e++) {
            dvmh_hip_fill((char *)r->gpuArr + e * count * r->elemSize, r->localArr.data() + e * r->elemSize, r->elemSize, count,
                    sl->stream);
            if (r->isLoc())
                dvmh_hip_fill((char *)r->gpuLoc + e * count * r->locSize, r->localLoc.data() + e * r->locSize, r->locSize, count,
                        sl->stream);
        }
    }
}

// DvmhReductionCuda::finish (rts/dvmh/loop.cpp:301-408): fold the partials on
// the device and merge the result into the variable.  The finish requests of a
// handler are collected and run together when the handler returns: one launch
// folds every reduction of an element type and writes the results straight
// into mapped pinned host memory (no separate copy).
void gpuRedFinish(SpecLoop *sl, const std::shared_ptr<Reduction> &r) {
    DVMH_CHECK(r->gpuArr != nullptr, "dvmh_loop_cuda_red_finish_C without dvmh_loop_cuda_red_prepare_C");
    DVMH_CHECK(r->elemCount >= 1, "Reduction of %ld elements", r->elemCount);
    DVMH_CHECK(r->rtType == rt_CHAR || r->rtType == rt_UCHAR || r->rtType == rt_INT || r->rtType == rt_UINT ||
                    r->rtType == rt_LOGICAL || r->rtType == rt_LONG || r->rtType == rt_ULONG || r->rtType == rt_LLONG ||
                    r->rtType == rt_ULLONG || r->rtType == rt_FLOAT || r->rtType == rt_DOUBLE,
            "Reduction element type %d is not supported on the GPU", r->rtType);
    for (auto &q : sl->finishQueue)
        DVMH_CHECK(q != r, "dvmh_loop_cuda_red_finish_C called twice for one reduction");
    // the pointers registered by dvmh_loop_cuda_register_red_C are the handler's
    // locals: clear them now, while the handler is still running
    if (r->gpuArrPtr) *r->gpuArrPtr = nullptr;
    if (r->gpuLocPtr) *r->gpuLocPtr = nullptr;
    r->gpuArrPtr = nullptr;
    r->gpuLocPtr = nullptr;
    sl->finishQueue.push_back(r);
}

void gpuRedFinishFlush(SpecLoop *sl) {
    auto &q = sl->finishQueue;
    if (q.empty()) return;
    // One folded record per reduction, posted in the order the handler asked
    // for the finishes (a captured graph's replays re-post them in that order):
    // its L values, then (16-byte aligned) its L loc records.
    struct Entry {
        Reduction *r;
        long e;
        char *outVal, *outLoc;
    };
    std::vector<Entry> entries;
    for (auto &r : q) {
        const long L = r->elemCount;
        const size_t locOff = ((size_t)(L * r->elemSize) + 15) & ~(size_t)15;
        const size_t recBytes = r->isLoc() ? locOff + (size_t)(L * r->locSize) : (size_t)(L * r->elemSize);
        // a captured finish writes into a slot owned by the graph
        char *host = graphCapturing() ? graphPinnedSlot(recBytes) : pinnedPool.take(recBytes);
        char *hostLoc = r->isLoc() ? host + locOff : nullptr;
        if (graphCapturing()) graphNotePost(r, host, hostLoc);
        char *dev = mappedDevicePtr(host);
        for (long e = 0; e < L; e++)
            entries.push_back(Entry{r.get(), e, dev + e * r->elemSize, r->isLoc() ? dev + locOff + e * r->locSize : nullptr});
        addPendingPost(r, host, hostLoc);
    }
    // every element of every reduction is folded by one multi-finish entry;
    // a launch takes up to DVMH_MULTI_MAX entries of one element type
    std::vector<int> types;
    for (auto &r : q)
        if (std::find(types.begin(), types.end(), r->rtType) == types.end()) types.push_back(r->rtType);
    std::vector<std::vector<const Entry *>> launches;
    for (int type : types) {
        std::vector<const Entry *> cur;
        for (const Entry &en : entries) {
            if (en.r->rtType != type) continue;
            cur.push_back(&en);
            if ((int)cur.size() == DVMH_MULTI_MAX) { launches.push_back(cur); cur.clear(); }
        }
        if (!cur.empty()) launches.push_back(cur);
    }
    for (size_t li = 0; li < launches.size(); li++) {
        const auto &group = launches[li];
        DvmhMultiRed d[DVMH_MULTI_MAX];
        std::vector<void *> slots;
        for (size_t i = 0; i < group.size(); i++) {
            const Entry &en = *group[i];
            Reduction *r = en.r;
            void *sv = gpu.alloc((size_t)dvmh_hip_multi_slot_bytes(r->items));
            slots.push_back(sv);
            d[i].func = r->func;
            d[i].n = r->items;
            d[i].vals = (const char *)r->gpuArr + en.e * r->items * r->elemSize;
            d[i].locs = r->isLoc() ? (const char *)r->gpuLoc + en.e * r->items * r->locSize : nullptr;
            d[i].locSize = r->isLoc() ? r->locSize : 0;
            d[i].outVal = en.outVal;
            d[i].outLoc = en.outLoc;
            d[i].slots = sv;
        }
        // the loop's last finish launch publishes a completion word (not in a
        // captured graph, whose replays would all publish the same value)
        const bool signal = li + 1 == launches.size() && !graphCapturing();
        const unsigned long long seq = signal ? ++gpu.doneSeq : 0;
        dvmh_hip_reduce_finish_multi(group[0]->r->rtType, (int)group.size(), d, gpu.finishCounters, signal ? gpu.doneDev : nullptr,
                seq, sl->stream);
        if (signal) pendingDone = seq;
        for (void *p : slots) gpu.release(p);
    }
    for (auto &r : q) {
        gpu.release(r->gpuArr);
        r->gpuArr = nullptr;
        if (r->gpuLoc) { gpu.release(r->gpuLoc); r->gpuLoc = nullptr; }
    }
    q.clear();
}

// ------------------------------------------------------------ layout policy
// Arrays keep a transformed device layout between loops (like the reference's
// persistent transform state); loops whose handlers index only permuted
// layouts get the diagonal undone, PARALLEL loops get the natural layout back.
static void prepareLayouts(Loop *loop) {
    if (!loop->region) return;
    int depCount = 0;
    for (long m = loop->depMask; m; m &= m - 1) depCount++;
    for (auto &ra : loop->region->arrays) {
        Data *d = ra.data;
        if (!d->hasDev || !d->dev.isTransformed(d->rank)) continue;
        if (depCount >= 2) continue;
        if (depCount == 1) {
            if (d->dev.diagState) d->setLayout(d->dev.axisPerm, 0);
        } else {
            d->setLayout(nullptr, 0);
        }
    }
}

void prepareLoopLayouts(Loop *loop) { prepareLayouts(loop); }

static void restoreNaturalLayouts(Loop *loop) {
    if (!loop->region) return;
    for (auto &ra : loop->region->arrays) {
        Data *d = ra.data;
        if (d->hasDev && d->dev.isTransformed(d->rank)) d->setLayout(nullptr, 0);
    }
}

// ------------------------------------------------------------ perform

// After the handler: allreduce is trivial for one process; then the
// variable's pre-loop value is folded in (rts/dvmh/dvmh2.cpp:3947-3953).  A GPU
// result still in flight gets the initial value folded in after it is posted.
static void finishLoopReductions(Loop *loop) {
    for (auto &r : loop->reds) {
        bool inFlight = false;
        for (auto it = pending.rbegin(); it != pending.rend(); ++it)
            if (it->red == r) { it->addInitial = true; inFlight = true; break; }
        if (!inFlight) r->postValues(r->initArr.data(), r->isLoc() ? r->initLoc.data() : nullptr, false);
    }
}

void finishLoopReductionsExt(Loop *loop) { finishLoopReductions(loop); }

// dvmh_loop_perform_C (rts/dvmh/dvmh2.cpp:3769-3969): decide where the loop
// runs; GPU loops of a lazy region or of an open batch are queued.
void performLoop(Loop *loop) {
    loop->computeDependencyMask();
    bool onGpu = loop->region && loop->region->onGpu && loop->hasGpu;
    if (!onGpu) {
        DVMH_CHECK(!(settings.requireGpu && loop->region), "Loop at %s:%ld would run on the host but DVMH_AMD_REQUIRE_GPU is set",
                loop->file.c_str(), loop->line);
        DVMH_CHECK(loop->hasHost, "No host handler is registered for the loop at %s:%ld", loop->file.c_str(), loop->line);
    }
    if (onGpu && loop->iterations() > 0) {
        if (regionIsLazy(loop->region)) {
            loop->queued = true;
            loop->region->deferred.push_back(loop);
            return;
        }
        if (batchQueue(loop)) return;
    }
    // loops queued before this one run first (a loop without iterations
    // touches no array: it only posts its reductions' initial values)
    if (loop->iterations() > 0) batchDrain();
    performLoopNow(loop);
}

void performLoopNow(Loop *loop) {
    for (auto &r : loop->reds) r->initValues(r->arrayAddr, r->locAddr);
    bool onGpu = loop->region && loop->region->onGpu && loop->hasGpu;
    if (loop->iterations() == 0) {
        finishLoopReductions(loop);
        return;
    }
    if (onGpu) prepareLayouts(loop);
    SpecLoop sl(loop, onGpu ? 1 : 0);
    sl.stream = gpu.stream;
    loop->spec = &sl;
    DvmType ref = (DvmType)&sl;
    if (onGpu) {
        callHandler(loop->gpuH, &ref);
        gpuRedFinishFlush(&sl);
        rt.loopsOnGpu++;
        // executions the tuner explores are timed from its pick to here; settled
        // sites run without events or synchronisation
        if (sl.explore && sl.tuneStart && !graphCapturing()) {
            // a candidate that re-lays out arrays is charged the round trip: the
            // arrays return to the natural layout before the end mark, as the
            // next loop that needs that layout would do
            if (sl.tune[DVMH_AMD_T_TRANSFORM]) restoreNaturalLayouts(loop);
            hipEvent_t ev1 = tunerEvent(1);
            HIP_CHECK(hipEventRecord(ev1, sl.stream));
            HIP_CHECK(hipEventSynchronize(ev1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, sl.tuneStart, ev1));
            rt.gpuLoopSeconds += ms * 1e-3;
            if (sl.candidate >= 0) tunerRecord(&sl, ms * 1e-3);
        }
        for (void *p : sl.toFree) gpu.release(p);
        finishLoopReductions(loop);
        if (syncAtLoopEnd(loop->region) && !graphCapturing()) flushDeferredPosts();
    } else {
        callHandler(loop->hostH, &ref);
        rt.loopsOnHost++;
        finishLoopReductions(loop);
    }
    loop->spec = nullptr;
}

}  // namespace dvmh

using namespace dvmh;

namespace dvmh {
// Arguments of a direct device fold (dvmh_amd_reduce_partials_C and its
// variant without allocation); a bad argument stops the program.
void checkFoldArgs(const char *where, long func, long rtType, long n, const void *vals, const void *locs, long locSize,
        const void *outVal, const void *outLoc) {
    DVMH_CHECK(n >= 1 && vals && outVal, "%s: %ld partials, values %p, result %p", where, n, vals, outVal);
    DVMH_CHECK(func >= rf_SUM && func <= rf_MINLOC, "%s: reduction function %ld", where, func);
    DVMH_CHECK(rtType == rt_CHAR || rtType == rt_UCHAR || rtType == rt_INT || rtType == rt_UINT || rtType == rt_LOGICAL ||
                    rtType == rt_LONG || rtType == rt_ULONG || rtType == rt_LLONG || rtType == rt_ULLONG || rtType == rt_FLOAT ||
                    rtType == rt_DOUBLE,
            "%s: element type %ld has no GPU fold", where, rtType);
    const bool loc = func == rf_MAXLOC || func == rf_MINLOC;
    DVMH_CHECK(!loc || (locs && outLoc && locSize > 0), "%s: MAXLOC/MINLOC need loc records", where);
}
}  // namespace dvmh

// A reduction's partial values folded on the device, outside a loop: the same
// deterministic fold the loop finishes use, for handlers and tools that keep
// their own partial buffers.  Scratch for the two-pass fold of more than
// 16384 partials comes from the runtime's device cache.
extern "C" DvmType dvmh_amd_reduce_partials_C(DvmType func, DvmType rtType, DvmType n, const void *vals, const void *locs,
        DvmType locSize, void *outVal, void *outLoc) {
    runtimeInit();
    if (!gpu.present) return -1;
    checkFoldArgs("dvmh_amd_reduce_partials_C", (long)func, (long)rtType, (long)n, vals, locs, (long)locSize, outVal, outLoc);
    const bool loc = func == rf_MAXLOC || func == rf_MINLOC;
    const long sb = dvmh_hip_reduce_scratch_bytes((long)n);
    void *scratch = sb > 0 ? gpu.alloc((size_t)sb) : nullptr;
    dvmh_hip_reduce_finish((int)func, (int)rtType, (long)n, vals, loc ? locs : nullptr, loc ? (long)locSize : 0, outVal,
            loc ? outLoc : nullptr, scratch, gpu.stream);
    HIP_CHECK(hipStreamSynchronize(gpu.stream));
    if (scratch) gpu.release(scratch);
    return 0;
}
