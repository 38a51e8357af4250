// Batching of many small loops into one fused launch per loop site, and the
// host slot count used by host handlers.
//
// Between dvmh_amd_batch_begin_C and dvmh_amd_batch_end_C, GPU loops that
// registered a batch handler are queued.  At the end the queue is split into
// groups, in program order: a loop joins the most recent group of its batch
// handler when it shares no array with any loop of that group or of a later
// group (so moving it earlier cannot change results); otherwise it starts a new
// group.  Each group runs once: a lone loop through its ordinary handler, a
// larger group through the batch handler, which sees every loop's bounds and
// headers and launches one persistent kernel over all of their tiles.
//
// A loop registered with a chain handler instead joins the last group when
// that group is a chain of the same handler, whatever arrays they share: the
// chain's loops stay in program order and the chain handler runs them with
// their dependencies kept (consecutive ACROSS sweeps of one array, pipelined
// a few tile-planes apart).  A chain whose loop site the tuner is still
// exploring runs loop by loop, so that every execution is timed.
//
// Two consecutive queued loops of one region whose handlers have a registered
// pair handler (a stencil loop and the loop after it, fused by the emitter)
// form a pair; pairs of the same pair handler group like loops of one batch
// handler, and a group of pairs -- even a single pair -- runs through the
// pair handler.  Lazy (graph) regions run their consecutive pairs the same
// way.
//
// Anything that needs the queued loops' results or would run before them
// (a loop without a batch handler, get_actual, actual, a host region) first
// runs the queue (batchDrain); the batch stays open.  Scalars a queued loop
// reads are copied to the device when its region starts, so the loops of one
// batch must see the same value of a scalar they share.
#include <algorithm>
#include <cstring>
#include <map>

#include "rt.h"
#include "../kernels/launch.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace dvmh {

long graphReplays = 0, batchGroups = 0, batchLoops = 0;
static long pairCalls = 0, pairsRun = 0;
static bool batching = false;
static bool draining = false;
static std::vector<Loop *> queued;
static long batchEpoch = 0;   // number of the batch being grouped (Data::batchEpoch)

constexpr uint32_t BATCH_MAGIC = 0x42415443u;   // "BATC"

struct BatchCtx {
    uint32_t magic = BATCH_MAGIC;
    std::vector<Loop *> loops;
    std::vector<SpecLoop *> specs;
    std::vector<Loop *> loops2;       // a pair handler's call: the second loop of each pair
    std::vector<SpecLoop *> specs2;
    DvmHandlerFunc func = nullptr;    // the pair handler (tags its persisted copies)
    bool declined = false;            // the pair handler ran nothing: the loops run one by one
    std::vector<void *> scratch;
};

struct PairReg {
    DvmHandlerFunc first, second, pair;
};
// filled by the emitted libraries' constructors, before the runtime starts
static std::vector<PairReg> &pairRegs() {
    static std::vector<PairReg> v;
    return v;
}

static bool pairMember(const Loop *loop) {
    if (!settings.fuse) return false;
    for (const PairReg &p : pairRegs())
        if (p.first == loop->gpuH.func || p.second == loop->gpuH.func) return true;
    return false;
}

DvmHandlerFunc pairHandlerFor(const Loop *a, const Loop *b) {
    if (!settings.fuse || !a || !b || !a->region || a->region != b->region || !a->region->onGpu) return nullptr;
    if (!a->hasGpu || !b->hasGpu || a->rank != b->rank || a->userBlockFlag || b->userBlockFlag) return nullptr;
    if (!a->acrosses.empty() || !b->acrosses.empty() || a->iterations() == 0) return nullptr;
    for (int i = 0; i < a->rank; i++)
        if (a->lo[i] != b->lo[i] || a->hi[i] != b->hi[i] || a->step[i] != 1 || b->step[i] != 1) return nullptr;
    for (auto &r : a->reds)
        if (r->elemCount != 1) return nullptr;
    for (auto &r : b->reds)
        if (r->elemCount != 1) return nullptr;
    for (const PairReg &p : pairRegs())
        if (p.first == a->gpuH.func && p.second == b->gpuH.func) return p.pair;
    return nullptr;
}

static BatchCtx *asBatch(DvmType ref, const char *where) {
    BatchCtx *b = (BatchCtx *)ref;
    DVMH_CHECK(b && b->magic == BATCH_MAGIC, "Invalid batch reference passed to %s", where);
    return b;
}

static bool pairMember(const Loop *loop);

// GPU loops with a batch handler, or that may be half of a fused pair, wait in
// the queue of an open batch.
bool batchQueue(Loop *loop) {
    if (!batching || !(loop->hasBatch || pairMember(loop))) return false;
    loop->queued = true;
    if (loop->region) loop->region->pendingBatch++;
    queued.push_back(loop);
    return true;
}

bool loopIsQueued(Loop *loop) { return loop->queued; }

// Called by dvmh_region_end_C: a region with queued loops lives until they ran.
bool batchKeepsRegion(Region *r) {
    if (r->pendingBatch == 0) return false;
    r->endedInBatch = true;
    return true;
}

int hostSlotCount() {
    if (settings.hostThreads > 0) return settings.hostThreads;
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Last group of this batch holding a loop that shares an array with l (-1 if none).
static int lastGroupOf(const Loop *l) {
    int g = -1;
    if (l->region)
        for (auto &ra : l->region->arrays)
            if (ra.data->batchEpoch == batchEpoch && ra.data->batchGroup > g) g = ra.data->batchGroup;
    return g;
}

static void markGroup(const Loop *l, int g) {
    if (l->region)
        for (auto &ra : l->region->arrays) {
            ra.data->batchEpoch = batchEpoch;
            ra.data->batchGroup = g;
        }
}

void finishLoopReductionsExt(Loop *loop);

static void runBatchGroup(std::vector<Loop *> &group) {
    BatchCtx ctx;
    ctx.loops = group;
    for (Loop *l : group) {
        for (auto &r : l->reds) r->initValues(r->arrayAddr, r->locAddr);
        SpecLoop *sl = new SpecLoop(l, 1);
        sl->stream = gpu.stream;
        l->spec = sl;
        ctx.specs.push_back(sl);
    }
    DvmType ref = (DvmType)&ctx;
    callHandler(group[0]->batchH, &ref);
    for (SpecLoop *sl : ctx.specs) gpuRedFinishFlush(sl);   // finishes requested per loop, if any
    // a configuration the tuner explores on this group (picked through the
    // first loop's reference) is timed from its pick to here, per loop
    SpecLoop *s0 = ctx.specs[0];
    if (s0->explore && s0->tuneStart && !graphCapturing()) {
        hipEvent_t ev1 = tunerEvent(1);
        HIP_CHECK(hipEventRecord(ev1, s0->stream));
        HIP_CHECK(hipEventSynchronize(ev1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, s0->tuneStart, ev1));
        if (s0->candidate >= 0) tunerRecord(s0, ms * 1e-3 / (double)group.size());
    }
    rt.loopsOnGpu += (long)group.size();
    batchGroups++;
    batchLoops += (long)group.size();
    for (size_t i = 0; i < group.size(); i++) {
        Loop *l = group[i];
        for (void *p : ctx.specs[i]->toFree) gpu.release(p);
        finishLoopReductionsExt(l);
        l->spec = nullptr;
        delete ctx.specs[i];
    }
    for (void *p : ctx.scratch) gpu.release(p);
    ctx.magic = 0;
}

// Runs pairs (firsts[i], seconds[i]) through their pair handler; a handler
// that declines leaves them to run loop by loop, in program order.
void runPairGroup(const std::vector<Loop *> &firsts, const std::vector<Loop *> &seconds, DvmHandlerFunc func) {
    const size_t n = firsts.size();
    DVMH_INTERNAL(n > 0 && seconds.size() == n, "pair group of %zu / %zu loops", n, seconds.size());
    BatchCtx ctx;
    ctx.loops = firsts;
    ctx.loops2 = seconds;
    ctx.func = func;
    for (size_t i = 0; i < n; i++) {
        for (Loop *l : {firsts[i], seconds[i]}) {
            for (auto &r : l->reds) r->initValues(r->arrayAddr, r->locAddr);
            prepareLoopLayouts(l);
        }
        SpecLoop *a = new SpecLoop(firsts[i], 1), *b = new SpecLoop(seconds[i], 1);
        a->stream = b->stream = gpu.stream;
        firsts[i]->spec = a;
        seconds[i]->spec = b;
        ctx.specs.push_back(a);
        ctx.specs2.push_back(b);
    }
    HandlerDesc h;
    h.func = func;
    DvmType ref = (DvmType)&ctx;
    callHandler(h, &ref);
    if (!ctx.declined) {
        for (SpecLoop *sl : ctx.specs) gpuRedFinishFlush(sl);
        for (SpecLoop *sl : ctx.specs2) gpuRedFinishFlush(sl);
        // an explored configuration of the pair site (picked through the first
        // loop of the first pair) is timed from its pick to here, per pair
        SpecLoop *s0 = ctx.specs[0];
        if (s0->explore && s0->tuneStart && !graphCapturing()) {
            hipEvent_t ev1 = tunerEvent(1);
            HIP_CHECK(hipEventRecord(ev1, s0->stream));
            HIP_CHECK(hipEventSynchronize(ev1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, s0->tuneStart, ev1));
            if (s0->candidate >= 0) tunerRecord(s0, ms * 1e-3 / (double)n);
        }
        rt.loopsOnGpu += 2 * (long)n;
        pairCalls++;
        pairsRun += (long)n;
    }
    for (size_t i = 0; i < n; i++)
        for (int m = 0; m < 2; m++) {
            SpecLoop *sl = m ? ctx.specs2[i] : ctx.specs[i];
            Loop *l = m ? seconds[i] : firsts[i];
            for (void *p : sl->toFree) gpu.release(p);
            if (!ctx.declined) finishLoopReductionsExt(l);
            l->spec = nullptr;
            delete sl;
        }
    for (void *p : ctx.scratch) gpu.release(p);
    ctx.magic = 0;
    if (ctx.declined)
        for (size_t i = 0; i < n; i++) {
            performLoopNow(firsts[i]);
            performLoopNow(seconds[i]);
        }
}

static void releaseLoop(Loop *l) {
    Region *r = l->region;
    delete l;
    if (r && --r->pendingBatch == 0 && r->endedInBatch) delete r;
}

long tunerSitePendingExcept(const std::string &file, long line, int kind);
long tunerSitePending(const std::string &file, long line);

// A queue may be captured as a graph when graphs are on and no tuner site of
// its loops is still being explored (every execution of an explored site is
// timed, which a replay cannot do).
static bool queueSettled(const std::vector<Loop *> &loops) {
    if (settings.tuneMode == 0 || settings.tuneMode == 3) return true;
    std::vector<std::pair<const std::string *, long>> seen;
    for (Loop *l : loops) {
        bool known = false;
        for (auto &s : seen)
            if (s.second == l->line && *s.first == l->file) { known = true; break; }
        if (known) continue;
        if (tunerSitePending(l->file, l->line) > 0) return false;
        seen.push_back({&l->file, l->line});
    }
    return true;
}

static void runQueue() {
    struct Group {
        DvmHandlerFunc func;
        bool chain;
        std::vector<Loop *> loops;
        bool explore = false;
        bool pair = false;
        std::vector<Loop *> loops2;   // pair groups: the second loop of each pair
    };
    std::vector<Loop *> all;
    all.swap(queued);
    bool settled = settings.graph && queueSettled(all);
    // the groups are formed only when the queue runs loop by loop or is
    // captured: a replay of its graph needs none of them
    std::vector<Group> groups;
    auto buildGroups = [&]() {
        batchEpoch++;
        for (size_t k = 0; k < all.size(); k++) {
            Loop *l = all[k];
            int target = -1;
            DvmHandlerFunc pf = !l->batchChain && k + 1 < all.size() ? pairHandlerFor(l, all[k + 1]) : nullptr;
            if (pf) {
                // a pair: joins the last group of its pair handler under the same
                // rule as a loop (both loops belong to one region)
                for (int g = (int)groups.size() - 1; g >= 0; g--)
                    if (groups[g].pair && groups[g].func == pf) { target = g; break; }
                if (target >= 0 && lastGroupOf(l) >= target) target = -1;
                if (target < 0) {
                    groups.push_back(Group{pf, false, {}, false, true, {}});
                    target = (int)groups.size() - 1;
                }
                groups[target].loops.push_back(l);
                groups[target].loops2.push_back(all[k + 1]);
                markGroup(l, target);
                k++;
                continue;
            }
            if (!l->hasBatch) {
                // queued as a possible pair member, unpaired: runs on its own
                groups.push_back(Group{nullptr, false, {l}, false, false, {}});
                markGroup(l, (int)groups.size() - 1);
                continue;
            }
            if (l->batchChain) {
                if (!groups.empty() && groups.back().chain && groups.back().func == l->batchH.func) target = (int)groups.size() - 1;
            } else {
                for (int g = (int)groups.size() - 1; g >= 0; g--)
                    if (!groups[g].chain && !groups[g].pair && groups[g].func == l->batchH.func) { target = g; break; }
                // joining group `target` moves l before the groups after it: allowed
                // when none of target's or the later groups' loops shares an array
                if (target >= 0 && lastGroupOf(l) >= target) target = -1;
            }
            if (target < 0) {
                groups.push_back(Group{l->batchH.func, l->batchChain, {}, false, false, {}});
                target = (int)groups.size() - 1;
            }
            groups[target].loops.push_back(l);
            markGroup(l, target);
        }
        for (auto &g : groups)
            g.explore = g.chain && tunerSitePendingExcept(g.loops[0]->file, g.loops[0]->line, DVMH_AMD_KIND_CHAIN) > 0;
    };
    auto runGroups = [&]() {
        buildGroups();
        for (auto &g : groups) {
            if (g.pair) {
                for (Loop *l : g.loops) l->queued = false;
                for (Loop *l : g.loops2) l->queued = false;
                runPairGroup(g.loops, g.loops2, g.func);
                bool sync = false;
                for (Loop *l : g.loops) sync = sync || syncAtLoopEnd(l->region);
                for (Loop *l : g.loops) releaseLoop(l);
                for (Loop *l : g.loops2) releaseLoop(l);
                if (sync && !graphCapturing()) flushDeferredPosts();
                continue;
            }
            if (g.loops.size() == 1 || g.explore) {
                for (Loop *l : g.loops) {
                    l->queued = false;
                    performLoopNow(l);
                }
            } else {
                runBatchGroup(g.loops);
            }
            bool sync = false;
            for (Loop *l : g.loops) sync = sync || syncAtLoopEnd(l->region);
            for (Loop *l : g.loops) releaseLoop(l);
            if (sync && !graphCapturing()) flushDeferredPosts();
        }
    };
    bool sync = false;
    for (Loop *l : all) sync = sync || syncAtLoopEnd(l->region);
    if (batchGraphRun(all, settled, runGroups))
        for (Loop *l : all) releaseLoop(l);
    // records of a replay or of a captured queue's first launch (an eager run
    // has posted its own group by group)
    if (sync) flushDeferredPosts();
}

// Runs what is queued so far (the batch stays open); see the file comment.
void batchDrain() {
    if (!batching || draining || queued.empty()) return;
    draining = true;
    runQueue();
    draining = false;
}

bool batchOpen() { return batching; }

}  // namespace dvmh

using namespace dvmh;

extern "C" void dvmh_amd_batch_begin_C(void) {
    DVMH_CHECK(!batching, "dvmh_amd_batch_begin_C: a batch is already open");
    batching = true;
}

extern "C" void dvmh_amd_batch_end_C(void) {
    DVMH_CHECK(batching, "dvmh_amd_batch_end_C without dvmh_amd_batch_begin_C");
    batching = false;
    runQueue();
}

extern "C" DvmType dvmh_amd_batch_count_C(DvmType batchRef) {
    return (DvmType)asBatch(batchRef, "dvmh_amd_batch_count_C")->loops.size();
}

extern "C" void dvmh_amd_register_pair_C(DvmHandlerFunc first, DvmHandlerFunc second, DvmHandlerFunc pair) {
    DVMH_CHECK(first && second && pair, "dvmh_amd_register_pair_C: null handler");
    pairRegs().push_back(PairReg{first, second, pair});
}

extern "C" void dvmh_amd_set_fuse_C(DvmType enable) {
    runtimeInit();
    settings.fuse = enable != 0;
}

extern "C" void dvmh_amd_pair_stats_C(DvmType out[2]) {
    out[0] = pairCalls;
    out[1] = pairsRun;
}

namespace dvmh {
void resetPairStats() { pairCalls = pairsRun = 0; }
}

static BatchCtx *asPair(DvmType ref, DvmType member, DvmType i, const char *where) {
    BatchCtx *b = asBatch(ref, where);
    DVMH_CHECK(!b->loops2.empty(), "%s: not called from a pair handler", where);
    DVMH_CHECK(member == 0 || member == 1, "%s: pair member %ld", where, (long)member);
    DVMH_CHECK(i >= 0 && i < (DvmType)b->loops.size(), "%s: pair index %ld out of range", where, (long)i);
    return b;
}

extern "C" DvmType dvmh_amd_pair_loop_C(DvmType batchRef, DvmType member, DvmType i) {
    BatchCtx *b = asPair(batchRef, member, i, "dvmh_amd_pair_loop_C");
    return (DvmType)(member ? b->specs2[i] : b->specs[i]);
}

extern "C" void *dvmh_amd_pair_param_C(DvmType batchRef, DvmType member, DvmType i, DvmType p) {
    BatchCtx *b = asPair(batchRef, member, i, "dvmh_amd_pair_param_C");
    const auto &params = (member ? b->loops2[i] : b->loops[i])->gpuH.params;
    DVMH_CHECK(p >= 0 && p < (DvmType)params.size(), "Pair parameter index %ld out of range", (long)p);
    return params[p];
}

extern "C" void dvmh_amd_pair_decline_C(DvmType batchRef) {
    BatchCtx *b = asPair(batchRef, 0, 0, "dvmh_amd_pair_decline_C");
    b->declined = true;
}

extern "C" DvmType dvmh_amd_batch_loop_C(DvmType batchRef, DvmType i) {
    BatchCtx *b = asBatch(batchRef, "dvmh_amd_batch_loop_C");
    DVMH_CHECK(i >= 0 && i < (DvmType)b->loops.size(), "Batch loop index %ld out of range", i);
    return (DvmType)b->specs[i];
}

extern "C" void *dvmh_amd_batch_param_C(DvmType batchRef, DvmType i, DvmType p) {
    BatchCtx *b = asBatch(batchRef, "dvmh_amd_batch_param_C");
    DVMH_CHECK(i >= 0 && i < (DvmType)b->loops.size(), "Batch loop index %ld out of range", i);
    const auto &params = b->loops[i]->gpuH.params;
    DVMH_CHECK(p >= 0 && p < (DvmType)params.size(), "Batch parameter index %ld out of range", p);
    return params[p];
}

extern "C" void *dvmh_amd_batch_scratch_C(DvmType batchRef, DvmType bytes) {
    BatchCtx *b = asBatch(batchRef, "dvmh_amd_batch_scratch_C");
    void *p = gpu.alloc((size_t)(bytes > 0 ? bytes : 8));
    b->scratch.push_back(p);
    return p;
}

// Pinned staging of batch uploads.  Copies are queued on the runtime's stream
// without waiting; staging space is handed out in order and only reused after
// a stream synchronisation, which happens when the space runs out (so every
// copy staged before it has completed).
namespace {
struct UploadRing {
    char *buf = nullptr;
    size_t cap = 0, used = 0;
    char *take(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (used + bytes > cap) {
            if (buf) HIP_CHECK(hipStreamSynchronize(gpu.stream));
            if (bytes > cap) {
                if (buf) HIP_CHECK(hipHostFree(buf));
                cap = std::max(bytes, (size_t)8 << 20);
                HIP_CHECK(hipHostMalloc((void **)&buf, cap, hipHostMallocDefault));
            }
            used = 0;
        }
        char *p = buf + used;
        used += bytes;
        return p;
    }
};
UploadRing uploadRing;

// Device copies a batch handler keeps from one execution to the next, per
// (tag, slot): re-uploaded only when the host bytes differ from the last ones.
struct Persisted {
    void *dev = nullptr;
    size_t cap = 0;
    std::vector<char> last;
};
std::map<std::pair<const void *, long>, Persisted> persisted;
}  // namespace

static void uploadAsync(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    if (graphCapturing()) {
        // What a handler uploads while a graph is captured is fixed for the
        // graph's life (its queue's signature covers the values), so it is
        // copied now, outside the graph, instead of on every replay.  The
        // destination is device memory the graph owns, allocated during the
        // capture (nothing else uses it; the stream was idle when the capture
        // began), and the copy ends before the capture goes on.
        static hipStream_t side = nullptr;
        if (!side) HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, side));
        HIP_CHECK(hipStreamSynchronize(side));
        return;
    }
    char *stage = uploadRing.take(bytes);
    memcpy(stage, src, bytes);
    HIP_CHECK(hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, gpu.stream));
}

extern "C" void dvmh_amd_batch_upload_C(DvmType batchRef, void *dst, const void *src, DvmType bytes) {
    asBatch(batchRef, "dvmh_amd_batch_upload_C");
    uploadAsync(dst, src, (size_t)bytes);
}

extern "C" void *dvmh_amd_batch_persist_C(DvmType batchRef, const void *tag, DvmType slot, const void *src, DvmType bytes) {
    asBatch(batchRef, "dvmh_amd_batch_persist_C");
    DVMH_CHECK(bytes >= 0, "dvmh_amd_batch_persist_C: %ld bytes", (long)bytes);
    const size_t n = (size_t)bytes;
    if (graphCapturing()) {
        // a graph keeps its own copy: later executions of other queues may
        // change the persisted one while the graph is still replayed
        void *d = gpu.alloc(std::max(n, (size_t)64));
        graphNoteAlloc(d);
        uploadAsync(d, src, n);
        return d;
    }
    Persisted &p = persisted[std::make_pair(tag, (long)slot)];
    if (p.dev && p.last.size() == n && (n == 0 || memcmp(p.last.data(), src, n) == 0)) return p.dev;
    if (!p.dev || p.cap < n) {
        if (p.dev) gpu.release(p.dev);   // stream-ordered: earlier launches reading it come first
        p.cap = std::max(n, (size_t)64);
        p.dev = gpu.alloc(p.cap);
    }
    p.last.assign((const char *)src, (const char *)src + n);
    uploadAsync(p.dev, src, n);
    return p.dev;
}

// Folds each loop's partial segment and posts the results (one kernel, one D2H).
static void redFinish(DvmType batchRef, BatchCtx *b, const std::vector<Loop *> &loops, const void *tag, long slot, DvmType redIndex,
        const void *partials, const void *locPartials, const DvmType *offsets) {
    long nl = (long)loops.size();
    DVMH_CHECK(nl > 0, "empty batch");
    DVMH_CHECK(redIndex >= 1 && redIndex <= (DvmType)loops[0]->reds.size(), "Batch reduction index %ld out of range", redIndex);
    auto &r0 = loops[0]->reds[redIndex - 1];
    long es = r0->elemSize, ls = r0->isLoc() ? r0->locSize : 0;
    for (Loop *l : loops) {
        DVMH_CHECK(redIndex >= 1 && redIndex <= (DvmType)l->reds.size(), "Batch reduction index %ld out of range", redIndex);
        auto &r = l->reds[redIndex - 1];
        DVMH_CHECK(r->func == r0->func && r->rtType == r0->rtType && r->elemCount == 1 && (r->isLoc() ? r->locSize : 0) == ls,
                "Batched loops disagree on reduction %ld", redIndex);
    }
    // the segment offsets usually repeat from one execution of the handler to the next
    void *dOff = dvmh_amd_batch_persist_C(batchRef, tag, slot, offsets, (nl + 1) * (long)sizeof(long));
    // the folded records go straight into mapped pinned host memory (a graph's
    // own slot when captured): no separate copy
    size_t total = (size_t)(((nl * es + 15) & ~15L) + nl * ls + 64);
    char *host = graphCapturing() ? graphPinnedSlot(total) : pinnedTake(total);
    char *dRes = mappedDevicePtr(host);
    char *dLoc = ls ? dRes + ((nl * es + 15) & ~15L) : nullptr;
    dvmh_hip_reduce_segments(r0->func, r0->rtType, nl, (const long *)dOff, partials, locPartials, ls, dRes, dLoc, gpu.stream);
    for (long i = 0; i < nl; i++) {
        auto &r = loops[i]->reds[redIndex - 1];
        char *val = host + i * es;
        char *loc = ls ? host + ((nl * es + 15) & ~15L) + i * ls : nullptr;
        if (graphCapturing()) graphNotePost(r, val, loc);
        addPendingPost(r, val, loc, false);
    }
    (void)b;
}

extern "C" void dvmh_amd_batch_red_finish_C(DvmType batchRef, DvmType redIndex, const void *partials, const void *locPartials,
        const DvmType *offsets) {
    BatchCtx *b = asBatch(batchRef, "dvmh_amd_batch_red_finish_C");
    DVMH_CHECK(!b->loops.empty(), "empty batch");
    redFinish(batchRef, b, b->loops, (const void *)b->loops[0]->batchH.func, 1000 + (long)redIndex, redIndex, partials, locPartials,
            offsets);
}

extern "C" void dvmh_amd_pair_red_finish_C(DvmType batchRef, DvmType member, DvmType redIndex, const void *partials,
        const void *locPartials, const DvmType *offsets) {
    BatchCtx *b = asPair(batchRef, member, 0, "dvmh_amd_pair_red_finish_C");
    redFinish(batchRef, b, member ? b->loops2 : b->loops, (const void *)b->func, 2000 + 100 * (long)member + (long)redIndex, redIndex,
            partials, locPartials, offsets);
}
