// Lazy regions and HIP-graph replay of repeating time steps.
//
// A region created with REGION_AMD_GRAPH queues its GPU loops instead of
// running them; dvmh_region_end_C then runs the queue.  The queue has a
// signature: every loop's site, bounds, handler and parameters, reduction
// variables and their initial bytes, the arrays' device addresses and layouts,
// and the tuner's choice.  The first time a signature is seen the loops run
// eagerly; if that run changed no array layout (steady state), the next
// occurrence is captured from the runtime's stream into a graph (kernels,
// reduction finishes, the D2H of each reduction record into pinned slots owned
// by the graph) and every later occurrence replays the graph with one launch.
// Reduction records are posted to the variables after the replay like any
// deferred GPU result.  Nothing in a captured region may allocate through
// hipMalloc or record hipMemsetAsync: device buffers used while capturing are
// set aside for the graph and never returned to the cache.
#include <cstring>
#include <functional>
#include <sstream>
#include <unordered_map>

#include "rt.h"

namespace dvmh {

struct GraphPost {
    std::shared_ptr<Reduction> red;   // the reduction as registered when captured
    char *val, *loc;                  // pinned slots the graph writes
};

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<void *> devBlocks;
    std::vector<char *> pinned;
    std::vector<GraphPost> posts;
    int eagerStable = 0;     // eager runs that changed no layout
    bool noCapture = false;
};

static std::map<std::string, GraphEntry> graphs;
static bool capturing = false;
static std::vector<void *> *captureBlocks = nullptr;
static std::vector<GraphPost> *capturePosts = nullptr;
static std::vector<char *> *capturePinned = nullptr;

bool graphCapturing() { return capturing; }

// Device blocks handed out while capturing belong to the graph being built.
void graphNoteAlloc(void *p) {
    if (capturing && captureBlocks) captureBlocks->push_back(p);
}

// While capturing, a reduction's D2H goes to a pinned slot owned by the graph.
char *graphPinnedSlot(size_t bytes) {
    char *p = nullptr;
    HIP_CHECK(hipHostMalloc((void **)&p, bytes < 64 ? 64 : bytes, hipHostMallocDefault));
    capturePinned->push_back(p);
    return p;
}

void graphNotePost(const std::shared_ptr<Reduction> &r, char *val, char *loc) {
    capturePosts->push_back(GraphPost{r, val, loc});
}

bool regionIsLazy(const Region *r) { return r && r->onGpu && settings.graph && (r->flags & REGION_AMD_GRAPH); }

static void hashBytes(std::ostringstream &o, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) o << std::hex << (int)b[i];
    o << std::dec << ';';
}

static std::string signature(Region *region) {
    std::ostringstream o;
    o << "F" << settings.fuse << ";";   // pairs run fused or not
    for (auto &ra : region->arrays) {
        Data *d = ra.data;
        o << "A" << (const void *)d << ":" << d->dev.ptr << ":" << d->dev.diagState;
        for (int i = 0; i < d->rank; i++) o << "," << d->dev.axisPerm[i];
        o << ";";
    }
    for (Loop *l : region->deferred) {
        o << "L" << l->file << ":" << l->line << ":" << l->rank;
        for (int i = 0; i < l->rank; i++) o << "," << l->lo[i] << "," << l->hi[i] << "," << l->step[i];
        o << "|h" << (const void *)l->gpuH.func;
        for (void *p : l->gpuH.params) {
            o << "," << p;
            // array headers: the object they name; scalars: their current bytes are
            // read by the kernel from device memory, so only the address matters
        }
        for (auto &r : l->reds) {
            o << "|r" << r->func << ":" << r->rtType << ":" << r->arrayAddr << ":" << r->locAddr << ":";
            hashBytes(o, r->initArr.data(), r->initArr.size());
            hashBytes(o, r->initLoc.data(), r->initLoc.size());
        }
        for (auto &a : l->acrosses) {
            o << "|x" << (const void *)a.data;
            for (int i = 0; i < a.data->rank; i++) o << "," << a.low[i] << "," << a.high[i];
        }
        o << "|u" << l->userBlock[0] << "," << l->userBlock[1] << "," << l->userBlock[2] << ";";
    }
    return o.str();
}

long tunerSitePending(const std::string &file, long line);

// No loop of the region is still being swept by the tuner.
static bool tunerSettled(Region *region) {
    if (settings.tuneMode == 0 || settings.tuneMode == 3) return true;
    for (Loop *l : region->deferred)
        if (tunerSitePending(l->file, l->line) > 0) return false;
    return true;
}

// The region's queued loops in program order; consecutive loops with a
// registered pair handler run fused.
static void runDeferred(Region *region) {
    std::vector<Loop *> &q = region->deferred;
    for (size_t k = 0; k < q.size(); k++) {
        DvmHandlerFunc pf = k + 1 < q.size() ? pairHandlerFor(q[k], q[k + 1]) : nullptr;
        if (!pf) {
            performLoopNow(q[k]);
            delete q[k];
            continue;
        }
        runPairGroup({q[k]}, {q[k + 1]}, pf);
        if (syncAtLoopEnd(region) && !graphCapturing()) flushDeferredPosts();
        delete q[k];
        delete q[k + 1];
        k++;
    }
    q.clear();
}

static void runEager(Region *region) { runDeferred(region); }

// Called by dvmh_region_end_C for a lazy region with queued loops.
void graphRunRegion(Region *region) {
    if (region->deferred.empty()) return;
    if (settings.graphNoCapture) {
        // deferred, fused where pairs allow, but launched one by one (no replay)
        runEager(region);
        return;
    }
    std::string key = rt.curFile + "#" + signature(region);
    GraphEntry &g = graphs[key];
    if (g.exec && !g.noCapture) {
        HIP_CHECK(hipGraphLaunch(g.exec, gpu.stream));
        graphReplays++;
        rt.loopsOnGpu += (long)region->deferred.size();
        // the graph's reduction records, in capture order = loop order x reduction order;
        // each variable is reset to its identity / pre-loop value as perform would do
        size_t k = 0;
        for (Loop *l : region->deferred)
            for (auto &r : l->reds) {
                DVMH_INTERNAL(k < g.posts.size(), "graph reduction slots do not match the region");
                r->initValues(r->arrayAddr, r->locAddr);
                addPendingPost(r, g.posts[k].val, g.posts[k].loc, true);
                k++;
            }
        for (Loop *l : region->deferred) delete l;
        region->deferred.clear();
        if (syncAtLoopEnd(region)) flushDeferredPosts();
        return;
    }
    if (g.noCapture || g.eagerStable < 1 || !tunerSettled(region)) {
        // layout changes while the tuner still explores (transform candidates)
        // say nothing about the settled region
        const bool settled = tunerSettled(region);
        long before = layoutChanges();
        runEager(region);
        if (!settled) g.eagerStable = 0;
        else if (layoutChanges() != before) g.noCapture = true;
        else g.eagerStable++;
        return;
    }
    // capture this occurrence, then launch it
    flushDeferredPosts();
    capturing = true;
    captureBlocks = &g.devBlocks;
    capturePosts = &g.posts;
    capturePinned = &g.pinned;
    int savedSync = settings.syncMode;
    settings.syncMode = 1;
    long before = layoutChanges();
    gpu.sync();   // uploads made while capturing go to idle memory (see uploadAsync)
    HIP_CHECK(hipStreamBeginCapture(gpu.stream, hipStreamCaptureModeRelaxed));
    size_t nLoops = region->deferred.size();
    runDeferred(region);
    HIP_CHECK(hipStreamEndCapture(gpu.stream, &g.graph));
    capturing = false;
    settings.syncMode = savedSync;
    DVMH_INTERNAL(layoutChanges() == before, "array layout changed inside a captured region");
    HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(g.exec, gpu.stream));
    (void)nLoops;
    // A graph launch costs the host more than launching a few kernels one by
    // one (measured: a three-kernel step 6 us slower replayed than eager):
    // a region whose graph is that small runs eagerly from now on.
    size_t nodes = 0;
    HIP_CHECK(hipGraphGetNodes(g.graph, nullptr, &nodes));
    if ((long)nodes < settings.graphMinNodes) g.noCapture = true;
    // the captured loops' records were queued as pending posts by the finishes
    if (syncAtLoopEnd(region)) flushDeferredPosts();
}

// ------------------------------------------------------------ batch queues
// The loops of a batch (dvmh_amd_batch_begin_C .. _end_C) form one time step
// of a program that repeats it: when DVMH_AMD_GRAPH is on, every tuner site
// involved is settled and the queue equals an earlier one word for word (its
// signature below), the whole queue (all fused groups, lone loops, reduction
// finishes and their D2H copies) is captured once and then replayed with one
// graph launch.  Device memory the captured handlers upload to (descriptor
// tables) is allocated for the graph and filled from pinned copies the graph
// owns, so graphs of different queues never share mutable buffers.

struct BatchGraph {
    std::vector<long> sig;
    GraphEntry g;
    std::vector<std::pair<long, long>> postAt;   // (queue position, reduction index) of each captured post
};
static std::unordered_map<size_t, BatchGraph> batchGraphs;
static BatchGraph *lastBatchGraph = nullptr;   // the entry of the last batch queue (element addresses are stable)
constexpr size_t MAX_BATCH_GRAPHS = 16;

static void destroyEntry(GraphEntry &g) {
    if (g.exec) HIP_CHECK(hipGraphExecDestroy(g.exec));
    if (g.graph) HIP_CHECK(hipGraphDestroy(g.graph));
    for (char *p : g.pinned) HIP_CHECK(hipHostFree(p));
    for (void *p : g.devBlocks) gpu.release(p);
    g = GraphEntry();
}

// Everything a captured queue depends on, as words: per loop its site, bounds,
// handlers and parameter addresses, the arrays its region registers (object,
// device address, layout), reductions (variables, initial bytes) and ACROSS
// widths.  Scalars are read by the kernels from device memory, so only their
// addresses enter.
template <class Emit>
static void signatureWords(const std::vector<Loop *> &loops, Emit &&emit) {
    emit(settings.fuse);   // pairs run fused or not
    auto bytes = [&emit](const std::vector<char> &v) {
        emit((long)v.size());
        for (size_t i = 0; i < v.size(); i += sizeof(long)) {
            long x = 0;
            memcpy(&x, v.data() + i, std::min(sizeof(long), v.size() - i));
            emit(x);
        }
    };
    for (Loop *l : loops) {
        emit((long)std::hash<std::string>()(l->file));
        emit(l->line);
        emit(l->rank);
        for (int i = 0; i < l->rank; i++) {
            emit(l->lo[i]);
            emit(l->hi[i]);
            emit(l->step[i]);
        }
        emit((long)l->batchH.func);
        emit((long)l->gpuH.func);
        emit(l->batchChain);
        for (void *p : l->gpuH.params) emit((long)p);
        if (l->region) {
            for (auto &ra : l->region->arrays) {
                const Data *d = ra.data;
                emit((long)d);
                emit((long)d->dev.ptr);
                emit(d->dev.diagState);
                for (int i = 0; i < d->rank; i++) emit(d->dev.axisPerm[i]);
            }
            for (auto &rs : l->region->scalars) {
                emit((long)rs.data);
                emit((long)rs.data->dev.ptr);
            }
        }
        for (auto &r : l->reds) {
            emit(r->func);
            emit(r->rtType);
            emit((long)r->arrayAddr);
            emit((long)r->locAddr);
            bytes(r->initArr);
            bytes(r->initLoc);
        }
        for (auto &a : l->acrosses) {
            emit((long)a.data);
            for (int i = 0; i < a.data->rank; i++) {
                emit(a.low[i]);
                emit(a.high[i]);
            }
        }
        emit(l->userBlock[0]);
        emit(l->userBlock[1]);
        emit(l->userBlock[2]);
        emit(-1);
    }
}

static std::vector<long> batchSignature(const std::vector<Loop *> &loops) {
    std::vector<long> w;
    w.reserve(loops.size() * 32);
    signatureWords(loops, [&w](long x) { w.push_back(x); });
    return w;
}

// The queue's signature equals `sig`, compared word by word as it is formed
// (no vector built, no hash).
static bool signatureEquals(const std::vector<long> &sig, const std::vector<Loop *> &loops) {
    size_t k = 0;
    bool eq = true;
    signatureWords(loops, [&](long x) {
        eq = eq && k < sig.size() && sig[k] == x;
        k++;
    });
    return eq && k == sig.size();
}

static size_t hashWords(const std::vector<long> &w) {
    unsigned long long h = 1469598103934665603ULL;   // FNV-1a over the words
    for (long x : w) {
        h ^= (unsigned long long)x;
        h *= 1099511628211ULL;
    }
    return (size_t)h;
}

bool batchGraphRun(const std::vector<Loop *> &loops, bool eligible, const std::function<void()> &runGroups) {
    if (!eligible || loops.empty() || settings.graphNoCapture) {
        runGroups();
        return false;
    }
    // a program repeats one time step: the entry used last is tried first,
    // compared word by word as the signature is formed
    BatchGraph *found = lastBatchGraph && lastBatchGraph->g.exec && signatureEquals(lastBatchGraph->sig, loops) ? lastBatchGraph : nullptr;
    if (!found) {
        std::vector<long> sig = batchSignature(loops);
        const size_t h = hashWords(sig);
        auto it = batchGraphs.find(h);
        if (it != batchGraphs.end() && it->second.sig != sig) {
            gpu.sync();
            if (lastBatchGraph == &it->second) lastBatchGraph = nullptr;
            destroyEntry(it->second.g);
            batchGraphs.erase(it);
            it = batchGraphs.end();
        }
        if (it == batchGraphs.end()) {
            if (batchGraphs.size() >= MAX_BATCH_GRAPHS) {
                gpu.sync();
                for (auto &kv : batchGraphs) destroyEntry(kv.second.g);
                batchGraphs.clear();
                lastBatchGraph = nullptr;
            }
            it = batchGraphs.emplace(h, BatchGraph()).first;
            it->second.sig = std::move(sig);
        }
        found = &it->second;
    }
    lastBatchGraph = found;
    BatchGraph &bg = *found;
    GraphEntry &g = bg.g;
    if (g.exec) {
        HIP_CHECK(hipGraphLaunch(g.exec, gpu.stream));
        graphReplays++;
        rt.loopsOnGpu += (long)loops.size();
        // each variable is reset to its identity as perform would do, the
        // graph's records are posted in capture order, then the initial
        // values are folded in
        for (Loop *l : loops)
            for (auto &r : l->reds) r->initValues(r->arrayAddr, r->locAddr);
        for (size_t k = 0; k < g.posts.size(); k++) {
            auto &r = loops[bg.postAt[k].first]->reds[bg.postAt[k].second];
            addPendingPost(r, g.posts[k].val, g.posts[k].loc, false);
        }
        for (Loop *l : loops) finishLoopReductionsExt(l);
        return true;
    }
    if (g.noCapture || g.eagerStable < 1) {
        long before = layoutChanges();
        runGroups();
        if (layoutChanges() != before) g.noCapture = true;
        else g.eagerStable++;
        return false;
    }
    // capture this occurrence (the queue's loops map their reductions to
    // positions first: the loops are released group by group as they run)
    std::unordered_map<const Reduction *, std::pair<long, long>> where;
    for (size_t i = 0; i < loops.size(); i++)
        for (size_t k = 0; k < loops[i]->reds.size(); k++) where[loops[i]->reds[k].get()] = {(long)i, (long)k};
    flushDeferredPosts();
    capturing = true;
    captureBlocks = &g.devBlocks;
    capturePosts = &g.posts;
    capturePinned = &g.pinned;
    int savedSync = settings.syncMode;
    settings.syncMode = 1;
    long before = layoutChanges();
    gpu.sync();   // uploads made while capturing go to idle memory (see uploadAsync)
    HIP_CHECK(hipStreamBeginCapture(gpu.stream, hipStreamCaptureModeRelaxed));
    runGroups();
    HIP_CHECK(hipStreamEndCapture(gpu.stream, &g.graph));
    capturing = false;
    settings.syncMode = savedSync;
    DVMH_INTERNAL(layoutChanges() == before, "array layout changed inside a captured batch");
    HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(g.exec, gpu.stream));
    for (auto &p : g.posts) {
        auto w = where.find(p.red.get());
        DVMH_INTERNAL(w != where.end(), "a captured reduction record belongs to no queued loop");
        bg.postAt.push_back(w->second);
        p.red.reset();   // the replays post to the loops of their own queue
    }
    // the captured run's records were queued as pending posts by the finishes
    return false;
}

void graphReleaseAll() {
    gpu.sync();
    for (auto &kv : graphs) destroyEntry(kv.second);
    graphs.clear();
    for (auto &kv : batchGraphs) destroyEntry(kv.second.g);
    batchGraphs.clear();
    lastBatchGraph = nullptr;
}

}  // namespace dvmh
