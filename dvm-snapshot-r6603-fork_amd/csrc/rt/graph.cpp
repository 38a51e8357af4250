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

static void runEager(Region *region) {
    for (Loop *l : region->deferred) {
        performLoopNow(l);
        delete l;
    }
    region->deferred.clear();
}

// Called by dvmh_region_end_C for a lazy region with queued loops.
void graphRunRegion(Region *region) {
    if (region->deferred.empty()) return;
    std::string key = rt.curFile + "#" + signature(region);
    GraphEntry &g = graphs[key];
    if (g.exec) {
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
    HIP_CHECK(hipStreamBeginCapture(gpu.stream, hipStreamCaptureModeRelaxed));
    size_t nLoops = region->deferred.size();
    for (Loop *l : region->deferred) {
        performLoopNow(l);
        delete l;
    }
    region->deferred.clear();
    HIP_CHECK(hipStreamEndCapture(gpu.stream, &g.graph));
    capturing = false;
    settings.syncMode = savedSync;
    DVMH_INTERNAL(layoutChanges() == before, "array layout changed inside a captured region");
    HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(g.exec, gpu.stream));
    (void)nLoops;
    // the captured loops' records were queued as pending posts by the finishes
    if (syncAtLoopEnd(region)) flushDeferredPosts();
}

void graphReleaseAll() {
    gpu.sync();
    for (auto &kv : graphs) {
        GraphEntry &g = kv.second;
        if (g.exec) HIP_CHECK(hipGraphExecDestroy(g.exec));
        if (g.graph) HIP_CHECK(hipGraphDestroy(g.graph));
        for (char *p : g.pinned) HIP_CHECK(hipHostFree(p));
        for (void *p : g.devBlocks) gpu.release(p);
    }
    graphs.clear();
}

}  // namespace dvmh
