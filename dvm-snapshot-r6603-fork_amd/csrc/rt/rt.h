// Internal declarations of libdvmh_amd: the one-process, one-GPU DVMH runtime.
//
// Object model (single process, so every array is wholly local):
//   Data       a distributed array or a registered regular variable; owns a host
//              buffer and at most one device representative (DevBuffer) whose
//              layout may be axis-permuted or diagonalised.
//   Region     a "#pragma dvm region": the arrays/scalars it uses and where it runs.
//   Loop       a parallel loop as registered by converted code (bounds, mapping,
//              reductions, ACROSS widths, handlers).
//   SpecLoop   the loop as seen by a handler on one device ("loop_ref").
// Reference counterparts: rts/dvmh/dvmh_data.h (DvmhData/DvmhRepresentative),
// rts/dvmh/dvmh_buffer.h, rts/dvmh/region.h, rts/dvmh/loop.h (DvmhLoop,
// DvmhSpecLoop/DvmhLoopCuda, DvmhReduction).
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "dvmhlib2.h"
#include "dvmh_amd.h"

namespace dvmh {

constexpr int MAX_RANK = 8;         // rts/dvmh/dvmh_types.h:223
constexpr int MAX_LOOP_RANK = 12;   // rts/dvmh/dvmh_types.h:224
constexpr int MAX_PARAMS = 256;     // handler parameters (rts/dvmh/dvmh_types.h:228 MAX_PARAM_COUNT)

// ---------------------------------------------------------------- errors / log
[[noreturn]] void fatal(const char *kind, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void logMsg(int level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
#define DVMH_CHECK(cond, ...) do { if (!(cond)) ::dvmh::fatal("error", __VA_ARGS__); } while (0)
#define DVMH_INTERNAL(cond, ...) do { if (!(cond)) ::dvmh::fatal("internal error", __VA_ARGS__); } while (0)
#define HIP_CHECK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) \
    ::dvmh::fatal("HIP error", "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); } while (0)
enum LogLevel { LOG_FATAL = 0, LOG_ERROR = 1, LOG_WARN = 2, LOG_INFO = 3, LOG_DEBUG = 4, LOG_TRACE = 5 };

// ---------------------------------------------------------------- settings
struct Settings {
    int logLevel = LOG_WARN;
    int numGpus = 1;            // DVMH_NUM_CUDAS: 0 runs every region on the host
    int hostThreads = 0;        // DVMH_NUM_THREADS (0: OpenMP default)
    bool requireGpu = false;    // DVMH_AMD_REQUIRE_GPU: abort instead of running a region on the host
    int tuneMode = 0;           // DVMH_AMD_TUNE: 0 default config, 1 exhaustive sweep, 2 cost model, 3 fixed (from DB)
    int syncMode = 0;           // DVMH_AMD_SYNC: 0 synchronise after every loop, 1 only when results are needed
    bool allowAsync = false;    // DVMH_ALLOW_ASYNC: regions created with REGION_ASYNC synchronise only when results are needed
    bool graph = false;         // DVMH_AMD_GRAPH: capture repeating regions into HIP graphs
    bool graphNoCapture = false;   // lazy regions and batches run eagerly, never captured (kernel timing)
    long graphMinNodes = 6;        // DVMH_AMD_GRAPH_MIN_NODES: a region's graph with fewer nodes is not replayed
    int tuneRepeats = 1;        // DVMH_AMD_TUNE_REPEATS: timed executions per candidate
    bool pollDone = true;          // DVMH_AMD_POLL: wait for reduction results on the completion word, not a stream sync
    bool tuneTransforms = true;    // DVMH_AMD_TUNE_TRANSFORMS (default 1): candidates may re-lay out arrays of ACROSS loops
    bool fuse = true;              // DVMH_AMD_FUSE (default 1): registered loop pairs held back together run fused
    bool index64 = false;          // DVMH_AMD_INDEX64: every loop's kernels take 64-bit index arithmetic
    std::string tuneDb;         // DVMH_AMD_TUNEDB: file of recorded timings
    std::string modelFile;      // DVMH_AMD_MODEL: trained cost-model weights
    void load();
};
extern Settings settings;

// ---------------------------------------------------------------- device
struct DeviceProps {
    int cuCount = 0, waveSize = 64, maxThreadsPerBlock = 1024, maxBlock[3] = {1024, 1024, 1024};
    long ldsPerBlock = 65536, ldsPerCu = 163840, regsPerCu = 0, l2Bytes = 0;
    int clockKhz = 0, memClockKhz = 0, busWidth = 0;
    char name[256] = {0}, arch[64] = {0};
};

class Gpu {
public:
    bool present = false;
    int ordinal = 0;
    DeviceProps props;
    hipStream_t stream = nullptr;
    unsigned *finishCounters = nullptr;   // arrival counters of the multi-reduction finish (kept zero between launches)
    // completion word in mapped pinned host memory: a reduction finish sets it
    // to its sequence number once its results are written
    volatile unsigned long long *doneHost = nullptr;
    unsigned long long *doneDev = nullptr;
    unsigned long long doneSeq = 0;
    long h2dBytes = 0, d2hBytes = 0;   // array and variable bytes moved by h2d / d2h
    void waitDone(unsigned long long seq);   // until the word reaches seq (or the stream drains)
    void init();
    void *alloc(size_t bytes);          // size-keyed cache in front of hipMalloc
    void release(void *p);              // back to the cache
    void trimCache();
    void *pinnedStaging(size_t bytes);  // reusable pinned host buffer
    void h2d(void *dst, const void *src, size_t bytes);
    void d2h(void *dst, const void *src, size_t bytes);
    void sync();
private:
    std::map<size_t, std::vector<void *>> freeBlocks;
    std::map<void *, size_t> liveBlocks;
    void *staging = nullptr;
    size_t stagingBytes = 0;
};
extern Gpu gpu;

// ---------------------------------------------------------------- objects
enum class Kind : uint32_t { Data = 1, Template, Region, Loop, SpecLoop };
constexpr uint32_t OBJ_MAGIC = 0x44564D48u;   // "DVMH"

struct Object {
    uint32_t magic = OBJ_MAGIC;
    Kind kind;
    explicit Object(Kind k): kind(k) {}
    virtual ~Object() { magic = 0; }
};

struct Interval {
    long lo = 0, hi = -1;
    long size() const { return hi - lo + 1; }
};

// Device representative of an array.  Element (idx[0..rank-1]) of the natural
// layout lives at position p = sum over storage positions q of
// (idx[axisPerm[q]-1] - portion[axisPerm[q]-1].lo) * coll[q], except that with
// diagState != 0 the last two storage positions are one diagonalised plane
// (rts/dvmh/dvmh_buffer.cpp:149-230).
struct DevBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    int axisPerm[MAX_RANK] = {0};   // storage position -> original axis (1-based)
    int diagState = 0;              // 0 none, 1 diagonal, 2 slash diagonal
    bool isTransformed(int rank) const;
    bool isDiagonalized() const { return diagState != 0; }
};

// A set of disjoint boxes of array indices (inclusive bounds per axis): where
// a copy of an array is current (the role of the reference's DvmhPieces,
// rts/dvmh/dvmh_pieces.cpp, used by its actual-state tracking,
// rts/dvmh/dvmh_data.cpp:339-600).
struct Box {
    long lo[MAX_RANK], hi[MAX_RANK];
};

struct Pieces {
    int rank = 0;
    std::vector<Box> boxes;
    bool empty() const { return boxes.empty(); }
    void clear() { boxes.clear(); }
    void add(const Box &b);                     // union with b
    void remove(const Box &b);                  // difference with b
    std::vector<Box> missing(const Box &b) const;   // parts of b not in the set, disjoint
    bool covers(const Box &b) const { return missing(b).empty(); }
    bool holds(const long point[]) const;       // the point lies in one of the boxes
    long elements() const;
};

struct Data : Object {
    int rank = 0;
    long typeSize = 0;
    int rtType = rt_UNKNOWN;
    Interval space[MAX_RANK];
    long shdLow[MAX_RANK] = {0}, shdHigh[MAX_RANK] = {0};
    char *host = nullptr;
    bool ownsHost = false;
    bool isRegular = false;         // regular C variable (dvmh_data_enter_C / scalar)
    DvmType *userHeader = nullptr;
    DevBuffer dev;
    bool hasDev = false;
    // where each copy is current; every element is current in at least one
    // (a device copy's pieces are in natural indices whatever its layout)
    Pieces hostPieces, devPieces;
    // grouping of a closing batch: the last group (of batch number batchEpoch)
    // holding a loop whose region registers this array
    long batchEpoch = 0;
    int batchGroup = -1;
    Data(): Object(Kind::Data) {}
    ~Data() override;
    long elemCount() const;
    Box whole() const;
    bool hostWhole() const { return hostPieces.covers(whole()); }
    bool devWhole() const { return devPieces.covers(whole()); }
    void hostOwnsAll();             // host copy current everywhere, device copy nowhere
    void devOwnsAll();              // the reverse (a device write of the whole array)
    size_t bytes() const { return (size_t)elemCount() * (size_t)typeSize; }
    long hostStride(int axis /*1-based*/) const;
    void ensureDev();
    void toDevice();                // host -> device (natural layout), marks device actual
    void toHost();                  // device -> host (undoing any layout transform)
    void setLayout(const int newPerm[], int newDiag);   // re-layout the device copy
    void fillDevHeader(const void *base, DvmType header[], DvmType ext[]) const;
    void *devNaturalBase() const;
    void *hostNaturalBase() const;
};

struct Template : Object {
    int rank = 0;
    Interval space[MAX_RANK];
    Template(): Object(Kind::Template) {}
};

struct RegionArray {
    Data *data;
    long intent;
    std::string name;
    bool part = false;   // registered as a sub-array: only `box` is moved and owned
    Box box{};
};

struct Region : Object {
    long flags = 0;
    long targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;
    bool onGpu = false;
    bool async = false;                    // REGION_ASYNC with DVMH_ALLOW_ASYNC: no synchronisation at loop or region end
    std::vector<RegionArray> arrays;
    std::vector<RegionArray> scalars;
    std::vector<struct Loop *> deferred;   // GPU loops queued by a lazy (graph) region
    int pendingBatch = 0;                  // loops of this region queued in an open batch
    bool endedInBatch = false;             // ended while pendingBatch > 0: deleted with its last loop
    Region(): Object(Kind::Region) {}
};

struct Reduction {
    int func = 0;
    int rtType = rt_UNKNOWN;
    long elemCount = 1, elemSize = 0;
    void *arrayAddr = nullptr, *locAddr = nullptr;
    long locSize = 0;
    std::vector<char> initArr, initLoc;   // values at registration (rts/dvmh/loop.cpp:17-35)
    std::mutex lock;
    // per-execution device state (DvmhReductionCuda, rts/dvmh/loop.cpp:258-297)
    void **gpuArrPtr = nullptr, **gpuLocPtr = nullptr;
    void *gpuArr = nullptr, *gpuLoc = nullptr;
    long items = 0;
    std::vector<char> localArr, localLoc;
    void initValues(void *arr, void *loc) const;                               // :37-91
    // lowerLocOnTie: equal values take the lower scalar-integer loc (the merge
    // of partial results); false keeps the destination's (initial values)
    void performOperation(void *dArr, void *dLoc, const void *sArr, const void *sLoc, bool lowerLocOnTie) const;  // :125-254
    void postValues(const void *arr, const void *loc, bool lowerLocOnTie = true);   // :97-102
    bool locKeyLess(const void *a, const void *b) const;
    bool isLoc() const { return locAddr != nullptr; }
};

struct AcrossSpec {
    Data *data;
    bool isOut;
    long low[MAX_RANK], high[MAX_RANK];
};

struct HandlerDesc {
    DvmHandlerFunc func = nullptr;
    std::vector<void *> params;
    long type = 0;
};

struct Loop;

// The loop as seen from inside a handler (rts/dvmh/loop.h DvmhSpecLoop / DvmhLoopCuda).
struct SpecLoop : Object {
    Loop *loop;
    int device;                     // 0 host, 1 GPU
    hipStream_t stream = nullptr;
    int block[3] = {0, 0, 0};
    DvmType tune[DVMH_AMD_TUNE_N] = {0};
    int candidate = -1;             // index in the tuner's candidate list
    std::string siteKey;            // tuner site of this execution
    bool explore = false;           // the tuner times this execution (candidate still lacks repeats)
    hipEvent_t tuneStart = nullptr; // recorded at the pick when exploring
    std::vector<std::shared_ptr<Reduction>> finishQueue;   // red_finish requests, run when the handler returns
    std::vector<void *> toFree;
    SpecLoop(Loop *l, int dev): Object(Kind::SpecLoop), loop(l), device(dev) {}
};

struct Loop : Object {
    Region *region = nullptr;
    int rank = 0;
    long lo[MAX_LOOP_RANK], hi[MAX_LOOP_RANK], step[MAX_LOOP_RANK];
    // mapping: loop axis (1-based, signed by direction, 0 = none) tied to each axis of mapData
    Data *mapData = nullptr;
    Template *mapTempl = nullptr;
    int mapRank = 0;
    long mapAxis[MAX_RANK] = {0}, mapMult[MAX_RANK] = {0}, mapSum[MAX_RANK] = {0};
    std::vector<std::pair<Data *, std::vector<long>>> correspondences;
    std::vector<std::shared_ptr<Reduction>> reds;
    std::vector<AcrossSpec> acrosses;
    int userBlock[3] = {0, 0, 0};
    bool userBlockFlag = false;
    long stage = 0;
    HandlerDesc hostH, gpuH;
    bool hasHost = false, hasGpu = false;
    long depMask = 0;
    std::string file;
    long line = 0;
    SpecLoop *spec = nullptr;
    bool queued = false;            // owned by a lazy region or a batch until it runs
    HandlerDesc batchH;             // handler that runs many loops of this site at once
    bool hasBatch = false;
    bool batchChain = false;        // batchH runs consecutive loops of the site in order (shared arrays allowed)
    Loop(): Object(Kind::Loop) {}
    ~Loop() override;
    void computeDependencyMask();                       // rts/dvmh/loop.cpp:1117-1150
    std::vector<long> correspondence(const Data *d) const;   // array axis -> signed loop axis
    long iterations() const;
};

// ---------------------------------------------------------------- runtime state
struct Runtime {
    bool inited = false;
    std::string curFile = "?";
    long curLine = 0;
    Region *curRegion = nullptr;
    std::map<const void *, Data *> regularVars;     // address -> regular variable data
    std::vector<Data *> allData;
    // statistics of the current process (exported through dvmh_amd_stats_C)
    long loopsOnHost = 0, loopsOnGpu = 0, kernelLaunchEstimate = 0;
    double gpuLoopSeconds = 0.0;
};
extern Runtime rt;

Object *asObject(DvmType handle);
Data *asData(const DvmType desc[], const char *where);
Loop *asLoop(DvmType h, const char *where);
SpecLoop *asSpec(DvmType h, const char *where);
Loop *loopOf(DvmType h, const char *where);         // Loop or SpecLoop handle
long rtTypeSize(int rtType);

// handler trampoline (rts/dvmh/util.cpp:659-720)
void callHandler(const HandlerDesc &h, DvmType *pLoopRef, const std::vector<void *> *override = nullptr);

// loop execution
void performLoop(Loop *loop);

// reductions on the GPU (kernels in csrc/kernels/reduce.hip)
void gpuRedPrepare(SpecLoop *sl, Reduction *r, long count, bool fill);
void gpuRedFinish(SpecLoop *sl, const std::shared_ptr<Reduction> &r);
void gpuRedFinishFlush(SpecLoop *sl);
void flushDeferredPosts();
void flushPostsFor(const void *addr);   // when a pending post targets addr
long pendingPostCount();
// Whether loops of region r synchronise when they end: by default yes, not in
// DVMH_AMD_SYNC=1 and not in an async region (results merge at the next
// get_actual / actual or at a region that registers their variables).
inline bool syncAtLoopEnd(const Region *r) { return settings.syncMode == 0 && !(r && r->async); }

// tuner
struct TuneChoice {
    int block[3];
    DvmType tune[DVMH_AMD_TUNE_N];
    int candidate;
};
void tunerPick(SpecLoop *sl, long kind, TuneChoice &out);
void tunerRecord(SpecLoop *sl, double seconds);
hipEvent_t tunerEvent(int which);   // 0 start, 1 stop of an explored execution
void tunerShutdown();

// batching and graphs (batch.cpp)
extern long graphReplays, batchGroups, batchLoops;
bool batchQueue(Loop *loop);
void batchDrain();                  // run the loops queued in an open batch
bool batchOpen();
bool loopIsQueued(Loop *loop);
int hostSlotCount();
bool batchKeepsRegion(Region *r);
char *pinnedTake(size_t bytes);
char *mappedDevicePtr(char *hostPinned);   // device-side alias of pinnedTake / graph slot memory
void performLoopNow(Loop *loop);
long layoutChanges();

// lazy regions and graphs (graph.cpp)
bool regionIsLazy(const Region *r);
bool graphCapturing();
void graphNoteAlloc(void *p);
char *graphPinnedSlot(size_t bytes);
void graphNotePost(const std::shared_ptr<Reduction> &r, char *val, char *loc);
void graphRunRegion(Region *region);
// Runs a batch queue through runGroups, or replays the graph captured from an
// identical earlier queue (returns true then); see graph.cpp.
bool batchGraphRun(const std::vector<Loop *> &loops, bool eligible, const std::function<void()> &runGroups);
// fused loop pairs (batch.cpp): the pair handler for a followed by b, or null
DvmHandlerFunc pairHandlerFor(const Loop *a, const Loop *b);
void runPairGroup(const std::vector<Loop *> &firsts, const std::vector<Loop *> &seconds, DvmHandlerFunc func);
void prepareLoopLayouts(Loop *loop);
void resetPairStats();
void finishLoopReductionsExt(Loop *loop);
void graphReleaseAll();
void runtimeInit();   // settings, GPU and tuner set-up on first use
void noallocInit(int device);   // device address of the no-allocation fold's scratch (noalloc.hip)
void addPendingPost(const std::shared_ptr<Reduction> &r, char *val, char *loc, bool addInitial = false);

}  // namespace dvmh
