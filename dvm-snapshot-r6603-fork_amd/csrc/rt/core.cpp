// Runtime core of libdvmh_amd: settings, diagnostics, the GPU device, array
// storage with device-side layout transforms, and reduction semantics.
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <algorithm>

#include "rt.h"
#include "../kernels/launch.h"

namespace dvmh {

Settings settings;
Gpu gpu;
Runtime rt;

// ---------------------------------------------------------------- diagnostics
void fatal(const char *kind, const char *fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fprintf(stderr, "DVMH %s (%s:%ld): %s\n", kind, rt.curFile.c_str(), rt.curLine, buf);
    fflush(stderr);
    abort();
}

void logMsg(int level, const char *fmt, ...) {
    if (level > settings.logLevel) return;
    static const char *names[] = {"FATAL", "ERROR", "WARNING", "INFO", "DEBUG", "TRACE"};
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[dvmh %s] %s\n", names[level < 0 ? 0 : (level > 5 ? 5 : level)], buf);
}

// ---------------------------------------------------------------- settings
static long envLong(const char *name, long def) {
    const char *v = getenv(name);
    if (!v || !*v) return def;
    char *end = nullptr;
    long r = strtol(v, &end, 10);
    DVMH_CHECK(end && *end == 0, "Invalid value '%s' of environment variable %s (an integer is expected)", v, name);
    return r;
}

void Settings::load() {
    logLevel = (int)envLong("DVMH_LOGLEVEL", LOG_WARN);
    numGpus = (int)envLong("DVMH_NUM_CUDAS", 1);
    DVMH_CHECK(numGpus >= 0, "DVMH_NUM_CUDAS must be non-negative");
    hostThreads = (int)envLong("DVMH_NUM_THREADS", 0);
    requireGpu = envLong("DVMH_AMD_REQUIRE_GPU", 0) != 0;
    tuneMode = (int)envLong("DVMH_AMD_TUNE", envLong("DVMH_OPT_PARAMS", 0) ? 1 : 0);
    DVMH_CHECK(tuneMode >= 0 && tuneMode <= 3, "DVMH_AMD_TUNE must be 0..3");
    syncMode = (int)envLong("DVMH_AMD_SYNC", 0);
    allowAsync = envLong("DVMH_ALLOW_ASYNC", 0) != 0;   // rts/dvmh/settings.cpp:361
    pollDone = envLong("DVMH_AMD_POLL", 1) != 0;
    graph = envLong("DVMH_AMD_GRAPH", 0) != 0;
    graphMinNodes = envLong("DVMH_AMD_GRAPH_MIN_NODES", 6);
    tuneRepeats = (int)envLong("DVMH_AMD_TUNE_REPEATS", 1);
    if (tuneRepeats < 1) tuneRepeats = 1;
    tuneTransforms = envLong("DVMH_AMD_TUNE_TRANSFORMS", 1) != 0;
    fuse = envLong("DVMH_AMD_FUSE", 1) != 0;
    index64 = envLong("DVMH_AMD_INDEX64", 0) != 0;
    const char *db = getenv("DVMH_AMD_TUNEDB");
    tuneDb = db ? db : "";
    const char *mf = getenv("DVMH_AMD_MODEL");
    modelFile = mf ? mf : "";
}

// ---------------------------------------------------------------- GPU device
void Gpu::init() {
    present = false;
    if (settings.numGpus == 0) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return;
    }
    ordinal = 0;
    HIP_CHECK(hipSetDevice(ordinal));
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, ordinal));
    props.cuCount = p.multiProcessorCount;
    props.waveSize = p.warpSize;
    props.maxThreadsPerBlock = p.maxThreadsPerBlock;
    for (int i = 0; i < 3; i++) props.maxBlock[i] = p.maxThreadsDim[i];
    props.ldsPerBlock = (long)p.sharedMemPerBlock;
    props.regsPerCu = p.regsPerMultiprocessor;
    props.l2Bytes = p.l2CacheSize;
    props.clockKhz = p.clockRate;
    props.memClockKhz = p.memoryClockRate;
    props.busWidth = p.memoryBusWidth;
    snprintf(props.name, sizeof(props.name), "%s", p.name);
    snprintf(props.arch, sizeof(props.arch), "%.63s", p.gcnArchName);
    // Default-priority, non-blocking stream (never a prioritised one).
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    noallocInit(ordinal);
    // arrival counters of the multi-reduction finish: zeroed once here, every
    // launch leaves them zero again
    HIP_CHECK(hipMalloc((void **)&finishCounters, 64 * sizeof(unsigned)));
    HIP_CHECK(hipMemset(finishCounters, 0, 64 * sizeof(unsigned)));
    void *done = nullptr;
    HIP_CHECK(hipHostMalloc(&done, 64, hipHostMallocMapped));
    doneHost = (volatile unsigned long long *)done;
    *doneHost = 0;
    HIP_CHECK(hipHostGetDevicePointer((void **)&doneDev, done, 0));
    HIP_CHECK(hipDeviceSynchronize());
    present = true;
    logMsg(LOG_INFO, "GPU %s (%s): %d CUs, wave%d", props.name, props.arch, props.cuCount, props.waveSize);
}

// Waiting on the completion word skips the runtime's signal wait (a stream
// sync costs several microseconds more than a poll of host memory).  Every
// 4096 polls the stream is queried: once it has drained the word must be set.
void Gpu::waitDone(unsigned long long seq) {
    for (unsigned long spins = 1;; spins++) {
        if (__atomic_load_n(doneHost, __ATOMIC_ACQUIRE) >= seq) return;
        if ((spins & 4095) == 0) {
            hipError_t q = hipStreamQuery(stream);
            if (q == hipSuccess) {
                DVMH_INTERNAL(__atomic_load_n(doneHost, __ATOMIC_ACQUIRE) >= seq,
                        "reduction finish %llu drained without setting its completion word", seq);
                return;
            }
            if (q != hipErrorNotReady) HIP_CHECK(q);
        }
        __builtin_ia32_pause();
    }
}

void *Gpu::alloc(size_t bytes) {
    if (bytes == 0) bytes = 8;
    size_t rounded = (bytes + 255) & ~(size_t)255;
    auto it = freeBlocks.find(rounded);
    void *p = nullptr;
    if (it != freeBlocks.end() && !it->second.empty()) {
        p = it->second.back();
        it->second.pop_back();
    } else {
        hipError_t e = hipMalloc(&p, rounded);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trimCache();
            HIP_CHECK(hipMalloc(&p, rounded));
        }
    }
    liveBlocks[p] = rounded;
    return p;
}

void Gpu::release(void *p) {
    if (!p) return;
    auto it = liveBlocks.find(p);
    DVMH_INTERNAL(it != liveBlocks.end(), "releasing unknown device block %p", p);
    if (graphCapturing()) {
        // a graph under construction references it: it stays live with the graph
        graphNoteAlloc(p);
        return;
    }
    freeBlocks[it->second].push_back(p);
    liveBlocks.erase(it);
}

void Gpu::trimCache() {
    HIP_CHECK(hipStreamSynchronize(stream));
    for (auto &kv : freeBlocks)
        for (void *p : kv.second) HIP_CHECK(hipFree(p));
    freeBlocks.clear();
}

void *Gpu::pinnedStaging(size_t bytes) {
    if (bytes > stagingBytes) {
        if (staging) {
            HIP_CHECK(hipStreamSynchronize(stream));
            HIP_CHECK(hipHostFree(staging));
        }
        size_t n = std::max(bytes, (size_t)1 << 16);
        HIP_CHECK(hipHostMalloc(&staging, n, hipHostMallocDefault));
        stagingBytes = n;
    }
    return staging;
}

void Gpu::h2d(void *dst, const void *src, size_t bytes) {
    h2dBytes += (long)bytes;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
}

void Gpu::d2h(void *dst, const void *src, size_t bytes) {
    d2hBytes += (long)bytes;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
}

void Gpu::sync() {
    if (present) HIP_CHECK(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------- objects
Object *asObject(DvmType handle) {
    Object *o = (Object *)handle;
    DVMH_CHECK(o != nullptr, "NULL object handle");
    DVMH_CHECK(o->magic == OBJ_MAGIC, "Invalid or deleted object handle %p", (void *)o);
    return o;
}

Data *asData(const DvmType desc[], const char *where) {
    DVMH_CHECK(desc != nullptr, "NULL array descriptor passed to %s", where);
    Object *o = asObject(desc[0]);
    DVMH_CHECK(o->kind == Kind::Data, "Only an array can be passed to %s", where);
    return (Data *)o;
}

Loop *asLoop(DvmType h, const char *where) {
    Object *o = asObject(h);
    DVMH_CHECK(o->kind == Kind::Loop, "Incorrect loop reference is passed to %s", where);
    return (Loop *)o;
}

SpecLoop *asSpec(DvmType h, const char *where) {
    Object *o = asObject(h);
    DVMH_CHECK(o->kind == Kind::SpecLoop, "Incorrect loop reference is passed to %s (a handler loop_ref is expected)", where);
    return (SpecLoop *)o;
}

Loop *loopOf(DvmType h, const char *where) {
    Object *o = asObject(h);
    if (o->kind == Kind::SpecLoop) return ((SpecLoop *)o)->loop;
    DVMH_CHECK(o->kind == Kind::Loop, "Incorrect loop reference is passed to %s", where);
    return (Loop *)o;
}

long rtTypeSize(int t) {
    switch (t) {
        case rt_CHAR: case rt_UCHAR: return 1;
        case rt_SHORT: case rt_USHORT: return 2;
        case rt_INT: case rt_UINT: case rt_LOGICAL: case rt_FLOAT: return 4;
        case rt_LONG: case rt_ULONG: case rt_LLONG: case rt_ULLONG: case rt_DOUBLE: case rt_POINTER: return 8;
        case rt_FLOAT_COMPLEX: return 8;
        case rt_DOUBLE_COMPLEX: return 16;
        default: return 0;
    }
}

// ---------------------------------------------------------------- Data
Data::~Data() {
    if (hasDev && dev.ptr) gpu.release(dev.ptr);
    if (ownsHost) free(host);
    magic = 0;
}

// ------------------------------------------------------------ pieces
static bool disjoint(int rank, const Box &a, const Box &b) {
    for (int i = 0; i < rank; i++)
        if (a.hi[i] < b.lo[i] || b.hi[i] < a.lo[i]) return true;
    return false;
}

static bool boxEmpty(int rank, const Box &a) {
    for (int i = 0; i < rank; i++)
        if (a.hi[i] < a.lo[i]) return true;
    return false;
}

// x minus b as at most 2 * rank disjoint boxes: slabs below and above b on
// each axis in turn, the remaining core clipped to b on the axes done so far.
static void subtractBox(int rank, const Box &x, const Box &b, std::vector<Box> &out) {
    if (disjoint(rank, x, b)) {
        out.push_back(x);
        return;
    }
    Box cur = x;
    for (int a = 0; a < rank; a++) {
        if (cur.lo[a] < b.lo[a]) {
            Box s = cur;
            s.hi[a] = b.lo[a] - 1;
            out.push_back(s);
            cur.lo[a] = b.lo[a];
        }
        if (cur.hi[a] > b.hi[a]) {
            Box s = cur;
            s.lo[a] = b.hi[a] + 1;
            out.push_back(s);
            cur.hi[a] = b.hi[a];
        }
    }
}

void Pieces::remove(const Box &b) {
    if (boxEmpty(rank, b)) return;
    std::vector<Box> out;
    for (const Box &x : boxes) subtractBox(rank, x, b, out);
    boxes.swap(out);
}

void Pieces::add(const Box &b) {
    if (boxEmpty(rank, b)) return;
    remove(b);
    boxes.push_back(b);
    // merge boxes that differ on one axis only and touch there
    for (bool merged = true; merged;) {
        merged = false;
        for (size_t i = 0; i < boxes.size() && !merged; i++)
            for (size_t j = i + 1; j < boxes.size() && !merged; j++) {
                int diff = -1, ndiff = 0;
                for (int a = 0; a < rank; a++)
                    if (boxes[i].lo[a] != boxes[j].lo[a] || boxes[i].hi[a] != boxes[j].hi[a]) { diff = a; ndiff++; }
                if (ndiff != 1) continue;
                Box &x = boxes[i];
                const Box &y = boxes[j];
                if (x.hi[diff] + 1 == y.lo[diff]) x.hi[diff] = y.hi[diff];
                else if (y.hi[diff] + 1 == x.lo[diff]) x.lo[diff] = y.lo[diff];
                else continue;
                boxes.erase(boxes.begin() + (long)j);
                merged = true;
            }
    }
}

std::vector<Box> Pieces::missing(const Box &b) const {
    std::vector<Box> rest;
    if (boxEmpty(rank, b)) return rest;
    rest.push_back(b);
    for (const Box &x : boxes) {
        std::vector<Box> next;
        for (const Box &r : rest) subtractBox(rank, r, x, next);
        rest.swap(next);
        if (rest.empty()) break;
    }
    return rest;
}

long Pieces::elements() const {
    long n = 0;
    for (const Box &x : boxes) {
        long c = 1;
        for (int a = 0; a < rank; a++) c *= x.hi[a] - x.lo[a] + 1;
        n += c;
    }
    return n;
}

bool Pieces::holds(const long point[]) const {
    for (const Box &x : boxes) {
        int a = 0;
        while (a < rank && point[a] >= x.lo[a] && point[a] <= x.hi[a]) a++;
        if (a == rank) return true;
    }
    return false;
}

Box Data::whole() const {
    Box w;
    for (int a = 0; a < rank; a++) { w.lo[a] = space[a].lo; w.hi[a] = space[a].hi; }
    return w;
}

void Data::hostOwnsAll() {
    hostPieces.rank = devPieces.rank = rank;
    hostPieces.clear();
    hostPieces.add(whole());
    devPieces.clear();
}

void Data::devOwnsAll() {
    hostPieces.rank = devPieces.rank = rank;
    devPieces.clear();
    devPieces.add(whole());
    hostPieces.clear();
}

long Data::elemCount() const {
    long n = 1;
    for (int i = 0; i < rank; i++) n *= space[i].size();
    return n;
}

long Data::hostStride(int axis) const {
    long s = 1;
    for (int i = rank; i > axis; i--) s *= space[i - 1].size();
    return s;
}

bool DevBuffer::isTransformed(int rank) const {
    if (diagState) return true;
    for (int i = 0; i < rank; i++)
        if (axisPerm[i] != i + 1) return true;
    return false;
}

void Data::ensureDev() {
    DVMH_INTERNAL(gpu.present, "device buffer requested without a GPU");
    if (hasDev) return;
    dev.bytes = bytes();
    dev.ptr = gpu.alloc(dev.bytes);
    for (int i = 0; i < rank; i++) dev.axisPerm[i] = i + 1;
    dev.diagState = 0;
    hasDev = true;
    devPieces.rank = rank;
    devPieces.clear();
}

void Data::toDevice() {
    ensureDev();
    if (dev.isTransformed(rank)) {
        // a fresh natural copy replaces the transformed one
        for (int i = 0; i < rank; i++) dev.axisPerm[i] = i + 1;
        dev.diagState = 0;
    }
    gpu.h2d(dev.ptr, host, bytes());
    devPieces.rank = rank;
    devPieces.clear();
    devPieces.add(whole());
}

// Sizes of storage positions of the current device layout.
static void storageSizes(const Data *d, long out[]) {
    for (int q = 0; q < d->rank; q++) out[q] = d->space[d->dev.axisPerm[q] - 1].size();
}

static long layoutChangeCount = 0;
long layoutChanges() { return layoutChangeCount; }

// Re-lay the device copy.  newPerm[q] = original axis stored at position q.
void Data::setLayout(const int newPerm[], int newDiag) {
    DVMH_INTERNAL(hasDev, "setLayout without a device copy");
    if (rank <= 1) return;
    int perm[MAX_RANK];
    for (int i = 0; i < rank; i++) perm[i] = newPerm ? newPerm[i] : i + 1;
    // a diagonal over a degenerate plane is no diagonal (DvmhBuffer::canonicalizeTransformState)
    if (newDiag) {
        long Rx = space[perm[rank - 1] - 1].size(), Ry = space[perm[rank - 2] - 1].size();
        if ((newDiag == 1 && (Rx == 1 || Ry == 1)) || (newDiag == 2 && Ry == 1)) newDiag = 0;
    }
    bool same = newDiag == dev.diagState;
    for (int i = 0; i < rank && same; i++) same = perm[i] == dev.axisPerm[i];
    if (same) return;
    layoutChangeCount++;
    const size_t nb = bytes();
    long sizes[MAX_RANK];
    if (dev.diagState) {
        // undo the diagonal first (keeps the permutation)
        storageSizes(this, sizes);
        long Rx = sizes[rank - 1], Ry = sizes[rank - 2], Rz = 1;
        for (int q = 0; q < rank - 2; q++) Rz *= sizes[q];
        DVMH_INTERNAL((size_t)(Rx * Ry * Rz) * typeSize == nb, "diagonal planes do not cover the array");
        void *tmp = gpu.alloc(nb);
        dvmh_hip_diagonal(dev.ptr, tmp, typeSize, Rx, Ry, Rz, 1, dev.diagState == 2, gpu.stream);
        gpu.release(dev.ptr);
        dev.ptr = tmp;
        dev.diagState = 0;
    }
    bool permSame = true;
    for (int i = 0; i < rank; i++) permSame = permSame && perm[i] == dev.axisPerm[i];
    if (!permSame) {
        storageSizes(this, sizes);
        int dstFromSrc[MAX_RANK];
        for (int p = 0; p < rank; p++) {
            int srcPos = -1;
            for (int q = 0; q < rank; q++) if (dev.axisPerm[q] == perm[p]) srcPos = q;
            DVMH_INTERNAL(srcPos >= 0, "bad axis permutation");
            dstFromSrc[p] = srcPos;
        }
        void *tmp = gpu.alloc(nb);
        dvmh_hip_permute(dev.ptr, tmp, typeSize, rank, sizes, dstFromSrc, gpu.stream);
        gpu.release(dev.ptr);
        dev.ptr = tmp;
        for (int i = 0; i < rank; i++) dev.axisPerm[i] = perm[i];
    }
    if (newDiag) {
        storageSizes(this, sizes);
        long Rx = sizes[rank - 1], Ry = sizes[rank - 2], Rz = 1;
        for (int q = 0; q < rank - 2; q++) Rz *= sizes[q];
        DVMH_INTERNAL((size_t)(Rx * Ry * Rz) * typeSize == nb, "diagonal planes do not cover the array");
        void *tmp = gpu.alloc(nb);
        dvmh_hip_diagonal(dev.ptr, tmp, typeSize, Rx, Ry, Rz, 0, newDiag == 2, gpu.stream);
        gpu.release(dev.ptr);
        dev.ptr = tmp;
        dev.diagState = newDiag;
    }
}

void Data::toHost() {
    DVMH_INTERNAL(hasDev, "toHost without a device copy");
    if (!dev.isTransformed(rank)) {
        gpu.d2h(host, dev.ptr, bytes());
    } else {
        // bring a natural-layout copy back without disturbing the device layout
        int keepPerm[MAX_RANK], keepDiag = dev.diagState;
        for (int i = 0; i < rank; i++) keepPerm[i] = dev.axisPerm[i];
        void *keep = gpu.alloc(bytes());
        HIP_CHECK(hipMemcpyAsync(keep, dev.ptr, bytes(), hipMemcpyDeviceToDevice, gpu.stream));
        setLayout(nullptr, 0);
        gpu.d2h(host, dev.ptr, bytes());
        gpu.release(dev.ptr);
        dev.ptr = keep;
        for (int i = 0; i < rank; i++) dev.axisPerm[i] = keepPerm[i];
        dev.diagState = keepDiag;
    }
    hostPieces.rank = rank;
    hostPieces.clear();
    hostPieces.add(whole());
}

// DvmhBuffer::fillHeader (rts/dvmh/dvmh_buffer.cpp:149-182) for the device copy.
void Data::fillDevHeader(const void *base, DvmType header[], DvmType ext[]) const {
    long collector = 1, offset = 0;
    header[0] = (DvmType)this;
    header[rank + 2] = (DvmType)base;
    DVMH_INTERNAL((((long)dev.ptr - (long)base) % typeSize) == 0, "impossible to express the array offset from the base");
    for (int i = rank; i >= 1; i--) {
        int origAx = dev.axisPerm[i - 1];
        if (i >= rank - 1 && dev.diagState != 0) {
            header[origAx] = 0;
        } else {
            offset += collector * space[origAx - 1].lo;
            header[origAx] = collector;
        }
        collector *= space[origAx - 1].size();
    }
    header[rank + 1] = ((long)dev.ptr - (long)base) / typeSize - offset;
    if (ext && dev.diagState != 0) {
        int xAxis = dev.axisPerm[rank - 1], yAxis = dev.axisPerm[rank - 2];
        ext[0] = xAxis;
        ext[1] = space[xAxis - 1].lo;
        ext[2] = space[xAxis - 1].size();
        ext[3] = yAxis;
        ext[4] = space[yAxis - 1].lo;
        ext[5] = space[yAxis - 1].size();
        ext[6] = dev.diagState == 2;
    }
}

// DvmhBuffer::getNaturalBase (rts/dvmh/dvmh_buffer.cpp:184-195).
void *Data::devNaturalBase() const {
    long collector = 1, offset = 0;
    for (int i = rank; i >= 1; i--) {
        int origAx = dev.axisPerm[i - 1];
        if (dev.diagState == 0 || i < rank - 1) offset += collector * space[origAx - 1].lo;
        collector *= space[origAx - 1].size();
    }
    return (char *)dev.ptr - offset * typeSize;
}

void *Data::hostNaturalBase() const {
    long collector = 1, offset = 0;
    for (int i = rank; i >= 1; i--) {
        offset += collector * space[i - 1].lo;
        collector *= space[i - 1].size();
    }
    return host - offset * typeSize;
}

// ---------------------------------------------------------------- reductions
// Element operations of DvmhReduction (rts/dvmh/loop.cpp:37-254): SUM/PROD and
// the bit/logic family start from an identity, MAX/MIN from the variable's value
// at registration; MAX/MIN replace on strict compare and then copy the loc.
// Unsigned element types compare as unsigned (the reference folds only the
// signed ones, rts/dvmh/loop.cpp:125-254).
template <typename T>
static void initTyped(int func, void *p) {
    T v = (func == rf_PROD) ? (T)1 : (T)0;
    std::memcpy(p, &v, sizeof(T));
}

void Reduction::initValues(void *arr, void *loc) const {
    if (arr) {
        bool identity = func == rf_SUM || func == rf_PROD || func == rf_XOR || func == rf_EQU || func == rf_NE || func == rf_EQ;
        if (identity) {
            for (long j = 0; j < elemCount; j++) {
                char *p = (char *)arr + j * elemSize;
                if (func == rf_EQU || func == rf_EQ) {
                    std::memset(p, 0, (size_t)elemSize);
                    p[0] = 1;   // logical true in the element's low byte (little endian)
                    continue;
                }
                switch (rtType) {
                    case rt_CHAR: case rt_UCHAR: initTyped<char>(func, p); break;
                    case rt_INT: case rt_UINT: case rt_LOGICAL: initTyped<int>(func, p); break;
                    case rt_LONG: case rt_ULONG: initTyped<long>(func, p); break;
                    case rt_LLONG: case rt_ULLONG: initTyped<long long>(func, p); break;
                    case rt_FLOAT: initTyped<float>(func, p); break;
                    case rt_DOUBLE: initTyped<double>(func, p); break;
                    case rt_FLOAT_COMPLEX: { float v[2] = {func == rf_PROD ? 1.0f : 0.0f, 0.0f}; std::memcpy(p, v, 8); break; }
                    case rt_DOUBLE_COMPLEX: { double v[2] = {func == rf_PROD ? 1.0 : 0.0, 0.0}; std::memcpy(p, v, 16); break; }
                    default: fatal("internal error", "reduction type %d is not supported", rtType);
                }
            }
        } else {
            std::memcpy(arr, initArr.data(), (size_t)(elemSize * elemCount));
        }
    }
    if (isLoc() && loc) std::memcpy(loc, initLoc.data(), (size_t)(locSize * elemCount));
}

template <typename T>
static int opTyped(int func, void *d, const void *s) {
    T a, b;
    std::memcpy(&a, d, sizeof(T));
    std::memcpy(&b, s, sizeof(T));
    int replaced = 0;
    switch (func) {
        case rf_SUM: a = a + b; break;
        case rf_PROD: a = a * b; break;
        case rf_MAX: case rf_MAXLOC: if (a < b) { a = b; replaced = 1; } else if (a == b) replaced = 2; break;
        case rf_MIN: case rf_MINLOC: if (a > b) { a = b; replaced = 1; } else if (a == b) replaced = 2; break;
        default: return -1;
    }
    std::memcpy(d, &a, sizeof(T));
    return replaced;   // 2: equal values (MAX/MIN family)
}

template <typename T>
static int opBits(int func, void *d, const void *s) {
    T a, b;
    std::memcpy(&a, d, sizeof(T));
    std::memcpy(&b, s, sizeof(T));
    switch (func) {
        case rf_AND: a &= b; break;
        case rf_OR: a |= b; break;
        case rf_XOR: a ^= b; break;
        case rf_EQU: case rf_EQ: a = (a == b); break;
        case rf_NE: a = (a != b); break;
        default: return opTyped<T>(func, d, s);
    }
    std::memcpy(d, &a, sizeof(T));
    return 0;
}

template <typename T>
static int opComplex(int func, void *d, const void *s) {
    T a[2], b[2];
    std::memcpy(a, d, sizeof(a));
    std::memcpy(b, s, sizeof(b));
    if (func == rf_SUM) { a[0] = a[0] + b[0]; a[1] = a[1] + b[1]; }
    else if (func == rf_PROD) {
        T re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
        a[0] = re;
        a[1] = im;
    } else return -1;
    std::memcpy(d, a, sizeof(a));
    return 0;
}

void Reduction::performOperation(void *dArr, void *dLoc, const void *sArr, const void *sLoc, bool lowerLocOnTie) const {
    for (long i = 0; i < elemCount; i++) {
        void *d = (char *)dArr + elemSize * i;
        const void *s = (const char *)sArr + elemSize * i;
        int r = -1;
        switch (rtType) {
            case rt_CHAR: r = opBits<signed char>(func, d, s); break;
            case rt_UCHAR: r = opBits<unsigned char>(func, d, s); break;
            case rt_INT: case rt_LOGICAL: r = opBits<int>(func, d, s); break;
            case rt_UINT: r = opBits<unsigned int>(func, d, s); break;
            case rt_LONG: r = opBits<long>(func, d, s); break;
            case rt_ULONG: r = opBits<unsigned long>(func, d, s); break;
            case rt_LLONG: r = opBits<long long>(func, d, s); break;
            case rt_ULLONG: r = opBits<unsigned long long>(func, d, s); break;
            case rt_FLOAT: r = opTyped<float>(func, d, s); break;
            case rt_DOUBLE: r = opTyped<double>(func, d, s); break;
            case rt_FLOAT_COMPLEX: r = opComplex<float>(func, d, s); break;
            case rt_DOUBLE_COMPLEX: r = opComplex<double>(func, d, s); break;
            default: break;
        }
        DVMH_INTERNAL(r != -1, "No suitable reduction found for function %d and type %d", func, rtType);
        if (isLoc() && dLoc && sLoc &&
                (r == 1 || (r == 2 && lowerLocOnTie && locKeyLess((const char *)sLoc + locSize * i, (const char *)dLoc + locSize * i))))
            std::memcpy((char *)dLoc + locSize * i, (const char *)sLoc + locSize * i, (size_t)locSize);
    }
}

// Tie rule of MAXLOC/MINLOC merges: a loc record of 1, 2, 4 or 8 bytes is a
// scalar integer (read as signed) and the lower one wins, as in the GPU folds;
// other records keep the destination's.
bool Reduction::locKeyLess(const void *a, const void *b) const {
    switch (locSize) {
        case 1: return *(const signed char *)a < *(const signed char *)b;
        case 2: { short x, y; std::memcpy(&x, a, 2); std::memcpy(&y, b, 2); return x < y; }
        case 4: { int x, y; std::memcpy(&x, a, 4); std::memcpy(&y, b, 4); return x < y; }
        case 8: { long x, y; std::memcpy(&x, a, 8); std::memcpy(&y, b, 8); return x < y; }
        default: return false;
    }
}

void Reduction::postValues(const void *arr, const void *loc, bool lowerLocOnTie) {
    std::lock_guard<std::mutex> g(lock);
    performOperation(arrayAddr, locAddr, arr, loc, lowerLocOnTie);
}

}  // namespace dvmh
