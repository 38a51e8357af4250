// extern "C" entry points of the DVMH interface (include/dvmhlib2.h).
// Each function cites the reference implementation whose contract it keeps
// (rts/dvmh/dvmh2.cpp unless stated otherwise).
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <sys/time.h>
#include <unistd.h>

#include "rt.h"
#include "../kernels/launch.h"

using namespace dvmh;

namespace dvmh {
void tunerInit();
char *pinnedTake(size_t bytes);
}

// ------------------------------------------------------------ helper indices
// dvmh_alignment_linear_C, dvmh_distribution_*_C and dvmh_handler_func_C save
// their arguments and return a 1-based index; the enclosing top-level call
// consumes them and the vector is cleared when that call returns
// (rts/dvmh/dvmh2.cpp:94-137).
namespace {
union Saved {
    DvmType i;
    void *p;
};
std::vector<Saved> saved;

int saveInt(DvmType v) { Saved s; s.i = v; saved.push_back(s); return (int)saved.size(); }
int savePtr(const void *p) { Saved s; s.p = (void *)p; saved.push_back(s); return (int)saved.size(); }

struct Guard {
    ~Guard() { saved.clear(); }
};

struct AlignRule {
    long axis, mult, summand;
};
AlignRule loadAlign(DvmType idx) {
    DVMH_CHECK(idx >= 1 && idx + 2 <= (DvmType)saved.size(), "Invalid alignment helper index %ld", idx);
    return AlignRule{saved[idx - 1].i, saved[idx].i, saved[idx + 1].i};
}

HandlerDesc loadHandler(DvmType idx) {
    DVMH_CHECK(idx >= 1 && idx + 1 <= (DvmType)saved.size(), "Invalid handler helper index %ld", idx);
    HandlerDesc h;
    h.func = (DvmHandlerFunc)saved[idx - 1].p;
    long n = saved[idx].i;
    DVMH_CHECK(idx + 1 + n <= (DvmType)saved.size(), "Corrupted handler helper");
    h.params.resize((size_t)n);
    for (long i = 0; i < n; i++) h.params[(size_t)i] = saved[idx + 1 + i].p;
    return h;
}

// Per-array alignment (dvmh_align_C): target object and, per target axis, the
// array axis tied to it.
struct AlignInfo {
    const Object *target = nullptr;
    int targetRank = 0;
    AlignRule rules[MAX_RANK];
};
std::map<const Data *, AlignInfo> alignments;

std::map<const void *, DvmType *> genHeaders;   // regular variables: runtime-owned headers

double wallTime() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + tv.tv_usec * 1e-6;
}

void ensureInit() {
    if (rt.inited) return;
    settings.load();
    gpu.init();
    tunerInit();
    rt.inited = true;
}

void allocateHost(Data *d) {
    if (d->host) return;
    size_t nb = d->bytes();
    d->host = (char *)calloc(nb ? nb : 1, 1);
    DVMH_CHECK(d->host != nullptr, "Out of host memory allocating %zu bytes", nb);
    d->ownsHost = true;
    d->hostOwnsAll();
}

void fillUserHeader(Data *d) {
    // Host-side header of a DVM array: [0] handle, [1..rank] host strides,
    // [rank+1] offset from the natural base, [rank+2] natural base.
    DvmType *h = d->userHeader;
    if (!h) return;
    h[0] = (DvmType)d;
    for (int a = 1; a <= d->rank; a++) h[a] = d->hostStride(a);
    h[d->rank + 1] = 0;
    h[d->rank + 2] = (DvmType)(d->host ? d->hostNaturalBase() : nullptr);
}

Data *regularVarOf(const void *addr) {
    auto it = rt.regularVars.find(addr);
    return it == rt.regularVars.end() ? nullptr : it->second;
}
}  // namespace

namespace dvmh {
// Compose array axis -> target axis through alignment links up to `root`.
bool alignToRoot(const Data *d, const Object *root, long tAxisOfArrayAxis[], long sign[]) {
    long curAxis[MAX_RANK], curSign[MAX_RANK];
    for (int a = 0; a < d->rank; a++) { curAxis[a] = a + 1; curSign[a] = 1; }
    const Object *cur = d;
    for (int hop = 0; hop < 16; hop++) {
        if (cur == root) {
            for (int a = 0; a < d->rank; a++) { tAxisOfArrayAxis[a] = curAxis[a]; sign[a] = curSign[a]; }
            return true;
        }
        if (cur->kind != Kind::Data) return false;
        auto it = alignments.find((const Data *)cur);
        if (it == alignments.end()) return false;
        const AlignInfo &ai = it->second;
        // array axis currently mapped to axis x of `cur` -> the target axis t with rules[t].axis == x
        for (int a = 0; a < d->rank; a++) {
            if (curAxis[a] <= 0) continue;
            long found = 0, sg = 1;
            for (int t = 0; t < ai.targetRank; t++)
                if (ai.rules[t].axis == curAxis[a]) { found = t + 1; sg = ai.rules[t].mult < 0 ? -1 : 1; }
            curAxis[a] = found;
            curSign[a] *= sg;
        }
        cur = ai.target;
    }
    return false;
}

// Lazy initialisation for entry points that may be the first call into the library.
void runtimeInit() { ensureInit(); }
}  // namespace dvmh

// ============================================================ init / exit
extern "C" void dvmh_line_C(DvmType lineNumber, const char fileName[]) {
    rt.curLine = lineNumber;
    if (fileName) rt.curFile = fileName;
}

extern "C" void dvmh_init_C(DvmType flags, int *pArgc, char ***pArgv) {   // :411
    (void)flags; (void)pArgc; (void)pArgv;
    ensureInit();
}

extern "C" void dvmh_init_lib_C(DvmType flags) {
    (void)flags;
    ensureInit();
}

extern "C" void dvmh_exit_C(DvmType exitCode) {
    flushDeferredPosts();
    gpu.sync();
    tunerShutdown();
    fflush(stdout);
    exit((int)exitCode);
}

// ============================================================ arrays
static void createArray(DvmType dvmDesc[], DvmType rank, DvmType typeSize, va_list &ap, bool declareOnly) {
    ensureInit();
    DVMH_CHECK(dvmDesc != nullptr, "NULL descriptor passed to dvmh_array_create_C");
    DVMH_CHECK(rank >= 1 && rank <= MAX_RANK, "Array rank %ld is out of range 1..%d", rank, MAX_RANK);
    Data *d = new Data();
    d->rank = (int)rank;
    if (typeSize < 0) {
        d->rtType = (int)-typeSize;
        d->typeSize = rtTypeSize(d->rtType);
        DVMH_CHECK(d->typeSize > 0, "Unknown element type code %ld", -typeSize);
    } else {
        d->typeSize = typeSize;
        d->rtType = rt_UNKNOWN;
    }
    for (int a = 0; a < rank; a++) {
        DvmType size = va_arg(ap, DvmType);
        d->shdLow[a] = va_arg(ap, DvmType);
        d->shdHigh[a] = va_arg(ap, DvmType);
        DVMH_CHECK(declareOnly || size >= 1, "Array axis %d has non-positive size %ld", a + 1, size);
        d->space[a].lo = 0;
        d->space[a].hi = size - 1;
    }
    d->userHeader = dvmDesc;
    dvmDesc[0] = (DvmType)d;
    for (int a = 1; a < 64 && a <= rank + 2; a++) dvmDesc[a] = 0;
    rt.allData.push_back(d);
}

extern "C" void dvmh_array_declare_C(DvmType dvmDesc[], DvmType rank, DvmType typeSize, ...) {   // :29
    Guard g;
    va_list ap;
    va_start(ap, typeSize);
    createArray(dvmDesc, rank, typeSize, ap, true);
    va_end(ap);
}

extern "C" void dvmh_array_create_C(DvmType dvmDesc[], DvmType rank, DvmType typeSize, ...) {    // :30
    Guard g;
    va_list ap;
    va_start(ap, typeSize);
    createArray(dvmDesc, rank, typeSize, ap, false);
    va_end(ap);
}

extern "C" void dvmh_template_create_C(DvmType dvmDesc[], DvmType rank, ...) {   // :33
    Guard g;
    ensureInit();
    DVMH_CHECK(rank >= 1 && rank <= MAX_RANK, "Template rank %ld is out of range", rank);
    Template *t = new Template();
    t->rank = (int)rank;
    va_list ap;
    va_start(ap, rank);
    for (int a = 0; a < rank; a++) {
        DvmType size = va_arg(ap, DvmType);
        t->space[a].lo = 0;
        t->space[a].hi = size - 1;
    }
    va_end(ap);
    dvmDesc[0] = (DvmType)t;
}

extern "C" void dvmh_array_alloc_C(DvmType dvmDesc[], DvmType byteCount) {   // :35
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_array_alloc_C");
    long rest = d->typeSize;
    for (int a = 1; a < d->rank; a++) rest *= d->space[a].size();
    DVMH_CHECK(rest > 0 && byteCount % rest == 0, "dvmh_array_alloc_C: %ld bytes is not a whole number of rows", byteCount);
    d->space[0].lo = 0;
    d->space[0].hi = byteCount / rest - 1;
}

extern "C" DvmType dvmh_handler_func_C(DvmHandlerFunc handlerFunc, DvmType customParamCount, ...) {   // :697
    DvmType res = savePtr((const void *)handlerFunc);
    saveInt(customParamCount);
    va_list ap;
    va_start(ap, customParamCount);
    for (long i = 0; i < customParamCount; i++) savePtr(va_arg(ap, void *));
    va_end(ap);
    return res;
}

extern "C" DvmType dvmh_distribution_replicated_() { return saveInt(0); }
extern "C" DvmType dvmh_distribution_block_C(DvmType mpsAxis) { return saveInt(mpsAxis); }
extern "C" DvmType dvmh_distribution_wgtblock_C(DvmType mpsAxis, DvmType, const void *, DvmType) { return saveInt(mpsAxis); }
extern "C" DvmType dvmh_distribution_genblock_C(DvmType mpsAxis, DvmType, const void *) { return saveInt(mpsAxis); }
extern "C" DvmType dvmh_distribution_multblock_C(DvmType mpsAxis, DvmType) { return saveInt(mpsAxis); }
extern "C" DvmType dvmh_distribution_indirect_C(DvmType, DvmType, const void *) {
    fatal("error", "INDIRECT distribution needs several processes; this runtime is single-process");
}
extern "C" DvmType dvmh_distribution_derived_C(DvmType, DvmType, DvmType, DvmType) {
    fatal("error", "DERIVED distribution needs several processes; this runtime is single-process");
}

// One process owns every element, so distributing means allocating the whole array.
extern "C" void dvmh_distribute_C(DvmType dvmDesc[], DvmType rank, ...) {   // :61
    Guard g;
    Object *o = asObject(dvmDesc[0]);
    if (o->kind == Kind::Template) return;
    Data *d = asData(dvmDesc, "dvmh_distribute_C");
    DVMH_CHECK(rank == d->rank, "dvmh_distribute_C: rank %ld does not match the array rank %d", rank, d->rank);
    DVMH_CHECK(d->space[0].size() >= 1, "dvmh_distribute_C: array has no size (missing dvmh_array_alloc_C)");
    allocateHost(d);
    fillUserHeader(d);
}

// A single process keeps every element in place; an array declared without a
// distribution (`#pragma dvm array`) gets its storage at its first one.
extern "C" void dvmh_redistribute_C(DvmType dvmDesc[], DvmType rank, ...) {   // :63
    Guard g;
    Object *o = asObject(dvmDesc[0]);
    if (o->kind == Kind::Template) return;
    Data *d = asData(dvmDesc, "dvmh_redistribute_C");
    DVMH_CHECK(rank == d->rank, "dvmh_redistribute_C: rank %ld does not match the array rank %d", rank, d->rank);
    DVMH_CHECK(d->space[0].size() >= 1, "dvmh_redistribute_C: array has no size (missing dvmh_array_alloc_C)");
    allocateHost(d);
    fillUserHeader(d);
}

extern "C" DvmType dvmh_alignment_linear_C(DvmType axis, DvmType multiplier, DvmType summand) {   // :1364
    DvmType res = saveInt(axis);
    saveInt(multiplier);
    saveInt(summand);
    return res;
}

extern "C" void dvmh_align_C(DvmType dvmDesc[], const DvmType templDesc[], DvmType templRank, ...) {   // :67
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_align_C");
    const Object *t = asObject(templDesc[0]);
    AlignInfo ai;
    ai.target = t;
    ai.targetRank = (int)templRank;
    va_list ap;
    va_start(ap, templRank);
    for (int i = 0; i < templRank; i++) ai.rules[i] = loadAlign(va_arg(ap, DvmType));
    va_end(ap);
    alignments[d] = ai;
    DVMH_CHECK(d->space[0].size() >= 1, "dvmh_align_C: array has no size (missing dvmh_array_alloc_C)");
    allocateHost(d);
    fillUserHeader(d);
}

extern "C" void dvmh_realign_C(DvmType dvmDesc[], DvmType, const DvmType templDesc[], DvmType templRank, ...) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_realign_C");
    AlignInfo ai;
    ai.target = asObject(templDesc[0]);
    ai.targetRank = (int)templRank;
    va_list ap;
    va_start(ap, templRank);
    for (int i = 0; i < templRank; i++) ai.rules[i] = loadAlign(va_arg(ap, DvmType));
    va_end(ap);
    alignments[d] = ai;
    // an array declared without a distribution gets its storage here
    DVMH_CHECK(d->space[0].size() >= 1, "dvmh_realign_C: array has no size (missing dvmh_array_alloc_C)");
    allocateHost(d);
    fillUserHeader(d);
}

static void deleteData(Data *d) {
    // captured graphs may hold its device address (or reuse it for another
    // object later): they are dropped with it
    if (d->hasDev) graphReleaseAll();
    alignments.erase(d);
    rt.allData.erase(std::remove(rt.allData.begin(), rt.allData.end(), d), rt.allData.end());
    for (auto it = rt.regularVars.begin(); it != rt.regularVars.end();)
        if (it->second == d) it = rt.regularVars.erase(it); else ++it;
    delete d;
}

extern "C" void dvmh_array_free_C(DvmType dvmDesc[]) {   // :77
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_array_free_C");
    deleteData(d);
    dvmDesc[0] = 0;
}

extern "C" void dvmh_delete_object_(DvmType dvmDesc[]) {   // :78
    Guard g;
    if (!dvmDesc || !dvmDesc[0]) return;
    Object *o = asObject(dvmDesc[0]);
    if (o->kind == Kind::Data) deleteData((Data *)o);
    else delete o;
    dvmDesc[0] = 0;
}

extern "C" void dvmh_forget_header_(DvmType dvmDesc[]) {
    if (dvmDesc) dvmDesc[0] = 0;
}

static void fetchMissing(Data *d, const Box &b);

// ============================================================ regular variables
static Data *registerRegular(const void *addr, long rank, long typeSize, const long sizes[]) {
    Data *d = regularVarOf(addr);
    if (d) return d;
    d = new Data();
    d->rank = (int)rank;
    d->typeSize = typeSize;
    for (int a = 0; a < rank; a++) { d->space[a].lo = 0; d->space[a].hi = sizes[a] - 1; }
    d->host = (char *)addr;
    d->ownsHost = false;
    d->isRegular = true;
    d->hostOwnsAll();
    rt.regularVars[addr] = d;
    rt.allData.push_back(d);
    return d;
}

extern "C" DvmType *dvmh_variable_gen_header_C(const void *addr, DvmType rank, DvmType typeSize, ...) {   // :81
    Guard g;
    ensureInit();
    long sizes[MAX_RANK];
    va_list ap;
    va_start(ap, typeSize);
    for (int a = 0; a < rank; a++) sizes[a] = va_arg(ap, DvmType);
    va_end(ap);
    long ts = typeSize < 0 ? rtTypeSize((int)-typeSize) : typeSize;
    Data *d = registerRegular(addr, rank, ts, sizes);
    if (typeSize < 0) d->rtType = (int)-typeSize;
    DvmType *&h = genHeaders[addr];
    if (!h) h = (DvmType *)calloc(64, sizeof(DvmType));
    d->userHeader = h;
    fillUserHeader(d);
    return h;
}

extern "C" DvmType *dvmh_variable_get_header_C(const void *addr) {
    auto it = genHeaders.find(addr);
    return it == genHeaders.end() ? nullptr : it->second;
}

extern "C" void dvmh_data_enter_C(const void *addr, DvmType size) {   // :87
    Guard g;
    ensureInit();
    if (size <= 0) return;
    long sizes[1] = {size};
    registerRegular(addr, 1, 1, sizes);
}

extern "C" void dvmh_data_exit_C(const void *addr, DvmType saveFlag) {
    Guard g;
    Data *d = regularVarOf(addr);
    if (!d) return;
    if (saveFlag) fetchMissing(d, d->whole());
    deleteData(d);
}

extern "C" void *dvmh_malloc_C(size_t size) { return malloc(size); }
extern "C" void *dvmh_calloc_C(size_t nmemb, size_t size) { return calloc(nmemb, size); }
extern "C" void *dvmh_realloc_C(void *ptr, size_t size) { return realloc(ptr, size); }
extern "C" char *dvmh_strdup_C(const char *s) { return strdup(s); }
extern "C" void dvmh_free_C(void *ptr) { free(ptr); }

// ============================================================ addressing
extern "C" void *dvmh_get_natural_base_C(DvmType deviceNum, const DvmType dvmDesc[]) {   // :2190
    DVMH_CHECK(deviceNum == 0 || deviceNum == 1, "Illegal device number %ld passed to dvmh_get_natural_base_C", deviceNum);
    Data *d = asData(dvmDesc, "dvmh_get_natural_base_C");
    if (deviceNum == 0) return d->hostNaturalBase();
    return d->hasDev ? d->devNaturalBase() : nullptr;
}

extern "C" void *dvmh_get_device_addr_C(DvmType deviceNum, const void *addr) {   // :2205
    if (deviceNum == 0) return (void *)addr;
    Data *d = regularVarOf(addr);
    DVMH_CHECK(d != nullptr, "Can not get device address of unknown variable %p", addr);
    return d->hasDev ? d->dev.ptr : nullptr;
}

extern "C" DvmType dvmh_fill_header_C(DvmType deviceNum, const void *baseAddr, const DvmType dvmDesc[], DvmType devHeader[],
        DvmType extendedParams[]) {   // :2226
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_fill_header_C");
    DVMH_CHECK(devHeader != nullptr, "NULL header passed to dvmh_fill_header_C");
    if (deviceNum == 1 && d->hasDev) {
        d->fillDevHeader(baseAddr, devHeader, extendedParams);
        return (d->dev.isTransformed(d->rank) ? 1 : 0) + (d->dev.isDiagonalized() ? 1 : 0);
    }
    // host copy: natural layout
    long collector = 1, offset = 0;
    devHeader[0] = (DvmType)d;
    devHeader[d->rank + 2] = (DvmType)baseAddr;
    for (int i = d->rank; i >= 1; i--) {
        offset += collector * d->space[i - 1].lo;
        devHeader[i] = collector;
        collector *= d->space[i - 1].size();
    }
    devHeader[d->rank + 1] = ((long)d->host - (long)baseAddr) / d->typeSize - offset;
    return 0;
}

// ============================================================ actual state
static void moveBlock(Data *d, const long lo[], const long hi[], bool toHost);

static Box boxOf(const Data *d, const long lo[], const long hi[]) {
    Box b;
    for (int a = 0; a < d->rank; a++) { b.lo[a] = lo[a]; b.hi[a] = hi[a]; }
    return b;
}

// The parts of b whose host copy is stale come from the device (where they
// are current); b is then current on the host.
static void fetchMissing(Data *d, const Box &b) {
    std::vector<Box> need = d->hostPieces.missing(b);
    if (need.empty()) return;
    DVMH_INTERNAL(d->hasDev, "array elements current nowhere");
    bool whole = d->hostPieces.empty();
    const Box w = d->whole();
    for (int a = 0; a < d->rank && whole; a++) whole = b.lo[a] == w.lo[a] && b.hi[a] == w.hi[a];
    if (whole) {
        d->toHost();   // the whole array, in one copy
        return;
    }
    for (const Box &x : need) moveBlock(d, x.lo, x.hi, true);
    d->hostPieces.add(b);
}

// The device copy is made current everywhere: the whole array when the device
// holds none of it, else only the pieces the host wrote since.
static void uploadMissing(Data *d) {
    d->ensureDev();
    std::vector<Box> need = d->devPieces.missing(d->whole());
    if (need.empty()) return;
    if (d->devPieces.empty()) {
        d->toDevice();
        return;
    }
    for (const Box &x : need) moveBlock(d, x.lo, x.hi, false);
    d->devPieces.add(d->whole());
}

// The same for one box of the array.
static void uploadMissingBox(Data *d, const Box &b) {
    d->ensureDev();
    std::vector<Box> need = d->devPieces.missing(b);
    if (need.empty()) return;
    for (const Box &x : need) moveBlock(d, x.lo, x.hi, false);
    d->devPieces.add(b);
}

// Each array keeps the pieces where its host copy and its device copy are
// current (Data::hostPieces / devPieces); transfers move only the pieces a
// request finds stale, sub-arrays as packed blocks (rts/dvmh/dvmh_data.cpp:339-600,
// getActualBase / setActual on DvmhPieces).
static void getActual(Data *d) {
    batchDrain();
    flushDeferredPosts();
    fetchMissing(d, d->whole());
}

extern "C" void dvmh_get_actual_variable2_(void *addr) {   // :2325
    Guard g;
    batchDrain();
    flushDeferredPosts();
    Data *d = regularVarOf(addr);
    if (d) getActual(d);
}

extern "C" void dvmh_get_actual_subvariable_C(void *addr, DvmType rank, ...) {
    Guard g;
    (void)rank;
    dvmh_get_actual_variable2_(addr);
}

extern "C" void dvmh_get_actual_array2_(const DvmType dvmDesc[]) {
    Guard g;
    getActual(asData(dvmDesc, "dvmh_get_actual_array2_"));
}

// Sub-array actualisation packs the block on the device and copies only it.
static void subarrayBounds(Data *d, va_list &ap, long lo[], long hi[]) {
    for (int a = 0; a < d->rank; a++) {
        DvmType l = va_arg(ap, DvmType), h = va_arg(ap, DvmType);
        lo[a] = (l == UNDEF_BOUND) ? d->space[a].lo : l;
        hi[a] = (h == UNDEF_BOUND) ? d->space[a].hi : h;
        DVMH_CHECK(lo[a] >= d->space[a].lo && hi[a] <= d->space[a].hi && lo[a] <= hi[a] + 1,
                "Sub-array bounds [%ld, %ld] are out of the array axis %d", lo[a], hi[a], a + 1);
    }
}

static void moveBlock(Data *d, const long lo[], const long hi[], bool toHost) {
    long sizes[MAX_RANK], origin[MAX_RANK], bext[MAX_RANK];
    long count = 1;
    for (int a = 0; a < d->rank; a++) {
        sizes[a] = d->space[a].size();
        origin[a] = d->space[a].lo;
        bext[a] = hi[a] - lo[a] + 1;
        count *= bext[a];
    }
    if (count <= 0) return;
    if (d->dev.isTransformed(d->rank)) d->setLayout(nullptr, 0);
    size_t nb = (size_t)count * (size_t)d->typeSize;
    void *stage = gpu.alloc(nb);
    char *hstage = (char *)gpu.pinnedStaging(nb);
    if (toHost) {
        dvmh_hip_copy_block(d->dev.ptr, sizes, origin, stage, bext, lo, lo, hi, d->rank, d->typeSize, gpu.stream);
        gpu.d2h(hstage, stage, nb);
        // unpack on the host
        long idx[MAX_RANK];
        for (int a = 0; a < d->rank; a++) idx[a] = lo[a];
        long rowLen = bext[d->rank - 1];
        for (long r = 0; r < count / rowLen; r++) {
            long off = 0;
            for (int a = 0; a < d->rank; a++) off = off * sizes[a] + (idx[a] - origin[a]);
            memcpy(d->host + off * d->typeSize, hstage + r * rowLen * d->typeSize, (size_t)(rowLen * d->typeSize));
            for (int a = d->rank - 2; a >= 0; a--) {
                if (++idx[a] <= hi[a]) break;
                idx[a] = lo[a];
            }
        }
    } else {
        long idx[MAX_RANK];
        for (int a = 0; a < d->rank; a++) idx[a] = lo[a];
        long rowLen = bext[d->rank - 1];
        for (long r = 0; r < count / rowLen; r++) {
            long off = 0;
            for (int a = 0; a < d->rank; a++) off = off * sizes[a] + (idx[a] - origin[a]);
            memcpy(hstage + r * rowLen * d->typeSize, d->host + off * d->typeSize, (size_t)(rowLen * d->typeSize));
            for (int a = d->rank - 2; a >= 0; a--) {
                if (++idx[a] <= hi[a]) break;
                idx[a] = lo[a];
            }
        }
        gpu.h2d(stage, hstage, nb);
        dvmh_hip_copy_block(stage, bext, lo, d->dev.ptr, sizes, origin, lo, hi, d->rank, d->typeSize, gpu.stream);
        gpu.sync();
    }
    gpu.release(stage);
}

extern "C" void dvmh_get_actual_subarray_C(const DvmType dvmDesc[], DvmType rank, ...) {   // :112
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_get_actual_subarray_C");
    DVMH_CHECK(rank == d->rank, "dvmh_get_actual_subarray_C: rank mismatch");
    long lo[MAX_RANK], hi[MAX_RANK];
    va_list ap;
    va_start(ap, rank);
    subarrayBounds(d, ap, lo, hi);
    va_end(ap);
    batchDrain();
    flushDeferredPosts();
    fetchMissing(d, boxOf(d, lo, hi));
}

extern "C" void dvmh_get_actual_all2_() {
    Guard g;
    batchDrain();
    flushDeferredPosts();
    for (Data *d : rt.allData) fetchMissing(d, d->whole());
}

extern "C" void dvmh_actual_variable2_(const void *addr) {   // host copy is the newest
    Guard g;
    Data *d = regularVarOf(addr);
    if (d) d->hostOwnsAll();
}

extern "C" void dvmh_actual_subvariable_C(const void *addr, DvmType rank, ...) {
    Guard g;
    (void)rank;
    dvmh_actual_variable2_(addr);
}

extern "C" void dvmh_actual_array2_(const DvmType dvmDesc[]) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_actual_array2_");
    batchDrain();
    fetchMissing(d, d->whole());   // keep elements the host did not write
    d->hostOwnsAll();
}

extern "C" void dvmh_actual_subarray_C(const DvmType dvmDesc[], DvmType rank, ...) {   // :120
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_actual_subarray_C");
    DVMH_CHECK(rank == d->rank, "dvmh_actual_subarray_C: rank mismatch");
    batchDrain();
    long lo[MAX_RANK], hi[MAX_RANK];
    va_list ap;
    va_start(ap, rank);
    subarrayBounds(d, ap, lo, hi);
    va_end(ap);
    // the host wrote the block: its host copy is current there, the device's is
    // not; nothing moves until a region needs the block on the device
    const Box b = boxOf(d, lo, hi);
    d->hostPieces.add(b);
    d->devPieces.remove(b);
}

extern "C" void dvmh_actual_all2_() {
    Guard g;
    batchDrain();
    for (Data *d : rt.allData) {
        fetchMissing(d, d->whole());
        d->hostOwnsAll();
    }
}

// ============================================================ shadows and elements
// One process owns the whole array and shadow edges beyond the global bounds
// are not stored (rts/dvmh/dvmh_data.cpp:981, extendBlock), so a renew has no
// neighbour to exchange with; the widths are validated like the reference does.
extern "C" void dvmh_shadow_renew_C(const DvmType dvmDesc[], DvmType cornerFlag, DvmType specifiedRank, ...) {   // :2607
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_shadow_renew_C");
    (void)cornerFlag;
    DVMH_CHECK(specifiedRank == 0 || specifiedRank == d->rank, "dvmh_shadow_renew_C: rank %ld does not match %d", specifiedRank,
            d->rank);
    va_list ap;
    va_start(ap, specifiedRank);
    for (int a = 0; a < specifiedRank; a++) {
        DvmType l = va_arg(ap, DvmType), h = va_arg(ap, DvmType);
        DVMH_CHECK(l <= d->shdLow[a] && h <= d->shdHigh[a], "Shadow renew width (%ld, %ld) exceeds the declared shadow (%ld, %ld) on axis %d",
                l, h, d->shdLow[a], d->shdHigh[a], a + 1);
    }
    va_end(ap);
}

extern "C" void dvmh_indirect_shadow_renew_C(const DvmType[], DvmType, const char[]) {
    fatal("error", "Indirect shadows need several processes; this runtime is single-process");
}

// The box a remote-access or sub-array rule list names: a constant rule one
// index, any other rule the whole axis.
static Box ruleBox(Data *d, va_list &ap) {
    Box b = d->whole();
    for (int a = 0; a < d->rank; a++) {
        AlignRule r = loadAlign(va_arg(ap, DvmType));
        if (r.axis == 0) {
            DVMH_CHECK(r.summand >= d->space[a].lo && r.summand <= d->space[a].hi, "Remote index %ld is out of the array axis %d",
                    r.summand, a + 1);
            b.lo[a] = b.hi[a] = r.summand;
        }
    }
    return b;
}

// REMOTE_ACCESS outside loops (rts/dvmh/dvmh2.cpp dvmh_remote_access_C): with
// one process every element is local, so the "buffer" rmaDesc names is a
// read-only alias of the array's host copy, whose referenced elements are made
// current first.  Deleting the alias leaves the array alone.
extern "C" void dvmh_remote_access_C(DvmType rmaDesc[], const DvmType dvmDesc[], DvmType rank, ...) {   // :133
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_remote_access_C");
    DVMH_CHECK(rank == d->rank, "dvmh_remote_access_C: rank mismatch");
    va_list ap;
    va_start(ap, rank);
    const Box b = ruleBox(d, ap);
    va_end(ap);
    batchDrain();
    flushDeferredPosts();
    fetchMissing(d, b);
    Data *alias = new Data();
    alias->rank = d->rank;
    alias->typeSize = d->typeSize;
    alias->rtType = d->rtType;
    for (int a = 0; a < d->rank; a++) alias->space[a] = d->space[a];
    alias->host = d->host;
    alias->ownsHost = false;
    alias->hostOwnsAll();
    alias->userHeader = rmaDesc;
    fillUserHeader(alias);
    rt.allData.push_back(alias);
}

static long linearIndex(Data *d, va_list &ap, bool *inside, long *point = nullptr) {
    long off = 0;
    *inside = true;
    for (int a = 0; a < d->rank; a++) {
        DvmType idx = va_arg(ap, DvmType);
        if (point) point[a] = idx;
        if (idx < d->space[a].lo || idx > d->space[a].hi) *inside = false;
        off = off * d->space[a].size() + (idx - d->space[a].lo);
    }
    return off;
}

extern "C" DvmType dvmh_has_element_C(const DvmType dvmDesc[], DvmType rank, ...) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_has_element_C");
    DVMH_CHECK(rank == d->rank, "dvmh_has_element_C: rank mismatch");
    va_list ap;
    va_start(ap, rank);
    bool inside;
    linearIndex(d, ap, &inside);
    va_end(ap);
    return inside ? 1 : 0;
}

extern "C" DvmType dvmh_calc_linear_C(const DvmType dvmDesc[], DvmType rank, ...) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_calc_linear_C");
    va_list ap;
    va_start(ap, rank);
    bool inside;
    long off = linearIndex(d, ap, &inside);
    va_end(ap);
    return off;
}

static void *elementAddr(Data *d, va_list &ap) {
    bool inside;
    long point[MAX_RANK];
    long off = linearIndex(d, ap, &inside, point);
    DVMH_CHECK(inside, "Element index is out of the array bounds");
    // host code touches the element: when the host copy of that element is
    // stale (the program did not get_actual it), the stale pieces come down
    if (!d->hostPieces.holds(point)) fetchMissing(d, d->whole());
    return d->host + off * d->typeSize;
}

extern "C" void *dvmh_get_own_element_addr_C(const DvmType dvmDesc[], DvmType rank, ...) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_get_own_element_addr_C");
    DVMH_CHECK(rank == d->rank, "dvmh_get_own_element_addr_C: rank mismatch");
    va_list ap;
    va_start(ap, rank);
    void *p = elementAddr(d, ap);
    va_end(ap);
    return p;
}

extern "C" void *dvmh_get_element_addr_C(const DvmType dvmDesc[], DvmType rank, ...) {
    Guard g;
    Data *d = asData(dvmDesc, "dvmh_get_element_addr_C");
    DVMH_CHECK(rank == d->rank, "dvmh_get_element_addr_C: rank mismatch");
    va_list ap;
    va_start(ap, rank);
    void *p = elementAddr(d, ap);
    va_end(ap);
    return p;
}

// ============================================================ regions
extern "C" DvmType dvmh_region_create_C(DvmType regionFlags) {   // :3002
    Guard g;
    ensureInit();
    Region *r = new Region();
    r->flags = regionFlags;
    r->async = settings.allowAsync && (regionFlags & REGION_ASYNC);   // rts/dvmh/region.cpp:442
    rt.curRegion = r;
    return (DvmType)r;
}

static Region *asRegion(DvmType h, const char *where) {
    Object *o = asObject(h);
    DVMH_CHECK(o->kind == Kind::Region, "Incorrect region reference passed to %s", where);
    return (Region *)o;
}

extern "C" void dvmh_region_register_array_C(DvmType curRegion, DvmType intent, const DvmType dvmDesc[], const char varName[]) {
    Guard g;
    Region *r = asRegion(curRegion, "dvmh_region_register_array_C");
    Data *d = asData(dvmDesc, "dvmh_region_register_array_C");
    for (auto &ra : r->arrays)
        if (ra.data == d) { ra.intent |= intent; return; }
    r->arrays.push_back(RegionArray{d, intent, varName ? varName : ""});
}

// A sub-array in a region's clause (rts/dvmh/dvmh2.cpp:3080): only that box
// is made current where the region runs, and with an out intent only that box
// is owned by the device afterwards.
extern "C" void dvmh_region_register_subarray_C(DvmType curRegion, DvmType intent, const DvmType dvmDesc[], const char varName[],
        DvmType rank, ...) {
    Guard g;
    Region *r = asRegion(curRegion, "dvmh_region_register_subarray_C");
    Data *d = asData(dvmDesc, "dvmh_region_register_subarray_C");
    DVMH_CHECK(rank == d->rank, "dvmh_region_register_subarray_C: rank mismatch");
    long lo[MAX_RANK], hi[MAX_RANK];
    va_list ap;
    va_start(ap, rank);
    subarrayBounds(d, ap, lo, hi);
    va_end(ap);
    RegionArray ra{d, intent, varName ? varName : ""};
    ra.part = true;
    ra.box = boxOf(d, lo, hi);
    r->arrays.push_back(ra);
}

extern "C" void dvmh_region_register_scalar_C(DvmType curRegion, DvmType intent, const void *addr, DvmType typeSize,
        const char varName[]) {   // :3096
    Guard g;
    Region *r = asRegion(curRegion, "dvmh_region_register_scalar_C");
    long sizes[1] = {1};
    Data *d = registerRegular(addr, 0, typeSize, sizes);
    d->rank = 0;
    flushPostsFor(addr);   // a reduction result of an earlier async region lands first
    r->scalars.push_back(RegionArray{d, intent, varName ? varName : ""});
}

// Choose the device and make registered data present there (rts/dvmh/region.cpp:523).
extern "C" void dvmh_region_execute_on_targets_C(DvmType curRegion, DvmType deviceTypes) {   // :3108
    Guard g;
    Region *r = asRegion(curRegion, "dvmh_region_execute_on_targets_C");
    r->targets = deviceTypes;
    r->onGpu = gpu.present && (deviceTypes & DEVICE_TYPE_CUDA);
    if (!r->onGpu) {
        DVMH_CHECK(!settings.requireGpu, "Region at %s:%ld would run on the host but DVMH_AMD_REQUIRE_GPU is set", rt.curFile.c_str(),
                rt.curLine);
        DVMH_CHECK(deviceTypes & DEVICE_TYPE_HOST, "No usable device for the region (targets %ld, GPU %s)", deviceTypes,
                gpu.present ? "present" : "absent");
        // host execution: host copies must be current
        for (auto &ra : r->arrays) {
            if (!ra.part) {
                getActual(ra.data);
                continue;
            }
            batchDrain();
            flushDeferredPosts();
            fetchMissing(ra.data, ra.box);
        }
        for (auto &rs : r->scalars) getActual(rs.data);
        return;
    }
    for (auto &ra : r->arrays) {
        Data *d = ra.data;
        if (ra.intent & INTENT_LOCAL) { d->ensureDev(); continue; }
        if (ra.part) uploadMissingBox(d, ra.box);
        else uploadMissing(d);
    }
    for (auto &rs : r->scalars) {
        Data *d = rs.data;
        d->ensureDev();
        if (!d->devWhole() || d->hostWhole()) {
            gpu.h2d(d->dev.ptr, d->host, (size_t)d->typeSize);
            d->devPieces.add(d->whole());
        }
    }
}

extern "C" void dvmh_region_end_C(DvmType curRegion) {   // :3122
    Guard g;
    Region *r = asRegion(curRegion, "dvmh_region_end_C");
    if (r->onGpu) {
        for (auto &ra : r->arrays) {
            if (!(ra.intent & INTENT_OUT)) continue;
            if (!ra.part) {
                ra.data->devOwnsAll();   // written on the device
            } else {
                ra.data->devPieces.add(ra.box);
                ra.data->hostPieces.remove(ra.box);
            }
        }
        // a lazy region's loops run (or replay) now, before its results are read
        if (!r->deferred.empty()) graphRunRegion(r);
        bool needSync = false;
        for (auto &rs : r->scalars)
            if (rs.intent & INTENT_OUT) needSync = true;
        if (needSync) {
            // the region's loops still queued in an open batch write them first
            if (r->pendingBatch > 0) batchDrain();
            for (auto &rs : r->scalars)
                if (rs.intent & INTENT_OUT) gpu.d2h(rs.data->host, rs.data->dev.ptr, (size_t)rs.data->typeSize);
        }
        if (settings.syncMode == 1 && !r->async) flushDeferredPosts();
    }
    if (!r->onGpu) {
        // a region run on the host wrote its output arrays there
        for (auto &ra : r->arrays) {
            if (!(ra.intent & INTENT_OUT)) continue;
            if (!ra.part) {
                ra.data->hostOwnsAll();
            } else {
                ra.data->hostPieces.add(ra.box);
                ra.data->devPieces.remove(ra.box);
            }
        }
    }
    if (rt.curRegion == r) rt.curRegion = nullptr;
    if (!batchKeepsRegion(r)) delete r;   // else freed when its queued loops have run
}

// ============================================================ loops
extern "C" DvmType dvmh_loop_create_C(DvmType curRegion, DvmType rank, ...) {   // :3181
    Guard g;
    ensureInit();
    DVMH_CHECK(rank >= 0 && rank <= MAX_LOOP_RANK, "Loop rank %ld is out of range", rank);
    Loop *l = new Loop();
    l->region = curRegion ? asRegion(curRegion, "dvmh_loop_create_C") : nullptr;
    l->rank = (int)rank;
    l->file = rt.curFile;
    l->line = rt.curLine;
    va_list ap;
    va_start(ap, rank);
    for (int i = 0; i < rank; i++) {
        l->lo[i] = va_arg(ap, DvmType);
        l->hi[i] = va_arg(ap, DvmType);
        l->step[i] = va_arg(ap, DvmType);
        DVMH_CHECK(l->step[i] != 0, "Loop step on axis %d is zero", i + 1);
        // end adjusted onto the step grid (rts/dvmh/loop.cpp:556-565)
        if ((l->step[i] > 0 && l->hi[i] >= l->lo[i]) || (l->step[i] < 0 && l->hi[i] <= l->lo[i]))
            l->hi[i] = l->lo[i] + (l->hi[i] - l->lo[i]) / l->step[i] * l->step[i];
    }
    va_end(ap);
    return (DvmType)l;
}

extern "C" void dvmh_loop_map_C(DvmType curLoop, const DvmType templDesc[], DvmType templRank, ...) {   // :3297
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_map_C");
    Object *o = asObject(templDesc[0]);
    if (o->kind == Kind::Data) l->mapData = (Data *)o;
    else if (o->kind == Kind::Template) l->mapTempl = (Template *)o;
    else fatal("error", "dvmh_loop_map_C: mapping target is neither an array nor a template");
    l->mapRank = (int)templRank;
    va_list ap;
    va_start(ap, templRank);
    for (int t = 0; t < templRank; t++) {
        AlignRule r = loadAlign(va_arg(ap, DvmType));
        DVMH_CHECK(r.axis <= l->rank, "dvmh_loop_map_C: loop axis %ld is out of the loop rank %d", r.axis, l->rank);
        l->mapAxis[t] = r.axis;
        l->mapMult[t] = r.mult;
        l->mapSum[t] = r.summand;
    }
    va_end(ap);
}

extern "C" void dvmh_loop_set_cuda_block_C(DvmType curLoop, DvmType xSize, DvmType ySize, DvmType zSize) {
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_set_cuda_block_C");
    DVMH_CHECK(xSize >= 1 && ySize >= 1 && zSize >= 1, "Invalid CUDA block (%ld, %ld, %ld)", xSize, ySize, zSize);
    l->userBlock[0] = (int)xSize;
    l->userBlock[1] = (int)ySize;
    l->userBlock[2] = (int)zSize;
    l->userBlockFlag = true;
}

extern "C" void dvmh_loop_set_stage_C(DvmType curLoop, DvmType stage) {
    Guard g;
    asLoop(curLoop, "dvmh_loop_set_stage_C")->stage = stage;
}

extern "C" void dvmh_loop_reduction_C(DvmType curLoop, DvmType redType, void *arrayAddr, DvmType varType, DvmType arrayLength,
        void *locAddr, DvmType locSize) {   // :3343
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_reduction_C");
    DVMH_CHECK(redType >= rf_SUM && redType <= rf_MINLOC, "Unknown reduction function %ld", redType);
    DVMH_CHECK(arrayAddr != nullptr, "NULL reduction variable");
    auto r = std::make_shared<Reduction>();
    r->func = (int)redType;
    r->rtType = (int)varType;
    r->elemSize = rtTypeSize((int)varType);
    DVMH_CHECK(r->elemSize > 0, "Unsupported reduction variable type %ld", varType);
    r->elemCount = arrayLength;
    r->arrayAddr = arrayAddr;
    r->initArr.assign((char *)arrayAddr, (char *)arrayAddr + r->elemSize * r->elemCount);
    if (redType == rf_MAXLOC || redType == rf_MINLOC) {
        DVMH_CHECK(locAddr != nullptr && locSize > 0, "MAXLOC/MINLOC reduction needs a location variable");
    }
    if (locAddr && locSize > 0) {
        r->locAddr = locAddr;
        r->locSize = locSize;
        r->initLoc.assign((char *)locAddr, (char *)locAddr + locSize * r->elemCount);
    }
    l->reds.push_back(r);
}

extern "C" void dvmh_loop_across_C(DvmType curLoop, DvmType isOut, const DvmType dvmDesc[], DvmType rank, ...) {   // :3390
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_across_C");
    Data *d = asData(dvmDesc, "dvmh_loop_across_C");
    DVMH_CHECK(rank == d->rank, "dvmh_loop_across_C: rank %ld does not match the array rank %d", rank, d->rank);
    AcrossSpec a;
    a.data = d;
    a.isOut = isOut != 0;
    va_list ap;
    va_start(ap, rank);
    for (int i = 0; i < rank; i++) {
        a.low[i] = va_arg(ap, DvmType);
        a.high[i] = va_arg(ap, DvmType);
        DVMH_CHECK(a.low[i] >= 0 && a.high[i] >= 0, "Negative ACROSS width on axis %d", i + 1);
        DVMH_CHECK(a.low[i] <= d->shdLow[i] && a.high[i] <= d->shdHigh[i],
                "ACROSS width (%ld, %ld) exceeds the declared shadow (%ld, %ld) on axis %d", a.low[i], a.high[i], d->shdLow[i],
                d->shdHigh[i], i + 1);
    }
    va_end(ap);
    l->acrosses.push_back(a);
}

extern "C" void dvmh_loop_shadow_compute_C(DvmType curLoop, const DvmType[], DvmType, ...) { Guard g; (void)curLoop; }
extern "C" void dvmh_loop_shadow_compute_array_C(DvmType curLoop, const DvmType[]) { Guard g; (void)curLoop; }
extern "C" void dvmh_loop_consistent_C(DvmType curLoop, const DvmType[], DvmType, ...) { Guard g; (void)curLoop; }
// REMOTE_ACCESS of a parallel loop: every element is local to the one
// process, so only the arguments are checked.
extern "C" void dvmh_loop_remote_access_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...) {   // :177
    Guard g;
    asLoop(curLoop, "dvmh_loop_remote_access_C");
    Data *d = asData(dvmDesc, "dvmh_loop_remote_access_C");
    DVMH_CHECK(rank == d->rank, "dvmh_loop_remote_access_C: rank mismatch");
    va_list ap;
    va_start(ap, rank);
    for (int a = 0; a < rank; a++) loadAlign(va_arg(ap, DvmType));
    va_end(ap);
}

extern "C" void dvmh_loop_array_correspondence_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...) {   // :179
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_array_correspondence_C");
    Data *d = asData(dvmDesc, "dvmh_loop_array_correspondence_C");
    std::vector<long> c(d->rank, 0);
    va_list ap;
    va_start(ap, rank);
    for (int i = 0; i < rank && i < d->rank; i++) c[i] = va_arg(ap, DvmType);
    va_end(ap);
    l->correspondences.push_back(std::make_pair(d, c));
}

extern "C" void dvmh_loop_register_handler_C(DvmType curLoop, DvmType deviceType, DvmType handlerType, DvmType handlerHelper) {
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_register_handler_C");
    HandlerDesc h = loadHandler(handlerHelper);
    h.type = handlerType;
    if (deviceType == DEVICE_TYPE_HOST) { l->hostH = std::move(h); l->hasHost = true; }
    else if (deviceType == DEVICE_TYPE_CUDA) { l->gpuH = std::move(h); l->hasGpu = true; }
    else fatal("error", "Unknown device type %ld in dvmh_loop_register_handler_C", deviceType);
}

extern "C" void dvmh_amd_loop_register_batch_handler_C(DvmType curLoop, DvmType handlerHelper) {
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_amd_loop_register_batch_handler_C");
    l->batchH = loadHandler(handlerHelper);
    l->hasBatch = true;
    l->batchChain = false;
}

extern "C" void dvmh_amd_loop_register_chain_handler_C(DvmType curLoop, DvmType handlerHelper) {
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_amd_loop_register_chain_handler_C");
    l->batchH = loadHandler(handlerHelper);
    l->hasBatch = true;
    l->batchChain = true;
}

extern "C" void dvmh_loop_perform_C(DvmType curLoop) {   // :3769
    Guard g;
    Loop *l = asLoop(curLoop, "dvmh_loop_perform_C");
    saved.clear();
    performLoop(l);
    // a queued loop is owned by its lazy region or batch until it runs
    if (!l->queued) delete l;
}

// ============================================================ handler queries
extern "C" DvmType dvmh_loop_get_dependency_mask_C(DvmType curLoop) { return loopOf(curLoop, "dvmh_loop_get_dependency_mask_C")->depMask; }

extern "C" DvmType dvmh_loop_get_device_num_C(DvmType curLoop) {
    Object *o = asObject(curLoop);
    return o->kind == Kind::SpecLoop ? ((SpecLoop *)o)->device : 0;
}

extern "C" void dvmh_loop_fill_bounds_C(DvmType curLoop, DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[]) {   // :3977
    Loop *l = loopOf(curLoop, "dvmh_loop_fill_bounds_C");
    for (int i = 0; i < l->rank; i++) {
        if (boundsLow) boundsLow[i] = l->lo[i];
        if (boundsHigh) boundsHigh[i] = l->hi[i];
        if (loopSteps) loopSteps[i] = l->step[i];
    }
}

extern "C" DvmType dvmh_loop_get_slot_count_C(DvmType curLoop) {
    Object *o = asObject(curLoop);
    if (o->kind == Kind::SpecLoop && ((SpecLoop *)o)->device == 1) return 1;
    return hostSlotCount();
}

extern "C" void dvmh_loop_fill_local_part_C(DvmType curLoop, const DvmType dvmDesc[], DvmType part[]) {
    (void)curLoop;
    Data *d = asData(dvmDesc, "dvmh_loop_fill_local_part_C");
    for (int a = 0; a < d->rank; a++) { part[2 * a] = d->space[a].lo; part[2 * a + 1] = d->space[a].hi; }
}

extern "C" const void *dvmh_loop_cuda_get_local_part_C(DvmType curLoop, const DvmType dvmDesc[], DvmType indexType) {   // :4333
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_get_local_part_C");
    Data *d = asData(dvmDesc, "dvmh_loop_cuda_get_local_part_C");
    long n = 2 * d->rank;
    size_t es = (indexType == rt_INT) ? sizeof(int) : sizeof(long);
    char *hostPart = (char *)calloc((size_t)n, es);
    for (int a = 0; a < d->rank; a++) {
        if (es == sizeof(int)) { ((int *)hostPart)[2 * a] = (int)d->space[a].lo; ((int *)hostPart)[2 * a + 1] = (int)d->space[a].hi; }
        else { ((long *)hostPart)[2 * a] = d->space[a].lo; ((long *)hostPart)[2 * a + 1] = d->space[a].hi; }
    }
    void *dp = gpu.alloc((size_t)n * es);
    gpu.h2d(dp, hostPart, (size_t)n * es);
    free(hostPart);
    sl->toFree.push_back(dp);
    return dp;
}

extern "C" DvmType dvmh_loop_guess_index_type_C(DvmType curLoop) {   // :4097
    Loop *l = loopOf(curLoop, "dvmh_loop_guess_index_type_C");
    long maxElems = 0;
    if (l->region)
        for (auto &ra : l->region->arrays) maxElems = std::max(maxElems, ra.data->elemCount());
    for (int i = 0; i < l->rank; i++) maxElems = std::max(maxElems, std::max(std::labs(l->lo[i]), std::labs(l->hi[i])) + 1);
    // an index that does not fit int takes the 8-byte type, which is long on
    // LP64 (the reference tries int, then long, then long long)
    if (settings.index64 || maxElems >= 2147483647L) return rt_LONG;
    return rt_INT;
}

// Every later loop's handlers are told to use 64-bit indices (1) or to pick
// by the arrays' extents again (0); for tests of the 64-bit kernels on arrays
// small enough to check.
extern "C" void dvmh_amd_set_index64_C(DvmType on) {
    ensureInit();
    settings.index64 = on != 0;
}

extern "C" DvmType dvmh_loop_has_element_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...) {
    (void)curLoop;
    Data *d = asData(dvmDesc, "dvmh_loop_has_element_C");
    va_list ap;
    va_start(ap, rank);
    bool inside;
    linearIndex(d, ap, &inside);
    va_end(ap);
    return inside ? 1 : 0;
}

// A workgroup's waves are dealt over the compute unit's 4 SIMDs; each SIMD has
// 512 vector registers per lane, given to a wave in granules of 8.
extern "C" DvmType dvmh_amd_block_fits_registers_C(const DvmType block[3], DvmType regsPerThread) {
    if (regsPerThread <= 0) return 1;
    long threads = (long)block[0] * block[1] * block[2];
    long waves = (threads + 63) / 64;
    long perSimd = (waves + 3) / 4;
    long alloc = (regsPerThread + 7) / 8 * 8;
    return perSimd * alloc <= 512 ? 1 : 0;
}

// dvmh_loop_cuda_get_config_C (:4281) -> DvmhLoopCuda::pickBlock (rts/dvmh/loop.cpp:2134-2199)
extern "C" void dvmh_loop_cuda_get_config_C(DvmType curLoop, DvmType sharedPerThread, DvmType regsPerThread, void *inOutThreads,
        void *outStream, DvmType *outSharedPerBlock) {
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_get_config_C");
    Loop *l = sl->loop;
    unsigned *thr = (unsigned *)inOutThreads;
    int block[3];
    if (sl->candidate < 0 && sl->block[0] == 0) {
        TuneChoice c;
        tunerPick(sl, DVMH_AMD_KIND_PARALLEL, c);
        for (int i = 0; i < 3; i++) sl->block[i] = c.block[i];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) sl->tune[i] = c.tune[i];
        sl->candidate = c.candidate;
    }
    for (int i = 0; i < 3; i++) block[i] = sl->block[i];
    if (thr && thr[0] >= 1 && thr[1] >= 1 && thr[2] >= 1 && sl->candidate < 0 && !l->userBlockFlag) {
        block[0] = (int)thr[0]; block[1] = (int)thr[1]; block[2] = (int)thr[2];   // handler-suggested default
    }
    int maxBlockRank = l->rank - (l->depMask > 0 ? 1 : 0);
    int blockRank = 3;
    while (blockRank > 0 && block[blockRank - 1] == 1) blockRank--;
    DVMH_CHECK(blockRank <= std::max(maxBlockRank, 1),
            "Launch block of %d dimensions exceeds the %d the loop can spread over threads (its rank, less one when it has a dependency)", blockRank, std::max(maxBlockRank, 1));
    static const char *axisNames[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; i++)
        DVMH_CHECK(block[i] <= gpu.props.maxBlock[i], "Block extent %d along %s exceeds the device limit %d", block[i],
                axisNames[i], gpu.props.maxBlock[i]);
    DVMH_CHECK(block[0] * block[1] * block[2] <= gpu.props.maxThreadsPerBlock, "Block of %d threads exceeds the device limit of %d per workgroup",
            block[0] * block[1] * block[2], gpu.props.maxThreadsPerBlock);
    if (sharedPerThread < 0) sharedPerThread = 0;
    long needed = (long)block[0] * block[1] * block[2] * sharedPerThread;
    if (needed > gpu.props.ldsPerBlock) {
        DVMH_CHECK(!l->userBlockFlag, "The requested block needs %ld bytes of LDS, more than the %ld a workgroup can have", needed, gpu.props.ldsPerBlock);
        block[0] = std::max(1, (int)std::min<long>(256, gpu.props.ldsPerBlock / sharedPerThread) / 64 * 64);
        block[1] = block[2] = 1;
    }
    if (regsPerThread > 0) {
        const DvmType b[3] = {block[0], block[1], block[2]};
        if (!dvmh_amd_block_fits_registers_C(b, regsPerThread)) {
            DVMH_CHECK(!l->userBlockFlag, "Register count is insufficient: %ld registers per thread do not fit a block of %d threads",
                    (long)regsPerThread, block[0] * block[1] * block[2]);
            long perSimdWaves = 512 / ((regsPerThread + 7) / 8 * 8);
            block[0] = (int)std::max<long>(64, std::min<long>(256, 4 * 64 * perSimdWaves) / 64 * 64);
            block[1] = block[2] = 1;
        }
    }
    if (thr) { thr[0] = block[0]; thr[1] = block[1]; thr[2] = block[2]; }
    if (outStream) *(hipStream_t *)outStream = sl->stream;
    if (outSharedPerBlock) *outSharedPerBlock = (DvmType)block[0] * block[1] * block[2] * sharedPerThread;
}

extern "C" void dvmh_amd_loop_get_tuning_C(DvmType curLoop, DvmType kind, DvmType tune[]) {
    SpecLoop *sl = asSpec(curLoop, "dvmh_amd_loop_get_tuning_C");
    if (sl->candidate < 0 && sl->block[0] == 0) {
        TuneChoice c;
        tunerPick(sl, kind, c);
        for (int i = 0; i < 3; i++) sl->block[i] = c.block[i];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) sl->tune[i] = c.tune[i];
        sl->candidate = c.candidate;
    }
    for (int i = 0; i < DVMH_AMD_TUNE_N; i++) tune[i] = sl->tune[i];
}

// ------------------------------------------------------------ reductions in handlers
static std::shared_ptr<Reduction> redOf(Loop *l, DvmType redIndex, const char *where) {
    DVMH_CHECK(redIndex >= 1 && redIndex <= (DvmType)l->reds.size(), "Invalid reduction index %ld passed to %s", redIndex, where);
    return l->reds[redIndex - 1];
}

extern "C" void dvmh_loop_cuda_register_red_C(DvmType curLoop, DvmType redIndex, void **arrayAddrPtr, void **locAddrPtr) {   // :4207
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_register_red_C");
    auto r = redOf(sl->loop, redIndex, "dvmh_loop_cuda_register_red_C");
    r->gpuArrPtr = arrayAddrPtr;
    r->gpuLocPtr = locAddrPtr;
    r->localArr.assign((size_t)(r->elemSize * r->elemCount), 0);
    r->localLoc.assign((size_t)(r->locSize * r->elemCount), 0);
    r->initValues(r->localArr.data(), r->isLoc() ? r->localLoc.data() : nullptr);
}

// Handler-local initial values (identity for SUM, the pre-loop value for MAX ...).
extern "C" void dvmh_loop_red_init_C(DvmType curLoop, DvmType redIndex, void *arrayAddr, void *locAddr) {   // :4224
    Loop *l = loopOf(curLoop, "dvmh_loop_red_init_C");
    auto r = redOf(l, redIndex, "dvmh_loop_red_init_C");
    r->initValues(arrayAddr, locAddr);
}

extern "C" void dvmh_loop_cuda_red_init_C(DvmType curLoop, DvmType redIndex, void **devArrayAddrPtr, void **devLocAddrPtr) {
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_red_init_C");
    auto r = redOf(sl->loop, redIndex, "dvmh_loop_cuda_red_init_C");
    if (r->localArr.empty()) {
        r->localArr.assign((size_t)(r->elemSize * r->elemCount), 0);
        r->localLoc.assign((size_t)(r->locSize * r->elemCount), 0);
        r->initValues(r->localArr.data(), r->isLoc() ? r->localLoc.data() : nullptr);
    }
    void *dv = gpu.alloc((size_t)(r->elemSize * r->elemCount));
    gpu.h2d(dv, r->localArr.data(), (size_t)(r->elemSize * r->elemCount));
    sl->toFree.push_back(dv);
    if (devArrayAddrPtr) *devArrayAddrPtr = dv;
    if (devLocAddrPtr && r->isLoc()) {
        void *dl = gpu.alloc((size_t)(r->locSize * r->elemCount));
        gpu.h2d(dl, r->localLoc.data(), (size_t)(r->locSize * r->elemCount));
        sl->toFree.push_back(dl);
        *devLocAddrPtr = dl;
    }
}

extern "C" void dvmh_loop_cuda_red_prepare_C(DvmType curLoop, DvmType redIndex, DvmType count, DvmType fillFlag) {   // :4317
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_red_prepare_C");
    auto r = redOf(sl->loop, redIndex, "dvmh_loop_cuda_red_prepare_C");
    if (r->localArr.empty()) {
        r->localArr.assign((size_t)(r->elemSize * r->elemCount), 0);
        r->localLoc.assign((size_t)(r->locSize * r->elemCount), 0);
        r->initValues(r->localArr.data(), r->isLoc() ? r->localLoc.data() : nullptr);
    }
    gpuRedPrepare(sl, r.get(), count, fillFlag != 0);
}

extern "C" void dvmh_loop_cuda_red_finish_C(DvmType curLoop, DvmType redIndex) {   // :4373
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_cuda_red_finish_C");
    auto r = redOf(sl->loop, redIndex, "dvmh_loop_cuda_red_finish_C");
    gpuRedFinish(sl, r);
}

extern "C" void dvmh_loop_red_post_C(DvmType curLoop, DvmType redIndex, const void *arrayAddr, const void *locAddr) {   // :4386
    Loop *l = loopOf(curLoop, "dvmh_loop_red_post_C");
    redOf(l, redIndex, "dvmh_loop_red_post_C")->postValues(arrayAddr, locAddr);
}

// ------------------------------------------------------------ intervals (statistics hooks)
extern "C" void dvmh_par_interval_start_C() {}
extern "C" void dvmh_seq_interval_start_C() {}
extern "C" void dvmh_sp_interval_end_() {}
extern "C" void dvmh_usr_interval_start_C(DvmType) {}
extern "C" void dvmh_usr_interval_end_() {}

// ------------------------------------------------------------ bounds permutation (:4540-4583)
static void applyPerm(long rank, DvmType vals[], const DvmType idxPerm[]) {
    DvmType tmp[MAX_LOOP_RANK];
    for (long i = 0; i < rank; i++) tmp[i] = vals[idxPerm[i] - 1];
    memcpy(vals, tmp, sizeof(DvmType) * (size_t)rank);
}

// idxPerm[i]: original (1-based, outermost first) loop axis placed at position i.
// Variant 1 (ref fillAxisPerm) puts the independent axes first.
extern "C" void dvmh_change_filled_bounds_C(DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[], DvmType rank,
        DvmType depMask, DvmType idxPerm[]) {
    Guard g;
    int depCount = __builtin_popcountl((unsigned long)depMask);
    int h = 0, hd = (int)rank - depCount;
    for (long i = 1; i <= rank; i++) {
        if ((depMask >> (rank - i)) & 1) idxPerm[hd++] = i;
        else idxPerm[h++] = i;
    }
    if (depCount > 0 && depCount < rank) {
        applyPerm(rank, boundsLow, idxPerm);
        applyPerm(rank, boundsHigh, idxPerm);
        applyPerm(rank, loopSteps, idxPerm);
    }
}

// Variant 2 puts the dependent axes first and returns a 0-based permutation.
extern "C" void dvmh_change_filled_bounds2_C(DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[], DvmType rank,
        DvmType depMask, DvmType idxPerm[]) {
    Guard g;
    int depCount = __builtin_popcountl((unsigned long)depMask);
    int h = 0, hd = depCount;
    for (long i = 1; i <= rank; i++) {
        if ((depMask >> (rank - i)) & 1) idxPerm[h++] = i;
        else idxPerm[hd++] = i;
    }
    if (depCount > 0 && depCount < rank) {
        applyPerm(rank, boundsLow, idxPerm);
        applyPerm(rank, boundsHigh, idxPerm);
        applyPerm(rank, loopSteps, idxPerm);
    }
    for (long i = 0; i < rank; i++) idxPerm[i]--;
}

// ------------------------------------------------------------ autotransform (rts/dvmh/dvmh_rts.cpp:1404-1481)
static int lowestSetBit(unsigned long v, int n) {
    // index (from 0) of the n-th set bit counting from the lowest
    for (int b = 0; b < 64; b++)
        if ((v >> b) & 1) { if (--n == 0) return b; }
    return -1;
}

extern "C" DvmType dvmh_loop_autotransform_C(DvmType curLoop, DvmType dvmDesc[]) {   // :4011
    Guard g;
    SpecLoop *sl = asSpec(curLoop, "dvmh_loop_autotransform_C");
    Loop *l = sl->loop;
    Data *d = asData(dvmDesc, "dvmh_loop_autotransform_C");
    if (sl->device != 1 || !d->hasDev || d->rank <= 0) return 0;
    unsigned long depMask = (unsigned long)l->depMask;
    int depCount = __builtin_popcountl(depMask);
    std::vector<long> corrByData = l->correspondence(d);
    // loop axis (1-based) -> signed data axis
    std::vector<long> corr(l->rank, 0);
    for (int a = 0; a < d->rank; a++) {
        long la = corrByData[a];
        if (la != 0) corr[std::labs(la) - 1] = la > 0 ? a + 1 : -(a + 1);
    }
    bool ok = false;
    int perm[MAX_RANK];
    for (int i = 0; i < d->rank; i++) perm[i] = d->dev.axisPerm[i];
    if (l->rank > depCount && depCount <= 1) {
        unsigned long indep = ~depMask & ((l->rank >= 64) ? ~0UL : ((1UL << l->rank) - 1));
        int loopFastest = l->rank - lowestSetBit(indep, 1);
        long dataFastest = std::labs(corr[loopFastest - 1]);
        if (dataFastest > 0) {
            int oldPlace = (int)(std::find(perm, perm + d->rank, (int)dataFastest) - perm);
            std::swap(perm[oldPlace], perm[d->rank - 1]);
            d->setLayout(perm, 0);
            ok = true;
        }
    } else if (depCount >= 2) {
        int loopFastest = l->rank - lowestSetBit(depMask, 1);
        int loopPrev = l->rank - lowestSetBit(depMask, 2);
        long dataFastest = std::labs(corr[loopFastest - 1]), dataPrev = std::labs(corr[loopPrev - 1]);
        if (dataFastest > 0 && dataPrev > 0) {
            int oldPlace = (int)(std::find(perm, perm + d->rank, (int)dataFastest) - perm);
            std::swap(perm[oldPlace], perm[d->rank - 1]);
            oldPlace = (int)(std::find(perm, perm + d->rank, (int)dataPrev) - perm);
            std::swap(perm[oldPlace], perm[d->rank - 2]);
            int dirFastest = corr[loopFastest - 1] > 0 ? 1 : -1, dirPrev = corr[loopPrev - 1] > 0 ? 1 : -1;
            int slash = dirFastest * dirPrev < 0;
            d->setLayout(perm, 1 + slash);
            ok = true;
        }
    }
    if (!ok && depCount <= 1 && d->dev.diagState) d->setLayout(d->dev.axisPerm, 0);
    return (d->dev.isTransformed(d->rank) ? 1 : 0) + (d->dev.isDiagonalized() ? 1 : 0);
}

// ------------------------------------------------------------ runtime API
extern "C" void dvmh_barrier() {
    flushDeferredPosts();
    gpu.sync();
}
extern "C" int dvmh_get_total_num_procs() { return 1; }
extern "C" int dvmh_get_num_proc_axes() { return 1; }
extern "C" int dvmh_get_num_procs(int axis) { (void)axis; return 1; }
extern "C" double dvmh_wtime() { return wallTime(); }
extern "C" int dvmh_remove_local(const char *filename) { return remove(filename); }
extern "C" int dvmh_rename_local(const char *oldFN, const char *newFN) { return rename(oldFN, newFN); }
extern "C" FILE *dvmh_tmpfile_local(void) { return tmpfile(); }
extern "C" char *dvmh_tmpnam_local(char *s) {
    static char buf[64];
    char *out = s ? s : buf;
    snprintf(out, 64, "/tmp/dvmh_tmp_%d_XXXXXX", (int)getpid());
    int fd = mkstemp(out);
    if (fd >= 0) close(fd);
    return out;
}

// ------------------------------------------------------------ statistics
extern "C" void dvmh_amd_get_stats_C(DvmType out[8]) {
    ensureInit();
    out[0] = rt.loopsOnHost;
    out[1] = rt.loopsOnGpu;
    out[2] = (DvmType)(rt.gpuLoopSeconds * 1e9);
    out[3] = gpu.present ? 1 : 0;
    out[4] = gpu.props.cuCount;
    out[5] = graphReplays;
    out[6] = batchGroups;
    out[7] = batchLoops;
}

extern "C" DvmType dvmh_amd_layout_changes_C(void) { return layoutChanges(); }

extern "C" DvmType dvmh_amd_pending_posts_C(void) {
    ensureInit();
    return pendingPostCount();
}

extern "C" void dvmh_amd_set_allow_async_C(DvmType allow) {
    ensureInit();
    settings.allowAsync = allow != 0;
}

extern "C" void dvmh_amd_transfer_stats_C(DvmType out[2]) {
    ensureInit();
    out[0] = gpu.h2dBytes;
    out[1] = gpu.d2hBytes;
}

// The piece algebra the actual-state tracking uses, on its own: `nops`
// operations (code 1 add, 2 remove, then lo[rank], hi[rank]) applied to an
// empty set, then the parts of `query` the set misses, written as boxes
// (lo[rank], hi[rank]) into outBoxes; returns their count (written up to maxBoxes).
extern "C" DvmType dvmh_amd_pieces_eval_C(DvmType rank, const DvmType *ops, DvmType nops, const DvmType *query, DvmType *outBoxes,
        DvmType maxBoxes) {
    DVMH_CHECK(rank >= 0 && rank <= MAX_RANK, "dvmh_amd_pieces_eval_C: rank %ld", (long)rank);
    Pieces p;
    p.rank = (int)rank;
    const long rec = 1 + 2 * rank;
    for (long i = 0; i < nops; i++) {
        const DvmType *o = ops + i * rec;
        Box b;
        for (int a = 0; a < rank; a++) { b.lo[a] = o[1 + a]; b.hi[a] = o[1 + rank + a]; }
        if (o[0] == 1) p.add(b);
        else if (o[0] == 2) p.remove(b);
        else DVMH_CHECK(false, "dvmh_amd_pieces_eval_C: operation %ld", (long)o[0]);
    }
    Box q;
    for (int a = 0; a < rank; a++) { q.lo[a] = query[a]; q.hi[a] = query[rank + a]; }
    std::vector<Box> miss = p.missing(q);
    for (long i = 0; i < (long)miss.size() && i < maxBoxes; i++)
        for (int a = 0; a < rank; a++) {
            outBoxes[i * 2 * rank + a] = miss[i].lo[a];
            outBoxes[i * 2 * rank + rank + a] = miss[i].hi[a];
        }
    return (DvmType)miss.size();
}

extern "C" DvmType dvmh_amd_host_slots_C(void) {
    ensureInit();
    return hostSlotCount();
}

extern "C" void dvmh_amd_reset_stats_C(void) {
    gpu.h2dBytes = gpu.d2hBytes = 0;
    rt.loopsOnHost = rt.loopsOnGpu = 0;
    rt.gpuLoopSeconds = 0;
    graphReplays = batchGroups = batchLoops = 0;
    resetPairStats();
}

extern "C" const char *dvmh_amd_device_name_C(void) {
    ensureInit();
    static char buf[320];
    if (!gpu.present) return "";
    snprintf(buf, sizeof(buf), "%s (%s)", gpu.props.name, gpu.props.arch);
    return buf;
}
