// Kernel timing marks used by emitted handlers, and execution options.
#include <map>
#include <string>
#include <vector>

#include "rt.h"
#include "../kernels/launch.h"

namespace dvmh {

struct SiteTiming {
    long launches = 0;
    double seconds = 0;
};
static bool timingOn = false;
static long timingPeriod = 1;   // time one launch in this many, per site
static std::map<std::string, SiteTiming> siteTimes;
static std::map<std::string, long> siteMarks;   // phase-0 marks seen per site while timing
// events recorded but not yet read back: (site, start, stop)
struct PendingEvents {
    std::string site;
    hipEvent_t a, b;
};
static std::vector<PendingEvents> inflight;
static std::vector<hipEvent_t> eventPool;
static hipEvent_t openStart = nullptr;

static hipEvent_t takeEvent() {
    if (!eventPool.empty()) {
        hipEvent_t e = eventPool.back();
        eventPool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
}

static std::string siteOf(const Loop *l) { return l->file + ":" + std::to_string(l->line); }

// Reads back every recorded pair (synchronising on the last one).
void drainKernelTimes() {
    if (inflight.empty()) return;
    HIP_CHECK(hipEventSynchronize(inflight.back().b));
    for (auto &p : inflight) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        SiteTiming &t = siteTimes[p.site];
        t.launches++;
        t.seconds += ms * 1e-3;
        eventPool.push_back(p.a);
        eventPool.push_back(p.b);
    }
    inflight.clear();
}

bool kernelTimingOn() { return timingOn; }

}  // namespace dvmh

using namespace dvmh;

extern "C" void dvmh_amd_mark_kernel_C(DvmType curLoop, DvmType phase) {
    runtimeInit();
    if (!timingOn) return;
    SpecLoop *sl = asSpec(curLoop, "dvmh_amd_mark_kernel_C");
    if (sl->device != 1) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_CHECK(hipStreamIsCapturing(sl->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return;   // events are not recorded into graphs
    if (phase == 0) {
        if (timingPeriod > 1 && siteMarks[siteOf(sl->loop)]++ % timingPeriod != 0) return;   // not sampled
        openStart = takeEvent();
        HIP_CHECK(hipEventRecord(openStart, sl->stream));
    } else if (openStart) {
        hipEvent_t b = takeEvent();
        HIP_CHECK(hipEventRecord(b, sl->stream));
        inflight.push_back(PendingEvents{siteOf(sl->loop), openStart, b});
        openStart = nullptr;
        if (inflight.size() > 4096) drainKernelTimes();
    }
}

extern "C" void dvmh_amd_set_kernel_timing_C(DvmType enable) {
    runtimeInit();
    if (!enable) drainKernelTimes();
    timingOn = enable != 0;
    timingPeriod = enable > 1 ? (long)enable : 1;
    siteMarks.clear();
}

extern "C" DvmType dvmh_amd_kernel_stats_C(const char *file, DvmType line, DvmType out[2]) {
    runtimeInit();
    drainKernelTimes();
    auto it = siteTimes.find(std::string(file) + ":" + std::to_string(line));
    if (it == siteTimes.end()) return 0;
    out[0] = it->second.launches;
    out[1] = (DvmType)(it->second.seconds * 1e9);
    return 1;
}

extern "C" void dvmh_amd_reset_kernel_stats_C(void) {
    runtimeInit();
    drainKernelTimes();
    siteTimes.clear();
}

extern "C" void dvmh_amd_set_sync_mode_C(DvmType mode) {
    runtimeInit();
    flushDeferredPosts();
    settings.syncMode = (int)mode;
}

extern "C" void dvmh_amd_set_graph_min_nodes_C(DvmType nodes) { settings.graphMinNodes = nodes < 0 ? 6 : (long)nodes; }

extern "C" void dvmh_amd_set_graph_C(DvmType enable) {
    settings.graph = enable != 0;
    settings.graphNoCapture = enable == 2;
}

// Streaming kernels of known traffic, timed with HIP events on the runtime's
// stream over `reps` back-to-back launches (after one untimed launch); the
// buffers are allocated here for the measurement and freed after it.
// Returns seconds per launch, or -1 without a GPU.
static double streamSeconds(long bytes, long reps, bool copy) {
    runtimeInit();
    if (!gpu.present || bytes <= 0 || reps <= 0) return -1.0;
    void *a = nullptr, *b = nullptr;
    uint64_t *sink = nullptr;
    HIP_CHECK(hipMalloc(&a, (size_t)bytes));
    if (copy) HIP_CHECK(hipMalloc(&b, (size_t)bytes));
    else HIP_CHECK(hipMalloc((void **)&sink, (size_t)dvmh_hip_stream_sink_bytes(bytes) + 64));
    HIP_CHECK(hipMemsetAsync(a, 1, (size_t)bytes, gpu.stream));
    auto launch = [&]() {
        if (copy) dvmh_hip_stream_copy(b, a, bytes, gpu.stream);
        else dvmh_hip_stream_read(a, bytes, sink, gpu.stream);
    };
    launch();
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, gpu.stream));
    for (long r = 0; r < reps; r++) launch();
    HIP_CHECK(hipEventRecord(e1, gpu.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    HIP_CHECK(hipFree(a));
    if (b) HIP_CHECK(hipFree(b));
    if (sink) HIP_CHECK(hipFree(sink));
    return ms * 1e-3 / (double)reps;
}

extern "C" double dvmh_amd_stream_copy_seconds_C(DvmType bytes, DvmType reps) { return streamSeconds(bytes, reps, true); }
extern "C" double dvmh_amd_stream_read_seconds_C(DvmType bytes, DvmType reps) { return streamSeconds(bytes, reps, false); }

// The layout-transform kernels on an FP64 array of n0 x n1 x n2 elements,
// timed with HIP events on the runtime's stream (`reps` round trips after one
// untimed round trip): which 0 exchanges the two fastest axes (the tiled
// transpose), 1 goes to the diagonal layout of the (n1, n2) planes, 2 rotates
// the axes to n2 x n0 x n1 (the general gather).  out[0] / out[1] receive the
// seconds of one forward / backward transform; returns 0, or -1 without a GPU.
extern "C" DvmType dvmh_amd_layout_transform_seconds_C(DvmType which, DvmType n0, DvmType n1, DvmType n2, DvmType reps,
        double out[2]) {
    runtimeInit();
    if (!gpu.present || which < 0 || which > 2 || n0 <= 0 || n1 <= 0 || n2 <= 0 || reps <= 0) return -1;
    const size_t bytes = (size_t)n0 * n1 * n2 * sizeof(double);
    void *a = nullptr, *b = nullptr;
    HIP_CHECK(hipMalloc(&a, bytes));
    HIP_CHECK(hipMalloc(&b, bytes));
    HIP_CHECK(hipMemsetAsync(a, 1, bytes, gpu.stream));
    const long fwdSizes[3] = {n0, n1, n2};
    const long swapBack[3] = {n0, n2, n1}, rotBack[3] = {n2, n0, n1};
    const int swapPerm[3] = {0, 2, 1}, rotPerm[3] = {2, 0, 1}, rotInv[3] = {1, 2, 0};
    auto step = [&](bool back) {
        const void *src = back ? b : a;
        void *dst = back ? a : b;
        if (which == 0) dvmh_hip_permute(src, dst, 8, 3, back ? swapBack : fwdSizes, swapPerm, gpu.stream);
        else if (which == 1) dvmh_hip_diagonal(src, dst, 8, n2, n1, n0, back ? 1 : 0, 0, gpu.stream);
        else dvmh_hip_permute(src, dst, 8, 3, back ? rotBack : fwdSizes, back ? rotInv : rotPerm, gpu.stream);
    };
    step(false);
    step(true);
    hipEvent_t ev[3];
    for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    float fwd = 0, bwd = 0;
    for (long r = 0; r < reps; r++) {
        float ms0 = 0, ms1 = 0;
        HIP_CHECK(hipEventRecord(ev[0], gpu.stream));
        step(false);
        HIP_CHECK(hipEventRecord(ev[1], gpu.stream));
        step(true);
        HIP_CHECK(hipEventRecord(ev[2], gpu.stream));
        HIP_CHECK(hipEventSynchronize(ev[2]));
        HIP_CHECK(hipEventElapsedTime(&ms0, ev[0], ev[1]));
        HIP_CHECK(hipEventElapsedTime(&ms1, ev[1], ev[2]));
        fwd += ms0;
        bwd += ms1;
    }
    for (auto &e : ev) HIP_CHECK(hipEventDestroy(e));
    HIP_CHECK(hipFree(a));
    HIP_CHECK(hipFree(b));
    out[0] = fwd * 1e-3 / (double)reps;
    out[1] = bwd * 1e-3 / (double)reps;
    return 0;
}
