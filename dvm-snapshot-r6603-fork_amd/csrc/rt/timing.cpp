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

extern "C" void dvmh_amd_set_graph_C(DvmType enable) { settings.graph = enable != 0; }

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
