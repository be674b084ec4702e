// kf_runtime.hip -- error state, device workspace cache, library info.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <string>

#include "kf_common.hpp"

namespace kf {

static thread_local std::string g_last_error;

void set_last_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

namespace {
constexpr int kMaxDevices = 16;
struct Slot {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
Slot g_ws[kMaxDevices][WS_NSLOTS];
Slot g_host[WS_NSLOTS];
}  // namespace

// Grow-only: a slot is reallocated (after draining the stream) only when a call
// needs more than it holds.  Sizes are rounded up to 2 MiB so that slowly
// growing requests do not reallocate every call.
void *workspace(int slot, size_t bytes, hipStream_t stream) {
    if (bytes == 0) bytes = 256;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Slot &s = g_ws[dev][slot];
    if (s.bytes >= bytes) return s.ptr;
    if (s.ptr) {
        if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
        (void)hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    const size_t want = (bytes + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    s.ptr = p;
    s.bytes = want;
    return p;
}

// One extra stream and a few events per device, created on first use and kept for the
// process: a call that overlaps two dependency chains (the Cholesky look-ahead) forks
// onto the aux stream and joins back through events.  Default priority, non-blocking
// (ordering with the caller's stream is only through the events).
hipStream_t aux_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    static hipStream_t streams[kMaxDevices] = {};
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    if (!streams[dev] && hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        streams[dev] = nullptr;
    }
    return streams[dev];
}

hipEvent_t aux_event(int i) {
    int dev = 0;
    if (i < 0 || i >= kAuxEvents) return nullptr;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    static hipEvent_t events[kMaxDevices][kAuxEvents] = {};
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    hipEvent_t &e = events[dev][i];
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        e = nullptr;
    }
    return e;
}

void *host_staging(int slot, size_t bytes) {
    if (bytes == 0) bytes = 64;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Slot &s = g_host[slot];
    if (s.bytes >= bytes) return s.ptr;
    if (s.ptr) (void)hipHostFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    s.ptr = p;
    s.bytes = bytes;
    return p;
}

namespace {
struct DeviceGate {
    std::recursive_mutex m;
    int depth = 0;  // nesting of entry points on the owning thread
    hipStream_t last = nullptr;
    hipEvent_t done = nullptr;  // recorded at the end of the last outermost call
    bool recorded = false;
};
DeviceGate g_gate[kMaxDevices];

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}
}  // namespace

EntryGuard::EntryGuard(hipStream_t s) : stream_(s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return;
    }
    dev_ = dev;
    DeviceGate &g = g_gate[dev];
    g.m.lock();
    if (g.depth++ == 0 && g.recorded && g.last != s && !capturing(s))
        (void)hipStreamWaitEvent(s, g.done, 0);
}

EntryGuard::~EntryGuard() {
    if (dev_ < 0) return;
    DeviceGate &g = g_gate[dev_];
    if (--g.depth == 0 && !capturing(stream_)) {
        if (!g.done && hipEventCreateWithFlags(&g.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            g.done = nullptr;
        }
        if (g.done && hipEventRecord(g.done, stream_) == hipSuccess) {
            g.recorded = true;
            g.last = stream_;
        } else {
            (void)hipGetLastError();
        }
    }
    g.m.unlock();
}

// Copy Q (nm x ma) and N (nm) from the device to host vectors (synchronises the stream).
int read_metadata(const int *Q, const int *N, int nm, int ma, std::vector<int> &hQ,
                  std::vector<int> &hN, hipStream_t stream, int slot) {
    hQ.resize((size_t)nm * ma);
    hN.resize(nm);
    int *stage = static_cast<int *>(host_staging(slot, sizeof(int) * ((size_t)nm * ma + nm)));
    KF_REQUIRE(stage, KF_ERR_NOMEM, "pinned staging allocation failed");
    if (nm * (size_t)ma)
        KF_HIP_TRY(hipMemcpyAsync(stage, Q, sizeof(int) * (size_t)nm * ma, hipMemcpyDeviceToHost,
                                  stream));
    if (nm)
        KF_HIP_TRY(hipMemcpyAsync(stage + (size_t)nm * ma, N, sizeof(int) * nm,
                                  hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    std::copy(stage, stage + (size_t)nm * ma, hQ.begin());
    std::copy(stage + (size_t)nm * ma, stage + (size_t)nm * ma + nm, hN.begin());
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Kernel timers
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tm_mutex;
bool g_tm_enabled = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_tm_events[TM_COUNT];
}  // namespace

ScopedTimer::ScopedTimer(int id_, hipStream_t s) : id(id_), stream(s), handle(nullptr) {
    std::lock_guard<std::mutex> lock(g_tm_mutex);
    if (!g_tm_enabled || id < 0 || id >= TM_COUNT) return;  // id -1: untimed
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
    (void)hipEventRecord(a, stream);
    g_tm_events[id].push_back({a, b});
    handle = reinterpret_cast<void *>(b);
}
void ScopedTimer::stop() {
    if (handle) (void)hipEventRecord(reinterpret_cast<hipEvent_t>(handle), stream);
    handle = nullptr;
}

}  // namespace kf

extern "C" int kf_timer_enable(int on) {
    std::lock_guard<std::mutex> lock(kf::g_tm_mutex);
    kf::g_tm_enabled = on != 0;
    return KF_OK;
}

extern "C" int kf_timer_reset(void) {
    std::lock_guard<std::mutex> lock(kf::g_tm_mutex);
    for (auto &v : kf::g_tm_events) {
        for (auto &e : v) {
            (void)hipEventSynchronize(e.second);
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        v.clear();
    }
    return KF_OK;
}

extern "C" int kf_timer_read(int id, double *total_ms, int *count) {
    KF_REQUIRE(id >= 0 && id < kf::TM_COUNT && total_ms && count, KF_ERR_ARG, "bad timer id");
    std::lock_guard<std::mutex> lock(kf::g_tm_mutex);
    double tot = 0.0;
    for (auto &e : kf::g_tm_events[id]) {
        KF_HIP_TRY(hipEventSynchronize(e.second));
        float ms = 0.f;
        KF_HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    *total_ms = tot;
    *count = (int)kf::g_tm_events[id].size();
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Device-memory plumbing for FFI callers that hold no GPU allocator of their own.
// ---------------------------------------------------------------------------
extern "C" int kf_malloc(void **ptr, size_t bytes) {
    KF_REQUIRE(ptr, KF_ERR_ARG, "null out-pointer");
    *ptr = nullptr;
    if (bytes == 0) return KF_OK;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        kf::set_last_error("hipMalloc of %zu bytes failed", bytes);
        return KF_ERR_NOMEM;
    }
    return KF_OK;
}

extern "C" int kf_free(void *ptr) {
    if (ptr) KF_HIP_TRY(hipFree(ptr));
    return KF_OK;
}

// kind: 0 host->device, 1 device->host, 2 device->device.  Host<->device copies
// complete before the call returns; device->device copies are stream-ordered.
extern "C" int kf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    if (bytes == 0) return KF_OK;
    KF_REQUIRE(dst && src && kind >= 0 && kind <= 2, KF_ERR_ARG, "bad kf_memcpy arguments");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                      : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    hipStream_t s = kf::as_stream(stream);
    KF_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, s));
    if (kind != 2) KF_HIP_TRY(hipStreamSynchronize(s));
    return KF_OK;
}

extern "C" int kf_memset_zero(void *dst, size_t bytes, void *stream) {
    if (bytes == 0) return KF_OK;
    KF_REQUIRE(dst, KF_ERR_ARG, "null pointer");
    KF_HIP_TRY(hipMemsetAsync(dst, 0, bytes, kf::as_stream(stream)));
    return KF_OK;
}

extern "C" int kf_stream_synchronize(void *stream) {
    KF_HIP_TRY(hipStreamSynchronize(kf::as_stream(stream)));
    return KF_OK;
}

extern "C" int kf_version(void) { return 1; }

extern "C" const char *kf_last_error(void) { return kf::g_last_error.c_str(); }

extern "C" int kf_device_name(char *buf, int len) {
    int dev = 0;
    KF_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    KF_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (buf && len > 0) {
        std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                      prop.multiProcessorCount);
    }
    return KF_OK;
}
