// runtime.cpp -- C-ABI runtime entry points and the plumbing in common.hpp.
#include <cstdarg>
#include <cstdio>
#include <atomic>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace th {

static thread_local char g_err[512] = "";
static thread_local hipStream_t g_stream = nullptr;

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail_hip(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? TIRA_ERR_OUT_OF_MEMORY : TIRA_ERR_HIP;
}

hipStream_t stream() { return g_stream; }

bool is_device(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    std::memset(&a, 0, sizeof(a));
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory reports an error on some runtimes
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

int cu_count() {
    static constexpr int kMaxDev = 64;
    static std::atomic<int> cache[kMaxDev];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
    if (dev < kMaxDev) {
        const int c = cache[dev].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    if (dev < kMaxDev) cache[dev].store(n, std::memory_order_relaxed);
    return n;
}

int use_device(int device) {
    if (device < 0) {
        set_error("device %d: the CPU path lives in tira::cpu; this library runs on a GPU", device);
        return TIRA_ERR_INVALID_ARG;
    }
    int n = 0;
    TH_HIP(hipGetDeviceCount(&n));
    if (device >= n) {
        set_error("device %d out of range (%d devices)", device, n);
        return TIRA_ERR_NO_DEVICE;
    }
    int cur = -1;
    TH_HIP(hipGetDevice(&cur));
    if (cur != device) TH_HIP(hipSetDevice(device));
    return TIRA_OK;
}

struct Workspace {
    void* ptr[WS_NSLOTS] = {};
    size_t cap[WS_NSLOTS] = {};
};
static constexpr int kMaxDevices = 16;
// Deliberately not freed at thread exit: the HIP runtime may already be torn
// down then. tira_hip_release_workspace() frees explicitly.
static thread_local Workspace* g_ws[kMaxDevices] = {};
// Bumped whenever a workspace buffer is allocated or released, so that
// callers caching the contents of a slot can tell it is still the same buffer.
static thread_local unsigned long long g_ws_gen = 1;

unsigned long long ws_generation() { return g_ws_gen; }

int ws_get(int slot, size_t bytes, void** p) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices || slot < 0 || slot >= WS_NSLOTS) {
        set_error("workspace: bad device %d / slot %d", dev, slot);
        return TIRA_ERR_INVALID_ARG;
    }
    if (!g_ws[dev]) g_ws[dev] = new Workspace();
    Workspace& w = *g_ws[dev];
    if (bytes == 0) bytes = 256;
    if (w.cap[slot] < bytes) {
        if (w.ptr[slot]) {
            TH_HIP(hipStreamSynchronize(g_stream));
            TH_HIP(hipFree(w.ptr[slot]));
            w.ptr[slot] = nullptr;
            w.cap[slot] = 0;
        }
        size_t b = (bytes + 4095) & ~size_t(4095);
        g_ws_gen++;
        TH_HIP(hipMalloc(&w.ptr[slot], b));
        w.cap[slot] = b;
    }
    *p = w.ptr[slot];
    return TIRA_OK;
}

size_t ws_capacity(int slot) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || slot < 0 || slot >= WS_NSLOTS ||
        !g_ws[dev])
        return 0;
    return g_ws[dev]->cap[slot];
}

int stage_in(const void* src, size_t bytes, int slot, const void** dev) {
    if (!src) { *dev = nullptr; return TIRA_OK; }
    if (is_device(src)) { *dev = src; return TIRA_OK; }
    void* d = nullptr;
    TH_TRY(ws_get(slot, bytes, &d));
    TH_HIP(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, g_stream));
    *dev = d;
    return TIRA_OK;
}

int stage_out(void* dst, size_t bytes, int slot, void** dev) {
    if (!dst) { *dev = nullptr; return TIRA_OK; }
    if (is_device(dst)) { *dev = dst; return TIRA_OK; }
    return ws_get(slot, bytes, dev);
}

int finish_out(void* dst, const void* dev, size_t bytes) {
    if (dst && dst != dev) {
        TH_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, g_stream));
    }
    TH_HIP(hipStreamSynchronize(g_stream));
    return TIRA_OK;
}

// ---------------------------------------------------------------- profiling
struct ProfRecord {
    const char* name;
    hipEvent_t a, b;
};
struct ProfStat {
    std::string name;
    double total_ms = 0;
    long count = 0;
};
static thread_local bool g_prof_on = false;
static thread_local std::vector<ProfRecord> g_prof_pending;
static thread_local std::vector<ProfStat> g_prof_stats;

ProfScope::ProfScope(const char* name, hipStream_t st) : st_(st) {
    if (!g_prof_on) return;
    ProfRecord r{name, nullptr, nullptr};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
    hipEventRecord(r.a, st);
    g_prof_pending.push_back(r);
    slot_ = (int)g_prof_pending.size() - 1;
}

ProfScope::~ProfScope() {
    if (slot_ >= 0) hipEventRecord(g_prof_pending[slot_].b, st_);
}

}  // namespace th

using namespace th;

extern "C" {

int tira_hip_abi_version(void) { return TIRA_HIP_ABI_VERSION; }
const char* tira_hip_last_error(void) { return g_err; }

int tira_hip_device_count(int* count) {
    TH_ARG(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
    *count = n;
    return n > 0 ? TIRA_OK : TIRA_ERR_NO_DEVICE;
}

int tira_hip_set_device(int device) { return use_device(device); }

int tira_hip_malloc(void** ptr, size_t bytes) {
    TH_ARG(ptr, "ptr is NULL");
    TH_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return TIRA_OK;
}

int tira_hip_free(void* ptr) {
    if (ptr) TH_HIP(hipFree(ptr));
    return TIRA_OK;
}

int tira_hip_memcpy(void* dst, const void* src, size_t bytes) {
    if (!bytes) return TIRA_OK;
    TH_ARG(dst && src, "NULL pointer");
    TH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_stream));
    TH_HIP(hipStreamSynchronize(g_stream));
    return TIRA_OK;
}

int tira_hip_memset(void* dst, int value, size_t bytes) {
    if (!bytes) return TIRA_OK;
    TH_ARG(dst, "NULL pointer");
    TH_HIP(hipMemsetAsync(dst, value, bytes, g_stream));
    TH_HIP(hipStreamSynchronize(g_stream));
    return TIRA_OK;
}

int tira_hip_synchronize(void) {
    TH_HIP(hipDeviceSynchronize());
    return TIRA_OK;
}

int tira_hip_stream_create(void** s) {
    TH_ARG(s, "stream is NULL");
    hipStream_t st;
    TH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (void*)st;
    return TIRA_OK;
}

int tira_hip_stream_destroy(void* s) {
    if (s) TH_HIP(hipStreamDestroy((hipStream_t)s));
    if ((hipStream_t)s == g_stream) g_stream = nullptr;
    return TIRA_OK;
}

int tira_hip_set_stream(void* s) {
    g_stream = (hipStream_t)s;
    return TIRA_OK;
}

int tira_hip_stream_synchronize(void* s) {
    TH_HIP(hipStreamSynchronize((hipStream_t)s));
    return TIRA_OK;
}

int tira_hip_event_create(void** e) {
    TH_ARG(e, "event is NULL");
    hipEvent_t ev;
    TH_HIP(hipEventCreate(&ev));
    *e = (void*)ev;
    return TIRA_OK;
}

int tira_hip_event_destroy(void* e) {
    if (e) TH_HIP(hipEventDestroy((hipEvent_t)e));
    return TIRA_OK;
}

int tira_hip_event_record(void* e, void* s) {
    TH_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return TIRA_OK;
}

int tira_hip_event_elapsed_ms(void* a, void* b, float* ms) {
    TH_ARG(ms, "ms is NULL");
    TH_HIP(hipEventSynchronize((hipEvent_t)b));
    TH_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return TIRA_OK;
}

int tira_hip_profile_begin(void) {
    g_prof_on = true;
    g_prof_pending.clear();
    g_prof_stats.clear();
    return TIRA_OK;
}

int tira_hip_profile_end(void) {
    g_prof_on = false;
    for (auto& r : g_prof_pending) {
        TH_HIP(hipEventSynchronize(r.b));
        float ms = 0;
        TH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        ProfStat* st = nullptr;
        for (auto& s : g_prof_stats)
            if (s.name == r.name) st = &s;
        if (!st) { g_prof_stats.push_back(ProfStat{r.name, 0, 0}); st = &g_prof_stats.back(); }
        st->total_ms += ms;
        st->count += 1;
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    g_prof_pending.clear();
    return TIRA_OK;
}

int tira_hip_profile_get(const char* name, double* total_ms, long* count) {
    TH_ARG(name && total_ms && count, "NULL pointer");
    *total_ms = 0; *count = 0;
    for (auto& s : g_prof_stats)
        if (s.name == name) { *total_ms = s.total_ms; *count = s.count; }
    return TIRA_OK;
}

int tira_hip_copy_async(void* dst, const void* src, size_t bytes) {
    TH_ARG(dst && src && bytes % 16 == 0, "copy needs device pointers and a multiple of 16 bytes");
    return th::launch_copy(dst, src, bytes, g_stream);
}

int tira_hip_fma_loop_async(float* out, unsigned blocks, int iters) {
    TH_ARG(out && blocks > 0 && iters > 0, "fma loop needs a device buffer of `blocks` floats");
    return th::launch_fma_loop(out, blocks, iters, g_stream);
}

int tira_hip_pointer_is_device(const void* ptr, int* is_dev) {
    TH_ARG(is_dev, "NULL pointer");
    *is_dev = th::is_device(ptr) ? 1 : 0;
    return TIRA_OK;
}

int tira_hip_release_workspace(void) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < kMaxDevices && g_ws[dev]) {
        g_ws_gen++;
        TH_HIP(hipStreamSynchronize(g_stream));
        for (int s = 0; s < WS_NSLOTS; s++)
            if (g_ws[dev]->ptr[s]) TH_HIP(hipFree(g_ws[dev]->ptr[s]));
        delete g_ws[dev];
        g_ws[dev] = nullptr;
    }
    return TIRA_OK;
}

}  // extern "C"
