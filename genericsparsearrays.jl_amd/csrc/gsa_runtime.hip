// gsa_runtime.hip — error reporting, device set-up, memory/stream helpers.
#include "gsa_common.h"
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace gsa {

static thread_local std::string g_err;

void set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = std::string(buf) + " (status " + std::to_string(code) + ")";
}

int hip_fail(hipError_t e, const char* what) {
    int code = (e == hipErrorOutOfMemory) ? GSA_ERR_OOM : GSA_ERR_HIP;
    set_error(code, "%s failed: %s", what, hipGetErrorString(e));
    return code;
}

static std::once_flag g_once;
static int g_init_status = GSA_OK;

int ensure_init() {
    std::call_once(g_once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) {
            g_init_status = (e != hipSuccess) ? hip_fail(e, "hipGetDeviceCount") : GSA_ERR_HIP;
            if (e == hipSuccess) set_error(GSA_ERR_HIP, "no HIP device visible");
            return;
        }
    });
    return g_init_status;
}

// Scratch memory: a cache of device blocks per (device, stream).  A block
// released on a stream is handed again only to work queued later on that same
// stream, so stream order makes the reuse safe without events; a miss takes a
// new block from hipMalloc.  (Stream-ordered pool allocation costs about
// 0.1 ms per call on this runtime, several times per structural operation.)
namespace {
struct WsBlock {
    void* p;
    size_t bytes;
};
std::mutex g_ws_mu;
std::map<std::pair<int, hipStream_t>, std::vector<WsBlock>> g_ws_free;
std::unordered_map<void*, std::pair<int, size_t>> g_ws_live;  // block -> (device, bytes)
// Blocks taken while the stream was being captured into a graph: stream-ordered
// allocations (graph nodes), released the same way and never cached.
std::unordered_map<void*, int> g_ws_captured;
size_t g_ws_cached = 0;
constexpr size_t WS_CACHE_MAX = 32ull << 30;  // cached bytes kept at most

size_t ws_round(size_t b) {
    if (b < 256) b = 256;
    const size_t g = b > (1u << 20) ? (1u << 20) : 256;
    return (b + g - 1) / g * g;
}

// Release every cached block (after the device has drained: a cached block may
// still be read by queued work).
void ws_trim() {
    std::vector<std::pair<int, void*>> out;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        for (auto& kv : g_ws_free)
            for (auto& b : kv.second) out.push_back({kv.first.first, b.p});
        g_ws_free.clear();
        g_ws_cached = 0;
    }
    if (out.empty()) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& q : out) {
        (void)hipSetDevice(q.first);
        (void)hipDeviceSynchronize();
        (void)hipFree(q.second);
    }
    (void)hipSetDevice(cur);
}
}  // namespace

int ws_alloc(void** p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    bytes = ws_round(bytes);
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) == hipSuccess && cap == hipStreamCaptureStatusActive) {
        hipError_t e = hipMallocAsync(p, bytes, s);
        if (e != hipSuccess) {
            *p = nullptr;
            return hip_fail(e, "hipMallocAsync(scratch)");
        }
        std::lock_guard<std::mutex> lk(g_ws_mu);
        g_ws_captured[*p] = dev;
        return GSA_OK;
    }
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto it = g_ws_free.find({dev, s});
        if (it != g_ws_free.end()) {
            // smallest cached block that fits and wastes at most its own size
            auto& v = it->second;
            size_t best = v.size();
            for (size_t i = 0; i < v.size(); ++i)
                if (v[i].bytes >= bytes && v[i].bytes <= 2 * bytes + (1u << 20) &&
                    (best == v.size() || v[i].bytes < v[best].bytes))
                    best = i;
            if (best != v.size()) {
                const WsBlock b = v[best];
                v[best] = v.back();
                v.pop_back();
                g_ws_cached -= b.bytes;
                g_ws_live[b.p] = {dev, b.bytes};
                *p = b.p;
                return GSA_OK;
            }
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        ws_trim();
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc(scratch)");
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws_live[*p] = {dev, bytes};
    return GSA_OK;
}

void ws_free(void* p, hipStream_t s) {
    if (!p) return;
    bool trim = false;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto ct = g_ws_captured.find(p);
        if (ct != g_ws_captured.end()) {
            g_ws_captured.erase(ct);
            (void)hipFreeAsync(p, s);
            return;
        }
        auto it = g_ws_live.find(p);
        if (it == g_ws_live.end()) return;
        const int dev = it->second.first;
        const size_t bytes = it->second.second;
        g_ws_live.erase(it);
        g_ws_free[{dev, s}].push_back({p, bytes});
        g_ws_cached += bytes;
        trim = g_ws_cached > WS_CACHE_MAX;
    }
    if (trim) ws_trim();
}

}  // namespace gsa

using namespace gsa;

extern "C" {

int gsa_init(void) { return ensure_init(); }

const char* gsa_last_error(void) { return g_err.c_str(); }

int gsa_version(void) { return 100; }  // 0.1.0

int gsa_malloc(void** ptr, int64_t bytes) {
    GSA_TRY(ensure_init());
    GSA_CHECK(ptr != nullptr && bytes >= 0, GSA_ERR_INVALID, "gsa_malloc: bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return GSA_OK;
    GSA_HIP(hipMalloc(ptr, (size_t)bytes));
    return GSA_OK;
}

int gsa_free(void* ptr) {
    if (ptr) GSA_HIP(hipFree(ptr));
    return GSA_OK;
}

int gsa_memcpy(void* dst, const void* src, int64_t bytes, int kind, void* stream) {
    GSA_TRY(ensure_init());
    GSA_CHECK(bytes >= 0 && kind >= 0 && kind <= 2, GSA_ERR_INVALID, "gsa_memcpy: bad arguments");
    if (bytes == 0) return GSA_OK;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    GSA_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, as_stream(stream)));
    if (kind == 1) GSA_HIP(hipStreamSynchronize(as_stream(stream)));
    return GSA_OK;
}

int gsa_memset(void* dst, int value, int64_t bytes, void* stream) {
    GSA_TRY(ensure_init());
    if (bytes <= 0) return GSA_OK;
    GSA_HIP(hipMemsetAsync(dst, value, (size_t)bytes, as_stream(stream)));
    return GSA_OK;
}

int gsa_stream_create(void** stream) {
    GSA_TRY(ensure_init());
    hipStream_t s;
    GSA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GSA_OK;
}

int gsa_stream_destroy(void* stream) {
    if (stream) GSA_HIP(hipStreamDestroy(as_stream(stream)));
    return GSA_OK;
}

int gsa_stream_sync(void* stream) {
    GSA_HIP(hipStreamSynchronize(as_stream(stream)));
    return GSA_OK;
}

int gsa_ctx_destroy(gsa_ctx ctx) {
    if (!ctx) return GSA_OK;
    for (void* b : ctx->buf) ws_free(b, ctx->stream);
    delete ctx;
    return GSA_OK;
}

}  // extern "C"
