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
// Blocks taken while the stream was being captured into a graph.  Each is plain
// device memory owned by the graph being captured (a user object of that
// graph), so the graph holds no allocation node and a replay costs nothing
// extra; when the graph and its executables are gone the block moves to the
// graveyard, freed at the next trim (after a device-wide synchronisation).
// Fallback when the graph cannot take the block: stream-ordered allocation
// (graph nodes), released the same way.
struct GraphBlock {
    int dev;
    void* p;
    size_t bytes;
};
std::unordered_map<void*, int> g_ws_captured;  // block -> 0: graph-owned, 1: allocation node
std::vector<GraphBlock> g_ws_graveyard;
size_t g_ws_graveyard_bytes = 0;
int64_t g_ws_graph_live = 0;  // blocks held by live graphs
size_t g_ws_cached = 0;
constexpr size_t WS_CACHE_MAX = 32ull << 30;     // cached bytes kept at most
constexpr size_t WS_GRAVEYARD_MAX = 1ull << 30;  // released graph blocks kept at most

// User-object destructor (an internal runtime thread; no HIP calls here).
void graph_block_release(void* arg) {
    auto* b = static_cast<GraphBlock*>(arg);
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        g_ws_graveyard.push_back(*b);
        g_ws_graveyard_bytes += b->bytes;
        --g_ws_graph_live;
    }
    delete b;
}

// A block for the graph being captured on s, owned by that graph.
hipError_t graph_block(void** p, size_t bytes, hipStream_t s, int dev) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamGetCaptureInfo_v2(s, &st, nullptr, &g, nullptr, nullptr);
    if (e != hipSuccess || st != hipStreamCaptureStatusActive || !g)
        return e != hipSuccess ? e : hipErrorUnknown;
    // plain allocation while this thread captures: relaxed mode for the call
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    e = hipMalloc(p, bytes);
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    auto* b = new GraphBlock{dev, *p, bytes};
    hipUserObject_t obj = nullptr;
    e = hipUserObjectCreate(&obj, b, graph_block_release, 1, hipUserObjectNoDestructorSync);
    if (e == hipSuccess) {
        {
            std::lock_guard<std::mutex> lk(g_ws_mu);
            ++g_ws_graph_live;
        }
        e = hipGraphRetainUserObject(g, obj, 1, hipGraphUserObjectMove);
        if (e != hipSuccess) (void)hipUserObjectRelease(obj, 1);  // runs the destructor
        return e;  // on failure the block sits in the graveyard
    }
    delete b;
    mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    (void)hipFree(*p);
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    *p = nullptr;
    return e;
}

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
        for (auto& b : g_ws_graveyard) out.push_back({b.dev, b.p});
        g_ws_graveyard.clear();
        g_ws_graveyard_bytes = 0;
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
        int kind = 0;
        if (graph_block(p, bytes, s, dev) != hipSuccess) {
            (void)hipGetLastError();
            kind = 1;
            hipError_t e = hipMallocAsync(p, bytes, s);
            if (e != hipSuccess) {
                *p = nullptr;
                return hip_fail(e, "hipMallocAsync(scratch)");
            }
        }
        std::lock_guard<std::mutex> lk(g_ws_mu);
        g_ws_captured[*p] = kind;
        return GSA_OK;
    }
    bool trim_graveyard = false;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        trim_graveyard = g_ws_graveyard_bytes > WS_GRAVEYARD_MAX;
    }
    if (trim_graveyard) ws_trim();
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
            const int kind = ct->second;
            g_ws_captured.erase(ct);
            if (kind == 1) (void)hipFreeAsync(p, s);  // an allocation node: its free node
            return;                                   // graph-owned: stays with the graph
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

void ws_release_stream(hipStream_t s) {
    std::vector<std::pair<int, void*>> out;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        for (auto it = g_ws_free.begin(); it != g_ws_free.end();) {
            if (it->first.second == s) {
                for (auto& b : it->second) {
                    out.push_back({it->first.first, b.p});
                    g_ws_cached -= b.bytes;
                }
                it = g_ws_free.erase(it);
            } else {
                ++it;
            }
        }
    }
    if (out.empty()) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& q : out) {
        (void)hipSetDevice(q.first);
        (void)hipFree(q.second);
    }
    (void)hipSetDevice(cur);
}

int64_t ws_graph_blocks() {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    return g_ws_graph_live;
}

size_t ws_cached_bytes(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    size_t t = 0;
    for (auto& kv : g_ws_free)
        if (kv.first.second == s)
            for (auto& b : kv.second) t += b.bytes;
    return t;
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
    if (!stream) return GSA_OK;
    hipStream_t s = as_stream(stream);
    // The stream's cached scratch goes with it: wait for its queued work (which
    // may still read those blocks), then free them, so neither the memory leaks
    // nor a later stream that receives the same handle inherits blocks in use.
    GSA_HIP(hipStreamSynchronize(s));
    ws_release_stream(s);
    GSA_HIP(hipStreamDestroy(s));
    return GSA_OK;
}

int64_t gsa_scratch_cached(void* stream) { return (int64_t)ws_cached_bytes(as_stream(stream)); }

int64_t gsa_scratch_graph_blocks(void) { return ws_graph_blocks(); }

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
