// gsa_runtime.hip — error reporting, device set-up, memory/stream helpers.
#include "gsa_common.h"
#include <cstdarg>
#include <cstdio>
#include <mutex>

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
        // Keep freed scratch memory cached in every device's default pool so
        // stream-ordered scratch allocations cost no driver round trip.
        int cur = 0;
        (void)hipGetDevice(&cur);
        for (int d = 0; d < n; ++d) {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, d) == hipSuccess) {
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            }
        }
        (void)hipSetDevice(cur);
    });
    return g_init_status;
}

int ws_alloc(void** p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMallocAsync(p, bytes, s);
    if (e != hipSuccess) { *p = nullptr; return hip_fail(e, "hipMallocAsync(scratch)"); }
    return GSA_OK;
}

void ws_free(void* p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
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
