// runtime.hip -- library state: status, stream, workspace pool.
#include "common.hpp"
#include <stdio.h>
#include <stdlib.h>
#include <mutex>

namespace tdx {

static thread_local int g_last_status = TDX_OK;
static hipStream_t g_stream = nullptr;

struct Slot {
    void *ptr = nullptr;
    size_t bytes = 0;
};
static Slot g_ws[WS_COUNT];
static std::mutex g_ws_mutex;

int record(int status) {
    g_last_status = status;
    return status;
}

int record_hip(hipError_t e, const char *what, const char *file, int line) {
    if (getenv("TDX_VERBOSE"))
        fprintf(stderr, "tdx: %s failed: %s (%s:%d)\n", what, hipGetErrorString(e), file, line);
    return record(e == hipErrorOutOfMemory ? TDX_ERR_ALLOC : TDX_ERR_HIP);
}

hipStream_t stream() { return g_stream; }

int workspace(int slot, size_t bytes, void **ptr) {
    if (slot < 0 || slot >= WS_COUNT) return record(TDX_ERR_ARG);
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    Slot &s = g_ws[slot];
    if (s.bytes < bytes) {
        if (s.ptr) {
            // the previous buffer may still be read by queued work on the library stream
            TDX_HIP(hipStreamSynchronize(g_stream));
            TDX_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
        TDX_HIP(hipMalloc(&s.ptr, want));
        s.bytes = want;
    }
    *ptr = s.ptr;
    return TDX_OK;
}

void release_workspace() {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    (void)hipStreamSynchronize(g_stream);
    for (auto &s : g_ws) {
        if (s.ptr) (void)hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
}

int device_cus() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
    if (dev >= 64) {
        int c = 0;
        return (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0) ? c
                                                                                                              : 256;
    }
    if (cache[dev] <= 0) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        cache[dev] = c;
    }
    return cache[dev];
}

}  // namespace tdx

extern "C" {

int tdx_init(void) {
    int dev = 0;
    TDX_HIP(hipGetDevice(&dev));
    TDX_HIP(hipFree(nullptr));
    // the no-allocation MTTKRP's scratch address on this device
    int st = tdx::noalloc_scratch(nullptr);
    if (st) return st;
    return tdx::record(TDX_OK);
}

int tdx_set_stream(void *s) {
    tdx::g_stream = (hipStream_t)s;
    return tdx::record(TDX_OK);
}

void *tdx_get_stream(void) { return (void *)tdx::g_stream; }

int tdx_last_status(void) { return tdx::g_last_status; }

const char *tdx_status_string(int st) {
    switch (st) {
        case TDX_OK: return "ok";
        case TDX_ERR_ARG: return "invalid argument";
        case TDX_ERR_ALLOC: return "device allocation failed";
        case TDX_ERR_HIP: return "HIP runtime error";
        case TDX_WARN_SINGULAR: return "singular normal equations: pseudo-inverse used";
        case TDX_ERR_NONFINITE: return "non-finite value detected";
        case TDX_ERR_UNSUPPORTED: return "unsupported shape or dtype";
        case TDX_ERR_TOO_LARGE: return "problem larger than the fixed 64 MiB scratch of the no-allocation entry point";
        case TDX_ERR_IO: return "tensor file I/O failed or the file is not a valid .tdxq tensor";
        default: return "unknown status";
    }
}

int tdx_synchronize(void) {
    TDX_HIP(hipStreamSynchronize(tdx::g_stream));
    return tdx::record(TDX_OK);
}

int tdx_copy_to_host(void *dst, const void *src, size_t bytes) {
    if (!dst || !src) return tdx::record(TDX_ERR_ARG);
    TDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, tdx::g_stream));
    TDX_HIP(hipStreamSynchronize(tdx::g_stream));
    return tdx::record(TDX_OK);
}

int tdx_copy_to_device(void *dst, const void *src, size_t bytes) {
    if (!dst || !src) return tdx::record(TDX_ERR_ARG);
    TDX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, tdx::g_stream));
    TDX_HIP(hipStreamSynchronize(tdx::g_stream));
    return tdx::record(TDX_OK);
}

int tdx_release_workspace(void) {
    tdx::release_cp_plans();
    tdx::release_workspace();
    return tdx::record(TDX_OK);
}

}  // extern "C"
