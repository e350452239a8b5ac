// Context management, error reporting and workspace for libngcasa_hip.
#include "ngc_internal.hpp"

namespace ngc {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int workspace(ngc_context *ctx, const char *name, size_t bytes, void **out) {
    if (bytes == 0) bytes = 256;
    Buffer &b = ctx->ws[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // the buffer may still be read by queued work on the last stream
            if (ctx->has_last_stream) NGC_HIP(hipStreamSynchronize(ctx->last_stream));
            NGC_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t grow = bytes + bytes / 8;
        hipError_t e = hipMalloc(&b.ptr, grow);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            e = hipMalloc(&b.ptr, bytes);
            grow = bytes;
        }
        if (e != hipSuccess) {
            set_error(std::string("workspace '") + name + "' (" +
                      std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
            b.ptr = nullptr;
            return NGC_ERR_HIP;
        }
        b.bytes = grow;
    }
    *out = b.ptr;
    return NGC_OK;
}

int enter_stream(ngc_context *ctx, hipStream_t stream) {
    NGC_HIP(hipSetDevice(ctx->device));
    if (ctx->has_last_stream && ctx->last_stream != stream)
        NGC_HIP(hipStreamSynchronize(ctx->last_stream));
    ctx->last_stream = stream;
    ctx->has_last_stream = true;
    return NGC_OK;
}

int upload(ngc_context *ctx, hipStream_t stream, const char *name,
           const void *host, size_t bytes, void **dev) {
    NGC_TRY(workspace(ctx, name, bytes, dev));
    // pageable source: the copy is staged before the call returns, so the
    // host buffer may be released afterwards
    NGC_HIP(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, stream));
    return NGC_OK;
}

}  // namespace ngc

extern "C" {

const char *ngc_last_error(void) { return ngc::g_last_error.c_str(); }

const char *ngc_version(void) { return "ngcasa-mi355x 0.1 (gfx950)"; }

int ngc_context_create(int device, ngc_context **ctx) {
    if (!ctx) {
        ngc::set_error("ctx is NULL");
        return NGC_ERR_ARG;
    }
    int n = 0;
    NGC_HIP(hipGetDeviceCount(&n));
    NGC_CHECK(device >= 0 && device < n, NGC_ERR_ARG, "device index out of range");
    NGC_HIP(hipSetDevice(device));
    ngc_context *c = new ngc_context();
    c->device = device;
    hipError_t e = hipMalloc(&c->d_error, sizeof(int));
    if (e != hipSuccess) {
        delete c;
        ngc::set_error(std::string("hipMalloc(error flag): ") + hipGetErrorString(e));
        return NGC_ERR_HIP;
    }
    *ctx = c;
    return NGC_OK;
}

int ngc_context_destroy(ngc_context *ctx) {
    if (!ctx) return NGC_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    ngc::fft_destroy_plans(ctx);
    for (auto &kv : ctx->ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->d_error) (void)hipFree(ctx->d_error);
    delete ctx;
    return NGC_OK;
}

}  // extern "C"
