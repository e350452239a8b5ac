// Context management, error reporting and workspace for libngcasa_hip.
#include <cstring>

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

static hipEvent_t pool_event(ngc_context *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void profile_new_call(ngc_context *ctx, hipStream_t s) {
    if (!ctx->profile) return;
    for (auto &m : ctx->marks) {
        ctx->event_pool.push_back(m.start);
        ctx->event_pool.push_back(m.stop);
    }
    ctx->marks.clear();
    ctx->profile_stream = s;
}

void phase_begin(ngc_context *ctx, hipStream_t s, const char *name) {
    if (!ctx->profile) return;
    ngc_profile_mark m;
    m.name = name;
    m.start = pool_event(ctx);
    m.stop = pool_event(ctx);
    if (m.start) (void)hipEventRecord(m.start, s);
    ctx->marks.push_back(m);
}

void phase_end(ngc_context *ctx, hipStream_t s) {
    if (!ctx->profile || ctx->marks.empty()) return;
    if (ctx->marks.back().stop) (void)hipEventRecord(ctx->marks.back().stop, s);
}

}  // namespace ngc

extern "C" {

int ngc_set_tile_rows(ngc_context *ctx, int tile_rows) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(tile_rows == 0 || tile_rows == 8 || tile_rows == 16 || tile_rows == 32, NGC_ERR_ARG,
              "tile_rows must be 0 (by size), 8, 16 or 32");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tile_rows = tile_rows;
    return NGC_OK;
}

int ngc_set_bin_lanes(ngc_context *ctx, int lanes) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8, NGC_ERR_ARG,
              "lanes must be 1, 2, 4 or 8");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->bin_chan_lanes = lanes;
    return NGC_OK;
}

int ngc_set_tile_waves(ngc_context *ctx, int waves) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(waves == 0 || waves == 1 || waves == 2 || waves == 4, NGC_ERR_ARG,
              "waves must be 0 (default), 1, 2 or 4");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tile_waves = waves;
    return NGC_OK;
}

int ngc_profile_enable(ngc_context *ctx, int enable) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->profile = enable != 0;
    return NGC_OK;
}

int ngc_profile_read(ngc_context *ctx, char *names, int64_t names_len, double *ms, int64_t max) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->marks.empty()) {
        if (names && names_len > 0) names[0] = 0;
        return 0;
    }
    NGC_HIP(hipSetDevice(ctx->device));
    NGC_HIP(hipStreamSynchronize(ctx->profile_stream));
    std::string all;
    int64_t n = 0;
    for (auto &m : ctx->marks) {
        if (n >= max) break;
        float t = 0.f;
        if (m.start && m.stop) NGC_HIP(hipEventElapsedTime(&t, m.start, m.stop));
        if (ms) ms[n] = (double)t;
        if (!all.empty()) all += ";";
        all += m.name;
        n++;
    }
    if (names && names_len > 0) {
        size_t k = std::min<size_t>(all.size(), (size_t)names_len - 1);
        memcpy(names, all.data(), k);
        names[k] = 0;
    }
    return (int)n;
}


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
    NGC_TRY(ngc::noalloc_scratch_setup(device));
    ngc_context *c = new ngc_context();
    c->device = device;
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        n_cu > 0)
        c->n_cu = n_cu;
    hipError_t e = hipMalloc(&c->d_error, 4 * sizeof(int));
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
    for (auto &m : ctx->marks) {
        if (m.start) (void)hipEventDestroy(m.start);
        if (m.stop) (void)hipEventDestroy(m.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &kv : ctx->ws)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->d_error) (void)hipFree(ctx->d_error);
    delete ctx;
    return NGC_OK;
}

}  // extern "C"
