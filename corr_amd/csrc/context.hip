// Context, status and device-memory entry points of the C-ABI.
#include "common.hpp"
#include <cstring>

namespace corr {

int set_error(corr_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int scratch(corr_ctx* ctx, const char* name, size_t bytes, void** out) {
    corr_ctx::Buf& b = ctx->scratch[name];
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (ctx->scratch_limit) {
            size_t held = 0;
            for (const auto& kv : ctx->scratch) held += kv.second.bytes;
            if (held - b.bytes + bytes > ctx->scratch_limit)
                return set_error(ctx, CORR_ENOMEM,
                                 std::string("scratch '") + name + "' of " + std::to_string(bytes) +
                                     " bytes would exceed the context's scratch limit of " +
                                     std::to_string(ctx->scratch_limit) + " bytes (" + std::to_string(held) +
                                     " held)");
        }
        if (b.p) {
            // The buffer may still be in use by queued work.
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return set_error(ctx, CORR_EHIP, std::string("sync: ") + hipGetErrorString(e));
            (void)hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(ctx, CORR_ENOMEM,
                             std::string("device allocation of ") + std::to_string(bytes) +
                                 " bytes for scratch '" + name + "' failed");
        }
        b.p = p;
        b.bytes = bytes;
    }
    *out = b.p;
    return CORR_OK;
}

}  // namespace corr

extern "C" {

int corr_version(void) { return 100; }

int corr_ctx_create(int device, corr_ctx** out) {
    if (!out) return CORR_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { (void)hipGetLastError(); return CORR_ENODEV; }
    if (device < 0 || device >= n) return CORR_EINVAL;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CORR_EHIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CORR_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return CORR_EHIP;
    corr_ctx* c = new corr_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CORR_EHIP;
    }
    c->stream = c->own_stream;
    // Device address of the allocation-free entry point's static scratch on
    // this device (current here); a failed lookup only disables that entry.
    (void)corr::noalloc_register_device(device);
    *out = c;
    return CORR_OK;
}

int corr_ctx_trim(corr_ctx* ctx) {
    CORR_CHECK_CTX(ctx);
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->scratch)
        if (kv.second.p) (void)hipFree(kv.second.p);
    ctx->scratch.clear();
    return CORR_OK;
}

int corr_blas_release(corr_ctx* ctx);  // metric.hip

int corr_ctx_destroy(corr_ctx* ctx) {
    if (!ctx) return CORR_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& t : ctx->pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto& kv : ctx->scratch)
        if (kv.second.p) (void)hipFree(kv.second.p);
    corr_blas_release(ctx);
    if (ctx->aux_stream) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        (void)hipStreamDestroy(ctx->aux_stream);
    }
    if (ctx->part_stream) {
        (void)hipStreamSynchronize(ctx->part_stream);
        (void)hipStreamDestroy(ctx->part_stream);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_part) (void)hipEventDestroy(ctx->ev_part);
    if (ctx->ev_switch) (void)hipEventDestroy(ctx->ev_switch);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CORR_OK;
}

int corr_ctx_set_stream(corr_ctx* ctx, void* hip_stream) {
    CORR_CHECK_CTX(ctx);
    const hipStream_t next = (hipStream_t)hip_stream;
    if (next == ctx->stream) return CORR_OK;
    // Work still queued on the old stream may use the context's scratch:
    // the new stream waits for it before anything enqueued from now on.
    if (!ctx->ev_switch)
        CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_switch, hipEventDisableTiming));
    CORR_HIP(ctx, hipEventRecord(ctx->ev_switch, ctx->stream));
    CORR_HIP(ctx, hipStreamWaitEvent(next, ctx->ev_switch, 0));
    ctx->stream = next;
    return CORR_OK;
}

int corr_ctx_set_scratch_limit(corr_ctx* ctx, int64_t bytes) {
    CORR_CHECK_CTX(ctx);
    CORR_REQUIRE(ctx, bytes >= 0, "corr_ctx_set_scratch_limit: negative limit");
    ctx->scratch_limit = (size_t)bytes;
    return CORR_OK;
}

int corr_ctx_synchronize(corr_ctx* ctx) {
    CORR_CHECK_CTX(ctx);
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

const char* corr_ctx_last_error(const corr_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int corr_ctx_enable_timing(corr_ctx* ctx, int on) {
    CORR_CHECK_CTX(ctx);
    ctx->timing = on != 0;
    return CORR_OK;
}

static int drain_timers(corr_ctx* ctx) {
    if (ctx->pending.empty()) return CORR_OK;
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& t : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto& acc = ctx->kernel_ms[t.name];
            acc.first += ms;
            acc.second += 1;
        }
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    ctx->pending.clear();
    return CORR_OK;
}

int corr_ctx_kernel_time(corr_ctx* ctx, const char* name, double* h_total_ms, int64_t* h_launches) {
    CORR_CHECK_CTX(ctx);
    CORR_TRY(drain_timers(ctx));
    if (!name) {
        ctx->kernel_ms.clear();
        return CORR_OK;
    }
    auto it = ctx->kernel_ms.find(name);
    if (h_total_ms) *h_total_ms = it == ctx->kernel_ms.end() ? 0.0 : it->second.first;
    if (h_launches) *h_launches = it == ctx->kernel_ms.end() ? 0 : it->second.second;
    return CORR_OK;
}

int corr_dmalloc(corr_ctx* ctx, void** d_ptr, int64_t bytes) {
    CORR_CHECK_CTX(ctx);
    CORR_REQUIRE(ctx, d_ptr && bytes >= 0, "corr_dmalloc: bad arguments");
    *d_ptr = nullptr;
    if (hipMalloc(d_ptr, bytes > 0 ? (size_t)bytes : 256) != hipSuccess) {
        (void)hipGetLastError();
        return corr::set_error(ctx, CORR_ENOMEM, "corr_dmalloc: " + std::to_string(bytes) + " bytes");
    }
    return CORR_OK;
}

int corr_dfree(corr_ctx* ctx, void* d_ptr) {
    CORR_CHECK_CTX(ctx);
    if (d_ptr) CORR_HIP(ctx, hipFree(d_ptr));
    return CORR_OK;
}

int corr_h2d(corr_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes) {
    CORR_CHECK_CTX(ctx);
    CORR_REQUIRE(ctx, bytes >= 0 && (bytes == 0 || (d_dst && h_src)), "corr_h2d: bad arguments");
    if (bytes) {
        CORR_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CORR_OK;
}

int corr_d2h(corr_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes) {
    CORR_CHECK_CTX(ctx);
    CORR_REQUIRE(ctx, bytes >= 0 && (bytes == 0 || (h_dst && d_src)), "corr_d2h: bad arguments");
    if (bytes) {
        CORR_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CORR_OK;
}

}  // extern "C"
