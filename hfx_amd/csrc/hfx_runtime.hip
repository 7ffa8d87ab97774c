/*
 * hfx_runtime.hip -- context, error reporting and memory/stream helpers.
 *
 * Replaces the parts of the Haskell CUDA binding HFX's search path uses
 * (packages/cuda-0.4.1.0: Runtime/Marshal.chs malloc/free/peek/poke/memset,
 * Runtime/Stream.chs create/destroy/block, Driver/Marshal.chs getMemInfo).
 */
#include "hfx_common.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

static thread_local char g_err[1024] = "";

void hfx_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *hfx_last_error(void) { return g_err; }

#define HFX_SCRATCH_SLOTS 16

struct hfx_ctx {
    int device;
    int num_cus;
    void *scratch[HFX_SCRATCH_SLOTS];
    size_t scratch_bytes[HFX_SCRATCH_SLOTS];
    void *pinned;
    size_t pinned_bytes;
};

extern "C" int hfx_ctx_create(int device, hfx_ctx **out)
{
    HFX_REQUIRE(out != nullptr, HFX_ERR_ARG, "hfx_ctx_create: null out");
    HFX_CHECK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    HFX_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    hfx_ctx *c = (hfx_ctx *)calloc(1, sizeof(hfx_ctx));
    HFX_REQUIRE(c != nullptr, HFX_ERR_NOMEM, "hfx_ctx_create: out of host memory");
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    *out = c;
    return HFX_OK;
}

extern "C" int hfx_ctx_destroy(hfx_ctx *ctx)
{
    if (!ctx) return HFX_OK;
    for (int i = 0; i < HFX_SCRATCH_SLOTS; ++i)
        if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    free(ctx);
    return HFX_OK;
}

int hfx_ctx_scratch(hfx_ctx *ctx, int slot, size_t bytes, void **out)
{
    HFX_REQUIRE(slot >= 0 && slot < HFX_SCRATCH_SLOTS, HFX_ERR_ARG, "scratch slot %d", slot);
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            HFX_CHECK_HIP(hipDeviceSynchronize());
            HFX_CHECK_HIP(hipFree(ctx->scratch[slot]));
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        size_t sz = bytes + bytes / 4;
        sz = (sz + 255) & ~(size_t)255;
        hipError_t e = hipMalloc(&ctx->scratch[slot], sz);
        if (e != hipSuccess) {
            ctx->scratch[slot] = nullptr;
            hfx_set_error("scratch allocation of %zu bytes failed: %s", sz, hipGetErrorString(e));
            return HFX_ERR_NOMEM;
        }
        ctx->scratch_bytes[slot] = sz;
    }
    *out = ctx->scratch[slot];
    return HFX_OK;
}

int hfx_ctx_pinned(hfx_ctx *ctx, size_t bytes, void **out)
{
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) {
            HFX_CHECK_HIP(hipDeviceSynchronize());
            HFX_CHECK_HIP(hipHostFree(ctx->pinned));
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
        }
        size_t sz = (bytes + 4095) & ~(size_t)4095;
        HFX_CHECK_HIP(hipHostMalloc(&ctx->pinned, sz, hipHostMallocDefault));
        ctx->pinned_bytes = sz;
    }
    *out = ctx->pinned;
    return HFX_OK;
}

int hfx_ctx_num_cus(hfx_ctx *ctx) { return ctx->num_cus; }

/* The reference ABI has no context argument: those entry points share one
 * per-process context on the current device, created on first use. */
hfx_ctx *hfx_default_ctx(void)
{
    static std::once_flag once;
    static hfx_ctx *ctx = nullptr;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hfx_ctx_create(dev, &ctx) != HFX_OK) ctx = nullptr;
    });
    return ctx;
}

extern "C" int hfx_malloc(void **d_ptr, size_t bytes)
{
    HFX_REQUIRE(d_ptr != nullptr, HFX_ERR_ARG, "hfx_malloc: null out");
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        hfx_set_error("hfx_malloc(%zu): %s", bytes, hipGetErrorString(e));
        return HFX_ERR_NOMEM;
    }
    return HFX_OK;
}

extern "C" int hfx_free(void *d_ptr)
{
    if (d_ptr) HFX_CHECK_HIP(hipFree(d_ptr));
    return HFX_OK;
}

extern "C" int hfx_host_alloc(void **h_ptr, size_t bytes)
{
    HFX_CHECK_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return HFX_OK;
}

extern "C" int hfx_host_free(void *h_ptr)
{
    if (h_ptr) HFX_CHECK_HIP(hipHostFree(h_ptr));
    return HFX_OK;
}

extern "C" int hfx_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream)
{
    HFX_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream)
{
    HFX_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
    HFX_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_memset(void *d_dst, int value, size_t bytes, void *stream)
{
    HFX_CHECK_HIP(hipMemsetAsync(d_dst, value, bytes, hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_stream_create(void **stream)
{
    hipStream_t s;
    HFX_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return HFX_OK;
}

extern "C" int hfx_stream_destroy(void *stream)
{
    if (stream) HFX_CHECK_HIP(hipStreamDestroy(hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_stream_sync(void *stream)
{
    HFX_CHECK_HIP(hipStreamSynchronize(hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_device_sync(void)
{
    HFX_CHECK_HIP(hipDeviceSynchronize());
    return HFX_OK;
}

extern "C" int hfx_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    HFX_CHECK_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return HFX_OK;
}

extern "C" int hfx_device_count(int *count)
{
    HFX_REQUIRE(count != nullptr, HFX_ERR_ARG, "hfx_device_count: null out");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        hfx_set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
        return HFX_ERR_HIP;
    }
    return HFX_OK;
}

/* Events (Runtime/Event.chs of the CUDA binding): kernel timing on a stream. */
extern "C" int hfx_event_create(void **ev)
{
    hipEvent_t e;
    HFX_CHECK_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return HFX_OK;
}

extern "C" int hfx_event_destroy(void *ev)
{
    if (ev) HFX_CHECK_HIP(hipEventDestroy((hipEvent_t)ev));
    return HFX_OK;
}

extern "C" int hfx_event_record(void *ev, void *stream)
{
    HFX_CHECK_HIP(hipEventRecord((hipEvent_t)ev, hfx_stream(stream)));
    return HFX_OK;
}

extern "C" int hfx_event_elapsed_ms(float *ms, void *start, void *stop)
{
    HFX_CHECK_HIP(hipEventSynchronize((hipEvent_t)stop));
    HFX_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return HFX_OK;
}
