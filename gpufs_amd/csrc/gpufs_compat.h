// gpufs_compat.h -- source compatibility for existing GPUfs host drivers.
//
// A reference driver (samples/*/fs.cu) does
//     volatile GPUGlobals* gpuGlobals;
//     initializer(&gpuGlobals);                                     // fs_initializer.cu:116
//     kernel<<<g, b, 0, gpuGlobals->streamMgr->kernelStream>>>(...);  // samples/grep_text/fs.cu:95
//     run_gpufs_handler(gpuGlobals, device);                        // host_loop.h:559
//     delete gpuGlobals;                                            // fs_initializer.cu.h:180-198
// This header keeps exactly those names on top of the C ABI (gpufs_amd.h):
// initializer() creates a context with the default configuration (1 GiB
// cache, 256 KiB pages) on the current device and binds the including
// translation unit's device API to it (other translation units that launch
// device-API kernels call gpufs_bind(gpuGlobals->ctx) themselves);
// run_gpufs_handler() serves the GPU's file requests until the context
// stream is idle; deleting the object destroys the context.  The reference's
// compile-time sizes (fs_constants.h) are gpufs_config fields here: set
// GPUFS_COMPAT_CACHE_BYTES / GPUFS_COMPAT_PAGE_SIZE in the environment to
// change them for a driver built against this header.
#pragma once
#include <cstdio>
#include <cstdlib>
#include "gpufs.h"

struct GpufsStreamMgr {
    hipStream_t kernelStream;
};

struct GPUGlobals {
    gpufs_ctx* ctx = nullptr;
    GpufsStreamMgr* streamMgr = nullptr;
    ~GPUGlobals() {
        if (ctx) gpufs_destroy(ctx);
        delete streamMgr;
    }
};
typedef volatile GPUGlobals* GPUGlobals_ptr;

// Fills *out, or sets it to nullptr and prints the reason (the reference's
// signature has no status to return).
static inline void initializer(GPUGlobals_ptr* out) {
    *out = nullptr;
    gpufs_config cfg;
    gpufs_config_default(&cfg);
    if (const char* v = getenv("GPUFS_COMPAT_CACHE_BYTES")) cfg.cache_bytes = strtoull(v, nullptr, 0);
    if (const char* v = getenv("GPUFS_COMPAT_PAGE_SIZE")) cfg.page_size = (uint32_t)strtoul(v, nullptr, 0);
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) cfg.device = dev;
    GPUGlobals* g = new GPUGlobals();
    int rc = gpufs_init(&cfg, &g->ctx);
    if (rc == 0) rc = gpufs_bind(g->ctx);
    if (rc != 0) {
        fprintf(stderr, "gpufs initializer: %s\n", gpufs_strerror(rc));
        delete g;
        return;
    }
    g->streamMgr = new GpufsStreamMgr{(hipStream_t)gpufs_stream(g->ctx)};
    *out = g;
}

// Serves file requests until the context stream is idle; returns the
// handler's status (the reference returns void and exits on errors).
static inline int run_gpufs_handler(GPUGlobals_ptr g, int gpuid) {
    (void)gpuid;   // the context already names its device
    if (!g) return -EINVAL;
    return gpufs_run_handler(const_cast<GPUGlobals*>(g)->ctx);
}
