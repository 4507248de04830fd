// runtime.hip -- host runtime and lifecycle C ABI (gpufs_init / gpufs_destroy /
// gpufs_run_handler / gpufs_host_start / ...).
//
// Reference counterparts: initializer() and GPUGlobals
// (libgpufs/fs_initializer.cu:116-167, fs_initializer.cu.h:110-199),
// run_gpufs_handler (libgpufs/host_loop.h:559-656).  The reference fixes every
// size at compile time and initialises the pool with a <<<1,1>>> kernel; here
// sizes come from gpufs_config and all state starts as zeroed memory, so init
// is a handful of allocations and memsets.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>
#include "../../include/gpufs_amd.h"
#include "gfs_layout.h"
#include "fstate.h"
#include "host_service.h"

struct gpufs_ctx {
    gpufs_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    gfs_dev host{};          // host mirror of the device struct (pointers and geometry)
    gfs_dev* dev = nullptr;  // device copy
    gfs_host::Service service;
    std::mutex mu;
    std::atomic<int> resident{0};   // resident (serving) kernels running on the context
};

// Resident kernels hold every compute unit, so a device copy issued beside
// them (done by a blit kernel) would wait until they exit.
__attribute__((visibility("hidden"))) void gpufs_resident_add(gpufs_ctx* c, int d) { c->resident += d; }
// Device address of the no-allocation grep scan's state array on `device`
// (k_grep.hip), looked up once per device.
int grep_scratch_setup(int device);
#define REFUSE_IF_RESIDENT(c) do { if ((c)->resident.load() > 0) return -EBUSY; } while (0)

#define HIP_TRY(x)                                                        \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) {                                           \
            fprintf(stderr, "gpufs: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return -EIO;                                                  \
        }                                                                 \
    } while (0)

static bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
static uint32_t pow2_at_least(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return (uint32_t)p;
}
static uint32_t ilog2(uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

static const size_t kPad = 4096;   // tail padding: copies may read whole dwords at a buffer's end

static gfs_host::RingView ring_view(gpufs_ctx* c) {
    gfs_host::RingView rv;
    rv.slots = c->host.slots;
    rv.sq = c->host.sq;
    rv.cq = c->host.cq;
    rv.staging = c->host.staging;
    rv.closeq = c->host.closeq;
    rv.sq_depth = c->host.sq_depth;
    rv.ring_depth = c->host.ring_depth;
    rv.log_page = c->host.log_page;
    return rv;
}

static void free_all(gpufs_ctx* c) {
    gfs_dev& h = c->host;
    (void)hipFree(h.pool);
    (void)hipFree(h.meta);
    (void)hipFree(h.fkey);
    (void)hipFree(h.fbytes);
    (void)hipFree(h.fref);
    (void)hipFree(h.nodes);
    (void)hipFree(h.files);
    (void)hipFree(h.name_index);
    (void)hipFree(h.file_free);
    (void)hipFree(h.slot_free);
    (void)hipHostFree(h.slots);
    (void)hipHostFree(h.sq);
    (void)hipHostFree(h.closeq);
    (void)hipHostFree(h.cq);
    (void)hipHostFree(h.staging);
    (void)hipFree(c->dev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    memset(&h, 0, sizeof(h));
    c->dev = nullptr;
    c->stream = nullptr;
}

extern "C" int gpufs_config_default(gpufs_config* cfg) {
    if (!cfg) return -EINVAL;
    memset(cfg, 0, sizeof(*cfg));
    cfg->cache_bytes = 1ull << 30;
    cfg->page_size = 256u << 10;
    cfg->max_files = 1u << 14;
    cfg->ring_depth = 512;
    unsigned hc = std::thread::hardware_concurrency();
    cfg->host_threads = std::max(1u, std::min(16u, hc ? hc : 4u));
    cfg->tree_nodes = 0;
    cfg->device = 0;
    return 0;
}

static int validate(const gpufs_config* c) {
    if (!is_pow2(c->page_size) || c->page_size < 4096 || c->page_size > (4u << 20)) return -EINVAL;
    if (c->cache_bytes < c->page_size) return -EINVAL;
    if ((c->cache_bytes / c->page_size) >= 0xFFFFFFF0ull) return -EINVAL;
    if (c->max_files == 0 || c->max_files > (1u << 24)) return -EINVAL;
    if (c->ring_depth == 0 || c->ring_depth > 65535) return -EINVAL;
    if (c->host_threads == 0 || c->host_threads > 256) return -EINVAL;
    if (c->flags != 0) return -EINVAL;
    return 0;
}

static int reset_state(gpufs_ctx* c) {
    gfs_dev& h = c->host;
    HIP_TRY(hipMemsetAsync(h.meta, 0, sizeof(uint64_t) * h.n_frames, c->stream));
    HIP_TRY(hipMemsetAsync(h.fkey, 0, sizeof(uint64_t) * h.n_frames, c->stream));
    HIP_TRY(hipMemsetAsync(h.fbytes, 0, sizeof(uint32_t) * h.n_frames, c->stream));
    HIP_TRY(hipMemsetAsync(h.fref, 0, h.n_frames, c->stream));
    HIP_TRY(hipMemsetAsync(h.nodes, 0, (size_t)h.node_cap * GFS_FANOUT * sizeof(uint64_t), c->stream));
    return 0;
}

extern "C" int gpufs_init(const gpufs_config* cfg_in, gpufs_ctx** out) {
    if (!out) return -EINVAL;
    *out = nullptr;
    gpufs_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else gpufs_config_default(&cfg);
    int v = validate(&cfg);
    if (v) return v;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return -ENODEV;
    if (cfg.device < 0 || cfg.device >= ndev) return -ENODEV;
    HIP_TRY(hipSetDevice(cfg.device));
    v = grep_scratch_setup(cfg.device);
    if (v) return v;

    gpufs_ctx* c = new (std::nothrow) gpufs_ctx();
    if (!c) return -ENOMEM;
    c->cfg = cfg;
    c->device = cfg.device;
    gfs_dev& h = c->host;
    h.log_page = ilog2(cfg.page_size);
    h.page_size = cfg.page_size;
    h.n_frames = (uint32_t)(cfg.cache_bytes / cfg.page_size);
    h.max_files = cfg.max_files;
    h.ring_depth = cfg.ring_depth;
    h.sq_depth = pow2_at_least(std::max<uint64_t>((uint64_t)cfg.ring_depth + cfg.max_files, 64));
    h.name_cap = pow2_at_least(2ull * cfg.max_files);
    h.slot_words = (cfg.ring_depth + 63) / 64;
    uint64_t auto_nodes = 4096ull + 2ull * cfg.max_files + 4ull * ((uint64_t)h.n_frames / GFS_FANOUT + 1);
    h.node_cap = cfg.tree_nodes ? cfg.tree_nodes : (uint32_t)std::min<uint64_t>(auto_nodes, 1ull << 26);

    int rc = 0;
    auto fail = [&](int r) {
        free_all(c);
        delete c;
        return r;
    };
#define TRY_ALLOC(x)                                                      \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) {                                           \
            fprintf(stderr, "gpufs: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return fail(-ENOMEM);                                         \
        }                                                                 \
    } while (0)
    TRY_ALLOC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    TRY_ALLOC(hipMalloc(&h.pool, ((size_t)h.n_frames << h.log_page) + kPad));
    TRY_ALLOC(hipMalloc(&h.meta, sizeof(uint64_t) * h.n_frames));
    TRY_ALLOC(hipMalloc(&h.fkey, sizeof(uint64_t) * h.n_frames));
    TRY_ALLOC(hipMalloc(&h.fbytes, sizeof(uint32_t) * h.n_frames));
    TRY_ALLOC(hipMalloc(&h.fref, h.n_frames));
    TRY_ALLOC(hipMalloc(&h.nodes, (size_t)h.node_cap * GFS_FANOUT * sizeof(uint64_t)));
    TRY_ALLOC(hipMalloc(&h.files, sizeof(gfs_file) * (size_t)h.max_files));
    TRY_ALLOC(hipMalloc(&h.name_index, sizeof(uint32_t) * (size_t)h.name_cap));
    TRY_ALLOC(hipMalloc(&h.file_free, sizeof(uint32_t) * (size_t)h.max_files));
    TRY_ALLOC(hipMalloc(&h.slot_free, sizeof(uint64_t) * h.slot_words * GFS_SLOT_WORD_STRIDE));
    const unsigned hflags = hipHostMallocMapped | hipHostMallocCoherent;
    TRY_ALLOC(hipHostMalloc((void**)&h.slots, sizeof(gfs_rpc_slot) * h.ring_depth, hflags));
    TRY_ALLOC(hipHostMalloc((void**)&h.sq, sizeof(uint64_t) * h.sq_depth, hflags));
    TRY_ALLOC(hipHostMalloc((void**)&h.closeq, sizeof(int64_t) * h.sq_depth, hflags));
    TRY_ALLOC(hipHostMalloc((void**)&h.cq, sizeof(uint64_t) * h.ring_depth, hflags));
    TRY_ALLOC(hipHostMalloc((void**)&h.staging, ((size_t)h.ring_depth << h.log_page) + kPad, hflags));
    memset(h.slots, 0, sizeof(gfs_rpc_slot) * h.ring_depth);
    memset(h.sq, 0xFF, sizeof(uint64_t) * h.sq_depth);
    memset(h.cq, 0, sizeof(uint64_t) * h.ring_depth);
    {
        std::vector<uint64_t> bits((size_t)h.slot_words * GFS_SLOT_WORD_STRIDE, 0);
        for (uint32_t s = 0; s < h.ring_depth; ++s) bits[(size_t)(s >> 6) * GFS_SLOT_WORD_STRIDE] |= 1ull << (s & 63);
        TRY_ALLOC(hipMemcpy(h.slot_free, bits.data(), bits.size() * 8, hipMemcpyHostToDevice));
    }
    TRY_ALLOC(hipMemset(h.files, 0, sizeof(gfs_file) * (size_t)h.max_files));
    TRY_ALLOC(hipMemset(h.name_index, 0, sizeof(uint32_t) * (size_t)h.name_cap));
    rc = reset_state(c);
    if (rc) return fail(rc);
    h.fresh_next = 0;
    h.clock_hand = 0;
    h.node_next = 1;
    h.files_next = 0;
    h.sq_tail = 0;
    h.free_taken = 0;
    h.free_n = 0;
    h.reclaim_hand = 0;
    h.open_gate = 0;
    h.node_free = 0;
    h.dirty_frames = 0;
    int wall_khz = 100000;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || wall_khz <= 0)
        wall_khz = 100000;
    const char* tmo = getenv("GPUFS_WAIT_TIMEOUT_S");
    double tsec = tmo ? atof(tmo) : 30.0;
    if (tsec <= 0) tsec = 30.0;
    h.wait_ticks = (uint64_t)(tsec * 1000.0 * wall_khz);
    const char* tune = getenv("GPUFS_TUNE");
    // default: per-page barrier on (measured faster for the sequential warm read)
    h.tune_flags = tune ? (uint32_t)strtoul(tune, nullptr, 0) : GFS_TUNE_PAGE_SYNC;
    memset(&h.ctr, 0, sizeof(h.ctr));
    TRY_ALLOC(hipMalloc(&c->dev, sizeof(gfs_dev)));
    TRY_ALLOC(hipMemcpy(c->dev, &h, sizeof(gfs_dev), hipMemcpyHostToDevice));
    TRY_ALLOC(hipStreamSynchronize(c->stream));
#undef TRY_ALLOC
    *out = c;
    return 0;
}

extern "C" int gpufs_destroy(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    // a resident server's workgroups wait for requests until gpufs_serve_stop
    REFUSE_IF_RESIDENT(c);
    c->service.stop();
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return 0;
}

extern "C" void* gpufs_stream(gpufs_ctx* c) { return c ? (void*)c->stream : nullptr; }
extern "C" void* gpufs_device_state(gpufs_ctx* c) { return c ? (void*)c->dev : nullptr; }
extern "C" int gpufs_get_config(gpufs_ctx* c, gpufs_config* out) {
    if (!c || !out) return -EINVAL;
    *out = c->cfg;
    return 0;
}

extern "C" int gpufs_host_start(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->service.start(ring_view(c), c->cfg.host_threads);
    return 0;
}
extern "C" int gpufs_host_kick(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    c->service.kick();
    return 0;
}
extern "C" int gpufs_host_stop(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    std::lock_guard<std::mutex> g(c->mu);
    c->service.stop();
    return 0;
}

// Serve RPCs until the context stream is idle (libgpufs/host_loop.h:339-344
// semantics), then return the stream's status.
extern "C" int gpufs_run_handler(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    bool started_here = false;
    {
        std::lock_guard<std::mutex> g(c->mu);
        if (!c->service.running()) {
            c->service.start(ring_view(c), c->cfg.host_threads);
            started_here = true;
        }
    }
    (void)hipSetDevice(c->device);
    hipError_t e;
    while ((e = hipStreamQuery(c->stream)) == hipErrorNotReady) std::this_thread::sleep_for(std::chrono::microseconds(20));
    if (started_here) {
        std::lock_guard<std::mutex> g(c->mu);
        c->service.stop();
    }
    if (e != hipSuccess) {
        fprintf(stderr, "gpufs: kernel stream error: %s\n", hipGetErrorString(e));
        return -EIO;
    }
    return 0;
}

__global__ void gfs_reset_files_kernel(gfs_file* files, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) files[i].root = 0;
}

extern "C" int gpufs_cache_drop(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    REFUSE_IF_RESIDENT(c);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    gfs_dev cur;
    HIP_TRY(hipMemcpy(&cur, c->dev, sizeof(gfs_dev), hipMemcpyDeviceToHost));
    gfs_dev& h = c->host;
    // refuse to discard dirty pages
    std::vector<uint64_t> meta(h.n_frames);
    HIP_TRY(hipMemcpy(meta.data(), h.meta, 8ull * h.n_frames, hipMemcpyDeviceToHost));
    for (uint64_t m : meta) {
        if (m & (GFS_FL_DIRTY | GFS_FL_WB)) return -EBUSY;
        if (m & GFS_REF_MASK) return -EBUSY;
    }
    uint32_t nf = (uint32_t)std::min<unsigned long long>(cur.files_next, h.max_files);
    if (nf) {
        hipLaunchKernelGGL(gfs_reset_files_kernel, dim3((nf + 255) / 256), dim3(256), 0, c->stream, h.files, nf);
        HIP_TRY(hipGetLastError());
    }
    int rc = reset_state(c);
    if (rc) return rc;
    cur.fresh_next = 0;
    cur.clock_hand = 0;
    cur.node_next = 1;
    cur.node_free = 0;   // the nodes were zeroed: no tree and no free list remains
    cur.dirty_frames = 0;
    HIP_TRY(hipMemcpyAsync(c->dev, &cur, sizeof(gfs_dev), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int gpufs_stats_get(gpufs_ctx* c, gpufs_stats* o) {
    if (!c || !o) return -EINVAL;
    REFUSE_IF_RESIDENT(c);
    HIP_TRY(hipSetDevice(c->device));
    gfs_dev cur;
    HIP_TRY(hipMemcpy(&cur, c->dev, sizeof(gfs_dev), hipMemcpyDeviceToHost));
    memset(o, 0, sizeof(*o));
    o->page_hits = cur.ctr.page_hits;
    o->page_misses = cur.ctr.page_misses;
    o->page_waits = cur.ctr.page_waits;
    o->evictions = cur.ctr.evictions;
    o->alloc_fail = cur.ctr.alloc_fail;
    o->rpc_reads = cur.ctr.rpc_reads;
    o->rpc_writes = cur.ctr.rpc_writes;
    o->rpc_opens = cur.ctr.rpc_opens;
    o->rpc_closes = cur.ctr.rpc_closes;
    o->bytes_filled = cur.ctr.bytes_filled;
    o->bytes_written = cur.ctr.bytes_written;
    o->nodes_used = std::min<unsigned long long>(cur.node_next, c->host.node_cap) - 1;
    {
        const uint64_t nf = std::min<unsigned long long>(cur.files_next, c->host.max_files);
        const uint64_t left = cur.free_n > cur.free_taken ? cur.free_n - cur.free_taken : 0;
        o->files_used = nf - left;
    }
    auto& hc = c->service.counters();
    o->host_requests = hc.requests.load();
    o->host_bytes_read = hc.bytes_read.load();
    o->host_bytes_written = hc.bytes_written.load();
    o->frames = c->host.n_frames;
    std::vector<uint64_t> meta(c->host.n_frames);
    HIP_TRY(hipMemcpy(meta.data(), c->host.meta, 8ull * meta.size(), hipMemcpyDeviceToHost));
    uint64_t valid = 0;
    for (uint64_t m : meta) valid += ((m >> GFS_ST_SHIFT) & 0xF) == GFS_F_VALID;
    o->frames_valid = valid;
    o->timeouts = cur.ctr.timeouts;
    o->host_busy_ns = hc.busy_ns.load();
    o->vm_errors = cur.ctr.vm_errors;
    o->slots_lost = cur.ctr.slots_lost;
    o->files_reclaimed = cur.ctr.files_reclaimed;
    o->nodes_recycled = cur.ctr.nodes_recycled;
    o->spec_parts = cur.ctr.spec_parts;
    o->spec_rescans = cur.ctr.spec_rescans;
    o->spec_changed = cur.ctr.spec_changed;
    return 0;
}

extern "C" int gpufs_stats_reset(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    REFUSE_IF_RESIDENT(c);
    HIP_TRY(hipSetDevice(c->device));
    gfs_counters z;
    memset(&z, 0, sizeof(z));
    HIP_TRY(hipMemcpy((char*)c->dev + offsetof(gfs_dev, ctr), &z, sizeof(z), hipMemcpyHostToDevice));
    return 0;
}

// Host-side audit of the page cache (idle context only).
extern "C" int gpufs_check_invariants(gpufs_ctx* c) {
    if (!c) return -EINVAL;
    REFUSE_IF_RESIDENT(c);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    gfs_dev cur;
    HIP_TRY(hipMemcpy(&cur, c->dev, sizeof(gfs_dev), hipMemcpyDeviceToHost));
    const gfs_dev& h = c->host;
    std::vector<uint64_t> meta(h.n_frames), fkey(h.n_frames);
    HIP_TRY(hipMemcpy(meta.data(), h.meta, 8ull * h.n_frames, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(fkey.data(), h.fkey, 8ull * h.n_frames, hipMemcpyDeviceToHost));
    uint64_t nn = std::min<uint64_t>(cur.node_next, h.node_cap);
    std::vector<uint64_t> nodes(nn * GFS_FANOUT);
    if (nn) HIP_TRY(hipMemcpy(nodes.data(), h.nodes, nodes.size() * 8, hipMemcpyDeviceToHost));
    uint32_t nf = (uint32_t)std::min<unsigned long long>(cur.files_next, h.max_files);
    std::vector<gfs_file> files(nf);
    if (nf) HIP_TRY(hipMemcpy(files.data(), h.files, sizeof(gfs_file) * nf, hipMemcpyDeviceToHost));
    {
        // every ring slot is free but those quarantined after a timed-out
        // request (the host can still write their staging pages)
        std::vector<uint64_t> bits((size_t)h.slot_words * GFS_SLOT_WORD_STRIDE);
        HIP_TRY(hipMemcpy(bits.data(), h.slot_free, bits.size() * 8, hipMemcpyDeviceToHost));
        uint64_t nfree = 0;
        for (uint32_t w = 0; w < h.slot_words; ++w) nfree += (uint64_t)__builtin_popcountll(bits[(size_t)w * GFS_SLOT_WORD_STRIDE]);
        if (nfree + cur.ctr.slots_lost != h.ring_depth) return -1010;
    }
    for (uint32_t f = 0; f < h.n_frames; ++f)
        if (meta[f] & GFS_REF_MASK) return -1001;   // a frame is still pinned
    for (uint32_t f = 0; f < h.n_frames; ++f)
        if (meta[f] & GFS_FL_WB) return -1015;   // a write-back never ended
    for (uint32_t f = 0; f < h.n_frames; ++f)
        if (((meta[f] >> GFS_ST_SHIFT) & 0xF) == GFS_F_FILLING) return -1002;   // a fill never finished
    {
        long long nd = 0;
        for (uint32_t f = 0; f < h.n_frames; ++f) nd += (meta[f] & GFS_FL_DIRTY) ? 1 : 0;
        if (nd != cur.dirty_frames) return -1014;
    }
    std::vector<uint8_t> seen(h.n_frames, 0);
    std::vector<uint8_t> nseen(nn, 0);   // nodes in a tree or on the free list
    for (uint32_t fi = 0; fi < nf; ++fi) {
        uint64_t rw = files[fi].root;
        if (!rw) continue;
        uint32_t height = (uint32_t)(rw >> 40);
        uint64_t root = rw & ((1ull << 40) - 1);
        if (height == 0 || height > GFS_MAX_HEIGHT || root >= nn) return -1003;
        // iterative DFS: (node, level, page base)
        struct It { uint64_t node; uint32_t lvl; uint64_t base; };
        std::vector<It> stack{{root, height - 1, 0}};
        while (!stack.empty()) {
            It it = stack.back();
            stack.pop_back();
            if (nseen[it.node]) return -1013;   // a node in two trees
            nseen[it.node] = 1;
            for (uint32_t k = 0; k < GFS_FANOUT; ++k) {
                uint64_t v = nodes[it.node * GFS_FANOUT + k];
                if (!v) continue;
                uint64_t page = it.base + ((uint64_t)k << (GFS_FANOUT_LOG * it.lvl));
                if (it.lvl > 0) {
                    if (v >= nn) return -1004;
                    stack.push_back({v, it.lvl - 1, page});
                    continue;
                }
                if (v == GFS_LEAF_RESERVED) return -1009;                // a miss claimed the page and never finished
                uint32_t fr = (uint32_t)(v & 0xFFFFFFFFu) - 1u;
                uint32_t gen = (uint32_t)(v >> 32);
                if (fr >= h.n_frames) return -1005;
                uint64_t m = meta[fr];
                if ((uint32_t)(m >> GFS_GEN_SHIFT) != gen) continue;   // stale leaf: frame was recycled
                if (((m >> GFS_ST_SHIFT) & 0xF) != GFS_F_VALID) return -1006;
                if (fkey[fr] != gfs_key(fi, page)) return -1007;      // live leaf points at a frame of another key
                if (seen[fr]) return -1008;                            // two live leaves share a frame
                seen[fr] = 1;
            }
        }
    }
    // radix-node free list: zeroed nodes, each once, in no tree
    {
        uint64_t n = cur.node_free & ((1ull << 40) - 1);
        for (uint64_t steps = 0; n; ++steps) {
            if (n >= nn || steps >= nn || nseen[n]) return -1013;
            nseen[n] = 1;
            for (uint32_t k = 1; k < GFS_FANOUT; ++k)
                if (nodes[n * GFS_FANOUT + k]) return -1013;
            n = nodes[n * GFS_FANOUT] & ((1ull << 40) - 1);
        }
    }
    // file table and name index: every entry in use is found by probing from
    // its home slot, free entries are in no slot, and nobody is inside the gate
    {
        if (cur.open_gate != 0) return -1011;
        std::vector<uint32_t> index(h.name_cap);
        HIP_TRY(hipMemcpy(index.data(), h.name_index, 4ull * h.name_cap, hipMemcpyDeviceToHost));
        const uint32_t mask = h.name_cap - 1;
        uint64_t in_index = 0, live = 0;
        for (uint32_t i = 0; i < h.name_cap; ++i) {
            const uint32_t v = index[i];
            if (!v) continue;
            if (v == GFS_NAME_RESERVED || v > nf) return -1012;
            if (fst_state(files[v - 1].st) == GFS_FS_FREE) return -1012;
            ++in_index;
        }
        for (uint32_t fi = 0; fi < nf; ++fi) {
            const uint32_t st = fst_state(files[fi].st);
            if (st == GFS_FS_FREE) {
                if (files[fi].root) return -1012;
                continue;
            }
            if (st == GFS_FS_EMPTY) continue;
            ++live;
            uint32_t pos = files[fi].name_hash & mask;
            bool found = false;
            for (uint32_t probe = 0; probe <= mask; ++probe, pos = (pos + 1) & mask) {
                if (index[pos] == fi + 1) {
                    found = true;
                    break;
                }
                if (index[pos] == 0) break;
            }
            if (!found) return -1012;
        }
        if (live != in_index) return -1012;
        const uint32_t fn = cur.free_n;
        if (fn > h.max_files) return -1012;
        std::vector<uint32_t> fl(fn);
        if (fn) HIP_TRY(hipMemcpy(fl.data(), h.file_free, 4ull * fn, hipMemcpyDeviceToHost));
        std::vector<uint8_t> listed(nf, 0);
        for (uint64_t i = std::min<uint64_t>(cur.free_taken, fn); i < fn; ++i) {
            const uint32_t m = fl[i];
            if (m >= nf || listed[m] || fst_state(files[m].st) != GFS_FS_FREE) return -1012;
            listed[m] = 1;
        }
    }
    return 0;
}

extern "C" const char* gpufs_strerror(int code) {
    switch (code) {
    case 0: return "success";
    case -1: return "file table full";
    case -2: return "host open failed";
    case -1001: return "invariant: frame still pinned";
    case -1002: return "invariant: frame left FILLING";
    case -1003: return "invariant: bad radix root";
    case -1004: return "invariant: bad radix child";
    case -1005: return "invariant: leaf frame out of range";
    case -1006: return "invariant: live leaf frame not VALID";
    case -1007: return "invariant: frame key mismatch";
    case -1008: return "invariant: frame shared by two leaves";
    case -1009: return "invariant: leaf left reserved";
    case -1010: return "invariant: ring slot not returned";
    case -1011: return "invariant: open gate not clear";
    case -1012: return "invariant: file table and name index disagree";
    case -1013: return "invariant: radix node lost, shared or on the free list twice";
    case -1014: return "invariant: dirty-frame count differs from the frames marked dirty";
    case -1015: return "invariant: a write-back never ended";
    default: break;
    }
    if (code < 0 && code > -4096) return strerror(-code);
    return "unknown error";
}
