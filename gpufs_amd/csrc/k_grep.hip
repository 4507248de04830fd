// k_grep.hip -- grep_text-style whole-word counting over files read with gread.
//
// Scan semantics follow samples/grep_text/grep_text_gpu.cu:
//   * word characters: alpha() = [A-Za-z_] (:79);
//   * a query word q (32-byte record, C string up to the first NUL) matches a
//     maximal run of word characters equal to q (match_string, :139-160;
//     an empty q never matches, :144);
//   * chunk loop (:298-318): the window at `cursor` is chunk+32 bytes (or the
//     rest of the file when fewer than chunk+32 remain: last window); a NUL is
//     placed after it; a non-last window is scanned over chunk+ov+1 bytes where
//     ov = number of word characters (max 32) starting at byte `chunk`, and the
//     cursor advances by chunk+ov; the last window is scanned over all its bytes
//     plus the NUL and the cursor advances by chunk, so a last window longer
//     than `chunk` is followed by one more window over its final bytes.
// The reference gives each thread one query word and walks the window once per
// word (O(words x bytes)).  Here the window sits in LDS, each thread tokenises
// a slice of it, and every token of at most 32 bytes is looked up in an LDS
// hash table of the dictionary (O(bytes)); equal query words share a counter.
#include "grep_scan.h"

#include <atomic>
#include <cstring>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

// Workgroup b scans part b / n_files of file b % n_files (grep_file_split): a
// launch over fewer files than compute units still fills the device.  Part 0
// of every file comes first in dispatch order (it decides the file's split
// and the other parts wait for it), and consecutive files go to different
// XCDs (workgroups are dealt to the 8 XCDs round robin), so the files a cold
// launch reads with part 0 alone spread over the whole device.
__global__ __launch_bounds__(1024) void k_grep(const char* names, uint32_t stride, uint32_t n_files,
                                               const char* dict, uint32_t nw, uint32_t chunk, uint32_t* counts_out,
                                               int32_t* status, uint32_t parts, uint64_t min_part,
                                               grep_split_state* fstate, uint32_t epoch) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    const uint32_t f = blockIdx.x % n_files, j = blockIdx.x / n_files;
    // Part 0 builds the dictionary once it has published its decision, which
    // the other parts wait for; they build theirs meanwhile.
    GFS_TRACE(40);
    if (j != 0) grep_build_dict(S, dict, nw);
    GFS_TRACE(41);
    [[clang::always_inline]] grep_file_split(S, names + (size_t)f * stride, nw, chunk, counts_out + (size_t)f * nw,
                                             status + f, j, parts, min_part, fstate ? fstate + f : nullptr, parts,
                                             epoch, false, true,
                                             [&](uint32_t) {
                                                 if (j == 0) grep_build_dict(S, dict, nw);
                                             },
                                             [] {});
}

// ---------------------------------------------------------------------------
// Batched form: one persistent kernel over many files of uneven size.  Each
// workgroup builds the dictionary table once and then takes work items until
// none is left: a new file from the file queue (a device counter), or an extra
// part of a file another workgroup has opened.  The workgroup that takes a
// file is its part 0: it opens it, learns its part count K_f (the split of
// grep_file_split) and publishes parts 1..K_f-1 in the part ring, where other
// workgroups take them, so a large file is scanned by several workgroups while
// small files take one each.  Work is over when the file queue is exhausted,
// no file taken is still publishing its parts, and the ring is empty.
struct grep_batch_ctl {
    uint32_t next_file;   // file queue: next file index
    uint32_t expanding;   // files taken whose extra parts are not published yet
    uint32_t head, tail;  // part ring: items [head, tail) reserved or published
    uint32_t pad[60];     // to 256 bytes
};
// Workspace: ctl | per-file split state | part ring (8-byte items:
// ready << 63 | file << 24 | K_f << 8 | part).
static uint64_t batch_ws_bytes(uint32_t n_files, uint32_t parts) {
    return sizeof(grep_batch_ctl) + sizeof(grep_split_state) * (uint64_t)n_files +
           8ull * (uint64_t)n_files * (parts > 1 ? parts - 1 : 1);
}

#define BATCH_STOP 0xFFFFFFFFFFFFFFFFull

__global__ __launch_bounds__(1024) void k_grep_batched(const char* names, uint32_t stride, uint32_t nfiles,
                                                       const char* dict, uint32_t nw, uint32_t chunk,
                                                       uint32_t* counts_out, int32_t* status, uint8_t* ws,
                                                       uint32_t parts, uint64_t min_part, uint64_t* times) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    __shared__ uint64_t item;
    grep_batch_ctl* q = (grep_batch_ctl*)ws;
    grep_split_state* fs = (grep_split_state*)(ws + sizeof(grep_batch_ctl));
    uint64_t* ring = (uint64_t*)(fs + nfiles);
    grep_build_dict(S, dict, nw);
    const uint32_t t = gfs::tid();
    gfs_dev* g = gfs_g;
    // Every item is taken once and a workgroup handles one per round: at most
    // nfiles * parts + 1 rounds.
    const uint64_t rounds = (uint64_t)nfiles * parts + 1;
    for (uint64_t round = 0; round < rounds; ++round) {
        if (t == 0) {
            uint64_t it = BATCH_STOP;
            const uint64_t t0 = gfs::now_ticks();
            // Relaxed polls throughout (an acquire load at agent scope
            // invalidates this XCD's L2 each time); one acquire below once an
            // item is taken.
            for (uint32_t spins = 0;; ++spins) {
                // an extra part first: its file is open already
                uint32_t h = __hip_atomic_load(&q->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t tl = __hip_atomic_load(&q->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (h < tl) {
                    if (__hip_atomic_compare_exchange_strong(&q->head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        // reserved by a running part 0, published right after: bounded wait
                        uint64_t v;
                        while (!((v = __hip_atomic_load(&ring[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 63) &&
                               gfs::now_ticks() - t0 < g->wait_ticks)
                            __builtin_amdgcn_s_sleep(2);
                        it = v;   // not ready after the wait limit: stop this workgroup
                        break;
                    }
                    continue;
                }
                if (__hip_atomic_load(&q->next_file, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nfiles) {
                    // count the file as expanding before taking it, so no workgroup
                    // stops while its parts may still be published
                    atomicAdd(&q->expanding, 1u);
                    const uint32_t f = atomicAdd(&q->next_file, 1u);
                    if (f < nfiles) {
                        it = (1ull << 63) | ((uint64_t)f << 24);   // part 0, K_f not known yet
                        break;
                    }
                    atomicSub(&q->expanding, 1u);
                    continue;
                }
                if (__hip_atomic_load(&q->expanding, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                    __hip_atomic_load(&q->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                        __hip_atomic_load(&q->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;   // nothing left
                if (gfs::now_ticks() - t0 > g->wait_ticks) break;
                if (spins < 64) __builtin_amdgcn_s_sleep(2);
                else __builtin_amdgcn_s_sleep(32);
            }
            gfs::fence_acq_agent();   // the part item's publication (a release store)
            item = it;
        }
        __syncthreads();
        const uint64_t it = gfs::uni((uint64_t)item);
        __syncthreads();
        if (it == BATCH_STOP || !(it >> 63)) break;
        const uint32_t f = (uint32_t)((it >> 24) & 0x7FFFFFFFFFull);
        const uint32_t j = (uint32_t)(it & 0xFF), kfi = (uint32_t)((it >> 8) & 0xFFFF);
        if (f >= nfiles) break;
        if (j == 0 && t == 0 && times) times[2 * (size_t)f] = gfs::now_ticks();
        [[clang::always_inline]] grep_file_split(
            S, names + (size_t)f * stride, nw, chunk, counts_out + (size_t)f * nw, status + f, j, parts, min_part,
            fs + f, j == 0 ? 0u : kfi, 0u, false, false, [&](uint32_t kf) {
                if (j != 0 || t != 0) return;
                if (kf > 1) {
                    const uint32_t r0 = atomicAdd(&q->tail, kf - 1);
                    for (uint32_t x = 1; x < kf; ++x)
                        __hip_atomic_store(&ring[r0 + x - 1],
                                           (1ull << 63) | ((uint64_t)f << 24) | ((uint64_t)kf << 8) | x,
                                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
                __hip_atomic_fetch_sub(&q->expanding, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            },
            [] {});
        if (t == 0 && times) times[2 * (size_t)f + 1] = gfs::now_ticks();   // the last part's end is kept
        // A barrier between this thread-0 region and the next round's keeps the
        // compiler from threading the `tid == 0` test across the back-edge.
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned int* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Epoch of a split launch (grep_split_state::tag): distinct for every launch
// of the process, never 0 (the batched form's workspace is zeroed per launch
// and uses epoch 0).
static uint32_t grep_next_epoch() {
    static std::atomic<uint32_t> e{0};
    uint32_t v;
    while ((v = e.fetch_add(1, std::memory_order_relaxed) + 1) == 0) {
    }
    return v;
}

// Parts per file when the caller leaves it to the launcher: enough workgroups
// to give every compute unit one (each holds a whole compute unit's LDS), at
// most 16 per file.
static uint32_t grep_auto_parts(uint32_t n_files) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    uint32_t k = ((uint32_t)cus + n_files - 1) / n_files;
    return k < 1 ? 1 : (k > 16 ? 16 : k);
}

static int launch_grep_split(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                             const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes, uint32_t* counts_dev,
                             int32_t* status_dev, uint32_t parts, uint64_t min_part, grep_split_state* fs,
                             uint32_t threads, hipStream_t s) {
    if (parts == 1) {
        // one workgroup per file: no shared state
        hipLaunchKernelGGL(k_grep, dim3(n_files), dim3(threads), 0, s, names_dev, name_stride, n_files, dict_dev, n_words,
                           chunk_bytes, counts_dev, status_dev, 1u, min_part, (grep_split_state*)nullptr, 0u);
        return hipGetLastError() == hipSuccess ? 0 : -EIO;
    }
    // nothing to clear first: part 0 of each file zeroes its row and state
    hipLaunchKernelGGL(k_grep, dim3(n_files * parts), dim3(threads), 0, s, names_dev, name_stride, n_files, dict_dev, n_words,
                       chunk_bytes, counts_dev, status_dev, parts, min_part, fs, grep_next_epoch());
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

extern "C" int gpufs_launch_grep_split(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                       uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                       uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                       uint32_t parts_per_file, uint64_t min_part_bytes, uint32_t threads,
                                       void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev) return -EINVAL;
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    if (parts_per_file == 0) parts_per_file = grep_auto_parts(n_files);
    if (parts_per_file > 64 || (uint64_t)n_files * parts_per_file > 0x7FFFFFFFull) return -EINVAL;
    if (min_part_bytes == 0) min_part_bytes = 4ull * chunk_bytes > (256u << 10) ? 4ull * chunk_bytes : (256u << 10);
    if (min_part_bytes < 256) return -EINVAL;
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    if (parts_per_file == 1)
        return launch_grep_split(ctx, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev,
                                 status_dev, 1, min_part_bytes, nullptr, threads, s);
    grep_split_state* fs = nullptr;
    if (hipMallocAsync((void**)&fs, sizeof(grep_split_state) * (size_t)n_files, s) != hipSuccess) return -ENOMEM;
    v = launch_grep_split(ctx, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev,
                          status_dev, parts_per_file, min_part_bytes, fs, threads, s);
    (void)hipFreeAsync(fs, s);
    return v;
}

// ---------------------------------------------------------------------------
// The split scan without any allocation (for callers that cannot allocate:
// code under graph capture, serving loops).  Its per-file state lives in
// `scratch`, a file-scope device array; every GPU holds its own copy at its
// own address, which gpufs_init looks up once per device
// (grep_scratch_setup).  Calls on one GPU share the array: an entry left by
// an earlier call is never taken for this call's (each launch has its own
// epoch, and part 0 of a file resets the file's entry before the other parts
// read it), and calls that overlap in time on one GPU race.
#define GREP_NOALLOC_MAX_FILES 65536u
#define GREP_MAX_DEVICES 64
static __device__ __align__(256) grep_split_state scratch[GREP_NOALLOC_MAX_FILES];
static grep_split_state* scratch_dev[GREP_MAX_DEVICES];   // device address of `scratch` per device

// Called by gpufs_init with `device` current.  In host code the array's name
// is only a host-side stand-in, so its device address is taken here once and
// used everywhere the host passes the array on.
__attribute__((visibility("hidden"))) int grep_scratch_setup(int device) {
    if (device < 0 || device >= GREP_MAX_DEVICES) return -ENODEV;
    if (__atomic_load_n(&scratch_dev[device], __ATOMIC_ACQUIRE)) return 0;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -EIO;
    __atomic_store_n(&scratch_dev[device], (grep_split_state*)p, __ATOMIC_RELEASE);
    return 0;
}

extern "C" int gpufs_launch_grep_noalloc(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                         uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                         uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                         uint32_t parts_per_file, uint64_t min_part_bytes, uint32_t threads,
                                         void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev) return -EINVAL;
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    if (n_files > GREP_NOALLOC_MAX_FILES) return -E2BIG;
    if (parts_per_file == 0) parts_per_file = grep_auto_parts(n_files);
    if (parts_per_file > 64) return -EINVAL;
    if (min_part_bytes == 0) min_part_bytes = 4ull * chunk_bytes > (256u << 10) ? 4ull * chunk_bytes : (256u << 10);
    if (min_part_bytes < 256) return -EINVAL;
    gpufs_config cc;
    gpufs_get_config(ctx, &cc);
    if (cc.device < 0 || cc.device >= GREP_MAX_DEVICES) return -ENODEV;
    grep_split_state* fs = __atomic_load_n(&scratch_dev[cc.device], __ATOMIC_ACQUIRE);
    if (!fs) return -ENODEV;   // gpufs_init did not set up this device
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    return launch_grep_split(ctx, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev,
                             status_dev, parts_per_file, min_part_bytes, fs, threads, s);
}

extern "C" int gpufs_launch_grep(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                                 const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes,
                                 uint32_t* counts_dev, int32_t* status_dev, uint32_t threads, void* stream) {
    return gpufs_launch_grep_split(ctx, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev,
                                   status_dev, 0, 0, threads, stream);
}

// Zeroes the batched form's workspace (control words, split state, part ring)
// and the count rows: a fill kernel, so the launch can be captured in a graph.
__global__ void k_batch_clear(uint64_t* ws, uint64_t ws_words, uint32_t* counts, uint64_t n_counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ws_words) ws[i] = 0;
    if (i < n_counts) counts[i] = 0;
}

static int launch_batched(hipStream_t s, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                          const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes, uint32_t* counts_dev,
                          int32_t* status_dev, void* ws_dev, uint32_t workgroups, uint32_t threads,
                          uint32_t parts, uint64_t min_part, uint64_t* times_dev) {
    const uint64_t words = batch_ws_bytes(n_files, parts) / 8, nc = (uint64_t)n_files * n_words;
    const uint64_t n = words > nc ? words : nc;
    hipLaunchKernelGGL(k_batch_clear, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, (uint64_t*)ws_dev, words,
                       counts_dev, nc);
    hipLaunchKernelGGL(k_grep_batched, dim3(workgroups), dim3(threads), 0, s, names_dev, name_stride, n_files,
                       dict_dev, n_words, chunk_bytes, counts_dev, status_dev, (uint8_t*)ws_dev, parts, min_part,
                       times_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

static int batched_args(uint32_t n_files, uint32_t n_words, uint32_t chunk_bytes, uint32_t threads,
                        uint32_t workgroups, uint32_t* parts, uint64_t* min_part) {
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    if (workgroups == 0 || n_files >= (1u << 31)) return -EINVAL;
    if (*parts == 0) *parts = 8;
    if (*parts > 255) return -EINVAL;
    if (*min_part == 0) *min_part = 4ull * chunk_bytes > (1u << 20) ? 4ull * chunk_bytes : (1u << 20);
    if (*min_part < 256) return -EINVAL;
    return 0;
}

extern "C" uint64_t gpufs_grep_batched_workspace_bytes(uint32_t n_files, uint32_t parts_per_file) {
    if (parts_per_file == 0) parts_per_file = 8;
    return batch_ws_bytes(n_files, parts_per_file);
}

extern "C" int gpufs_launch_grep_batched(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                         uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                         uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                         void* workspace_dev, uint32_t workgroups, uint32_t threads,
                                         uint32_t parts_per_file, uint64_t min_part_bytes, uint64_t* file_times_dev,
                                         void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev || !workspace_dev) return -EINVAL;
    int v = batched_args(n_files, n_words, chunk_bytes, threads, workgroups, &parts_per_file, &min_part_bytes);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    return launch_batched(s, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev, status_dev,
                          workspace_dev, workgroups, threads, parts_per_file, min_part_bytes, file_times_dev);
}

// Capture (workspace clear + persistent kernel) into an executable hipGraph.
extern "C" int gpufs_grep_batched_graph(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                        uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                        uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                        void* workspace_dev, uint32_t workgroups, uint32_t threads,
                                        uint32_t parts_per_file, uint64_t min_part_bytes, uint64_t* file_times_dev,
                                        void** graph_exec_out) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev || !workspace_dev || !graph_exec_out)
        return -EINVAL;
    int v = batched_args(n_files, n_words, chunk_bytes, threads, workgroups, &parts_per_file, &min_part_bytes);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = (hipStream_t)gpufs_stream(ctx);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return -EIO;
    v = launch_batched(s, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev, status_dev,
                       workspace_dev, workgroups, threads, parts_per_file, min_part_bytes, file_times_dev);
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(s, &gr) != hipSuccess || !gr || v) {
        if (gr) (void)hipGraphDestroy(gr);
        return v ? v : -EIO;
    }
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (r != hipSuccess) return -EIO;
    *graph_exec_out = (void*)e;
    return 0;
}

extern "C" int gpufs_graph_launch(void* graph_exec, void* stream) {
    if (!graph_exec) return -EINVAL;
    return hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream) == hipSuccess ? 0 : -EIO;
}

extern "C" int gpufs_graph_destroy(void* graph_exec) {
    if (!graph_exec) return -EINVAL;
    return hipGraphExecDestroy((hipGraphExec_t)graph_exec) == hipSuccess ? 0 : -EIO;
}

// ---------------------------------------------------------------------------
// The grep_text sample end to end (samples/grep_text/grep_text_gpu.cu:193-387):
// three GPUfs files -- `src`, query words as 32-byte NUL-padded records
// (dict.gpu), `dbs`, a list of database file names, and `out`, opened
// O_GWRONCE.  Workgroup b takes query records [b*wpc, b*wpc+wpc) (the last
// workgroup also takes the remainder), reads the list, and scans every
// database in its chunk loop; after each window it formats a line
// "<word> <count> <database>\n" for every record whose count in that window
// is nonzero, reserves that many bytes of the output with a device atomic and
// gwrites them there.  The line order depends on timing, as in the
// reference; the set of lines does not.  Counts are printed in full decimal.
struct grep_text_args {
    const char* src;
    const char* out;
    const char* dbs;
    uint64_t total_words;
    uint32_t wpc;
    uint32_t chunk;
    uint8_t* scratch;            // per workgroup: words | list | name | lines
    uint64_t stride, off_list, off_name, off_lines, list_cap;
    unsigned long long* out_off;
    uint32_t* trunc;             // 0, 1 while the output is truncated, 2 done, 3 failed
    int32_t* status;
    uint32_t records;            // 0: text lines; 1: fixed 100-byte binary records
};

// Binary record of one (query record, window) match, the layout that the
// reference's record printer reads (samples/utils/display.cpp:16-25): the
// query's 32 bytes as stored, the count as a 32-bit integer, the database
// name NUL-padded to 64 bytes.
#define GREP_REC_BYTES 100u

// get_next() of grep_text_gpu.cu:163-186 on a NUL-terminated list: skips
// ' ', '\t', ',', '\r', '\n'; a name ends at '\n', '\r', ',' or NUL; returns
// its length (0 when the list is over or the name is 64 bytes or longer) and
// advances *pos past the byte that ended it.
__device__ inline uint32_t text_next_name(const char* list, uint64_t* pos, char* name) {
    uint64_t i = *pos;
    while (list[i] == ' ' || list[i] == '\t' || list[i] == ',' || list[i] == '\r' || list[i] == '\n') ++i;
    const uint64_t beg = i;
    while (list[i] != '\n' && list[i] != '\r' && list[i] != ',' && list[i] != 0 && i < beg + 64) ++i;
    const uint64_t n = i - beg;
    if (n == 0 || n >= 64) return 0;
    for (uint64_t k = 0; k < n; ++k) name[k] = list[beg + k];
    for (uint64_t k = n; k < 128; ++k) name[k] = 0;
    *pos = list[i] ? i + 1 : i;
    return (uint32_t)n;
}

__global__ __launch_bounds__(1024) void k_grep_text(grep_text_args a) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    __shared__ uint32_t line_bytes, name_len;
    __shared__ int wst;   // first gwrite failure of the current database (thread 0 writes it)
    __shared__ uint64_t list_pos, out_at;
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    uint8_t* base = a.scratch + (size_t)blockIdx.x * a.stride;
    char* words = (char*)base;
    char* list = (char*)(base + a.off_list);
    char* name = (char*)(base + a.off_name);
    char* lines = (char*)(base + a.off_lines);
    const uint64_t first = (uint64_t)blockIdx.x * a.wpc;
    const uint32_t nw = first >= a.total_words ? 0u
                        : (blockIdx.x == gridDim.x - 1 ? (uint32_t)(a.total_words - first) : a.wpc);
    int st = 0;
    const int fd_dbs = gopen(a.dbs, O_GRDONLY);
    const int fd_o = gopen(a.out, O_GWRONCE);
    const int fd_src = gopen(a.src, O_GRDONLY);
    if (fd_dbs < 0 || fd_o < 0 || fd_src < 0) st = fd_dbs < 0 ? fd_dbs : (fd_o < 0 ? fd_o : fd_src);
    if (st == 0) {
        // The output starts empty: the first workgroup here truncates it from
        // one thread, as the reference does (grep_text_gpu.cu:258-268), and the
        // others wait for that before anything is written (bounded wait).
        if (t == 0) {
            int r = 0;
            if (atomicCAS(a.trunc, 0u, 1u) == 0u) {
                r = single_thread_ftruncate(fd_o, 0);
                __hip_atomic_store(a.trunc, r ? 3u : 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const uint64_t w0 = gfs::now_ticks();
                uint32_t v;
                while ((v = __hip_atomic_load(a.trunc, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) == 1u &&
                       gfs::now_ticks() - w0 < gfs_g->wait_ticks)
                    __builtin_amdgcn_s_sleep(8);
                r = v == 2u ? 0 : (v == 1u ? -ETIMEDOUT : -EIO);
            }
            wst = r;
        }
        __syncthreads();
        st = (int)gfs::uni((uint32_t)wst);
    }
    if (st == 0 && nw > 0) {
        size_t got = gread(fd_src, first * 32, (size_t)nw * 32, (uchar*)words);
        if (got != (size_t)nw * 32) st = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
    }
    if (st == 0 && nw > 0) {
        size_t lsz = fstat(fd_dbs);
        if (lsz > a.list_cap - 1) lsz = a.list_cap - 1;
        size_t got = lsz ? gread(fd_dbs, 0, lsz, (uchar*)list) : 0;
        if (got != lsz) st = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
        if (t == 0) {
            list[lsz] = 0;
            list_pos = 0;
            line_bytes = 0;
        }
    }
    if (st == 0 && nw > 0) {
        grep_build_dict(S, words, nw);
        for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
        for (;;) {
            if (t == 0) {
                uint64_t p = list_pos;
                name_len = text_next_name(list, &p, name);
                list_pos = p;
            }
            __syncthreads();
            const uint32_t nl = gfs::uni(name_len);
            if (nl == 0) break;
            const int fd = gopen(name, O_GRDONLY);
            if (fd < 0) { st = fd; break; }
            if (t == 0) wst = 0;
            const int fst = grep_file_windows(S, fd, a.chunk, [&] {
                // this window's lines, then the counts start again from zero
                for (uint32_t q = t; q < nw; q += T) {
                    const uint32_t k = S.canon[q];
                    const uint32_t c = k == 0xFFFF ? 0u : S.counts[k];
                    if (!c) continue;
                    if (a.records) {
                        char* o = lines + atomicAdd(&line_bytes, GREP_REC_BYTES);
                        for (uint32_t i = 0; i < 32; ++i) o[i] = words[(size_t)q * 32 + i];
                        for (uint32_t i = 0; i < 4; ++i) o[32 + i] = (char)(c >> (8 * i));
                        for (uint32_t i = 0; i < 64; ++i) o[36 + i] = i < nl ? name[i] : 0;
                        continue;
                    }
                    char num[10];
                    uint32_t nd = 0;
                    for (uint32_t v = c; v; v /= 10) num[nd++] = (char)('0' + v % 10);
                    const uint32_t wl = S.wlen[q];
                    const uint32_t len = wl + 1 + nd + 1 + nl + 1;
                    char* o = lines + atomicAdd(&line_bytes, len);
                    for (uint32_t i = 0; i < wl; ++i) o[i] = words[(size_t)q * 32 + i];
                    o[wl] = ' ';
                    for (uint32_t i = 0; i < nd; ++i) o[wl + 1 + i] = num[nd - 1 - i];
                    o[wl + 1 + nd] = ' ';
                    for (uint32_t i = 0; i < nl; ++i) o[wl + 2 + nd + i] = name[i];
                    o[len - 1] = '\n';
                }
                __threadfence();
                __syncthreads();
                const uint32_t total = gfs::uni(line_bytes);
                if (total) {
                    if (t == 0) out_at = atomicAdd(a.out_off, (unsigned long long)total);
                    __syncthreads();
                    const size_t w = gwrite(fd_o, gfs::uni(out_at), total, (uchar*)lines);
                    if (w != total && t == 0 && wst == 0) wst = (int64_t)w < 0 ? (int)(int64_t)w : -EIO;
                }
                for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
                if (t == 0) line_bytes = 0;
                __syncthreads();
            });
            const int c = gclose(fd);   // collective: its barriers also publish wst
            const int w = gfs::uni(wst);
            st = fst ? fst : (w ? w : c);
            if (st) break;
        }
    }
    if (fd_src >= 0) gclose(fd_src);
    if (fd_dbs >= 0) gclose(fd_dbs);
    const int co = fd_o >= 0 ? gclose(fd_o) : 0;
    if (t == 0 && (st || co)) atomicMin(a.status, st ? st : co);
}

static int grep_text_launch(gpufs_ctx* ctx, const char* src_path, const char* out_path, const char* dbs_path,
                            uint32_t chunk_bytes, uint32_t workgroups, uint32_t threads, int32_t* status_dev,
                            void* stream, uint32_t records) {
    if (!ctx || !src_path || !out_path || !dbs_path || !status_dev) return -EINVAL;
    if (chunk_bytes == 0 || chunk_bytes > GREP_MAX_CHUNK) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64 || workgroups == 0) return -EINVAL;
    const size_t ln = strlen(src_path), lo = strlen(out_path), ld = strlen(dbs_path);
    if (ln >= GFS_NAME_MAX || lo >= GFS_NAME_MAX || ld >= GFS_NAME_MAX) return -ENAMETOOLONG;
    struct stat sb;
    if (stat(src_path, &sb) != 0) return -errno;
    const uint64_t total = (uint64_t)sb.st_size / 32;
    if (total == 0) return -EINVAL;   // "empty input" in the reference
    if (stat(dbs_path, &sb) != 0) return -errno;
    const uint64_t list_cap = (uint64_t)sb.st_size + 1;
    // words per workgroup as the reference splits them; more workgroups when a
    // slice would not fit the 1024-word LDS table (the lines do not depend on
    // how the records are split)
    uint64_t grid = workgroups;
    for (;; ++grid) {
        uint64_t wpc = total / grid;
        if (wpc == 0) wpc = 1;
        const uint64_t last = total > wpc * (grid - 1) ? total - wpc * (grid - 1) : 0;
        if (wpc <= GREP_MAX_WORDS && last <= GREP_MAX_WORDS) break;
        if (grid > (1u << 20)) return -EINVAL;
    }
    uint64_t wpc = total / grid;
    if (wpc == 0) wpc = 1;
    if (grid > total) grid = total;   // every workgroup has a record
    int v = gpufs_bind(ctx);
    if (v) return v;
    const uint64_t words_b = (uint64_t)GREP_MAX_WORDS * 32;
    const uint64_t lines_b = (uint64_t)GREP_MAX_WORDS * (32 + 1 + 10 + 1 + 63 + 1);
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t off_list = al(words_b), off_name = al(off_list + list_cap + 64), off_lines = al(off_name + 128);
    const uint64_t stride = al(off_lines + lines_b);
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    uint8_t* scratch = nullptr;
    char* names = nullptr;
    unsigned long long* off = nullptr;
    // names | output offset (8 B) | truncation flag (8 B), zeroed by the copy
    char hn[3 * GFS_NAME_MAX + 16];
    auto release = [&](int r) {
        if (scratch) (void)hipFreeAsync(scratch, s);
        if (names) (void)hipFreeAsync(names, s);
        return r;
    };
    if (hipMallocAsync((void**)&scratch, stride * grid, s) != hipSuccess) return -ENOMEM;
    if (hipMallocAsync((void**)&names, sizeof(hn), s) != hipSuccess) {
        names = nullptr;
        return release(-ENOMEM);
    }
    off = (unsigned long long*)(names + 3 * GFS_NAME_MAX);
    memset(hn, 0, sizeof(hn));
    memcpy(hn, src_path, ln);
    memcpy(hn + GFS_NAME_MAX, out_path, lo);
    memcpy(hn + 2 * GFS_NAME_MAX, dbs_path, ld);
    if (hipMemcpyAsync(names, hn, sizeof(hn), hipMemcpyHostToDevice, s) != hipSuccess) return release(-EIO);
    if (hipStreamSynchronize(s) != hipSuccess) return release(-EIO);   // hn is a stack buffer
    grep_text_args a{names, names + GFS_NAME_MAX, names + 2 * GFS_NAME_MAX, total, (uint32_t)wpc, chunk_bytes,
                     scratch, stride, off_list, off_name, off_lines, list_cap, off,
                     (uint32_t*)(names + 3 * GFS_NAME_MAX + 8), status_dev, records};
    hipLaunchKernelGGL(k_grep_text, dim3((uint32_t)grid), dim3(threads), 0, s, a);
    const int lr = hipGetLastError() == hipSuccess ? (int)grid : -EIO;
    return release(lr);
}

extern "C" int gpufs_grep_text(gpufs_ctx* ctx, const char* src_path, const char* out_path, const char* dbs_path,
                               uint32_t chunk_bytes, uint32_t workgroups, uint32_t threads, int32_t* status_dev,
                               void* stream) {
    return grep_text_launch(ctx, src_path, out_path, dbs_path, chunk_bytes, workgroups, threads, status_dev, stream,
                            0);
}

extern "C" int gpufs_grep_text_records(gpufs_ctx* ctx, const char* src_path, const char* out_path,
                                       const char* dbs_path, uint32_t chunk_bytes, uint32_t workgroups,
                                       uint32_t threads, int32_t* status_dev, void* stream) {
    return grep_text_launch(ctx, src_path, out_path, dbs_path, chunk_bytes, workgroups, threads, status_dev, stream,
                            1);
}
