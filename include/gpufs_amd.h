/*
 * gpufs_amd.h -- C ABI of the MI355X-native GPUfs file I/O path.
 *
 * Host side of the drop-in boundary.  Device code (gopen/gread/gwrite/gmmap/
 * gmunmap/gfsync/gclose/...) is the header-only HIP API in
 * gpufs_amd/csrc/gpufs.h; it cannot cross a C ABI, so the entry points below
 * are the host init / host-loop calls of the reference plus the launchers of
 * the workload kernels that run on top of the device API.
 *
 * Conventions: every function returns 0 (or a non-negative count) on success
 * and a negative GPUFS_E* / -errno code on failure; nothing here calls exit().
 * Pointers named *_dev are device (HBM) addresses, e.g. a torch CUDA tensor's
 * data_ptr(); pointers named *_host are host addresses.  Streams are
 * hipStream_t passed as void*; NULL means the context's own stream.
 */
#ifndef GPUFS_AMD_H
#define GPUFS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes shared with the device API.  -1/-2 keep the values of the
 * reference's E_FSTABLE_FULL / E_IPC_OPEN_ERROR (libgpufs/fs_constants.h:50-51);
 * everything else is a negated errno. */
#define GPUFS_OK               0
#define GPUFS_E_FSTABLE_FULL  (-1)
#define GPUFS_E_IPC_OPEN      (-2)

typedef struct gpufs_ctx gpufs_ctx;

/* Runtime sizing of everything the reference fixes with #defines in
 * libgpufs/fs_constants.h:25-77 (FS_BLOCKSIZE, PPOOL_SIZE, MAX_NUM_FILES,
 * RW_IPC_SIZE, RW_HOST_WORKERS). */
typedef struct gpufs_config {
    uint64_t cache_bytes;   /* page-cache pool in HBM; frames = cache_bytes / page_size */
    uint32_t page_size;     /* power of two, 4 KiB .. 4 MiB */
    uint32_t max_files;     /* file-table capacity (distinct file names per context) */
    uint32_t ring_depth;    /* GPU->host RPC slots; each owns page_size bytes of pinned staging */
    uint32_t host_threads;  /* host service threads doing open/pread/pwrite */
    uint32_t tree_nodes;    /* radix-tree node pool (512 B each); 0 = sized automatically */
    int32_t  device;        /* HIP device ordinal */
    uint32_t flags;         /* reserved, must be 0 */
    uint32_t reserved;
} gpufs_config;

/* Counters exported by gpufs_stats_get (the reference's MALLOC_STATS /
 * PRINT_TIMES counters, libgpufs/fs_debug.cu.h:74-149, host_loop.h:617-655). */
typedef struct gpufs_stats {
    uint64_t page_hits;      /* page lookups served from a VALID frame */
    uint64_t page_misses;    /* page lookups that filled a frame */
    uint64_t page_waits;     /* lookups that found a frame being filled by another workgroup */
    uint64_t evictions;      /* VALID frames reclaimed by the clock sweep */
    uint64_t alloc_fail;     /* frame allocations that found nothing evictable */
    uint64_t rpc_reads;      /* READ requests posted to the host */
    uint64_t rpc_writes;     /* WRITE / DIFF requests posted to the host */
    uint64_t rpc_opens;      /* OPEN requests */
    uint64_t rpc_closes;     /* CLOSE requests */
    uint64_t bytes_filled;   /* bytes copied from pinned staging into frames */
    uint64_t bytes_written;  /* bytes copied from frames into staging for write-back */
    uint64_t nodes_used;     /* radix-tree nodes taken from the never-used range (high-water mark) */
    uint64_t files_used;     /* file-table entries in use (not free) */
    uint64_t host_requests;  /* requests served by the host service */
    uint64_t host_bytes_read;
    uint64_t host_bytes_written;
    uint64_t frames;         /* frames in the pool */
    uint64_t frames_valid;   /* frames VALID at the time of the call */
    uint64_t timeouts;       /* device waits that hit the wait limit (GPUFS_WAIT_TIMEOUT_S) */
    uint64_t host_busy_ns;   /* time the host workers spent serving requests, summed over workers */
    uint64_t vm_errors;      /* FatPointer dereferences whose page could not be mapped */
    uint64_t slots_lost;     /* ring slots kept out of the ring after a request timed out (the host
                                may still answer it); the ring is that much smaller until destroy */
    uint64_t files_reclaimed;  /* closed entries taken back for other names when the table was full
                                  (libgpufs/fs_structures.cu:364-369) */
    uint64_t nodes_recycled;   /* radix nodes returned to the free list (trees of reclaimed files and
                                  of files found changed on reopen) */
    uint64_t spec_parts;       /* workgroups (split grep parts, gread_ranges ranges) that read a cached
                                  file while its reopen was at the host */
    uint64_t spec_rescans;     /* of those, the ones that found a page gone from the cache and read
                                  their range again through the open file */
    uint64_t spec_changed;     /* of those, the ones that found the file changed by the reopen (read
                                  again, or recounted, through the open file) */
} gpufs_stats;

/* ---- lifecycle: replaces initializer(GPUGlobals_ptr*) (libgpufs/fs_initializer.cu:116)
 * and `delete GPUGlobals` (libgpufs/fs_initializer.cu.h:180-198). */
int   gpufs_config_default(gpufs_config* cfg);
int   gpufs_init(const gpufs_config* cfg, gpufs_ctx** out);
int   gpufs_destroy(gpufs_ctx* ctx);   /* -EBUSY while a server (gpufs_serve_start) runs on ctx */
/* The stream kernels that call the device API are launched on
 * (GPUGlobals::streamMgr->kernelStream, libgpufs/fs_initializer.cu.h:56-75). */
void* gpufs_stream(gpufs_ctx* ctx);
/* Device address of the context's device state; gpufs.h's gpufs_bind() stores it
 * in the translation unit that launches device-API kernels. */
void* gpufs_device_state(gpufs_ctx* ctx);
int   gpufs_get_config(gpufs_ctx* ctx, gpufs_config* out);

/* ---- host service: replaces run_gpufs_handler(volatile GPUGlobals*, int)
 * (libgpufs/host_loop.h:559-656).  gpufs_run_handler has the reference's
 * semantics: it serves RPCs until the context stream is idle, then returns.
 * gpufs_host_start/stop run the same service free-standing, independent of
 * kernel completion, for persistent, graph-captured and serving kernels. */
int   gpufs_run_handler(gpufs_ctx* ctx);
int   gpufs_host_start(gpufs_ctx* ctx);
int   gpufs_host_stop(gpufs_ctx* ctx);
/* No reference counterpart (the reference's handler polls continuously): wakes
 * the running service's idle worker threads and keeps them polling for about a
 * millisecond, so a kernel launched right after has its first requests (opens)
 * served without a thread wake-up delay.  gpufs.h's gpufs_bind() calls it;
 * a no-op when the service is not running. */
int   gpufs_host_kick(gpufs_ctx* ctx);

/* ---- cache control and introspection (no reference counterpart; used by the
 * cold-read measurements and the invariant tests). */
int   gpufs_cache_drop(gpufs_ctx* ctx);   /* empty the page cache; no kernel may be running */
int   gpufs_stats_get(gpufs_ctx* ctx, gpufs_stats* out);
int   gpufs_stats_reset(gpufs_ctx* ctx);
/* Checks page-cache invariants on the host: no two live frames own the same
 * (file, page) key, no frame is pinned, every live leaf points at a live frame
 * of the right key.  Returns 0, or a negative code naming the first violation. */
int   gpufs_check_invariants(gpufs_ctx* ctx);
const char* gpufs_strerror(int code);

/* ---- workload launchers (device API callers), one per benchmark shape.
 * Each launches on the context stream and returns without waiting; call
 * gpufs_run_handler (or keep the host service started) and synchronise. */

/* Sequential gread (samples/memcopy/cp.cu:73-108 pattern): workgroup w greads
 * [w*range_bytes, (w+1)*range_bytes) of `path_dev` (a NUL-terminated device
 * string) into dst_dev + w*range_bytes.  bytes_out_dev[w] = gread's return. */
int gpufs_launch_gread_ranges(gpufs_ctx* ctx, const char* path_dev, uint32_t n_ranges,
                              uint64_t range_bytes, void* dst_dev, int64_t* bytes_out_dev,
                              uint32_t threads, void* stream);

/* Sequential gwrite, the write side of the same pattern (the gwrite loop of
 * samples/memcopy/cp.cu:73-108 with one range per workgroup): workgroup w
 * gopens `path_dev` with open_flags (O_GWRONLY / O_GRDWR / O_GWRONCE, with
 * O_GCREAT as needed), gwrites src_dev + w*range_bytes at w*range_bytes,
 * gfsyncs when do_fsync, and gcloses.  bytes_out_dev[w] = bytes written, or
 * the first error of gwrite, gfsync, gclose.  -EINVAL for a read-only mode. */
int gpufs_launch_gwrite_ranges(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags,
                               uint32_t n_ranges, uint64_t range_bytes, const void* src_dev,
                               int32_t do_fsync, int64_t* bytes_out_dev, uint32_t threads,
                               void* stream);

/* grep_text-style scan (samples/grep_text/grep_text_gpu.cu:193-387, scan
 * semantics :79,139-160, stitching :298-318): file f is names_dev +
 * f*name_stride; its whole-word match counts of the n_words dictionary words
 * (dict_dev: n_words records of 32 bytes, NUL-padded) are written to
 * counts_dev[f*n_words + q] (uint32), exactly as the reference's chunk loop of
 * chunk_bytes windows with a 32-byte look-ahead counts them.  Each file is
 * gopen()ed and gread() into LDS by several workgroups (parts, chosen so the
 * launch fills the device); status_dev[f] = 0 or a negative error code. */
int gpufs_launch_grep(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                      uint32_t n_files, const char* dict_dev, uint32_t n_words,
                      uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                      uint32_t threads, void* stream);
/* The same scan with the split chosen by the caller: a file of S bytes is
 * scanned by min(parts_per_file, max(1, S / min_part_bytes)) workgroups, each
 * counting the tokens that start in its byte range; the workgroup that
 * finishes last adds what the reference's last-window rule counts again and,
 * when a boundary meets a run of 32 or more word bytes, recounts the file with
 * the reference's loop.  parts_per_file 0 = as gpufs_launch_grep (<= 64);
 * min_part_bytes 0 = max(256 KiB, 4 * chunk_bytes), else >= 256.  No reference
 * counterpart (the reference gives each file one thread block). */
int gpufs_launch_grep_split(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                            uint32_t n_files, const char* dict_dev, uint32_t n_words,
                            uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                            uint32_t parts_per_file, uint64_t min_part_bytes, uint32_t threads,
                            void* stream);
/* gpufs_launch_grep_split without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-file state lives in a fixed device array of the library (56 B x 65,536
 * files per GPU, held until the process ends); n_files above 65,536 returns
 * -E2BIG.  Calls on one GPU share that array (each call clears what it uses
 * first), so calls that overlap in time on one GPU race. */
int gpufs_launch_grep_noalloc(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                              uint32_t n_files, const char* dict_dev, uint32_t n_words,
                              uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                              uint32_t parts_per_file, uint64_t min_part_bytes, uint32_t threads,
                              void* stream);

/* Batched form of the grep scan over many files of uneven size: `workgroups`
 * persistent workgroups build the dictionary table once, then take work until
 * none is left: a new file from a device queue, or an extra part of a large
 * file another workgroup opened (a file of S bytes is scanned by
 * min(parts_per_file, max(1, S / min_part_bytes)) workgroups, as in
 * gpufs_launch_grep_split; parts_per_file 0 = 8, at most 255; min_part_bytes
 * 0 = max(1 MiB, 4 * chunk_bytes), else >= 256).  workspace_dev: device
 * memory of gpufs_grep_batched_workspace_bytes(n_files, parts_per_file)
 * bytes, cleared by the launch itself.  Same per-file counts and status as
 * gpufs_launch_grep.  When file_times_dev is not NULL it receives each file's
 * start and end wall-clock ticks (2 per file, 100 MHz).  No reference
 * counterpart (the reference launches one grid over a fixed file list). */
uint64_t gpufs_grep_batched_workspace_bytes(uint32_t n_files, uint32_t parts_per_file);
int gpufs_launch_grep_batched(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                              uint32_t n_files, const char* dict_dev, uint32_t n_words,
                              uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                              void* workspace_dev, uint32_t workgroups, uint32_t threads,
                              uint32_t parts_per_file, uint64_t min_part_bytes,
                              uint64_t* file_times_dev, void* stream);
/* The same two launches (workspace clear, persistent kernel) captured on the
 * context stream into an executable hipGraph; replay with gpufs_graph_launch.
 * The host service must be running (gpufs_host_start) while a replay runs. */
int gpufs_grep_batched_graph(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                             uint32_t n_files, const char* dict_dev, uint32_t n_words,
                             uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                             void* workspace_dev, uint32_t workgroups, uint32_t threads,
                             uint32_t parts_per_file, uint64_t min_part_bytes,
                             uint64_t* file_times_dev, void** graph_exec_out);
int gpufs_graph_launch(void* graph_exec, void* stream);
int gpufs_graph_destroy(void* graph_exec);

/* ---- serving form: a resident request loop (no reference counterpart).
 * gpufs_serve_start launches `workgroups` persistent workgroups on the context
 * stream and starts the host service; each request (file id) is an
 * open -> gread -> whole-word scan -> close of that file, with the grep
 * semantics above.  gpufs_serve_run drives an open-loop arrival schedule:
 * request i becomes eligible arrival_s[i] seconds after the start and is
 * admitted (FIFO) when a request slot is free and its pages fit the frame
 * budget; results land in counts_out[i*n_words ...], latency_s_out[i] is
 * arrival -> completion seen by the host.  Page cache and file table stay
 * resident across requests and runs; the workgroups wait for requests however
 * long the server idles, until gpufs_serve_stop ends them.  gpufs_serve_run
 * returns -EIO if the server kernel has ended (it cannot serve any more). */
typedef struct gpufs_server gpufs_server;
typedef struct gpufs_serve_config {
    uint32_t workgroups;    /* persistent server workgroups */
    uint32_t threads;       /* threads per workgroup (multiple of 64, <= 1024) */
    uint32_t slots;         /* request slots: most requests in flight at once */
    uint32_t chunk_bytes;   /* scan window */
    uint32_t n_words;       /* dictionary words (32-byte records) */
    uint32_t frame_budget;  /* frames in-flight requests may need; 0 = 90% of the pool */
    uint32_t parts_max;     /* a large file's request runs on up to this many workgroups (<= 64); 0 = 8 */
    uint32_t min_part_kib;  /* ... each part at least this long (>= 4); 0 = 512: smaller files take one */
} gpufs_serve_config;
int gpufs_serve_start(gpufs_ctx* ctx, const gpufs_serve_config* cfg, const char* names_dev,
                      uint32_t name_stride, uint32_t n_files, const uint64_t* file_sizes_host,
                      const char* dict_dev, gpufs_server** out);
int gpufs_serve_run(gpufs_server* srv, const uint32_t* file_ids_host, const double* arrival_s_host,
                    uint32_t n_requests, uint32_t* counts_out_host, int32_t* status_out_host,
                    double* latency_s_out_host, double* elapsed_s_out);
int gpufs_serve_stop(gpufs_server* srv);
/* Page lookups (cache hits, misses) of the requests completed since
 * gpufs_serve_start, counted by the serving workgroups themselves.  While a
 * server runs, gpufs_stats_get, gpufs_stats_reset, gpufs_cache_drop and
 * gpufs_check_invariants return -EBUSY: their device copies cannot run
 * beside workgroups that hold every compute unit. */
int gpufs_serve_pages(gpufs_server* srv, uint64_t* page_hits, uint64_t* page_misses);
/* Requests completed since gpufs_serve_start that ran on more than one
 * workgroup (a file of at least 2 x min_part_kib is split into parts handed to
 * waiting workgroups), and the parts they ran as.  No counterpart in the
 * reference, whose samples have no serving loop. */
int gpufs_serve_parts(gpufs_server* srv, uint64_t* split_requests, uint64_t* parts);

/* Record sums through ActivePointers (FatPointer / gvmmap with a per-workgroup
 * TLB of tlb_lines = 0, 8 or 32 lines; libgpufs/fat_pointer.cu.h, pattern of
 * samples/mosaic/mosaic_gpufs.cu:1010-1060): wave v reads the records
 * order_dev[v*per_wave + j] (record_elems uint32 each, a multiple of 64) and
 * stores each record's 64-bit element sum in sums_dev[v*per_wave + j].
 * order_dev holds workgroups * (threads / 64) * per_wave record ids.
 * *status_dev (caller-zeroed) receives the first failure. */
int gpufs_launch_vm_records(gpufs_ctx* ctx, const char* path_dev, uint64_t record_elems,
                            const uint32_t* order_dev, uint64_t per_wave, uint64_t* sums_dev,
                            int32_t* status_dev, uint32_t tlb_lines, uint32_t workgroups,
                            uint32_t threads, void* stream);

/* The grep_text sample end to end (samples/grep_text/grep_text_gpu.cu:193-387,
 * driver samples/grep_text/fs.cu): host path strings of the query records
 * (32-byte NUL-padded, dict.gpu format), the output file and the database list
 * (names < 64 bytes separated by newlines, commas, spaces or tabs).  The
 * output is truncated, then every workgroup scans every database for its
 * slice of the records and appends "<word> <count> <database>\n" for each
 * (record, window) with a nonzero count through gwrite (O_GWRONCE); line
 * order follows timing, as in the reference.  Launches on `stream` (NULL =
 * context stream); the caller then serves it (gpufs_run_handler).  Returns
 * the number of workgroups launched (more than asked when a slice would
 * exceed 1024 records) or a negative code; *status_dev (caller-zeroed)
 * receives the first failure. */
int gpufs_grep_text(gpufs_ctx* ctx, const char* src_path, const char* out_path, const char* dbs_path,
                    uint32_t chunk_bytes, uint32_t workgroups, uint32_t threads, int32_t* status_dev,
                    void* stream);
/* The same, writing a fixed 100-byte binary record per (record, window) match
 * instead of a text line: the query's 32 bytes as stored, the count as a
 * little-endian int32, the database name NUL-padded to 64 bytes: the record
 * layout the reference's printer reads (samples/utils/display.cpp:16-25;
 * gpufs_amd.display prints such a file the same way). */
int gpufs_grep_text_records(gpufs_ctx* ctx, const char* src_path, const char* out_path, const char* dbs_path,
                            uint32_t chunk_bytes, uint32_t workgroups, uint32_t threads, int32_t* status_dev,
                            void* stream);

/* Record checksums through the wave-level mapping gmmap_warp / gmunmap_warp
 * (samples/mosaic/mosaic_gpufs.cu:612-651 pattern; device API
 * libgpufs/fs_calls.cu:483-683): every wave maps fixed-size records
 * (record_bytes divides the page size, multiple of 8) and stores the 64-bit
 * wrapping sum of each record's 8-byte words in sums_dev[r]; status_dev[r]
 * = 0 or a negative code.  Waves of one workgroup map records concurrently. */
int gpufs_launch_mmap_warp_records(gpufs_ctx* ctx, const char* path_dev, uint64_t record_bytes,
                                   uint64_t n_records, uint64_t* sums_dev, int32_t* status_dev,
                                   uint32_t workgroups, uint32_t threads, void* stream);

/* y = A x with A (rows x cols fp32, row-major) gmmap()ed page by page from
 * `path_dev` (samples/bigmatrix_mmap/bigmatrix_mmap.cu:66-168 pattern): workgroup
 * w owns rows [w*rows_per_wg, (w+1)*rows_per_wg).  x_dev has cols floats in
 * device memory.  The reduction order is fixed, so y is reproducible.
 * checksum_dev[w] = sum mod 2^64 of every mapped 8-byte word of the
 * workgroup's rows (order-independent check of the mapped bytes);
 * status_dev[w] = 0 or a negative error code. */
int gpufs_launch_gemv_mmap(gpufs_ctx* ctx, const char* path_dev, uint64_t rows, uint64_t cols,
                           const float* x_dev, float* y_dev, uint32_t rows_per_wg,
                           int32_t* status_dev, uint64_t* checksum_dev, uint32_t threads,
                           void* stream);

/* FNV-1a-64 of n_ranges consecutive ranges of a device buffer (the per-range
 * checksum of the sequential-read workload; no reference counterpart). */
int gpufs_fnv1a64_ranges(const void* data_dev, uint32_t n_ranges, uint64_t range_bytes,
                         uint64_t* out_dev, void* stream);

/* Generic single-call launchers used by the API tests. */
int gpufs_launch_read(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags,
                      uint64_t offset, uint64_t length, void* dst_dev,
                      int64_t* result_dev, uint32_t threads, void* stream);
int gpufs_launch_write(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags,
                       uint64_t offset, uint64_t length, const void* src_dev,
                       int32_t do_fsync, int64_t* result_dev, uint32_t threads, void* stream);
/* A script of device API calls made by one workgroup on one file, for the API
 * tests: record k is ops_dev[4k..4k+3] = {op, a, b, c} and its result goes to
 * results_dev[k].  op: 1 gopen(flags=a), 2 gread(off=a, len=b, buf_dev+c),
 * 3 gwrite(same), 4 gfsync, 5 gftruncate(size=a), 6 gclose, 7 fstat,
 * 8 single_thread_open(flags=a), 9 single_thread_ftruncate(size=a),
 * 10 single_thread_fsync, 11 single_thread_close, 12 gmmap of [a, a+b) filled
 * with byte c, gmsync, gmunmap (result: gmsync's), 13 gunlink.  Calls 8-11 and
 * 13 are made by thread 0 alone (reference device API,
 * libgpufs/fs_calls.cu.h:34-65). */
int gpufs_launch_api_script(gpufs_ctx* ctx, const char* path_dev, const int64_t* ops_dev,
                            uint32_t n_ops, void* buf_dev, int64_t* results_dev, uint32_t threads,
                            void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GPUFS_AMD_H */
