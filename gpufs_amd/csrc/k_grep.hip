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

__global__ __launch_bounds__(1024) void k_grep(const char* names, uint32_t stride, const char* dict, uint32_t nw,
                                               uint32_t chunk, uint32_t* counts_out, int32_t* status) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    const uint32_t f = blockIdx.x;
    grep_build_dict(S, dict, nw);
    int st = grep_one_file(S, names + (size_t)f * stride, nw, chunk, counts_out + (size_t)f * nw);
    if (gfs::tid() == 0) status[f] = st;
}

// Persistent form: each workgroup builds the dictionary table once, then pulls
// file indices from a device counter until the list is exhausted.
__global__ __launch_bounds__(1024) void k_grep_batched(const char* names, uint32_t stride, uint32_t nfiles,
                                                       const char* dict, uint32_t nw, uint32_t chunk,
                                                       uint32_t* counts_out, int32_t* status,
                                                       unsigned int* queue, uint64_t* times) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    __shared__ unsigned int cur;
    grep_build_dict(S, dict, nw);
    for (;;) {
        if (gfs::tid() == 0) cur = atomicAdd(queue, 1u);
        __syncthreads();
        const unsigned int f = gfs::uni(cur);
        __syncthreads();
        if (f >= nfiles) break;
        const uint64_t t0 = gfs::now_ticks();
        int st = grep_one_file(S, names + (size_t)f * stride, nw, chunk, counts_out + (size_t)f * nw);
        if (gfs::tid() == 0) {
            status[f] = st;
            if (times) {
                times[2 * (size_t)f] = t0;
                times[2 * (size_t)f + 1] = gfs::now_ticks();
            }
        }
        // A barrier between this thread-0 region and the next iteration's keeps
        // the compiler from threading the loop-invariant `tid == 0` test across
        // the back-edge, which would send the other lanes round the loop
        // without thread 0.
        __syncthreads();
    }
}

__global__ void k_zero_u32(unsigned int* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" int gpufs_launch_grep(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                                 const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes,
                                 uint32_t* counts_dev, int32_t* status_dev, uint32_t threads, void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev) return -EINVAL;
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    v = gpufs_bind(ctx);
    if (v) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    hipLaunchKernelGGL(k_grep, dim3(n_files), dim3(threads), 0, s, names_dev, name_stride, dict_dev, n_words,
                       chunk_bytes, counts_dev, status_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

static void launch_batched(hipStream_t s, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                           const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes, uint32_t* counts_dev,
                           int32_t* status_dev, uint32_t* queue_dev, uint32_t workgroups, uint32_t threads,
                           uint64_t* times_dev) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, (unsigned int*)queue_dev, 1u);
    hipLaunchKernelGGL(k_grep_batched, dim3(workgroups), dim3(threads), 0, s, names_dev, name_stride, n_files,
                       dict_dev, n_words, chunk_bytes, counts_dev, status_dev, (unsigned int*)queue_dev, times_dev);
}

extern "C" int gpufs_launch_grep_batched(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                         uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                         uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                         uint32_t* queue_dev, uint32_t workgroups, uint32_t threads,
                                         uint64_t* file_times_dev, void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev || !queue_dev || workgroups == 0) return -EINVAL;
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    launch_batched(s, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev, status_dev,
                   queue_dev, workgroups, threads, file_times_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// Capture (queue reset + persistent kernel) into an executable hipGraph.
extern "C" int gpufs_grep_batched_graph(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride,
                                        uint32_t n_files, const char* dict_dev, uint32_t n_words,
                                        uint32_t chunk_bytes, uint32_t* counts_dev, int32_t* status_dev,
                                        uint32_t* queue_dev, uint32_t workgroups, uint32_t threads,
                                        uint64_t* file_times_dev, void** graph_exec_out) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev || !queue_dev || !graph_exec_out ||
        workgroups == 0)
        return -EINVAL;
    int v = grep_args_ok(n_files, n_words, chunk_bytes, threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipStream_t s = (hipStream_t)gpufs_stream(ctx);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return -EIO;
    launch_batched(s, names_dev, name_stride, n_files, dict_dev, n_words, chunk_bytes, counts_dev, status_dev,
                   queue_dev, workgroups, threads, file_times_dev);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s, &g) != hipSuccess || !g) return -EIO;
    hipGraphExec_t e = nullptr;
    hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
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
