// k_gread.hip -- gread/gwrite harness kernels and their C-ABI launchers.
//
// gpufs_launch_gread_ranges is the sequential-read workload (one gopen + one
// gread of a disjoint byte range per workgroup, the access pattern of
// samples/memcopy/cp.cu:73-108 scaled to large ranges); gpufs_launch_read /
// gpufs_launch_write run a single open/read|write/close sequence in one
// workgroup for the API tests; gpufs_launch_gwrite_ranges is the sequential
// write workload (one gopen + gwrite + gfsync + gclose per workgroup).
#include "gpufs.h"

static int check_threads(uint32_t threads) {
    if (threads < 64 || threads > 1024 || (threads % 64)) return -EINVAL;
    return 0;
}
static hipStream_t pick_stream(gpufs_ctx* ctx, void* stream) {
    return stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
}

//
// The reopen of a cached file overlaps the copy.  The workgroup that takes
// the closed entry to OPENING posts the host open and copies its range from
// the cached pages (copy_cached) before completing the open; the others find
// the entry OPENING, take no open, copy their ranges the same way and then
// confirm that an open succeeded after they looked and found the file as its
// last open left it (same version and size: a reopen that finds the file
// changed bumps the version).  Otherwise, or when a page was not cached, a
// workgroup reads its range the usual way (gopen, gread, gclose),
// overwriting what it copied.  So a launch costs one host open and close
// instead of 256 waits for the open and 512 atomics on one state word.
__global__ __launch_bounds__(1024) void k_gread_ranges(const char* path, uint64_t range, uint8_t* dst,
                                                       int64_t* out) {
    using namespace gfs;
    __shared__ gfs_open_pend pend;
    __shared__ uint64_t snap[2];   // the entry's size, version (as its last open left them)
    __shared__ uint32_t okf;
    gfs_dev* g = gfs_g;
    const uint32_t t = tid();
    const uint64_t off = (uint64_t)blockIdx.x * range;
    __syncthreads();
    if (t == 0) {
        gfs_hist.fd = -1;
        gfs_memo.fd = -1;
        pend.pending = 0;
        pend.peek_ok = 1;
        const int r0 = open_single(path, O_GRDONLY, &pend);
        if (r0 >= 0 && pend.pending) {
            gfs_file* e = &g->files[r0];
            snap[0] = e->size;   // stable while OPENING: only the opener writes it, at its end
            snap[1] = fst_version(ld_rlx(&e->st));
            if (pend.pending == 1) open_post(&pend);
        }
        gfs_ctl.ret = r0;
    }
    __syncthreads();
    int fd = (int)uni(gfs_ctl.ret);
    const uint32_t mode = fd >= 0 ? uni(pend.pending) : 0u;
    if (fd < 0) {
        if (t == 0) out[blockIdx.x] = fd;
        return;
    }
    size_t r = 0;
    bool done = false;
    if (mode) {
        const uint64_t S = uni(snap[0]);
        const uint64_t n = off < S ? (S - off < range ? S - off : range) : 0;
        // the opener completes its open halfway through its range (the host
        // has answered by then), so that it ends about when the others do,
        // which wait for that open at their end
        const uint64_t n1 = mode == 1 ? ((n >> 1) & ~(g->page_size - 1)) : n;
        int sst = copy_cached(g, fd, off, n1, dst + off);
        if (t == 0) {
            uint32_t ok = 0;
            gfs_file* e = &g->files[fd];
            if (mode == 1) {
                const int rr = open_finish(&pend);
                gfs_ctl.ret = rr;
                if (rr >= 0) ok = fst_version(ld_rlx(&e->st)) == (uint32_t)snap[1] && e->size == S;
            } else {
                // An open that succeeded after the look (bounded wait).  The
                // opener counts `opened` right after the state turns OPEN,
                // and long before its close: OPEN or CLOSING with the old
                // count is that moment, CLOSED with the old count (read after
                // the state, with an acquire) is a failed open.
                const uint64_t w0 = now_ticks();
                for (;;) {
                    uint32_t o = __hip_atomic_load(&e->opened, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bool seen = o != pend.opened0;
                    if (!seen) {
                        const uint32_t s = fst_state(ld_acq(&e->st));
                        if (s == GFS_FS_CLOSED || s == GFS_FS_FREE) {
                            o = __hip_atomic_load(&e->opened, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if (o == pend.opened0) break;   // that open failed
                            seen = true;
                        }
                    }
                    if (seen) {
                        fence_acq_agent();
                        ok = fst_version(ld_rlx(&e->st)) == (uint32_t)snap[1] && e->size == S;
                        break;
                    }
                    if (now_ticks() - w0 >= g->wait_ticks) break;
                    __builtin_amdgcn_s_sleep(8);
                }
                gfs_ctl.ret = -1;   // no open is held
            }
            okf = ok && sst == 0;
            count(&g->ctr.spec_parts, 1);
            if (!ok && (mode == 2 || (int)gfs_ctl.ret >= 0)) count(&g->ctr.spec_changed, 1);
            else if (ok && sst) count(&g->ctr.spec_rescans, 1);
        }
        __syncthreads();
        fd = (int)uni(gfs_ctl.ret);
        if (mode == 1 && fd < 0) {
            if (t == 0) out[blockIdx.x] = fd;
            return;
        }
        if (mode == 1 && uni(okf) && n1 < n) {
            sst = copy_cached(g, fd, off + n1, n - n1, dst + off + n1);
            if (t == 0) {
                okf = sst == 0;
                if (sst) count(&g->ctr.spec_rescans, 1);
            }
            __syncthreads();
        }
        if (uni(okf)) {
            r = n;
            done = true;
        }
    }
    if (!done) {
        if (fd < 0) fd = gopen(path, O_GRDONLY);
        if (fd < 0) {
            if (t == 0) out[blockIdx.x] = fd;
            return;
        }
        [[clang::always_inline]] r = gread(fd, off, range, dst + off);
    }
    const int c = fd >= 0 ? gclose(fd) : 0;
    if (t == 0) out[blockIdx.x] = (int64_t)r < 0 ? (int64_t)r : (c < 0 ? (int64_t)c : (int64_t)r);
}

extern "C" int gpufs_launch_gread_ranges(gpufs_ctx* ctx, const char* path_dev, uint32_t n_ranges,
                                         uint64_t range_bytes, void* dst_dev, int64_t* bytes_out_dev,
                                         uint32_t threads, void* stream) {
    if (!ctx || !path_dev || !dst_dev || !bytes_out_dev || n_ranges == 0) return -EINVAL;
    int v = check_threads(threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipLaunchKernelGGL(k_gread_ranges, dim3(n_ranges), dim3(threads), 0, pick_stream(ctx, stream), path_dev,
                       range_bytes, (uint8_t*)dst_dev, bytes_out_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// The write-side twin of k_gread_ranges: workgroup i opens the file, writes
// range i of `src` at offset i * range, optionally gfsyncs it, and closes it;
// out[i] = bytes written or the first error (gwrite, then gfsync, then gclose).
__global__ __launch_bounds__(1024) void k_gwrite_ranges(const char* path, int flags, uint64_t range,
                                                        const uint8_t* src, int do_fsync, int64_t* out) {
    int fd = gopen(path, flags);
    if (fd < 0) {
        if (gfs::tid() == 0) out[blockIdx.x] = fd;
        return;
    }
    const size_t r = gwrite(fd, (size_t)blockIdx.x * range, range, (uchar*)(src + (size_t)blockIdx.x * range));
    const int s = (do_fsync && (int64_t)r >= 0) ? gfsync(fd) : 0;
    const int c = gclose(fd);
    if (gfs::tid() == 0) out[blockIdx.x] = (int64_t)r < 0 ? (int64_t)r : (s < 0 ? (int64_t)s : (c < 0 ? (int64_t)c : (int64_t)r));
}

extern "C" int gpufs_launch_gwrite_ranges(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags,
                                          uint32_t n_ranges, uint64_t range_bytes, const void* src_dev,
                                          int32_t do_fsync, int64_t* bytes_out_dev, uint32_t threads,
                                          void* stream) {
    if (!ctx || !path_dev || !src_dev || !bytes_out_dev || n_ranges == 0) return -EINVAL;
    if ((open_flags & O_ACCMODE) == O_RDONLY) return -EINVAL;
    int v = check_threads(threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipLaunchKernelGGL(k_gwrite_ranges, dim3(n_ranges), dim3(threads), 0, pick_stream(ctx, stream), path_dev,
                       open_flags, range_bytes, (const uint8_t*)src_dev, do_fsync, bytes_out_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

__global__ __launch_bounds__(1024) void k_read_one(const char* path, int flags, uint64_t off, uint64_t len,
                                                   uint8_t* dst, int64_t* out) {
    int fd = gopen(path, flags);
    if (fd < 0) {
        if (gfs::tid() == 0) out[0] = fd;
        return;
    }
    size_t r = gread(fd, off, len, dst);
    size_t sz = fstat(fd);
    int c = gclose(fd);
    if (gfs::tid() == 0) {
        out[0] = (int64_t)r;
        out[1] = (int64_t)sz;
        out[2] = c;
    }
}

extern "C" int gpufs_launch_read(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags, uint64_t offset,
                                 uint64_t length, void* dst_dev, int64_t* result_dev, uint32_t threads,
                                 void* stream) {
    if (!ctx || !path_dev || !result_dev) return -EINVAL;
    int v = check_threads(threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipLaunchKernelGGL(k_read_one, dim3(1), dim3(threads), 0, pick_stream(ctx, stream), path_dev, open_flags,
                       offset, length, (uint8_t*)dst_dev, result_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

__global__ __launch_bounds__(1024) void k_write_one(const char* path, int flags, uint64_t off, uint64_t len,
                                                    const uint8_t* src, int do_fsync, int64_t* out) {
    int fd = gopen(path, flags);
    if (fd < 0) {
        if (gfs::tid() == 0) out[0] = fd;
        return;
    }
    size_t r = gwrite(fd, off, len, (uchar*)src);
    int s = do_fsync ? gfsync(fd) : 0;
    size_t sz = fstat(fd);
    int c = gclose(fd);
    if (gfs::tid() == 0) {
        out[0] = (int64_t)r;
        out[1] = (int64_t)sz;
        out[2] = c;
        out[3] = s;
    }
}

extern "C" int gpufs_launch_write(gpufs_ctx* ctx, const char* path_dev, int32_t open_flags, uint64_t offset,
                                  uint64_t length, const void* src_dev, int32_t do_fsync, int64_t* result_dev,
                                  uint32_t threads, void* stream) {
    if (!ctx || !path_dev || !result_dev) return -EINVAL;
    int v = check_threads(threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipLaunchKernelGGL(k_write_one, dim3(1), dim3(threads), 0, pick_stream(ctx, stream), path_dev, open_flags,
                       offset, length, (const uint8_t*)src_dev, do_fsync, result_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// FNV-1a-64 of each of n_ranges consecutive ranges of a device buffer (one
// lane per range; the hash is byte-serial).  Used to check multi-GiB gread
// outputs against the host without copying them back.
__global__ void k_fnv_ranges(const uint8_t* p, uint32_t n, uint64_t range, uint64_t* out) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint8_t* b = p + (uint64_t)r * range;
    uint64_t h = 0xcbf29ce484222325ull;
    uint64_t i = 0;
    const uint64_t words = ((uintptr_t)b & 7) ? 0 : range / 8;
    for (; i < words; ++i) {
        uint64_t w = ((const uint64_t*)b)[i];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            h ^= (w >> (8 * k)) & 0xFF;
            h *= 0x100000001b3ull;
        }
    }
    for (i *= 8; i < range; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    out[r] = h;
}

extern "C" int gpufs_fnv1a64_ranges(const void* data_dev, uint32_t n_ranges, uint64_t range_bytes,
                                    uint64_t* out_dev, void* stream) {
    if (!data_dev || !out_dev || n_ranges == 0) return -EINVAL;
    hipLaunchKernelGGL(k_fnv_ranges, dim3((n_ranges + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                       (const uint8_t*)data_dev, n_ranges, range_bytes, out_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// A script of device API calls run by one workgroup on one file, for the API
// tests: record k is ops[4k .. 4k+3] = {op, a, b, c}, its result goes to
// results[k].  Collective calls are made by every thread; the single-thread
// calls by thread 0 alone while the others wait at a barrier, the way the
// reference's samples call them (samples/grep_text/grep_text_gpu.cu:258-268).
// `buf` is the device buffer GREAD / GWRITE use at byte offset c.
enum api_op : int64_t {
    API_GOPEN = 1,        // a = flags -> fd (kept as the current descriptor)
    API_GREAD = 2,        // a = offset, b = length, c = buffer offset -> bytes
    API_GWRITE = 3,       // same -> bytes
    API_GFSYNC = 4,
    API_GFTRUNCATE = 5,   // a = size
    API_GCLOSE = 6,
    API_FSTAT = 7,
    API_ST_OPEN = 8,      // a = flags -> fd
    API_ST_FTRUNCATE = 9, // a = size
    API_ST_FSYNC = 10,
    API_ST_CLOSE = 11,
    API_GMMAP_FILL = 12,  // a = offset, b = length (inside one page), c = byte: gmmap, fill, gmsync, gmunmap -> gmsync's result
    API_GUNLINK = 13,
};

__global__ __launch_bounds__(1024) void k_api_script(const char* path, const int64_t* ops, uint32_t n_ops,
                                                     uint8_t* buf, int64_t* results) {
    __shared__ int64_t st_ret;
    int fd = -1;
    const uint32_t t = gfs::tid();
    for (uint32_t k = 0; k < n_ops; ++k) {
        // every thread reads the same record; readfirstlane makes the branches scalar
        const int64_t op = gfs::uni(ops[4 * k]), a = gfs::uni(ops[4 * k + 1]), b = gfs::uni(ops[4 * k + 2]),
                      c = gfs::uni(ops[4 * k + 3]);
        int64_t r = 0;
        switch (op) {
        case API_GOPEN: r = fd = gopen(path, (int)a); break;
        case API_GREAD: r = (int64_t)gread(fd, (size_t)a, (size_t)b, buf + c); break;
        case API_GWRITE: r = (int64_t)gwrite(fd, (size_t)a, (size_t)b, buf + c); break;
        case API_GFSYNC: r = gfsync(fd); break;
        case API_GFTRUNCATE: r = gftruncate(fd, (int)a); break;
        case API_GCLOSE: r = gclose(fd); break;
        case API_FSTAT: r = (int64_t)fstat(fd); break;
        case API_GMMAP_FILL: {
            volatile uint8_t* p = (volatile uint8_t*)gmmap(nullptr, (size_t)b, 0, 0, fd, (off_t)a);
            if (p == (volatile uint8_t*)GMAP_FAILED) {
                r = -EIO;
                break;
            }
            for (int64_t i = t; i < b; i += gfs::nthreads()) p[i] = (uint8_t)c;
            __syncthreads();
            r = gmsync(p, (size_t)b, 0);
            const int u = gmunmap(p, (size_t)b);
            if (r == 0) r = u;
            break;
        }
        default: {
            // single-thread calls: thread 0 only
            __syncthreads();
            if (t == 0) {
                int64_t v = -EINVAL;
                if (op == API_ST_OPEN) v = single_thread_open(path, (int)a);
                else if (op == API_ST_FTRUNCATE) v = single_thread_ftruncate(fd, (int)a);
                else if (op == API_ST_FSYNC) v = single_thread_fsync(fd);
                else if (op == API_ST_CLOSE) v = single_thread_close(fd);
                else if (op == API_GUNLINK) v = (int32_t)gunlink((char*)path);
                st_ret = v;
            }
            __syncthreads();
            r = gfs::uni((uint64_t)st_ret);
            if (op == API_ST_OPEN) fd = (int)r;
            __syncthreads();
        }
        }
        if (t == 0) results[k] = r;
        // keeps this thread-0 region apart from the next record's (see k_grep_batched)
        __syncthreads();
    }
}

extern "C" int gpufs_launch_api_script(gpufs_ctx* ctx, const char* path_dev, const int64_t* ops_dev,
                                       uint32_t n_ops, void* buf_dev, int64_t* results_dev, uint32_t threads,
                                       void* stream) {
    if (!ctx || !path_dev || !ops_dev || !results_dev || n_ops == 0) return -EINVAL;
    int v = check_threads(threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    hipLaunchKernelGGL(k_api_script, dim3(1), dim3(threads), 0, pick_stream(ctx, stream), path_dev, ops_dev, n_ops,
                       (uint8_t*)buf_dev, results_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}
