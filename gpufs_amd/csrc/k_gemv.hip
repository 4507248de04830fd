// k_gemv.hip -- y = A x with A gmmap()ed page by page through the page cache
// (the access pattern of samples/bigmatrix_mmap/bigmatrix_mmap.cu:66-168:
// each workgroup owns a block of rows, maps the page holding the next piece of
// its row, consumes it, unmaps it).  Differences by design: x lives in device
// memory, the reduction order is fixed (per-thread partials in a fixed column
// assignment, then a fixed butterfly and a fixed cross-wave sum) so y is
// reproducible run to run, and every mapped 8-byte word is added into a
// per-workgroup 64-bit checksum so the mapped bytes can be verified exactly.
#include "gpufs.h"
#include "gpufs_vm.h"

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(1024) void k_gemv_mmap(const char* path, uint64_t rows, uint64_t cols, const float* x,
                                                    float* y, uint32_t rows_per_wg, int32_t* status,
                                                    unsigned long long* wsum) {
    __shared__ float part[16];
    __shared__ unsigned long long upart[16];
    __shared__ int st;
    const uint32_t t = gfs::tid(), T = gfs::nthreads(), w = t >> 6, nw = (T + 63) >> 6;
    int fd = gopen(path, O_GRDONLY);
    if (fd < 0) {
        if (t == 0) status[blockIdx.x] = fd;
        return;
    }
    if (t == 0) st = (fstat(fd) < rows * cols * 4) ? -EINVAL : 0;
    __syncthreads();
    const uint64_t PS = gfs_g->page_size;
    const uint64_t rb = cols * 4;
    unsigned long long csum = 0;
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_wg;
    for (uint64_t r = r0; r < r0 + rows_per_wg && r < rows && gfs::uni(st) == 0; ++r) {
        float acc = 0.f;
        const uint64_t row_end = (r + 1) * rb;
        for (uint64_t off = r * rb; off < row_end;) {
            const uint64_t pstart = off & ~(PS - 1);
            const uint64_t len = (pstart + PS < row_end ? pstart + PS : row_end) - off;
            volatile void* m = gmmap(nullptr, len, 0, O_GRDONLY, fd, (off_t)off);
            if (m == GMAP_FAILED) {
                if (t == 0) st = -EIO;
                break;
            }
            const float* p = (const float*)(uintptr_t)m;
            const float* xx = x + (off - r * rb) / 4;
            const uint64_t n4 = len / 16;
            for (uint64_t i = t; i < n4; i += T) {
                float4 a = ((const float4*)p)[i];
                float4 b = ((const float4*)xx)[i];
                acc = fmaf(a.x, b.x, acc);
                acc = fmaf(a.y, b.y, acc);
                acc = fmaf(a.z, b.z, acc);
                acc = fmaf(a.w, b.w, acc);
                const unsigned long long* u = (const unsigned long long*)&a;
                csum += u[0] + u[1];
            }
            for (uint64_t i = n4 * 4 + t; i < len / 4; i += T) {
                acc = fmaf(p[i], xx[i], acc);
                csum += (unsigned long long)((const uint32_t*)p)[i];
            }
            gmunmap(m, len);
            off += len;
        }
        float s = wave_sum(acc);
        if ((t & 63) == 0) part[w] = s;
        __syncthreads();
        if (t == 0) {
            float tot = 0.f;
            for (uint32_t k = 0; k < nw; ++k) tot += part[k];
            y[r] = tot;
        }
        __syncthreads();
    }
    unsigned long long cs = wave_sum_u64(csum);
    if ((t & 63) == 0) upart[w] = cs;
    __syncthreads();
    int c = gclose(fd);
    if (t == 0) {
        unsigned long long tot = 0;
        for (uint32_t k = 0; k < nw; ++k) tot += upart[k];
        wsum[blockIdx.x] = tot;
        status[blockIdx.x] = st ? st : c;
    }
}

extern "C" int gpufs_launch_gemv_mmap(gpufs_ctx* ctx, const char* path_dev, uint64_t rows, uint64_t cols,
                                      const float* x_dev, float* y_dev, uint32_t rows_per_wg,
                                      int32_t* status_dev, uint64_t* checksum_dev, uint32_t threads,
                                      void* stream) {
    if (!ctx || !path_dev || !x_dev || !y_dev || !status_dev || !checksum_dev) return -EINVAL;
    if (rows == 0 || cols == 0 || rows_per_wg == 0 || (cols % 4)) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    int v = gpufs_bind(ctx);
    if (v) return v;
    const uint64_t nwg = (rows + rows_per_wg - 1) / rows_per_wg;
    if (nwg > 0x7FFFFFFF) return -EINVAL;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    hipLaunchKernelGGL(k_gemv_mmap, dim3((uint32_t)nwg), dim3(threads), 0, s, path_dev, rows, cols, x_dev, y_dev,
                       rows_per_wg, status_dev, (unsigned long long*)checksum_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// Record checksums through wave-level mappings, the access pattern of
// samples/mosaic/mosaic_gpufs.cu:612-651 (each warp maps one fixed-size record
// with gmmap_warp, reads it and unmaps it; waves of a workgroup work on
// different records at the same time and may meet on one page).  Record r
// covers bytes [r*R, r*R+R) of the file and is handled by the wave whose
// global index is r modulo the number of waves; its 64-bit wrapping sum of
// 8-byte words goes to sums[r], 0 or a negative code to status[r].
__global__ __launch_bounds__(1024) void k_mmap_warp_records(const char* path, uint64_t rec, uint64_t nrec,
                                                            unsigned long long* sums, int32_t* status) {
    const uint32_t ln = gfs::lane(), w = gfs::wave(), nw = (gfs::nthreads() + 63) >> 6;
    const uint64_t gw = (uint64_t)blockIdx.x * nw + w, tw = (uint64_t)gridDim.x * nw;
    int fd = gopen(path, O_GRDONLY);
    if (fd < 0) {
        for (uint64_t r = gw; r < nrec; r += tw)
            if (ln == 0) status[r] = fd;
        return;
    }
    for (uint64_t r = gw; r < nrec; r += tw) {
        volatile unsigned long long* m =
            (volatile unsigned long long*)gmmap_warp(nullptr, rec, 0, O_GRDONLY, fd, (off_t)(r * rec));
        if (m == (volatile unsigned long long*)GMAP_FAILED) {
            if (ln == 0) status[r] = -EIO;
            continue;
        }
        unsigned long long acc = 0;
        for (uint64_t k = ln; k < rec / 8; k += 64) acc += m[k];
        acc = wave_sum_u64(acc);
        gmunmap_warp((volatile void*)(m + ln), rec);   // any address inside the page
        if (ln == 0) {
            sums[r] = acc;
            status[r] = 0;
        }
    }
    gclose(fd);
}

extern "C" int gpufs_launch_mmap_warp_records(gpufs_ctx* ctx, const char* path_dev, uint64_t record_bytes,
                                              uint64_t n_records, uint64_t* sums_dev, int32_t* status_dev,
                                              uint32_t workgroups, uint32_t threads, void* stream) {
    if (!ctx || !path_dev || !sums_dev || !status_dev || n_records == 0 || workgroups == 0) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    gpufs_config cc;
    gpufs_get_config(ctx, &cc);
    if (record_bytes == 0 || record_bytes % 8 || record_bytes > cc.page_size || cc.page_size % record_bytes)
        return -EINVAL;
    int v = gpufs_bind(ctx);
    if (v) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    hipLaunchKernelGGL(k_mmap_warp_records, dim3(workgroups), dim3(threads), 0, s, path_dev, record_bytes, n_records,
                       (unsigned long long*)sums_dev, status_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}

// Record sums through ActivePointers, the pattern of
// samples/mosaic/mosaic_gpufs.cu:1010-1060 (one FatPointer per thread over
// the whole file; for each record: moveTo the record, += lane, then steps of
// 64 elements, dereferencing each).  Wave v handles order slots
// [v*per_wave, v*per_wave + per_wave); slot j's record is order[j], and the
// 64-bit sum of its uint32 elements goes to sums[j].  A TLB of N lines per
// workgroup (N = 0: every page mapped directly).
template <int N>
__global__ __launch_bounds__(1024) void k_vm_records(const char* path, uint64_t rec_elems, const uint32_t* order,
                                                     uint64_t per_wave, unsigned long long* sums, int32_t* status) {
    __shared__ TLB<N> tlb;
    tlb.init();
    const uint32_t ln = gfs::lane(), w = gfs::wave(), nw = (gfs::nthreads() + 63) >> 6;
    const uint64_t gw = (uint64_t)blockIdx.x * nw + w;
    int fd = gopen(path, O_GRDONLY);
    if (fd < 0) {
        if (gfs::tid() == 0) atomicMin(status, fd);
        return;
    }
    {
        FatPointer<uint32_t, N> p = gvmmap<uint32_t, N>(nullptr, ~0ull >> 1, 0, O_GRDONLY, fd, 0, &tlb);
        for (uint64_t j = 0; j < per_wave; ++j) {
            const uint64_t slot = gw * per_wave + j;
            const uint64_t r = order[slot];
            p.moveTo(r * rec_elems * sizeof(uint32_t));
            p += ln;
            unsigned long long acc = 0;
            for (uint64_t k = 0; k < rec_elems; k += 64) {
                if (k) p += 64;
                acc += *p;
            }
            acc = wave_sum_u64(acc);
            if (ln == 0) sums[slot] = acc;
        }
    }   // the pointer is gone: what it held is dropped
    tlb.release();
    const int c = gclose(fd);
    if (c && gfs::tid() == 0) atomicMin(status, c);
}

extern "C" int gpufs_launch_vm_records(gpufs_ctx* ctx, const char* path_dev, uint64_t record_elems,
                                       const uint32_t* order_dev, uint64_t per_wave, uint64_t* sums_dev,
                                       int32_t* status_dev, uint32_t tlb_lines, uint32_t workgroups,
                                       uint32_t threads, void* stream) {
    if (!ctx || !path_dev || !order_dev || !sums_dev || !status_dev || per_wave == 0 || workgroups == 0)
        return -EINVAL;
    if (record_elems == 0 || record_elems % 64) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    int v = gpufs_bind(ctx);
    if (v) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    unsigned long long* sums = (unsigned long long*)sums_dev;
    switch (tlb_lines) {
    case 0: hipLaunchKernelGGL(k_vm_records<0>, dim3(workgroups), dim3(threads), 0, s, path_dev, record_elems,
                               order_dev, per_wave, sums, status_dev); break;
    case 8: hipLaunchKernelGGL(k_vm_records<8>, dim3(workgroups), dim3(threads), 0, s, path_dev, record_elems,
                               order_dev, per_wave, sums, status_dev); break;
    case 32: hipLaunchKernelGGL(k_vm_records<32>, dim3(workgroups), dim3(threads), 0, s, path_dev, record_elems,
                                order_dev, per_wave, sums, status_dev); break;
    default: return -EINVAL;
    }
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}
