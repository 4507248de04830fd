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
#include "gpufs.h"

#define GREP_MAX_CHUNK 65536
#define GREP_MAX_WORDS 1024
#define GREP_HT 4096

struct grep_lds {
    uint8_t corpus[GREP_MAX_CHUNK + 64];
    uint8_t words[GREP_MAX_WORDS][32];
    uint32_t ht[GREP_HT];            // canonical word index + 1
    uint32_t counts[GREP_MAX_WORDS];
    uint16_t canon[GREP_MAX_WORDS];  // 0xFFFF: word can never match
    uint8_t wlen[GREP_MAX_WORDS];
    int ov;
    int status;
};

__device__ __forceinline__ bool g_alpha(uint32_t c) {
    return (c >= 65 && c <= 90) || (c >= 97 && c <= 122) || c == 95;
}

__device__ __forceinline__ uint32_t g_hash_step(uint32_t h, uint32_t c) { return (h ^ c) * 16777619u; }

__device__ inline void grep_build_dict(grep_lds& S, const char* dict, uint32_t nw) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    for (uint32_t i = t; i < GREP_HT; i += T) S.ht[i] = 0;
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t len = 0;
        bool ok = true;
        for (uint32_t i = 0; i < 32; ++i) {
            uint8_t c = (uint8_t)dict[(size_t)q * 32 + i];
            S.words[q][i] = c;
            if (len == i && c != 0) {
                len = i + 1;
                if (!g_alpha(c)) ok = false;
            }
        }
        // bytes after the first NUL are ignored (the C string ends there)
        for (uint32_t i = len; i < 32; ++i) S.words[q][i] = 0;
        S.wlen[q] = (uint8_t)len;
        S.counts[q] = 0;
        S.canon[q] = (ok && len > 0) ? (uint16_t)q : (uint16_t)0xFFFF;
    }
    __syncthreads();
    for (uint32_t q = t; q < nw; q += T) {
        if (S.canon[q] == 0xFFFF) continue;
        uint32_t h = 2166136261u;
        for (uint32_t i = 0; i < S.wlen[q]; ++i) h = g_hash_step(h, S.words[q][i]);
        uint32_t pos = h & (GREP_HT - 1);
        for (;;) {
            uint32_t old = atomicCAS(&S.ht[pos], 0u, q + 1);
            if (old == 0) break;
            uint32_t r = old - 1;
            bool same = S.wlen[r] == S.wlen[q];
            for (uint32_t i = 0; same && i < S.wlen[q]; ++i) same = S.words[r][i] == S.words[q][i];
            if (same) { S.canon[q] = (uint16_t)r; break; }
            pos = (pos + 1) & (GREP_HT - 1);
        }
    }
    __syncthreads();
    // a duplicate may have been linked to another duplicate that itself was
    // linked later; follow links to the word that owns the table entry
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t c = S.canon[q];
        if (c == 0xFFFF) continue;
        while (S.canon[c] != c) c = S.canon[c];
        S.canon[q] = (uint16_t)c;
    }
    __syncthreads();
}

// Count tokens of window [0, dsz) of S.corpus.  corpus[dsz-1] is never a word
// character, so every token ends inside the window.
__device__ inline void grep_scan_window(grep_lds& S, uint32_t dsz) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    const uint32_t per = (dsz + T - 1) / T;
    uint32_t p = t * per;
    const uint32_t end = min(dsz, p + per);
    const uint8_t* c = S.corpus;
    while (p < end) {
        uint32_t ch = c[p];
        if (!g_alpha(ch) || (p > 0 && g_alpha(c[p - 1]))) { ++p; continue; }
        uint32_t h = 2166136261u;
        uint32_t q = p;
        while (g_alpha(c[q])) {
            if (q - p < 32) h = g_hash_step(h, c[q]);
            ++q;
        }
        const uint32_t len = q - p;
        if (len <= 32) {
            uint32_t pos = h & (GREP_HT - 1);
            for (;;) {
                uint32_t e = S.ht[pos];
                if (!e) break;
                uint32_t w = e - 1;
                if (S.wlen[w] == len) {
                    bool same = true;
                    for (uint32_t i = 0; same && i < len; ++i) same = S.words[w][i] == c[p + i];
                    if (same) { atomicAdd(&S.counts[w], 1u); break; }
                }
                pos = (pos + 1) & (GREP_HT - 1);
            }
        }
        p = q;
    }
}

__global__ __launch_bounds__(1024) void k_grep(const char* names, uint32_t stride, const char* dict, uint32_t nw,
                                               uint32_t chunk, uint32_t* counts_out, int32_t* status) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    const uint32_t f = blockIdx.x;
    grep_build_dict(S, dict, nw);
    int fd = gopen(names + (size_t)f * stride, O_GRDONLY);
    if (fd < 0) {
        if (gfs::tid() == 0) status[f] = fd;
        return;
    }
    const uint64_t size = fstat(fd);
    if (gfs::tid() == 0) S.status = 0;
    uint64_t cursor = 0;
    while (cursor < size) {
        const bool last = size - cursor < (uint64_t)chunk + 32;
        const uint32_t left = last ? (uint32_t)(size - cursor) : chunk + 32;
        size_t got = gread(fd, cursor, left, S.corpus);
        if (got != left) {
            if (gfs::tid() == 0) S.status = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
            break;
        }
        if (gfs::tid() == 0) {
            S.corpus[left] = 0;
            int ov = 0;
            if (!last) while (ov < 32 && g_alpha(S.corpus[chunk + ov])) ++ov;
            S.ov = ov;
        }
        __syncthreads();
        const uint32_t ov = (uint32_t)S.ov;
        const uint32_t dsz = last ? left + 1 : chunk + ov + 1;
        cursor += (uint64_t)chunk + ov;
        grep_scan_window(S, dsz);
        __syncthreads();
    }
    __syncthreads();
    int c = gclose(fd);
    for (uint32_t q = gfs::tid(); q < nw; q += gfs::nthreads()) {
        uint32_t k = S.canon[q];
        counts_out[(size_t)f * nw + q] = k == 0xFFFF ? 0u : S.counts[k];
    }
    if (gfs::tid() == 0) status[f] = S.status ? S.status : c;
}

extern "C" int gpufs_launch_grep(gpufs_ctx* ctx, const char* names_dev, uint32_t name_stride, uint32_t n_files,
                                 const char* dict_dev, uint32_t n_words, uint32_t chunk_bytes,
                                 uint32_t* counts_dev, int32_t* status_dev, uint32_t threads, void* stream) {
    if (!ctx || !names_dev || !dict_dev || !counts_dev || !status_dev || n_files == 0) return -EINVAL;
    if (n_words == 0 || n_words > GREP_MAX_WORDS) return -EINVAL;
    if (chunk_bytes == 0 || chunk_bytes > GREP_MAX_CHUNK) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    int v = gpufs_bind(ctx);
    if (v) return v;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)gpufs_stream(ctx);
    hipLaunchKernelGGL(k_grep, dim3(n_files), dim3(threads), 0, s, names_dev, name_stride, dict_dev, n_words,
                       chunk_bytes, counts_dev, status_dev);
    return hipGetLastError() == hipSuccess ? 0 : -EIO;
}
