// grep_scan.h -- the whole-word scan shared by the plain, batched and serving
// forms (see k_grep.hip for the semantics and the reference lines).
#pragma once
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

// One file, the whole chunk loop; S.counts is zeroed here.  Returns 0 or a
// negative code (workgroup-uniform).
__device__ inline int grep_one_file(grep_lds& S, const char* name, uint32_t nw, uint32_t chunk, uint32_t* row) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
    int fd = gopen(name, O_GRDONLY);
    if (fd < 0) {
        for (uint32_t q = t; q < nw; q += T) row[q] = 0;
        return fd;
    }
    const uint64_t size = fstat(fd);
    if (t == 0) S.status = 0;
    uint64_t cursor = 0;
    while (cursor < size) {
        const bool last = size - cursor < (uint64_t)chunk + 32;
        const uint32_t left = last ? (uint32_t)(size - cursor) : chunk + 32;
        size_t got = gread(fd, cursor, left, S.corpus);
        if (got != left) {
            if (t == 0) S.status = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
            break;
        }
        if (t == 0) {
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
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t k = S.canon[q];
        row[q] = k == 0xFFFF ? 0u : S.counts[k];
    }
    int st = S.status ? S.status : c;
    __syncthreads();
    return st;
}

static int grep_args_ok(uint32_t n_files, uint32_t n_words, uint32_t chunk_bytes, uint32_t threads) {
    if (n_files == 0 || n_words == 0 || n_words > GREP_MAX_WORDS) return -EINVAL;
    if (chunk_bytes == 0 || chunk_bytes > GREP_MAX_CHUNK) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    return 0;
}

