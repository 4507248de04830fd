// grep_scan.h -- the whole-word scan shared by the plain, batched and serving
// forms (see k_grep.hip for the semantics and the reference lines).
//
// Window scan: each thread owns 64-byte segments of the LDS window.  A
// segment is read with four ds_read_b128, word bytes ([A-Za-z_]) are flagged
// four at a time with SWAR arithmetic into a 64-bit mask, token starts are
// mask & ~(mask << 1 | previous byte).  A token is skipped at once when no
// query word has its length; otherwise its first eight bytes (byte-aligned
// from three LDS dwords, bytes past the token zeroed) and its length are
// hashed with one 32-bit multiply.  The hash's top bits index a 2^18-bit
// membership filter of the dictionary: a token whose bit is clear is no query
// word (almost all tokens; measured: without the filter, walking the table's
// probe chains for them took a third of the scan).  Its low bits pick a slot
// of an LDS table whose entries carry the word's length; only a slot with the
// same length and the same first eight bytes leads to reading and comparing
// the token's remaining words.  A token's bytes and a query word's bytes
// compare equal iff the strings are equal, so the counts are exact.
#pragma once
#include <type_traits>
#include "gpufs.h"

#define GREP_MAX_CHUNK 65536
#define GREP_MAX_WORDS 1024
#define GREP_HT 4096
#define GREP_FILT_LOG 18   // membership filter bits (32 KiB of LDS)

struct grep_lds {
    uint8_t corpus[GREP_MAX_CHUNK + 128];    // alignment (< 16) + window + look-ahead + NUL + read padding
    uint64_t words[GREP_MAX_WORDS][4];       // query bytes, zero-padded, little-endian u64
    uint32_t ht[GREP_HT];                    // (canonical word index + 1) | length << 16; 0 = empty
    uint32_t filt[(1u << GREP_FILT_LOG) / 32];   // bit (hash >> (32 - GREP_FILT_LOG)) set for each query word
    uint32_t counts[GREP_MAX_WORDS];
    uint16_t canon[GREP_MAX_WORDS];          // 0xFFFF: word can never match
    uint8_t wlen[GREP_MAX_WORDS];
    uint32_t lenmask;                        // bit L-1: some query word has length L
    int status;
    uint64_t fsize;
};

__device__ __forceinline__ bool g_alpha(uint32_t c) {
    return (c >= 65 && c <= 90) || (c >= 97 && c <= 122) || c == 95;
}

// Word-byte flags of the four bytes of d (bit i = byte i), branch-free.
// Letters: d|0x20 lands in [0x61,0x7a] exactly for A-Z and a-z; '_' by a
// per-byte zero test of d^0x5f.  No carry crosses a byte boundary.
__device__ __forceinline__ uint32_t alpha4(uint32_t d) {
    const uint32_t x = d | 0x20202020u;
    const uint32_t x7 = x & 0x7F7F7F7Fu;
    const uint32_t lower = (x7 + 0x1F1F1F1Fu) & ~(x7 + 0x05050505u) & ~x & 0x80808080u;
    const uint32_t y = d ^ 0x5F5F5F5Fu;
    const uint32_t nz = ((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y;
    const uint32_t v = lower | (~nz & 0x80808080u);
    return ((v >> 7) & 1u) | ((v >> 14) & 2u) | ((v >> 21) & 4u) | ((v >> 28) & 8u);
}

// Hash of a token's first eight bytes and its length: one 32-bit multiply
// (a quarter-rate instruction: the three of a full 32-bit finaliser cost more
// than the rest of the hash).  The filter takes the product's high bits,
// which depend on every input bit; the table slot takes its low bits after
// the high half is folded into them.
__device__ __forceinline__ uint32_t tok_hash8(uint64_t w0, uint32_t len) {
    const uint32_t lo = (uint32_t)w0, hi = (uint32_t)(w0 >> 32);
    const uint32_t h = (lo ^ ((hi << 13) | (hi >> 19)) ^ (len << 26)) * 0x9E3779B1u;
    return h ^ (h >> 15);
}

// Eight bytes of the LDS window starting at byte p (two aligned loads).
__device__ __forceinline__ uint64_t lds_u64_at(const uint8_t* corpus, uint32_t p) {
    const uint64_t* W = (const uint64_t*)corpus;
    const uint32_t a = p >> 3, sh = (p & 7) * 8;
    const uint64_t lo = W[a], hi = W[a + 1];
    return sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
}
// The low n bytes of v (n may be <= 0 or >= 8).
__device__ __forceinline__ uint64_t keep_bytes(uint64_t v, int n) {
    return n >= 8 ? v : (n <= 0 ? 0ull : (v & ((1ull << (8 * n)) - 1)));
}
// Eight bytes of the LDS window starting at byte p, without a branch on the
// alignment: three aligned dwords and two byte-aligns.
__device__ __forceinline__ uint64_t lds_u64_nb(const uint8_t* corpus, uint32_t p) {
    const uint32_t* W = (const uint32_t*)corpus;
    const uint32_t a = p >> 2, sh = p & 3;
    const uint32_t w0 = W[a], w1 = W[a + 1], w2 = W[a + 2];
    const uint32_t lo = __builtin_amdgcn_alignbyte(w1, w0, sh);
    const uint32_t hi = __builtin_amdgcn_alignbyte(w2, w1, sh);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ inline void grep_build_dict(grep_lds& S, const char* dict, uint32_t nw) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    for (uint32_t i = t; i < GREP_HT; i += T) S.ht[i] = 0;
    for (uint32_t i = t; i < (1u << GREP_FILT_LOG) / 32; i += T) S.filt[i] = 0;
    if (t == 0) S.lenmask = 0;
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t len = 0;
        bool ok = true;
        uint8_t b[32];
        for (uint32_t i = 0; i < 32; ++i) {
            uint8_t c = (uint8_t)dict[(size_t)q * 32 + i];
            if (len == i && c != 0) {
                len = i + 1;
                if (!g_alpha(c)) ok = false;
            }
            b[i] = c;
        }
        // bytes after the first NUL are ignored (the C string ends there)
        for (uint32_t k = 0; k < 4; ++k) {
            uint64_t v = 0;
            for (uint32_t i = 0; i < 8; ++i)
                if (8 * k + i < len) v |= (uint64_t)b[8 * k + i] << (8 * i);
            S.words[q][k] = v;
        }
        S.wlen[q] = (uint8_t)len;
        S.counts[q] = 0;
        S.canon[q] = (ok && len > 0) ? (uint16_t)q : (uint16_t)0xFFFF;
    }
    __syncthreads();
    for (uint32_t q = t; q < nw; q += T) {
        if (S.canon[q] == 0xFFFF) continue;
        const uint32_t len = S.wlen[q];
        atomicOr(&S.lenmask, 1u << (len - 1));
        const uint32_t h = tok_hash8(S.words[q][0], len);
        const uint32_t fi = h >> (32 - GREP_FILT_LOG);
        atomicOr(&S.filt[fi >> 5], 1u << (fi & 31));
        uint32_t pos = h & (GREP_HT - 1);
        for (;;) {
            uint32_t old = atomicCAS(&S.ht[pos], 0u, (q + 1) | (len << 16));
            if (old == 0) break;
            uint32_t r = (old & 0xFFFFu) - 1;
            if (S.wlen[r] == S.wlen[q] && S.words[r][0] == S.words[q][0] && S.words[r][1] == S.words[q][1] &&
                S.words[r][2] == S.words[q][2] && S.words[r][3] == S.words[q][3]) {
                S.canon[q] = (uint16_t)r;
                break;
            }
            pos = (pos + 1) & (GREP_HT - 1);
        }
    }
    __syncthreads();
    // a duplicate links to the word that owns the table entry
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t c = S.canon[q];
        if (c == 0xFFFF) continue;
        while (S.canon[c] != c) c = S.canon[c];
        S.canon[q] = (uint16_t)c;
    }
    __syncthreads();
}

// Probe the dictionary for a token that passed the filter: table slots whose
// length and first eight bytes agree lead to comparing the remaining words.
__device__ inline void grep_probe(grep_lds& S, uint32_t p, uint32_t len, uint64_t t0, uint32_t h) {
    uint32_t pos = h & (GREP_HT - 1);
    for (;;) {
        const uint32_t e = S.ht[pos];
        if (!e) return;
        if ((e >> 16) == len) {
            const uint32_t w = (e & 0xFFFFu) - 1;
            if (S.words[w][0] == t0) {
                bool eq = true;
                for (uint32_t i = 1; 8 * i < len; ++i)
                    eq = eq && S.words[w][i] == keep_bytes(lds_u64_at(S.corpus, p + 8 * i), (int)(len - 8 * i));
                if (eq) {
                    atomicAdd(&S.counts[w], 1u);
                    return;
                }
            }
        }
        pos = (pos + 1) & (GREP_HT - 1);
    }
}

// Word-byte mask of the 16*nq bytes at `base` (nq quads), cut at the window end.
__device__ inline uint64_t grep_seg_mask(const uint8_t* corpus, uint32_t base, uint32_t dsz, int nq) {
    typedef __attribute__((address_space(3))) gfs::v4u lv4;
    const lv4* src = (const lv4*)(corpus + base);
    uint64_t am = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= nq) break;
        gfs::v4u v = src[k];
        am |= ((uint64_t)alpha4(v.x) | ((uint64_t)alpha4(v.y) << 4) | ((uint64_t)alpha4(v.z) << 8) |
               ((uint64_t)alpha4(v.w) << 12)) << (16 * k);
    }
    const uint32_t valid = dsz - base;
    if (valid < 64) am &= (1ull << valid) - 1;
    return am;
}

// Length of the token that starts at bit b of the segment mask `am`; a token
// that runs past the segment takes the rest of its length from the next
// segment's mask `nx`.
__device__ __forceinline__ uint32_t grep_tok_len(uint64_t am, uint64_t nx, uint32_t b) {
    const uint64_t rest = ~(am >> b);
    uint32_t len = rest ? (uint32_t)__ffsll((unsigned long long)rest) - 1 : 64;
    if (b + len >= 64) {
        const uint64_t nr = ~nx;
        len = (64 - b) + (nr ? (uint32_t)__ffsll((unsigned long long)nr) - 1 : 64);
    }
    return len;
}

// Count tokens of window [0, dsz) of S.corpus.  Bytes from dsz on count as
// separators, so every token ends inside the window.
// Each lane classifies one 64-byte segment and walks its tokens.  The loop is
// wave-uniform so that lanes can trade masks: a token that runs past its
// segment's end takes the rest of its length from the next segment's mask (the
// neighbour lane's, or lane 63's own classification of the next 32 bytes) and
// the byte before a segment comes from the previous lane's mask.  A byte-wise
// LDS walk there (one load and wait per byte, taken by almost every lane in the
// same iteration) cost a third of the scan.
// A lane takes GREP_UNROLL tokens per iteration: their LDS reads (first eight
// bytes, then the filter word) are independent and issue together, so one
// iteration waits for LDS about as long as a one-token iteration did.
// The window's first `skip` bytes (< 64) are not part of it: they are masked
// out as if they were separators.
#ifndef GREP_UNROLL
#define GREP_UNROLL 2
#endif
// Optionally, the next window is loaded meanwhile: `nunits` 16-byte units
// from file offset `wbase` (16-aligned), read straight from frames pinned by
// pin_pages (page pb + i in gfs_ctl.pinf[i]) into registers before the scan
// and stored into S.corpus after it (GREP_PF units per thread at most).
#define GREP_PF 5
typedef __attribute__((address_space(3))) grep_lds grep_lds_as3;
__device__ __noinline__ void grep_scan_window_lds(grep_lds_as3* SP, uint32_t dsz, uint32_t skip, uint64_t wbase,
                                                  uint64_t pb, uint32_t nunits) {
    grep_lds& S = *(grep_lds*)SP;
    constexpr int U = GREP_UNROLL;
    const uint32_t T = gfs::nthreads(), t = gfs::tid(), ln = gfs::lane();
    gfs::v4u pf[GREP_PF];
    if (nunits) {
        const gfs_dev* g = gfs_g;
        const uint8_t* pool = g->pool;
        const uint32_t lp = g->log_page;
        const uint64_t pm = (1ull << lp) - 1;
#pragma unroll
        for (int i = 0; i < GREP_PF; ++i) {
            const uint32_t u = t + T * (uint32_t)i;
            if (u < nunits) {
                const uint64_t pos = wbase + 16ull * u;
                const uint32_t fr = gfs_ctl.pinf[(pos >> lp) - pb];
                pf[i] = gfs::ld16_coh(pool + ((uint64_t)fr << lp) + (pos & pm));
            }
        }
    }
    const uint32_t nseg = (dsz + 63) >> 6;
    const uint32_t lm = S.lenmask;
    for (uint32_t sg0 = t - ln; sg0 < nseg; sg0 += T) {
        const uint32_t sg = sg0 + ln;
        const uint32_t base = sg << 6;
        uint64_t am = sg < nseg ? grep_seg_mask(S.corpus, base, dsz, 4) : 0ull;
        if (sg == 0) am &= ~0ull << skip;
        uint64_t nx = __shfl_down(am, 1);
        uint64_t pv = __shfl_up(am, 1) >> 63;
        if (sg >= nseg) continue;
        if (ln == 63) nx = sg + 1 < nseg ? grep_seg_mask(S.corpus, base + 64, dsz, 2) : 0ull;
        if (ln == 0) pv = (base > 0 && g_alpha(S.corpus[base - 1])) ? 1ull : 0ull;
        uint64_t starts = am & ~((am << 1) | pv);
        while (starts) {
            uint32_t p[U], len[U], h[U];
            uint64_t t0[U];
            uint32_t okb = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // a lane with fewer tokens left takes bit 0 again, not counted
                const uint32_t have = starts != 0 ? 1u : 0u;
                const uint32_t b = (uint32_t)__builtin_ctzll(starts | (1ull << 63)) & (have ? 63u : 0u);
                starts &= starts - 1;
                len[u] = grep_tok_len(am, nx, b);
                p[u] = base + b;
                // len in [1, 32] and some query word has this length
                okb |= (have & (uint32_t)(len[u] <= 32) & (lm >> ((len[u] - 1) & 31))) << u;
            }
            if (!okb) continue;
            // every token's reads are made whether or not it passed, so that
            // they issue together instead of one behind the other's branch
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t kb = len[u] < 8 ? len[u] : 8u;
                t0[u] = lds_u64_nb(S.corpus, p[u]) & (~0ull >> (64 - 8 * kb));
            }
#pragma unroll
            for (int u = 0; u < U; ++u) h[u] = tok_hash8(t0[u], len[u]);
            uint32_t hit = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t fw = S.filt[h[u] >> (32 - GREP_FILT_LOG + 5)];
                hit |= ((fw >> ((h[u] >> (32 - GREP_FILT_LOG)) & 31)) & 1u) << u;
            }
            hit &= okb;
            if (!hit) continue;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if ((hit >> u) & 1u) grep_probe(S, p[u], len[u], t0[u], h[u]);
        }
    }
    if (nunits) {
        __syncthreads();   // every wave is done with the current window
        typedef __attribute__((address_space(3))) gfs::v4u lv4;
        lv4* dst = (lv4*)S.corpus;
#pragma unroll
        for (int i = 0; i < GREP_PF; ++i) {
            const uint32_t u = t + T * (uint32_t)i;
            if (u < nunits) dst[u] = pf[i];
        }
        __syncthreads();
    }
}
// Out of line, with the window's LDS address space kept in the pointer type,
// so that the scan has the register file to itself instead of sharing it with
// the caller's read loop (ds_ instructions still, not flat ones).
__device__ __forceinline__ void grep_scan_window(grep_lds& S, uint32_t dsz, uint32_t skip = 0) {
    grep_scan_window_lds((grep_lds_as3*)&S, dsz, skip, 0, 0, 0);
}

// The chunk loop of one open file (grep_text_gpu.cu:298-318): windows of
// `chunk` bytes plus the look-ahead that completes a straddling word, each
// read with gread into LDS and scanned into S.counts; after every window all
// threads call on_window() (workgroup-collective).  Returns 0 or a negative
// code (workgroup-uniform).
template <class OnWindow>
__device__ inline int grep_file_windows(grep_lds& S, int fd, uint32_t chunk, OnWindow&& on_window) {
    const uint32_t t = gfs::tid();
    if (t == 0) {
        S.status = 0;
        S.fsize = fstat(fd);   // one read, broadcast: the loop bound must be workgroup-uniform
    }
    __syncthreads();
    const uint64_t size = gfs::uni(S.fsize);
    uint64_t cursor = 0;
    while (cursor < size) {
        const bool last = size - cursor < (uint64_t)chunk + 32;
        const uint32_t left = last ? (uint32_t)(size - cursor) : chunk + 32;
        // The read starts at the 16-byte boundary at or below the cursor, so
        // that it copies with aligned 16-byte loads (a window starting
        // anywhere else takes five dword loads per 16 bytes); the `mis` bytes
        // before the cursor are masked out of the scan.
        const uint32_t mis = (uint32_t)(cursor & 15);
        size_t got;
        // Inlined here: as an outlined call, gread's prologue and epilogue save and
        // restore about fifty registers through scratch on every window.  The
        // window is in LDS, so only the global->LDS copy path is compiled in.
        [[clang::always_inline]] got = gread<gfs::SPACE_LDS>(fd, cursor - mis, left + mis, S.corpus);
        if (got != left + mis) {
            if (t == 0) S.status = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
            break;
        }
        // The word that straddles the chunk end: its bytes in the look-ahead
        // (at most 32), found by one ballot over the look-ahead in every wave
        // (each reads the same bytes, so no barrier is needed to share it).
        // The scan stops before the byte after that word, which gread did not
        // necessarily write when the look-ahead is all word bytes.
        uint32_t ov = 0;
        {
            const uint32_t ln = gfs::lane();
            const bool stop = last || ln >= 32 || !g_alpha(S.corpus[mis + chunk + ln]);
            ov = gfs::uni((uint32_t)__builtin_ctzll(__ballot(stop) | (1ull << 32)));
        }
        const uint32_t dsz = mis + (last ? left : chunk + ov);
        cursor += (uint64_t)chunk + ov;
        grep_scan_window(S, dsz, mis);
        // A callback that reads the counts waits for every wave's scan.  With
        // none, the next gread's barrier keeps the window until every wave is
        // done with it, and thread 0's round trip for that read overlaps the
        // other waves' scans.
        if constexpr (!std::is_empty<typename std::remove_reference<OnWindow>::type>::value) __syncthreads();
        GFS_TRACE(7);
        on_window();
    }
    __syncthreads();
    const int st = gfs::uni(S.status);
    __syncthreads();
    return st;
}

// One file, the whole chunk loop; S.counts is zeroed here and the file's
// counts land in row[0..nw).  Returns 0 or a negative code (workgroup-uniform).
__device__ inline int grep_one_file(grep_lds& S, const char* name, uint32_t nw, uint32_t chunk, uint32_t* row) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
    int fd = gopen(name, O_GRDONLY);
    if (fd < 0) {
        for (uint32_t q = t; q < nw; q += T) row[q] = 0;
        return fd;
    }
    const int sst = grep_file_windows(S, fd, chunk, [] {});
    int c = gclose(fd);
    for (uint32_t q = t; q < nw; q += T) {
        uint32_t k = S.canon[q];
        row[q] = k == 0xFFFF ? 0u : S.counts[k];
    }
    int st = gfs::uni(sst ? sst : c);
    __syncthreads();
    return st;
}

// ---------------------------------------------------------------------------
// One file scanned by several workgroups (parts).  The reference's chunk loop
// (grep_text_gpu.cu:298-318) moves its window by chunk + ov, where ov is the
// run of word bytes (at most 32) at the chunk's end, so every window but the
// last ends right before a non-word byte whenever ov < 32: the windows then
// split no token, and the count of a non-last window is the count of the
// tokens that start in it.  So, unless some ov is 32,
//     counts(file) = tokens of the whole file
//                    + tokens of [c_L + i*chunk, size) for every i >= 1 with
//                      c_L + i*chunk < size,
// where c_L is the start of the reference's last window (that window is
// scanned to the end of the file and the cursor then moves by chunk only, so
// each later window re-scans the file's last bytes, from a word start).
// Each part counts the tokens that start in its byte range, with window
// boundaries of its own that also fall before non-word bytes; the last part
// to finish computes the reference's window chain from the cached pages,
// adds the re-scanned tail, and when a boundary (the parts' or the
// reference's) meets a run of 32 or more word bytes, rescans the whole file
// with the reference's own loop.  The result equals grep_file_windows.

// Run of word bytes at LDS offset x (at most 64), bytes from `avail` on
// counting as non-word; every wave computes it from the same bytes.
__device__ __forceinline__ uint32_t lds_word_run(const uint8_t* corpus, uint32_t x, uint32_t avail) {
    const uint32_t ln = gfs::lane();
    const bool stop = ln >= avail || !g_alpha(corpus[x + (ln < avail ? ln : 0)]);
    const uint64_t m = __ballot(stop);
    return m ? (uint32_t)__builtin_ctzll(m) : 64u;
}

// The tokens that start in [b, e) of the file, where b is the first non-word
// byte at or after p_begin (p_begin itself when it is 0) and e likewise for
// p_end (the file's end when p_end == size), counted into S.counts.  Windows of
// `chunk` bytes end at the first non-word byte after the chunk (the
// reference's look-ahead rule).  *longrun is set when a run of 32 or more word
// bytes meets one of these boundaries (a token might be split there).
// Returns 0 or a negative code (workgroup-uniform).
__device__ inline int grep_part_windows(grep_lds& S, int fd, uint32_t chunk, uint64_t size, uint64_t p_begin,
                                        uint64_t p_end, uint32_t* longrun) {
    const uint32_t t = gfs::tid();
    if (t == 0) {
        S.status = 0;
        // The range is read front to back: start gread's read history as if
        // two reads had led up to it, so that the first window's batch already
        // takes the read-ahead (otherwise the first two windows of a cold file
        // wait for the host one page at a time).  The history is advisory.
        const uint64_t p0 = p_begin >> gfs_g->log_page;
        gfs_hist.fd = fd;
        gfs_hist.streak = 1;
        gfs_hist.first = p0;
        gfs_hist.last = p0;
        gfs_hist.fetched = p0;
    }
    __syncthreads();
    uint64_t cursor = p_begin;
    bool first = p_begin > 0;
    uint32_t lr = 0;
    for (;;) {
        const bool final = p_end - cursor <= chunk;
        const uint64_t want = final ? p_end + 64 : cursor + chunk + 32;
        const uint64_t rend = want < size ? want : size;
        const uint32_t mis = (uint32_t)(cursor & 15);
        const uint32_t len = (uint32_t)(rend - cursor) + mis;
        size_t got;
        [[clang::always_inline]] got = gread<gfs::SPACE_LDS>(fd, cursor - mis, len, S.corpus);
        if (got != len) {
            if (t == 0) S.status = (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
            break;
        }
        uint32_t skip = mis;
        if (first) {
            // the token that runs into p_begin belongs to the previous part
            const uint32_t r = lds_word_run(S.corpus, mis, len - mis < 64 ? len - mis : 64);
            if (r >= 32) lr = 1;
            else skip = mis + r;
            first = false;
        }
        uint32_t dsz, ov = 0;
        if (final) {
            uint64_t e = size;
            if (p_end < size) {
                const uint32_t xo = mis + (uint32_t)(p_end - cursor);
                const uint32_t r = lds_word_run(S.corpus, xo, len - xo < 64 ? len - xo : 64);
                if (r >= 32) lr = 1;
                else e = p_end + r;
            }
            dsz = mis + (uint32_t)(e - cursor);
        } else {
            const uint32_t xo = mis + chunk;
            ov = lds_word_run(S.corpus, xo, len - xo < 32 ? len - xo : 32);
            if (ov >= 32) lr = 1;
            dsz = mis + chunk + ov;
        }
        grep_scan_window(S, dsz, skip);
        if (final) break;
        cursor += (uint64_t)chunk + ov;
        // a window whose look-ahead run reached p_end ended where this part
        // ends (the run at p_end is part of the same run, shorter than 32)
        if (cursor >= p_end) break;
    }
    __syncthreads();
    const int st = gfs::uni(S.status);
    __syncthreads();
    *longrun = lr;
    return st;
}

// grep_part_windows for a part whose pages are cached: the part's pages are
// pinned (at most 64 at a time) and each window is loaded from the frames
// into registers while the previous window is scanned (grep_scan_window_lds),
// instead of a gread per window between the scans.  Same windows, same
// results.  Needs a workgroup large enough that GREP_PF 16-byte units per
// thread cover the largest window; returns 1 (nothing done) otherwise.
// Whether T threads can hold the largest window of grep_part_pinned.
__device__ __forceinline__ bool grep_pinned_fits(uint32_t T, uint32_t chunk) {
    return (uint64_t)T * GREP_PF * 16 >= (uint64_t)chunk + 64 + 16 + 16;
}
// SPEC: pages are pinned with pin_cached (no host read, the entry need not be
// open); a page that is not cached ends the part with 1, its counts partial.
template <bool SPEC = false>
__device__ inline int grep_part_pinned(grep_lds& S, int fd, uint32_t chunk, uint64_t size, uint64_t p_begin,
                                       uint64_t p_end, uint32_t* longrun, bool held) {
    gfs_dev* g = gfs_g;
    const uint32_t T = gfs::nthreads();
    if (!grep_pinned_fits(T, chunk)) return 1;
    const uint32_t lp = g->log_page;
    const uint64_t last_page = (size - 1) >> lp;
    // window at cursor c: bytes [c - mis, rend) land at corpus[0, len)
    auto window = [&](uint64_t c, bool& final, uint64_t& wbase, uint32_t& len) {
        final = p_end - c <= chunk;
        const uint64_t want = final ? p_end + 64 : c + chunk + 32;
        const uint64_t rend = want < size ? want : size;
        wbase = c & ~15ull;
        len = (uint32_t)(rend - wbase);
    };
    uint64_t cursor = p_begin, wbase;
    bool final, first = p_begin > 0;
    uint32_t len, lr = 0;
    window(cursor, final, wbase, len);
    // `held`: every page of the file is pinned by the caller, page i's frame in
    // pinf[i], and stays pinned after the return
    uint64_t pb = held ? 0 : wbase >> lp;
    uint32_t cnt = (uint32_t)(last_page - pb + 1 < 64 ? last_page - pb + 1 : 64);
    int st = held ? 0 : (SPEC ? gfs::pin_cached(g, fd, pb, cnt) : gfs::pin_pages(g, fd, pb, cnt));
    if (st) {
        *longrun = 0;
        return st;
    }
    grep_scan_window_lds((grep_lds_as3*)&S, 0, 0, wbase, pb, (len + 15) >> 4);   // load only
    for (;;) {
        const uint32_t mis = (uint32_t)(cursor & 15);
        uint32_t skip = mis;
        if (first) {
            // the token that runs into p_begin belongs to the previous part
            const uint32_t r = lds_word_run(S.corpus, mis, len - mis < 64 ? len - mis : 64);
            if (r >= 32) lr = 1;
            else skip = mis + r;
            first = false;
        }
        uint32_t dsz, ov = 0;
        if (final) {
            uint64_t e = size;
            if (p_end < size) {
                const uint32_t xo = mis + (uint32_t)(p_end - cursor);
                const uint32_t r = lds_word_run(S.corpus, xo, len - xo < 64 ? len - xo : 64);
                if (r >= 32) lr = 1;
                else e = p_end + r;
            }
            dsz = mis + (uint32_t)(e - cursor);
        } else {
            const uint32_t xo = mis + chunk;
            ov = lds_word_run(S.corpus, xo, len - xo < 32 ? len - xo : 32);
            if (ov >= 32) lr = 1;
            dsz = mis + chunk + ov;
        }
        // the next window (a window whose look-ahead run reached p_end ended
        // where this part ends)
        const uint64_t ncur = cursor + (uint64_t)chunk + ov;
        const bool more = !final && ncur < p_end;
        bool nfinal = false;
        uint64_t nwbase = 0;
        uint32_t nlen = 0;
        if (more) {
            window(ncur, nfinal, nwbase, nlen);
            if (((nwbase + nlen - 1) >> lp) >= pb + cnt) {
                // it leaves the pinned pages; this window is in LDS already
                gfs::unpin_pages(g, cnt);
                pb = nwbase >> lp;
                cnt = (uint32_t)(last_page - pb + 1 < 64 ? last_page - pb + 1 : 64);
                st = SPEC ? gfs::pin_cached(g, fd, pb, cnt) : gfs::pin_pages(g, fd, pb, cnt);
                if (st) {
                    cnt = 0;
                    break;
                }
            }
        }
        GFS_TRACE(50);
        grep_scan_window_lds((grep_lds_as3*)&S, dsz, skip, nwbase, pb, more ? (nlen + 15) >> 4 : 0u);
        if (!more) break;
        cursor = ncur;
        final = nfinal;
        wbase = nwbase;
        len = nlen;
    }
    // every wave's counts are in S.counts before the caller reads them (the
    // last scan call has no barrier after it)
    __syncthreads();
    if (cnt && !held) gfs::unpin_pages(g, cnt);
    *longrun = lr;
    return st;
}

#define GREP_CHAIN_UNITS 4608   // 16-byte units whose word-byte masks one chain round keeps in LDS (9 KiB)
#define GREP_CHAIN_STEPS 64     // chain steps per round: 16 waves x 4

// Word-byte mask of 16 bytes, bit i = byte i.
__device__ __forceinline__ uint32_t alpha16(gfs::v4u v) {
    return alpha4(v.x) | (alpha4(v.y) << 4) | (alpha4(v.z) << 8) | (alpha4(v.w) << 12);
}
// Units of chain step j of a round: step j's look-ahead starts at most 31*j
// bytes after the 16-byte unit holding the round's j-th chunk end, and needs
// 32 bytes plus the unit offset (at most 126 units for j < 64).
__device__ __forceinline__ uint32_t chain_units(uint32_t j) { return ((31u * j + 15u) >> 4) + 3u; }

typedef __attribute__((address_space(3))) uint16_t lds_u16;

// The reference's window chain over a file of `size` bytes
// (grep_text_gpu.cu:298-318): c_0 = 0, c_{k+1} = c_k + chunk + ov_k, while at
// least chunk + 32 bytes remain after c_k.  *c_last = the last window's start;
// *cut = 1 when some ov_k is 32 (that window cuts a token).  Read from the
// cached pages, pinned a batch at a time.  A round covers up to 4 steps per
// wave (64 with 16 waves): wave w loads, for steps w + k * waves (k < 4), every 16-byte unit the step's
// look-ahead can fall in (all its loads issued together) and stores their
// word-byte masks in LDS; then wave 0 walks the steps over the masks.
// Workgroup-collective; returns 0 or a negative code.
__device__ __noinline__ int grep_chain_last(grep_lds& S, int fd, uint64_t size, uint32_t chunk, uint64_t* c_last,
                                            uint32_t* cut, bool held) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    // read once into scalars: reloaded inside the load loop, they made every
    // unit load wait for the one before it
    const uint32_t L = uni(g->log_page);
    const uint64_t PS = uni((uint64_t)g->page_size);
    uint8_t* const pool = (uint8_t*)uni((uint64_t)(uintptr_t)g->pool);
    const uint64_t npages = (size + PS - 1) >> L;
    uint32_t maxpin = uni(g->n_frames) / 8;
    maxpin = maxpin > 64 ? 64 : (maxpin < 2 ? 2 : maxpin);
    const uint32_t ln = lane(), w = wave(), nwaves = nthreads() >> 6;
    // the masks live in the window buffer, addressed as LDS (this code may be
    // compiled outside the kernel, where S is a generic reference)
    lds_u16* m = (lds_u16*)(S.corpus);
    uint64_t c = 0;
    uint32_t sp = 0;
    while (size - c >= (uint64_t)chunk + 32 && !sp) {
        // `held`: every page of the file is pinned already, page i's frame in pinf[i]
        const uint64_t P = held ? 0 : (c + chunk) >> L;
        const uint32_t cnt = (uint32_t)(npages - P < maxpin ? npages - P : maxpin);
        if (!held) {
            const int r = pin_pages(g, fd, P, cnt);
            if (r) return r;
        }
        const uint64_t hi0 = (P + cnt) << L;
        const uint64_t hi = hi0 < size ? hi0 : size;
        // this round's steps: as many as fit the unit budget and the pinned pages
        uint32_t R = 0, U = 0;
        const uint32_t rmax = 4 * nwaves < GREP_CHAIN_STEPS ? 4 * nwaves : GREP_CHAIN_STEPS;
        while (R < rmax) {
            const uint32_t n = chain_units(R);
            const uint64_t b = (c + (uint64_t)(R + 1) * chunk) & ~15ull;
            if (U + n > GREP_CHAIN_UNITS || b >= hi) break;
            U += n;
            ++R;
        }
        // wave w: steps w + 16k (k < 4), two units per lane per step
        v4u val[8];
        uint64_t ad[8];
        uint32_t dst[8];
        uint32_t s0 = 0, jn = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t j = w + (uint32_t)k * nwaves;
            while (jn < j) s0 += chain_units(jn++);   // s0 = first unit of step j
            const uint32_t n = chain_units(j);
            const uint64_t b = (c + (uint64_t)(j + 1) * chunk) & ~15ull;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t l = ln + 64u * (uint32_t)h;
                const bool on = j < R && l < n;
                ad[2 * k + h] = b + 16ull * l;
                dst[2 * k + h] = on ? s0 + l : 0xFFFFFFFFu;
                val[2 * k + h] = v4u{0, 0, 0, 0};
                if (on && ad[2 * k + h] < hi) {
                    const uint64_t a = ad[2 * k + h];
                    const uint32_t f = gfs_ctl.pinf[(a >> L) - P];
                    val[2 * k + h] = ld16_coh(pool + ((uint64_t)f << L) + (a & (PS - 1)));
                }
            }
        }
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            if (dst[x] == 0xFFFFFFFFu) continue;
            uint32_t mk = 0;
            if (ad[x] < hi) {
                mk = alpha16(val[x]);
                if (ad[x] + 16 > size) mk &= (1u << (uint32_t)(size - ad[x])) - 1;
            }
            m[dst[x]] = (uint16_t)mk;
        }
        __syncthreads();
        if (w == 0) {
            // the whole wave walks with the same values (they stay scalar); the
            // three masks of a step are one broadcast read each
            const uint64_t c0 = c;
            uint32_t su = 0;
            for (uint32_t j = 0; j < R; su += chain_units(j++)) {
                if (size - c < (uint64_t)chunk + 32) break;
                const uint64_t q = c + chunk;
                if (q + 32 > hi) break;
                const uint64_t off = q - ((c0 + (uint64_t)(j + 1) * chunk) & ~15ull);
                const uint32_t idx = su + (uint32_t)(off >> 4);
                const uint32_t m0 = m[idx], m1 = m[idx + 1], m2 = m[idx + 2];
                const uint64_t wd = (uint64_t)uni(m0) | ((uint64_t)uni(m1) << 16) | ((uint64_t)uni(m2) << 32);
                const uint32_t bits = (uint32_t)(wd >> (off & 15));
                const uint32_t run = bits == 0xFFFFFFFFu ? 32u : (uint32_t)__builtin_ctz(~bits);
                if (run == 32) {
                    sp = 1;
                    break;
                }
                c = q + run;
            }
            if (ln == 0) {
                S.fsize = c;
                S.status = (int)sp;
            }
        }
        __syncthreads();
        c = uni((uint64_t)S.fsize);
        sp = (uint32_t)uni(S.status);
        if (!held) unpin_pages(g, cnt);
    }
    *c_last = c;
    *cut = sp;
    return 0;
}

// The reference's windows after its last full one: the window at c_L is
// scanned to the end of the file and the cursor moves by `chunk` only, so
// every x = c_L + i*chunk below the end starts one more window [x, size), each
// scanned from a word start (several when chunk < 32).  [from, size) with
// from = c_L + chunk is shorter than 32 bytes; its windows' tokens are added
// into S.counts.  Returns 0 or a negative code.
__device__ __noinline__ int grep_tail(grep_lds& S, int fd, uint64_t from, uint64_t size, uint32_t chunk) {
    const uint32_t mis = (uint32_t)(from & 15);
    const uint32_t len = (uint32_t)(size - from) + mis;
    size_t got;
    [[clang::always_inline]] got = gread<gfs::SPACE_LDS>(fd, from - mis, len, S.corpus);
    if (got != len) return (int64_t)got < 0 ? (int)(int64_t)got : -EIO;
    for (uint64_t x = from; x < size; x += chunk)
        grep_scan_window(S, len, mis + (uint32_t)(x - from));
    __syncthreads();
    return 0;
}

// Per-file state of a split scan.  Nothing needs clearing before a launch:
// part 0 resets the other fields (and zeroes the file's count row) before it
// publishes its decision, and the decision carries the launch's epoch, so a
// decision left from an earlier launch in the same memory is never taken for
// this one's.
struct grep_split_state {
    uint32_t arrived;   // parts done
    uint32_t longrun;   // some part met a run of 32 or more word bytes at a boundary
    int32_t err;        // first error of a part
    uint32_t cut;       // part 0's chain: some reference window cuts a token
    uint64_t c_last;    // part 0's chain: start of the reference's last window
    uint64_t tag;       // part 0's decision: epoch << 32 | spec << 30 | (K_f + 1)
    uint64_t spec_size; // spec decision: the file's size as its last open left it
    uint32_t spec_ver;  // spec decision: the file's version then
    uint32_t spec_idx;  // spec decision: its file-table entry
    uint32_t p0_open;   // part 0 left its open for the last part to close
    uint32_t spec_opened;   // spec decision: the entry's count of successful opens before part 0's reopen
};

// Part boundaries: part 0 also walks the reference's window chain before its
// scan, so it takes 4/5 of the bytes of each other part.  Multiples of 64, and
// consecutive boundaries are at least min_part * 4/5 - 63 > 64 bytes apart
// (min_part >= 256).
__device__ __forceinline__ uint64_t split_point(uint64_t size, uint64_t j, uint64_t kf) {
    if (j == 0) return 0;
    if (j >= kf) return size;
    return (size * (5 * j - 1) / (5 * kf - 1)) & ~63ull;
}

// The reference's own loop over the whole file, the row overwritten with its
// counts (a split whose boundary met a run of 32 or more word bytes).
__device__ __noinline__ int grep_recount(grep_lds& S, int fd, uint32_t nw, uint32_t chunk, uint32_t* row) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
    __syncthreads();
    const int st = grep_file_windows(S, fd, chunk, [] {});
    for (uint32_t q = t; q < nw; q += T) {
        const uint32_t k = S.canon[q];
        row[q] = k == 0xFFFF ? 0u : S.counts[k];
    }
    __syncthreads();
    return st;
}

// Completes a split file (the workgroup whose part finished last; cold path,
// kept out of the part loop's registers): given the start c_last of the
// reference's last window, either the re-scanned tail added into the row, or,
// when a boundary met a run of 32 or more word bytes (redo), the reference's
// own loop over the whole file with the row overwritten.  Returns 0 or a
// negative code.
__device__ __noinline__ int grep_split_finish(grep_lds& S, int fd, uint32_t nw, uint32_t chunk, uint64_t size,
                                              uint32_t* row, uint32_t redo, uint64_t cl) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    int st = 0;
    if (redo) {
        st = grep_recount(S, fd, nw, chunk, row);
    } else if (size - cl > chunk) {
        for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
        __syncthreads();
        st = grep_tail(S, fd, cl + chunk, size, chunk);
        for (uint32_t q = t; q < nw; q += T) {
            const uint32_t k = S.canon[q];
            const uint32_t v = k == 0xFFFF ? 0u : S.counts[k];
            if (v) atomicAdd(&row[q], v);
        }
    }
    __syncthreads();
    return st;
}

// Part j of the file `name` (workgroup-collective).  The file is split into
// K_f = min(parts, size / min_part) parts (at least 1; min_part >= 256, so no
// part is empty) when it is cached, else into one; workgroups j >= K_f scan
// nothing.  Workgroup 0 also walks the reference's window chain
// (grep_chain_last).  Every one of the `parts` workgroups of the file arrives
// exactly once, after adding its counts into the caller-zeroed row (a
// workgroup whose gopen failed records the error); the last to arrive
// completes the row (grep_split_finish) and writes status[0].  `epoch`
// identifies the launch (see grep_split_state); part 0 zeroes the row.
//
// `target` is the number of arrivals that completes the file: `parts` when a
// workgroup is launched for every part slot (gpufs_launch_grep), or 0 for
// "the file's own part count" (K_f, 1 when gopen failed) when only the parts
// that exist are run (the batched form, whose part 0 learns K_f and hands the
// other parts out: on_size(K_f) is called by every thread once the size is
// known, or with 1 when gopen failed).
//
// With fs == nullptr the file is scanned by this workgroup alone (the
// reference's form, one workgroup per file): its counts stay in LDS, the tail
// is added there and the row is stored at the end, so nothing needs clearing
// before the launch and no state is shared.
//
// `split_cold`: a file that is not cached is split too (the serving form,
// where a large cold file read by one workgroup sets the tail latency).
// on_arrive() is called by every thread of each part right before the part
// counts itself arrived.  Returns true in the workgroup that completed the
// file (the last to arrive, or the only one) and wrote *status.
//
// `spec` (split form only): the scan of a cached file overlaps its reopen.
// When part 0's open is a reopen of a closed entry (a host round trip), it
// posts the open and, if every page is still cached, pins them all and
// publishes a decision from the size and version the entry's last open left,
// then walks the chain and scans its part while the host answers; the other
// parts scan their ranges from the cached pages (pin_cached: no host read)
// before they join the open.  Each part then checks that the open found the
// same file (same entry, version and size): a reopen that finds the file
// changed on the host bumps the version, and that part sets the long-run
// flag, so the last part recounts the whole file through the open file with
// the reference's loop.  A part that finds a page no longer cached scans its
// range again through the open file.  Results are those of opening first.
template <class OnSize, class OnArrive>
__device__ inline bool grep_file_split(grep_lds& S, const char* name, uint32_t nw, uint32_t chunk, uint32_t* row,
                                       int32_t* status, uint32_t j, uint32_t parts, uint64_t min_part,
                                       grep_split_state* fs, uint32_t target, uint32_t epoch, bool split_cold,
                                       bool spec, OnSize&& on_size, OnArrive&& on_arrive) {
    const uint32_t T = gfs::nthreads(), t = gfs::tid();
    __shared__ uint32_t fin_flag;
    __shared__ uint32_t spec_sh[3];   // spec decision: entry, version, open count
    __shared__ gfs::gfs_open_pend opend;
    const bool single = fs == nullptr;
    for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
    if (!single && j == 0) {
        // The other parts add into the row and touch the state only after
        // they have read the decision, which is published with a release
        // after gopen's barriers: each wave's stores are performed first.
        for (uint32_t q = t; q < nw; q += T) row[q] = 0;
        if (t == 0) {
            fs->arrived = 0;
            fs->longrun = 0;
            fs->err = 0;
            fs->cut = 0;
            fs->c_last = 0;
            fs->p0_open = 0;
        }
        __builtin_amdgcn_s_waitcnt(0);
    }
    const bool sp = spec && !single;
    int fd = -1;
    uint32_t pend = 0;   // part 0: its reopen is posted, not yet completed
    // part 0 after a posted reopen: the other parts join the open only after
    // their scans, later than part 0 ends, so part 0's close would be the last
    // one and close the file under them; it leaves its open to the last part
    bool defer_close = false;
    if (sp && j == 0) {
        __syncthreads();
        if (t == 0) {
            gfs_hist.fd = -1;
            gfs_memo.fd = -1;
            opend.pending = 0;
            opend.peek_ok = 0;
            gfs_ctl.ret = gfs::open_single(name, O_GRDONLY, &opend);
            fin_flag = opend.pending;
        }
        __syncthreads();
        fd = (int)gfs::uni(gfs_ctl.ret);
        pend = gfs::uni(fin_flag);
    } else if (!sp) {
        fd = gopen(name, O_GRDONLY);
    }   // the other parts of the split form open after the decision (below)
    GFS_TRACE(42);
    uint64_t size = 0, kf = 1, cl = 0;
    int st = fd < 0 && (!sp || j == 0) ? fd : 0;
    uint32_t lr = 0, cut = 0;
    uint32_t held = 0;   // part 0: every page of the file pinned (page i's frame in pinf[i])
    uint32_t sver = 0, sidx = 0, sop = 0;
    uint32_t maxpin = gfs_g->n_frames / 8;
    maxpin = maxpin > 64 ? 64 : (maxpin < 2 ? 2 : maxpin);
    if (pend) {
        // the entry is OPENING and this workgroup's until open_finish: its
        // size and version stay as the last open left them
        if (t == 0) {
            gfs_file* e = &gfs_g->files[fd];
            S.fsize = e->size;
            fin_flag = fst_version(gfs::ld_rlx(&e->st));
            spec_sh[2] = __hip_atomic_load(&e->opened, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint64_t ssz = gfs::uni((uint64_t)S.fsize);
        sver = gfs::uni(fin_flag);
        sop = gfs::uni(spec_sh[2]);
        const uint64_t np = (ssz + gfs_g->page_size - 1) >> gfs_g->log_page;
        // (the chain and the part scan then read only the pinned frames: no
        // gread, which needs the open completed)
        if (ssz && np <= maxpin && grep_pinned_fits(T, chunk) && gfs::pin_cached(gfs_g, fd, 0, (uint32_t)np) == 0) {
            held = (uint32_t)np;
        } else {
            // not all cached: the open first, then the usual path
            __syncthreads();
            if (t == 0) {
                gfs::open_post(&opend);
                gfs_ctl.ret = gfs::open_finish(&opend);
            }
            __syncthreads();
            fd = (int)gfs::uni(gfs_ctl.ret);
            if (fd < 0) st = fd;
            pend = 0;
        }
    }

    // Part 0 decides the split.  A file that is not cached (its last page is
    // not) is read by part 0 alone: one workgroup streams a cold file best
    // (its sequential read-ahead keeps a batch in flight while it scans; cold
    // reads measured slower the more parts share a file).  A cached file gets
    // K_f = min(parts, size / min_part) parts, and part 0 walks the chain first,
    // overlapped with the other parts' scans.  The other parts wait for the
    // decision: part 0 has the lower workgroup index, so it was dispatched
    // first, and it publishes a decision on every path (bounded wait).
    bool warm = true;
    if (t == 0) {
        uint32_t d = 0;
        if (fd >= 0) S.fsize = fstat(fd);   // with a posted reopen: the size the last open left
        if (single || j == 0) {
            uint64_t k = 1;
            if (fd >= 0 && !single) {
                const uint64_t sz = S.fsize;
                warm = sz == 0 || gfs::page_cached(gfs_g, fd, (sz - 1) >> gfs_g->log_page);
                k = (warm || split_cold) ? sz / min_part : 1;
                k = k < 1 ? 1 : (k > parts ? parts : k);
            } else if (fd >= 0) {
                warm = S.fsize == 0 || gfs::page_cached(gfs_g, fd, (S.fsize - 1) >> gfs_g->log_page);
            }
            d = (uint32_t)k + 1;
            if (held) {
                // a decision from the cached pages; the parts check it after their open
                d |= 0x40000000u;
                fs->spec_size = S.fsize;
                fs->spec_ver = sver;
                fs->spec_idx = (uint32_t)fd;
                fs->spec_opened = sop;
            }
            if (!single)
                __hip_atomic_store(&fs->tag, ((uint64_t)epoch << 32) | d, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            // the other parts scan while the host opens the file
            if (held) gfs::open_post(&opend);
        } else {
            const uint64_t w0 = gfs::now_ticks();
            // relaxed polls, one acquire once the decision is seen (an acquire
            // load at agent scope invalidates the poller's L2 on every poll)
            for (;;) {
                const uint64_t tg = __hip_atomic_load(&fs->tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                d = (tg >> 32) == epoch ? (uint32_t)tg : 0u;
                if (d || gfs::now_ticks() - w0 >= gfs_g->wait_ticks) break;
                __builtin_amdgcn_s_sleep(8);
            }
            gfs::fence_acq_agent();
            // no decision within the wait limit: this part scans nothing and
            // the file reports -ETIMEDOUT
            if (d & 0x40000000u) {
                S.fsize = fs->spec_size;
                spec_sh[0] = fs->spec_idx;
                spec_sh[1] = fs->spec_ver;
                spec_sh[2] = fs->spec_opened;
            }
        }
        fin_flag = d | (warm ? 0x80000000u : 0u);
    }
    __syncthreads();
    bool specd = false;   // a part other than 0 reading under a spec decision
    {
        const uint32_t d = gfs::uni(fin_flag);
        warm = (d >> 31) != 0;
        specd = j > 0 && ((d >> 30) & 1u);
        kf = (d & 0x3FFFFFFFu) ? (d & 0x3FFFFFFFu) - 1 : 0;
        if ((d & 0x3FFFFFFFu) == 0 && !st) st = -ETIMEDOUT;
        if (fd >= 0 || specd) size = gfs::uni((uint64_t)S.fsize);
        if (specd) {
            sidx = gfs::uni(spec_sh[0]);
            sver = gfs::uni(spec_sh[1]);
            sop = gfs::uni(spec_sh[2]);
        }
    }
    __syncthreads();
    if (sp && j > 0 && !specd && !st) {
        // the decision followed part 0's open: join it now
        fd = gopen(name, O_GRDONLY);
        if (t == 0 && fd >= 0) S.fsize = fstat(fd);
        __syncthreads();
        if (fd < 0) st = fd;
        else size = gfs::uni((uint64_t)S.fsize);
    }
    GFS_TRACE(43);
    on_size((uint32_t)kf);
    if (specd) {
        // This part's range from the cached pages, before the open; then part
        // 0's open is confirmed rather than joined: the entry counts its
        // successful host opens, part 0 published the count from before its
        // reopen, and part 0's open stays held until the last part closes
        // it, so this part holds none (no join, no close on the state word).
        int sst = 0;
        const bool mine = j < kf && size && !st;
        if (mine) {
            const uint64_t pb = split_point(size, j, kf), pe = split_point(size, j + 1, kf);
            sst = grep_part_pinned<true>(S, (int)sidx, chunk, size, pb, pe, &lr, false);
        }
        if (t == 0) {
            // 0: no open succeeded after the look (it failed); 1: the same
            // file; 2: the file changed.  OPEN or CLOSING with the old count is
            // the moment between the state and the count; CLOSED with the old
            // count (read after the state, with an acquire) is a failed open.
            gfs_file* e = &gfs_g->files[sidx];
            uint32_t res = 0;
            const uint64_t w0 = gfs::now_ticks();
            for (;;) {
                uint32_t o = __hip_atomic_load(&e->opened, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bool seen = o != sop;
                if (!seen) {
                    const uint32_t s = fst_state(gfs::ld_acq(&e->st));
                    if (s == GFS_FS_CLOSED || s == GFS_FS_FREE) {
                        o = __hip_atomic_load(&e->opened, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (o == sop) break;
                        seen = true;
                    }
                }
                if (seen) {
                    gfs::fence_acq_agent();
                    res = fst_version(gfs::ld_rlx(&e->st)) == sver && e->size == size ? 1u : 2u;
                    break;
                }
                if (gfs::now_ticks() - w0 >= gfs_g->wait_ticks) break;
                __builtin_amdgcn_s_sleep(8);
            }
            fin_flag = res;
            gfs::count(&gfs_g->ctr.spec_parts, 1);
            if (res == 2) gfs::count(&gfs_g->ctr.spec_changed, 1);
            else if (res == 1 && mine && sst) gfs::count(&gfs_g->ctr.spec_rescans, 1);
        }
        __syncthreads();
        const uint32_t res = gfs::uni(fin_flag);
        if (res == 2) {
            lr = 1;   // the file changed on the host: the last part recounts it
        } else if (res == 0 || (mine && sst)) {
            // part 0's open failed, or a page left the cache: an open of this
            // part's own, and the range again through it
            fd = gopen(name, O_GRDONLY);
            if (fd < 0) {
                st = fd;
            } else if (mine) {
                for (uint32_t q = t; q < nw; q += T) S.counts[q] = 0;
                lr = 0;
                if (t == 0) S.fsize = fstat(fd);
                __syncthreads();
                if (gfs::uni((uint64_t)S.fsize) != size) {
                    lr = 1;   // not the size the split was made for: the last part recounts
                } else {
                    const uint64_t pb = split_point(size, j, kf), pe = split_point(size, j + 1, kf);
                    st = grep_part_pinned(S, fd, chunk, size, pb, pe, &lr, false);
                    if (st == 1) st = grep_part_windows(S, fd, chunk, size, pb, pe, &lr);
                }
            }
        }
        if (mine && !st) {
            for (uint32_t q = t; q < nw; q += T) {
                const uint32_t k = S.canon[q];
                const uint32_t v = k == 0xFFFF ? 0u : S.counts[k];
                if (v) atomicAdd(&row[q], v);
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        GFS_TRACE(46);
    } else if (fd >= 0) {
        // Part 0 of a cached file that fits one pin batch pins all of it once
        // for both the chain and its own part (already, with a posted reopen).
        if (j == 0 && size && warm && !st && !held) {
            const uint64_t np = (size + gfs_g->page_size - 1) >> gfs_g->log_page;
            if (np <= maxpin) {
                st = gfs::pin_pages(gfs_g, fd, 0, (uint32_t)np);
                if (!st) held = (uint32_t)np;
            }
        }
        GFS_TRACE(44);
        if (j == 0 && size && warm && !st) {
            st = grep_chain_last(S, fd, size, chunk, &cl, &cut, held != 0);
            if (t == 0 && !single) {
                fs->c_last = cl;
                fs->cut = cut;
            }
        }
        GFS_TRACE(45);
        bool changed = false;
        if (pend) {
            // Part 0: the host's answer (it came during the chain), then the
            // same check as the other parts.  Completing the open here, before
            // this part's scan, lets the other parts join it without waiting
            // when their scans end.
            __syncthreads();
            if (t == 0) {
                const int r = gfs::open_finish(&opend);
                uint32_t ok = 0;
                if (r >= 0) {
                    gfs_file* e = &gfs_g->files[r];
                    ok = fst_version(gfs::ld_rlx(&e->st)) == sver && e->size == size;
                }
                gfs_ctl.ret = r;
                fin_flag = ok;
            }
            __syncthreads();
            fd = (int)gfs::uni(gfs_ctl.ret);
            if (fd < 0) {
                if (!st) st = fd;
            } else if (!gfs::uni(fin_flag)) {
                changed = true;
                lr = 1;   // the file changed on the host: the last part recounts it
            }
            if (t == 0) {
                gfs::count(&gfs_g->ctr.spec_parts, 1);
                if (changed) gfs::count(&gfs_g->ctr.spec_changed, 1);
            }
            defer_close = fd >= 0;
            pend = 0;
        }
        if (j < kf && size && !st && !changed) {
            const uint64_t pb = split_point(size, j, kf), pe = split_point(size, j + 1, kf);
            // a cached file is read from pinned frames, a cold one with gread
            // (whose read-ahead streams it from the host)
            st = warm ? grep_part_pinned(S, fd, chunk, size, pb, pe, &lr, held != 0) : 1;
            if (st == 1) st = grep_part_windows(S, fd, chunk, size, pb, pe, &lr);
            if (!single) {
                for (uint32_t q = t; q < nw; q += T) {
                    const uint32_t k = S.canon[q];
                    const uint32_t v = k == 0xFFFF ? 0u : S.counts[k];
                    if (v) atomicAdd(&row[q], v);
                }
                // The adds commute, so no fence orders them for the other
                // parts; each wave only waits for its adds to be performed, so
                // that a finisher that overwrites the row (grep_recount) comes
                // after them.  (A fence per thread here, an L2 write-back and
                // invalidate in every wave of 256 workgroups, cost ~40 us.)
                __builtin_amdgcn_s_waitcnt(0);
            }
        }
        GFS_TRACE(46);
        if (held) gfs::unpin_pages(gfs_g, held);
        if (fd >= 0 && j == 0 && size && !st && !warm) {
            st = grep_chain_last(S, fd, size, chunk, &cl, &cut, false);
            if (t == 0 && !single) {
                fs->c_last = cl;
                fs->cut = cut;
            }
        }
    }
    if (single) {
        // alone on the file: the counts stay in LDS until the row is stored
        if (fd >= 0 && size && !st && (lr || cut)) {
            st = grep_recount(S, fd, nw, chunk, row);
        } else {
            if (fd >= 0 && size && !st && size - cl > chunk) st = grep_tail(S, fd, cl + chunk, size, chunk);
            for (uint32_t q = t; q < nw; q += T) {
                const uint32_t k = S.canon[q];
                row[q] = k == 0xFFFF || fd < 0 ? 0u : S.counts[k];
            }
        }
        const int c = fd >= 0 ? gclose(fd) : 0;
        if (t == 0) *status = st ? st : c;
        __syncthreads();
        return true;
    }
    on_arrive();
    __syncthreads();
    if (t == 0) {
        if (lr) atomicOr(&fs->longrun, 1u);
        if (st) atomicCAS(&fs->err, 0, st);
        if (defer_close) fs->p0_open = 1;
        const uint32_t old = __hip_atomic_fetch_add(&fs->arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        fin_flag = old + 1 == (target ? target : (uint32_t)kf) ? 1u : 0u;
    }
    __syncthreads();
    GFS_TRACE(47);
    if (gfs::uni(fin_flag)) {
        // every part's counts are in the row: their adds precede their arrival
        if (t == 0) {
            S.status = __hip_atomic_load(&fs->err, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            fin_flag = __hip_atomic_load(&fs->longrun, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) | fs->cut;
            S.fsize = fs->c_last;
            spec_sh[0] = j != 0 && fs->p0_open;
        }
        __syncthreads();
        st = gfs::uni(S.status);
        const uint32_t redo = gfs::uni(fin_flag);
        const uint64_t cl = gfs::uni((uint64_t)S.fsize);
        const bool close_p0 = gfs::uni(spec_sh[0]) != 0;
        // a part that confirmed part 0's open holds none: it reads through
        // part 0's, which it closes after
        const int ffd = fd >= 0 ? fd : (close_p0 ? (int)sidx : -1);
        if (!st && size) st = ffd >= 0 ? grep_split_finish(S, ffd, nw, chunk, size, row, redo, cl) : -EBADF;
        GFS_TRACE(48);
        if (close_p0 && ffd >= 0) gclose(ffd);   // part 0's open (the same entry)
        const int c = fd >= 0 ? gclose(fd) : 0;
        if (t == 0) *status = st ? st : c;
        __syncthreads();
        GFS_TRACE(49);
        return true;
    }
    if (fd >= 0 && !defer_close) gclose(fd);
    GFS_TRACE(49);
    return false;
}

static int grep_args_ok(uint32_t n_files, uint32_t n_words, uint32_t chunk_bytes, uint32_t threads) {
    if (n_files == 0 || n_words == 0 || n_words > GREP_MAX_WORDS) return -EINVAL;
    if (chunk_bytes == 0 || chunk_bytes > GREP_MAX_CHUNK) return -EINVAL;
    if (threads < 64 || threads > 1024 || threads % 64) return -EINVAL;
    return 0;
}
