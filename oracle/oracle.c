/*
 * oracle.c -- CPU restatement of the reference's file-read path, used ONLY as
 * the checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in the product links or calls this code.
 *
 * Functions and the reference lines they restate:
 *   orc_scan_naive   samples/grep_text/grep_text_gpu.cu:79 (alpha), :139-160
 *                    (match_string, one pass per query word), :298-318 (chunk
 *                    loop with 32-byte look-ahead and find_overlap :89-101).
 *                    Literal form, O(words x bytes); small inputs only.
 *   orc_scan_fast    same semantics, one tokenising pass per window plus a hash
 *                    lookup per token (O(bytes)).  Pinned against orc_scan_naive
 *                    on adversarial inputs and against the generator's tally.
 *   orc_scan_files   the CPU path: pread() each window of each file (the gread
 *                    of :306) + orc_scan_fast's window scan, N threads.
 *   orc_fnv1a64      FNV-1a 64-bit (offset basis 0xcbf29ce484222325, prime
 *                    0x100000001b3), pinned by the FNV specification's test
 *                    vectors.
 *   orc_pread_ranges the CPU path of the sequential read: pread() of disjoint
 *                    ranges (libgpufs/host_loop.h:491 does the same pread per
 *                    page), optional FNV-1a-64 per range, N threads.
 *   orc_gemv_f64     y = A x accumulated in double (the error reference for the
 *                    mapped-matrix GEMV, samples/bigmatrix_mmap/bigmatrix_mmap.cu:24-41).
 *   orc_gemv_file    the same over a matrix file: threaded pread of row blocks,
 *                    y (double or float accumulation) and per-block word sums.
 *   orc_or_merge     the O_GWRONCE write-back merge, libgpufs/host_loop.h:76-137.
 *   orc_grep_text_lines  the grep_text sample's output lines (grep_text_gpu.cu:286-349).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ---------------------------------------------------------------- scan */

/* grep_text_gpu.cu:79 -- on a signed char, so bytes >= 0x80 are never alpha */
static inline int alpha(uint8_t c) { return (c >= 65 && c <= 90) || (c >= 97 && c <= 122) || c == 95; }

/* grep_text_gpu.cu:139-160.  `a` is a 33-byte buffer: the 32-byte record and a
 * terminating NUL (the reference reads a[32] from uninitialised shared memory;
 * it is defined as NUL here). */
static int match_string(const char* a, const uint8_t* data, int data_size) {
    int matches = 0;
    int sizecount = 0;
    int word_start = 1;
    if (a[0] == 0) return -1; /* LEN_ZERO */
    for (int i = 0; i < data_size; i++) {
        if (!alpha(data[i])) {
            if ((sizecount == 32 || a[sizecount] == '\0') && word_start) matches++;
            word_start = 1;
            sizecount = 0;
        } else {
            if ((uint8_t)a[sizecount] == data[i]) sizecount++;
            else {
                word_start = 0;
                sizecount = 0;
            }
        }
    }
    return matches;
}

typedef void (*window_fn)(void* st, const uint8_t* win, int dsz);

/* The chunk loop of grep_text_gpu.cu:298-318 over an in-memory file. */
static void chunk_loop(const uint8_t* data, size_t n, uint32_t chunk, uint8_t* buf, window_fn fn, void* st) {
    for (size_t cursor = 0; cursor < n;) {
        int last = (n - cursor) < (size_t)chunk + 32;
        size_t left = last ? n - cursor : (size_t)chunk + 32;
        buf[left] = 0;
        memcpy(buf, data + cursor, left);
        int overlap = 0;
        if (!last) {
            int i = 0;
            for (; i < 32 && alpha(buf[chunk + i]); i++);
            overlap = i;
            cursor += overlap;
        }
        cursor += chunk;
        int dsz = last ? (int)left + 1 : (int)chunk + overlap + 1;
        fn(st, buf, dsz);
    }
}

typedef struct {
    const char* dict;
    uint32_t nw;
    uint32_t* counts;
} naive_st;

static void naive_window(void* p, const uint8_t* win, int dsz) {
    naive_st* s = (naive_st*)p;
    char a[33];
    for (uint32_t q = 0; q < s->nw; q++) {
        memcpy(a, s->dict + (size_t)q * 32, 32);
        a[32] = 0;
        int r = match_string(a, win, dsz);
        if (r > 0) s->counts[q] += (uint32_t)r;
    }
}

int orc_scan_naive(const uint8_t* data, size_t n, uint32_t chunk, const char* dict, uint32_t nw, uint32_t* counts) {
    if (chunk == 0) return -EINVAL;
    uint8_t* buf = (uint8_t*)malloc((size_t)chunk + 64);
    if (!buf) return -ENOMEM;
    naive_st s = {dict, nw, counts};
    chunk_loop(data, n, chunk, buf, naive_window, &s);
    free(buf);
    return 0;
}

/* ---- fast form: dictionary hash (canonical word per distinct string) */
typedef struct {
    uint32_t cap;      /* power of two */
    int32_t* slot;     /* canonical word index or -1 */
    const char* dict;
    uint8_t* len;      /* C-string length of each record (0..32) */
    int32_t* canon;    /* canonical index, -1 = never matches */
    uint32_t nw;
} dict_map;

static uint32_t hash_bytes(const uint8_t* p, uint32_t n) {
    uint32_t h = 2166136261u;
    for (uint32_t i = 0; i < n; i++) h = (h ^ p[i]) * 16777619u;
    return h;
}

static int dict_build(dict_map* m, const char* dict, uint32_t nw) {
    m->nw = nw;
    m->dict = dict;
    m->cap = 16;
    while (m->cap < 2 * nw + 2) m->cap <<= 1;
    m->slot = (int32_t*)malloc(sizeof(int32_t) * m->cap);
    m->len = (uint8_t*)malloc(nw + 1);
    m->canon = (int32_t*)malloc(sizeof(int32_t) * (nw + 1));
    if (!m->slot || !m->len || !m->canon) return -ENOMEM;
    for (uint32_t i = 0; i < m->cap; i++) m->slot[i] = -1;
    for (uint32_t q = 0; q < nw; q++) {
        const uint8_t* w = (const uint8_t*)dict + (size_t)q * 32;
        uint32_t L = 0;
        while (L < 32 && w[L]) L++;
        m->len[q] = (uint8_t)L;
        int ok = L > 0;
        for (uint32_t i = 0; i < L; i++)
            if (!alpha(w[i])) ok = 0;
        m->canon[q] = -1;
        if (!ok) continue; /* empty (LEN_ZERO) or holds a non-word byte: can never match */
        uint32_t pos = hash_bytes(w, L) & (m->cap - 1);
        for (;;) {
            int32_t e = m->slot[pos];
            if (e < 0) {
                m->slot[pos] = (int32_t)q;
                m->canon[q] = (int32_t)q;
                break;
            }
            if (m->len[e] == L && memcmp(dict + (size_t)e * 32, w, L) == 0) {
                m->canon[q] = e;
                break;
            }
            pos = (pos + 1) & (m->cap - 1);
        }
    }
    return 0;
}
static void dict_free(dict_map* m) {
    free(m->slot);
    free(m->len);
    free(m->canon);
}
static inline int32_t dict_find(const dict_map* m, const uint8_t* tok, uint32_t L) {
    uint32_t pos = hash_bytes(tok, L) & (m->cap - 1);
    for (;;) {
        int32_t e = m->slot[pos];
        if (e < 0) return -1;
        if (m->len[e] == L && memcmp(m->dict + (size_t)e * 32, tok, L) == 0) return e;
        pos = (pos + 1) & (m->cap - 1);
    }
}

typedef struct {
    const dict_map* m;
    uint32_t* acc; /* per canonical index */
} fast_st;

/* One window: every maximal run of word bytes that starts inside the window
 * (the window start counts as a boundary) and is at most 32 bytes long is
 * looked up; the window always ends with a non-word byte. */
static void fast_window(void* p, const uint8_t* win, int dsz) {
    fast_st* s = (fast_st*)p;
    int i = 0;
    while (i < dsz) {
        if (!alpha(win[i])) {
            i++;
            continue;
        }
        int j = i;
        while (j < dsz && alpha(win[j])) j++;
        if (j < dsz && j - i <= 32) {
            int32_t e = dict_find(s->m, win + i, (uint32_t)(j - i));
            if (e >= 0) s->acc[e]++;
        }
        i = j;
    }
}

static void spread_counts(const dict_map* m, const uint32_t* acc, uint32_t* counts) {
    for (uint32_t q = 0; q < m->nw; q++)
        if (m->canon[q] >= 0) counts[q] += acc[m->canon[q]];
}

int orc_scan_fast(const uint8_t* data, size_t n, uint32_t chunk, const char* dict, uint32_t nw, uint32_t* counts) {
    if (chunk == 0) return -EINVAL;
    dict_map m;
    if (dict_build(&m, dict, nw)) return -ENOMEM;
    uint32_t* acc = (uint32_t*)calloc(nw + 1, sizeof(uint32_t));
    uint8_t* buf = (uint8_t*)malloc((size_t)chunk + 64);
    if (!acc || !buf) return -ENOMEM;
    fast_st s = {&m, acc};
    chunk_loop(data, n, chunk, buf, fast_window, &s);
    spread_counts(&m, acc, counts);
    free(acc);
    free(buf);
    dict_free(&m);
    return 0;
}

/* ---- the grep_text sample's output file (grep_text_gpu.cu:286-349):
 * database names come from a list in the format get_next() parses
 * (grep_text_gpu.cu:163-186: leading ' ', '\t', ',', '\r', '\n' skipped; a name
 * ends at '\n', '\r', ',' or NUL; a name of 64 or more bytes, or an empty one,
 * ends the list); for each database, each window of the chunk loop and each
 * query record q in order whose count in that window is nonzero, the line
 * "<word> <count> <database name>\n".  The reference emits these lines in an
 * order set by thread timing, so callers compare sorted lines.  Counts are
 * printed in full (the reference's print_uint, grep_text_gpu.cu:34-54, swaps
 * the middle digits of 4-digit counts and overflows its buffer at 5).
 * Returns the bytes of output (written up to `cap`) or -errno. */
static int64_t pread_all(int fd, uint8_t* buf, size_t len, uint64_t off);

typedef struct {
    const dict_map* m;
    uint32_t* acc;
    const char* dict;
    uint32_t nw;
    const char* name;
    int name_len;
    char* out;
    uint64_t cap, len;
} lines_st;

static void lines_window(void* p, const uint8_t* win, int dsz) {
    lines_st* s = (lines_st*)p;
    fast_st f = {s->m, s->acc};
    memset(s->acc, 0, sizeof(uint32_t) * (s->nw + 1));
    fast_window(&f, win, dsz);
    for (uint32_t q = 0; q < s->nw; q++) {
        if (s->m->canon[q] < 0) continue;
        uint32_t c = s->acc[s->m->canon[q]];
        if (!c) continue;
        char num[16];
        int nl = 0;
        char rev[16];
        do { rev[nl++] = (char)('0' + c % 10); c /= 10; } while (c);
        for (int i = 0; i < nl; i++) num[i] = rev[nl - 1 - i];
        const uint32_t wl = s->m->len[q];
        const uint64_t ll = (uint64_t)wl + 1 + (uint64_t)nl + 1 + (uint64_t)s->name_len + 1;
        if (s->len + ll <= s->cap) {
            char* o = s->out + s->len;
            memcpy(o, s->dict + (size_t)q * 32, wl);
            o[wl] = ' ';
            memcpy(o + wl + 1, num, (size_t)nl);
            o[wl + 1 + nl] = ' ';
            memcpy(o + wl + 2 + nl, s->name, (size_t)s->name_len);
            o[ll - 1] = '\n';
        }
        s->len += ll;
    }
}

int64_t orc_grep_text_lines(const char* dict, uint32_t nw, const char* dblist, uint64_t dblist_len, uint32_t chunk,
                            char* out, uint64_t cap) {
    if (chunk == 0) return -EINVAL;
    dict_map m;
    if (dict_build(&m, dict, nw)) return -ENOMEM;
    uint32_t* acc = (uint32_t*)calloc(nw + 1, sizeof(uint32_t));
    uint8_t* buf = (uint8_t*)malloc((size_t)chunk + 64);
    if (!acc || !buf) return -ENOMEM;
    lines_st st = {&m, acc, dict, nw, NULL, 0, out, cap, 0};
    uint64_t i = 0;
    int64_t rc = 0;
    for (;;) {
        while (i < dblist_len && (dblist[i] == ' ' || dblist[i] == '\t' || dblist[i] == ',' || dblist[i] == '\r' ||
                                  dblist[i] == '\n'))
            i++;
        const uint64_t beg = i;
        while (i < dblist_len && dblist[i] != '\n' && dblist[i] != '\r' && dblist[i] != ',' && dblist[i] != 0 && i < 64 + beg)
            i++;
        const uint64_t nlen = i - beg;
        if (nlen == 0 || nlen >= 64) break;
        char name[65];
        memcpy(name, dblist + beg, nlen);
        name[nlen] = 0;
        int fd = open(name, O_RDONLY);
        if (fd < 0) { rc = -errno; break; }
        off_t sz = lseek(fd, 0, SEEK_END);
        uint8_t* data = (uint8_t*)malloc((size_t)sz + 1);
        if (!data || pread_all(fd, data, (size_t)sz, 0) != (int64_t)sz) { close(fd); free(data); rc = -EIO; break; }
        close(fd);
        st.name = name;
        st.name_len = (int)nlen;
        chunk_loop(data, (size_t)sz, chunk, buf, lines_window, &st);
        free(data);
        if (i >= dblist_len) break;
        i++;   /* the byte that ended the name */
    }
    free(acc);
    free(buf);
    dict_free(&m);
    return rc ? rc : (int64_t)st.len;
}

/* ---- CPU path over files: pread windows + fast scan, one file per task */
typedef struct {
    const char* names;
    uint32_t stride, nfiles, chunk, nw;
    const char* dict;
    uint32_t* counts;
    int32_t* status;
    const dict_map* m;
    atomic_uint next;
} files_job;

static int64_t pread_all(int fd, uint8_t* buf, size_t len, uint64_t off) {
    size_t done = 0;
    while (done < len) {
        ssize_t r = pread(fd, buf + done, len - done, (off_t)(off + done));
        if (r < 0) {
            if (errno == EINTR) continue;
            return -errno;
        }
        if (r == 0) break;
        done += (size_t)r;
    }
    return (int64_t)done;
}

static void* files_worker(void* p) {
    files_job* j = (files_job*)p;
    uint8_t* buf = (uint8_t*)malloc((size_t)j->chunk + 64);
    uint32_t* acc = (uint32_t*)malloc(sizeof(uint32_t) * (j->nw + 1));
    for (;;) {
        uint32_t f = atomic_fetch_add(&j->next, 1);
        if (f >= j->nfiles) break;
        memset(acc, 0, sizeof(uint32_t) * (j->nw + 1));
        int fd = open(j->names + (size_t)f * j->stride, O_RDONLY | O_CLOEXEC);
        if (fd < 0) {
            j->status[f] = -errno;
            continue;
        }
        off_t size = lseek(fd, 0, SEEK_END);
        fast_st s = {j->m, acc};
        int st = 0;
        for (uint64_t cursor = 0; cursor < (uint64_t)size;) {
            int last = ((uint64_t)size - cursor) < (uint64_t)j->chunk + 32;
            size_t left = last ? (size_t)((uint64_t)size - cursor) : (size_t)j->chunk + 32;
            buf[left] = 0;
            int64_t r = pread_all(fd, buf, left, cursor);
            if (r != (int64_t)left) {
                st = r < 0 ? (int)r : -EIO;
                break;
            }
            int overlap = 0;
            if (!last) {
                int i = 0;
                for (; i < 32 && alpha(buf[j->chunk + i]); i++);
                overlap = i;
                cursor += overlap;
            }
            cursor += j->chunk;
            fast_window(&s, buf, last ? (int)left + 1 : (int)j->chunk + overlap + 1);
        }
        close(fd);
        j->status[f] = st;
        spread_counts(j->m, acc, j->counts + (size_t)f * j->nw);
    }
    free(buf);
    free(acc);
    return NULL;
}

int orc_scan_files(const char* names, uint32_t stride, uint32_t nfiles, uint32_t chunk, const char* dict, uint32_t nw,
                   uint32_t* counts, int32_t* status, int threads) {
    if (chunk == 0 || threads <= 0) return -EINVAL;
    dict_map m;
    if (dict_build(&m, dict, nw)) return -ENOMEM;
    files_job j = {names, stride, nfiles, chunk, nw, dict, counts, status, &m, 0};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, files_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
    dict_free(&m);
    return 0;
}

/* ---------------------------------------------------------------- FNV-1a */
uint64_t orc_fnv1a64_update(uint64_t h, const uint8_t* p, size_t n) {
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
uint64_t orc_fnv1a64(const uint8_t* p, size_t n) { return orc_fnv1a64_update(0xcbf29ce484222325ull, p, n); }

/* ---- CPU path of the sequential read */
typedef struct {
    int fd;
    uint32_t nranges;
    uint64_t range;
    uint8_t* dst;
    uint64_t* fnv;
    int64_t* got;
    atomic_uint next;
} ranges_job;

static void* ranges_worker(void* p) {
    ranges_job* j = (ranges_job*)p;
    const size_t piece = 4u << 20;
    uint8_t* scratch = j->dst ? NULL : (uint8_t*)malloc(piece);
    for (;;) {
        uint32_t r = atomic_fetch_add(&j->next, 1);
        if (r >= j->nranges) break;
        uint64_t off = (uint64_t)r * j->range;
        uint64_t h = 0xcbf29ce484222325ull;
        int64_t total = 0;
        for (uint64_t done = 0; done < j->range;) {
            size_t len = (size_t)((j->range - done) < piece ? (j->range - done) : piece);
            uint8_t* b = j->dst ? j->dst + off + done : scratch;
            int64_t g = pread_all(j->fd, b, len, off + done);
            if (g < 0) {
                total = g;
                break;
            }
            if (j->fnv) h = orc_fnv1a64_update(h, b, (size_t)g);
            total += g;
            done += len;
            if ((size_t)g < len) break;
        }
        if (j->fnv) j->fnv[r] = h;
        if (j->got) j->got[r] = total;
    }
    free(scratch);
    return NULL;
}

int orc_pread_ranges(const char* path, uint32_t nranges, uint64_t range, uint8_t* dst, uint64_t* fnv_out,
                     int64_t* got_out, int threads) {
    if (threads <= 0) return -EINVAL;
    int fd = open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return -errno;
    ranges_job j = {fd, nranges, range, dst, fnv_out, got_out, 0};
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, ranges_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
    close(fd);
    return 0;
}

/* ---------------------------------------------------------------- GEMV */
void orc_gemv_f64(const float* A, uint64_t rows, uint64_t cols, const float* x, double* y) {
    for (uint64_t r = 0; r < rows; r++) {
        const float* a = A + r * cols;
        double s = 0.0;
        for (uint64_t c = 0; c < cols; c++) s += (double)a[c] * (double)x[c];
        y[r] = s;
    }
}

/* CPU path of the mapped-matrix GEMV over a file (samples/bigmatrix_mmap:
 * bigmatrix_mmap.cu:66-168 reads rows page by page; bigmatrix_cpu.cpp:144-159
 * is the reference's own CPU GEMV).  Rows are pread() in blocks of
 * `rows_per_block`; each block gets y for its rows (accumulated in double when
 * `f64`, else in float in column order) and the wrapping sum of the block's
 * 8-byte words (the GPU kernel's per-workgroup checksum of the mapped bytes).
 * N threads take blocks in turn.  Returns 0 or -errno. */
struct gemv_job {
    int fd;
    uint64_t rows, cols, rpb, nblocks;
    const float* x;
    double* y;
    uint64_t* csum;
    int f64;
    atomic_uint_fast64_t next;
    atomic_int err;
};
static void* gemv_worker(void* arg) {
    struct gemv_job* j = (struct gemv_job*)arg;
    const uint64_t rb = j->cols * 4;
    float* buf = (float*)malloc(rb);
    if (!buf) { atomic_store(&j->err, -ENOMEM); return NULL; }
    for (;;) {
        uint64_t b = atomic_fetch_add(&j->next, 1);
        if (b >= j->nblocks) break;
        uint64_t cs = 0;
        for (uint64_t r = b * j->rpb; r < (b + 1) * j->rpb && r < j->rows; r++) {
            uint64_t done = 0;
            while (done < rb) {
                ssize_t n = pread(j->fd, (char*)buf + done, rb - done, (off_t)(r * rb + done));
                if (n <= 0) { atomic_store(&j->err, n < 0 ? -errno : -EIO); free(buf); return NULL; }
                done += (uint64_t)n;
            }
            if (j->f64) {
                double s = 0.0;
                for (uint64_t c = 0; c < j->cols; c++) s += (double)buf[c] * (double)j->x[c];
                j->y[r] = s;
            } else {
                float s = 0.0f;
                for (uint64_t c = 0; c < j->cols; c++) s += buf[c] * j->x[c];
                j->y[r] = (double)s;
            }
            const uint64_t* w = (const uint64_t*)buf;
            for (uint64_t k = 0; k < rb / 8; k++) cs += w[k];
        }
        j->csum[b] = cs;
    }
    free(buf);
    return NULL;
}
int orc_gemv_file(const char* path, uint64_t rows, uint64_t cols, const float* x, uint64_t rows_per_block,
                  double* y, uint64_t* csum, int f64, int threads) {
    if (rows_per_block == 0 || (cols * 4) % 8) return -EINVAL;
    int fd = open(path, O_RDONLY);
    if (fd < 0) return -errno;
    struct gemv_job j;
    j.fd = fd; j.rows = rows; j.cols = cols; j.rpb = rows_per_block;
    j.nblocks = (rows + rows_per_block - 1) / rows_per_block;
    j.x = x; j.y = y; j.csum = csum; j.f64 = f64;
    atomic_init(&j.next, 0);
    atomic_init(&j.err, 0);
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, gemv_worker, &j);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
    close(fd);
    return atomic_load(&j.err);
}

/* ---------------------------------------------------------------- write merge */
void orc_or_merge(uint8_t* page, const uint8_t* disk, uint64_t disk_bytes) {
    for (uint64_t i = 0; i < disk_bytes; i++) page[i] |= disk[i];
}
