// host_service.h -- host side of the GPU->host RPC ring (no HIP dependency).
//
// Replaces the reference's host loop (libgpufs/host_loop.h:156-557: open_loop,
// rw_task, mainLoop, async_close_loop).  Differences by design:
//  * one submission queue instead of per-worker slot scans: the GPU appends
//    (ticket << 16 | slot) entries; any idle worker claims the next ticket with
//    a CAS, so requests are served in ticket order by all workers at once;
//  * the worker pread()s straight into the slot's pinned staging buffer (the
//    GPU copies it into HBM), so there is no host-side batching thread and no
//    H2D memcpy on the host;
//  * completion = result field + a release OR of the request's bit into its
//    batch's 64-bit completion mask, which one thread of the workgroup polls,
//    so a workgroup can use each page as soon as it has arrived;
//  * the service runs independently of kernel completion (start/stop), which
//    persistent and graph-launched kernels need; run_handler adds the
//    reference's "until the stream is idle" behaviour on top.
// All ring memory is plain host memory here; the runtime allocates it as
// coherent pinned memory visible to the GPU.
#pragma once
#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <climits>
#include <fcntl.h>
#include <linux/futex.h>
#include <sys/prctl.h>
#include <sys/syscall.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>
#include "gfs_layout.h"
#if defined(__x86_64__)
#include <emmintrin.h>
#endif

namespace gfs_host {

struct RingView {
    gfs_rpc_slot* slots = nullptr;
    uint64_t* sq = nullptr;
    uint64_t* cq = nullptr;     // completion masks
    uint8_t* staging = nullptr;
    int64_t* closeq = nullptr;  // host fd of the slot-less close with ticket t at [t & (sq_depth - 1)]
    uint32_t sq_depth = 0;      // power of two
    uint32_t ring_depth = 0;
    uint32_t log_page = 0;
};

struct ServiceCounters {
    std::atomic<uint64_t> requests{0}, bytes_read{0}, bytes_written{0}, errors{0};
    std::atomic<uint64_t> busy_ns{0};   // summed over workers: time spent serving requests
};

static inline uint64_t ts_ns(const struct timespec& t) {
    return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

// pread/pwrite until `len` bytes moved or EOF / error.
static inline int64_t pread_full(int fd, uint8_t* buf, uint64_t len, uint64_t off) {
    uint64_t done = 0;
    while (done < len) {
        ssize_t r = ::pread(fd, buf + done, len - done, (off_t)(off + done));
        if (r < 0) {
            if (errno == EINTR) continue;
            return -(int64_t)errno;
        }
        if (r == 0) break;
        done += (uint64_t)r;
    }
    return (int64_t)done;
}
static inline int64_t pwrite_full(int fd, const uint8_t* buf, uint64_t len, uint64_t off) {
    uint64_t done = 0;
    while (done < len) {
        ssize_t r = ::pwrite(fd, buf + done, len - done, (off_t)(off + done));
        if (r < 0) {
            if (errno == EINTR) continue;
            return -(int64_t)errno;
        }
        done += (uint64_t)r;
    }
    return (int64_t)done;
}

// Per-worker buffers: the OR-merge's copy of the disk bytes, and the bounce
// buffer of streamed page reads.
struct WorkerBufs {
    std::vector<uint8_t> merge;
    std::vector<uint8_t> bounce;
};

// Page reads land in the staging page through a small bounce buffer that
// stays in the core's cache: pread() copies the file bytes there, and they
// are streamed on to staging with non-temporal stores.  The staging lines then
// sit in DRAM, not dirty in a CPU cache, when the GPU reads them over PCIe
// (measured on MI355X: cold sequential reads about 2.5 % faster than
// pread() straight into staging, with less host time per page).
static constexpr uint64_t kBounceBytes = 64 << 10;
static inline void stream_copy(uint8_t* dst, const uint8_t* src, uint64_t n) {
    uint64_t i = 0;
#if defined(__x86_64__)
    if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) {
        for (; i + 64 <= n; i += 64) {
            const __m128i a = _mm_load_si128((const __m128i*)(src + i));
            const __m128i b = _mm_load_si128((const __m128i*)(src + i + 16));
            const __m128i c = _mm_load_si128((const __m128i*)(src + i + 32));
            const __m128i d = _mm_load_si128((const __m128i*)(src + i + 48));
            _mm_stream_si128((__m128i*)(dst + i), a);
            _mm_stream_si128((__m128i*)(dst + i + 16), b);
            _mm_stream_si128((__m128i*)(dst + i + 32), c);
            _mm_stream_si128((__m128i*)(dst + i + 48), d);
        }
    }
#endif
    if (i < n) memcpy(dst + i, src + i, n - i);
}
// pread of [off, off+len) into `stage`: the bytes read, or -errno.  The
// streamed stores are fenced before the caller publishes the result.
static inline int64_t pread_staged(int fd, uint8_t* stage, uint64_t len, uint64_t off, std::vector<uint8_t>& bounce) {
    if (bounce.size() < kBounceBytes + 64) bounce.resize(kBounceBytes + 64);
    uint8_t* b = (uint8_t*)(((uintptr_t)bounce.data() + 63) & ~(uintptr_t)63);
    uint64_t done = 0;
    while (done < len) {
        const uint64_t want = len - done < kBounceBytes ? len - done : kBounceBytes;
        const int64_t r = pread_full(fd, b, want, off + done);
        if (r < 0) return r;
        stream_copy(stage + done, b, (uint64_t)r);
        done += (uint64_t)r;
        if ((uint64_t)r < want) break;   // end of file
    }
#if defined(__x86_64__)
    _mm_sfence();
#endif
    return (int64_t)done;
}

// O_GWRONCE write-back: new bytes are OR-ed with the bytes already on disk
// (libgpufs/host_loop.h:76-137).  The reference ORs whole 16-byte vectors
// against a buffer that may hold stale bytes past the read; here exactly the
// min(len, bytes on disk) overlapping bytes are merged.
static inline int64_t diff_merge_write(int fd, uint8_t* page, uint64_t len, uint64_t off,
                                       std::vector<uint8_t>& tmp) {
    tmp.resize(len);
    int64_t r = pread_full(fd, tmp.data(), len, off);
    if (r < 0) return r;
    for (int64_t i = 0; i < r; ++i) page[i] |= tmp[(size_t)i];
    return pwrite_full(fd, page, len, off);
}

// Serve one request in slot s.  Returns the result stored in the slot.
static inline int64_t serve_slot(const RingView& rv, uint32_t s, ServiceCounters& ctr, WorkerBufs& wb) {
    gfs_rpc_slot* r = &rv.slots[s];
    const uint32_t op = r->op;
    const int fd = r->host_fd;
    const uint64_t off = r->offset, len = r->length;
    uint8_t* stage = rv.staging + ((uint64_t)s << rv.log_page);
    int64_t res = -EINVAL;
    switch (op) {
    case GFS_OP_READ:
        res = pread_staged(fd, stage, len, off, wb.bounce);
        if (res > 0) ctr.bytes_read += (uint64_t)res;
        break;
    case GFS_OP_WRITE:
        res = pwrite_full(fd, stage, len, off);
        if (res > 0) ctr.bytes_written += (uint64_t)res;
        break;
    case GFS_OP_DIFF:
        res = diff_merge_write(fd, stage, len, off, wb.merge);
        if (res > 0) ctr.bytes_written += (uint64_t)res;
        break;
    case GFS_OP_OPEN: {
        char name[GFS_NAME_MAX + 1];
        memcpy(name, r->name, GFS_NAME_MAX);
        name[GFS_NAME_MAX] = 0;
        int flags = r->flags;
        // O_GWRONCE files are opened read-write so the OR-merge can read them
        // (libgpufs/host_loop.h:170-173)
        if ((flags & (O_CREAT | O_ACCMODE)) == (O_CREAT | O_WRONLY)) flags = O_RDWR | O_CREAT;
        int hfd = ::open(name, flags | O_CLOEXEC, 0644);
        if (hfd < 0) {
            res = -(int64_t)errno;
            break;
        }
        struct stat st;
        if (::fstat(hfd, &st) != 0) {
            res = -(int64_t)errno;
            ::close(hfd);
            break;
        }
        r->size = (uint64_t)st.st_size;
        r->ctime_ns = ts_ns(st.st_ctim);
        r->mtime_ns = ts_ns(st.st_mtim);
        r->ino = (uint64_t)st.st_ino;
        res = hfd;
        break;
    }
    case GFS_OP_CLOSE:
        res = ::close(fd) == 0 ? 0 : -(int64_t)errno;
        break;
    case GFS_OP_TRUNC:
        res = ::ftruncate(fd, (off_t)off) == 0 ? 0 : -(int64_t)errno;
        break;
    default:
        res = -EINVAL;
    }
    if (res < 0) ctr.errors++;
    ctr.requests++;
    // result first, then the completion bit; a reader of either sees the staging data
    const uint32_t cqw = r->cq;
    __atomic_store_n(&r->result, res, __ATOMIC_RELEASE);
    __atomic_fetch_or(&rv.cq[cqw & 0xFFFFu], 1ull << ((cqw >> 16) & 63u), __ATOMIC_RELEASE);
    return res;
}

// Try to take the next published request.  Returns true and sets *slot.
// Takes the next published request.  For a close posted without a slot, the
// host fd is read from closeq before the ticket is taken: while the ticket is
// still the head, no later ticket can reuse its closeq entry (the device never
// has more than sq_depth tickets ahead of the head), and a failed take drops
// the value.
static inline bool try_take(const RingView& rv, std::atomic<uint64_t>& head, uint32_t* slot, int64_t* close_fd) {
    uint64_t h = head.load(std::memory_order_relaxed);
    uint64_t e = __atomic_load_n(&rv.sq[h & (rv.sq_depth - 1)], __ATOMIC_ACQUIRE);
    if ((e >> 16) != (h & 0xFFFFFFFFFFFFull)) return false;
    const uint32_t s = (uint32_t)(e & 0xFFFF);
    const int64_t fd = s == GFS_SQ_NOSLOT ? __atomic_load_n(&rv.closeq[h & (rv.sq_depth - 1)], __ATOMIC_RELAXED) : -1;
    if (!head.compare_exchange_strong(h, h + 1, std::memory_order_acq_rel)) return false;
    *slot = s;
    *close_fd = fd;
    return true;
}

// A close posted without a slot (the device does not wait for it).
static inline void serve_slotless_close(int64_t fd, ServiceCounters& ctr) {
    if (fd < 0 || ::close((int)fd) != 0) ctr.errors.fetch_add(1, std::memory_order_relaxed);
    ctr.requests.fetch_add(1, std::memory_order_relaxed);
}

// Worker loop: spin while work arrives, back off to short sleeps when idle.
static inline uint64_t mono_ns() {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return ts_ns(t);
}

// Sleeping workers wait on `word` (a futex) with a timeout; a worker that
// takes a request while more are already queued wakes them, so a burst (the
// page reads of a large cold file) is served by all workers at once instead
// of by the one that was awake until the others' timers expire.
struct WakeState {
    std::atomic<uint32_t> word{0};
    std::atomic<uint32_t> sleepers{0};
    std::atomic<uint32_t> kicks{0};   // launches announced by the host (Service::kick)
};

static inline void wake_all(WakeState* w) {
    w->word.fetch_add(1, std::memory_order_release);
    syscall(SYS_futex, (uint32_t*)&w->word, FUTEX_WAKE_PRIVATE, INT_MAX, nullptr, nullptr, 0);
}

// True when the submission queue holds a published request at `head`.
static inline bool backlog(const RingView& rv, const std::atomic<uint64_t>& head) {
    const uint64_t h = head.load(std::memory_order_relaxed);
    const uint64_t e = __atomic_load_n(&rv.sq[h & (rv.sq_depth - 1)], __ATOMIC_ACQUIRE);
    return (e >> 16) == (h & 0xFFFFFFFFFFFFull);
}

// Worker `index` of the service.  Worker 0 polls the queue without sleeping
// while the service runs, so a lone request (an open between two launches)
// is picked up within a poll; the others spin for a while after their last
// request, then sleep on the wake futex (2 us steps, 50 us after a longer
// idle; timer slack lowered to 1 us so short sleeps are not stretched).
static inline void worker_loop(RingView rv, std::atomic<uint64_t>* head, std::atomic<bool>* stop,
                               ServiceCounters* ctr, WakeState* wk, unsigned index) {
    prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
    WorkerBufs wb;
    uint64_t idle = 0;
    uint32_t kicks = wk->kicks.load(std::memory_order_relaxed);
    for (;;) {
        uint32_t s;
        int64_t close_fd;
        if (try_take(rv, *head, &s, &close_fd)) {
            if (wk->sleepers.load(std::memory_order_relaxed) && backlog(rv, *head)) wake_all(wk);
            const uint64_t t0 = mono_ns();
            if (s == GFS_SQ_NOSLOT) serve_slotless_close(close_fd, *ctr);
            else serve_slot(rv, s, *ctr, wb);
            ctr->busy_ns.fetch_add(mono_ns() - t0, std::memory_order_relaxed);
            idle = 0;
            continue;
        }
        // stopping: only once every published request is taken, so closes the
        // device posted without waiting are not left behind
        if (stop->load(std::memory_order_relaxed)) break;
        ++idle;
        // a launch was announced: spin again, its first requests are close
        const uint32_t k = wk->kicks.load(std::memory_order_relaxed);
        if (k != kicks) {
            kicks = k;
            idle = 0;
        }
        if (index == 0 || idle < 20000) {
#if defined(__x86_64__)
            __builtin_ia32_pause();
#endif
        } else {
            const uint32_t seen = wk->word.load(std::memory_order_acquire);
            wk->sleepers.fetch_add(1, std::memory_order_acq_rel);
            if (!backlog(rv, *head)) {
                struct timespec ts = {0, idle < 200000 ? 2000 : 50000};
                syscall(SYS_futex, (uint32_t*)&wk->word, FUTEX_WAIT_PRIVATE, seen, &ts, nullptr, 0);
            }
            wk->sleepers.fetch_sub(1, std::memory_order_acq_rel);
        }
    }
}

class Service {
public:
    void start(const RingView& rv, unsigned nthreads) {
        if (running_) return;
        stop_.store(false);
        for (unsigned i = 0; i < nthreads; ++i)
            threads_.emplace_back(worker_loop, rv, &head_, &stop_, &ctr_, &wake_, i);
        running_ = true;
    }
    void stop() {
        if (!running_) return;
        stop_.store(true);
        wake_all(&wake_);
        for (auto& t : threads_) t.join();
        threads_.clear();
        running_ = false;
    }
    bool running() const { return running_; }
    // Wakes sleeping workers and keeps every worker spinning for its spin
    // period (20,000 polls, a millisecond or more), so that the requests of a
    // kernel launched right after are served without a wake-up delay.
    void kick() {
        if (!running_) return;
        wake_.kicks.fetch_add(1, std::memory_order_relaxed);
        wake_all(&wake_);
    }
    ServiceCounters& counters() { return ctr_; }
    uint64_t head() const { return head_.load(); }
    ~Service() { stop(); }

private:
    std::vector<std::thread> threads_;
    std::atomic<uint64_t> head_{0};
    std::atomic<bool> stop_{false};
    ServiceCounters ctr_;
    WakeState wake_;
    bool running_ = false;
};

} // namespace gfs_host
