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
#include <algorithm>
#include <climits>
#include <memory>
#include <string>
#include <unordered_map>
#include <sys/resource.h>
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

// Host descriptors kept across a file's close and reopen.  The device closes
// a file when its last user leaves and opens it again on the next gopen
// (libgpufs/fs_calls.cu:152-225: a reopen goes to the host, which opens the
// path and returns its size and ctime, so a changed file gets a new cache
// version).  Opening and closing cost 4-7 us of host time when 16 workers do
// it at once (contention in the kernel), against well under 1 us for a
// stat(): so a descriptor opened with no flag beyond its access mode is kept
// after its close, and a reopen of the same path with the same access mode
// reuses it when stat() of the path still finds the same file (same device
// and inode); its size and times then come from that stat(), as they would
// from a fresh open.  Anything else (O_CREAT, O_TRUNC, ..., a file replaced or
// gone, a descriptor in use) takes a real open.  The table is split into
// shards by path hash, each under a spin lock held for a lookup only (a
// blocking mutex parked the workers of a burst of opens in the kernel); at
// most `cap` descriptors are kept (per shard, the least recently closed eighth
// of the idle ones goes when a shard is full); all are closed with the service.
class DescCache {
public:
    DescCache() {
        struct rlimit rl;
        size_t lim = 1024;
        if (getrlimit(RLIMIT_NOFILE, &rl) == 0 && rl.rlim_cur != RLIM_INFINITY) lim = (size_t)rl.rlim_cur;
        const size_t cap = lim / 4 < 4096 ? lim / 4 : 4096;
        shard_cap_ = cap / kShards;
        for (auto& o : fd_shard_) o.store(0, std::memory_order_relaxed);
    }
    ~DescCache() { clear(); }
    // A host descriptor for `name` opened with `flags`, with its stat; the
    // descriptor or -errno.
    int64_t open(const char* name, int flags, struct stat* st) {
        const bool cacheable = (flags & ~O_ACCMODE) == 0 && shard_cap_ > 0;
        std::string key;
        Shard* sh = nullptr;
        if (cacheable) {
            key = key_of(name, flags);
            sh = &shards_[std::hash<std::string>()(key) % kShards];
            int fd = -1;
            dev_t dev = 0;
            ino_t ino = 0;
            sh->lock();
            auto it = sh->by_path.find(key);
            if (it != sh->by_path.end() && !it->second.busy) {
                it->second.busy = true;
                fd = it->second.fd;
                dev = it->second.dev;
                ino = it->second.ino;
            }
            sh->unlock();
            if (fd >= 0) {
                if (::stat(name, st) == 0 && st->st_dev == dev && st->st_ino == ino) return fd;
                sh->lock();   // replaced or gone: a real open decides
                sh->by_path.erase(key);
                sh->unlock();
                fd_shard_[fd].store(0, std::memory_order_release);
                ::close(fd);
            }
        }
        const int fd = ::open(name, flags | O_CLOEXEC, 0644);
        if (fd < 0) return -(int64_t)errno;
        if (::fstat(fd, st) != 0) {
            const int64_t e = -(int64_t)errno;
            ::close(fd);
            return e;
        }
        if (cacheable && fd < kMaxFd) {
            std::vector<int> evicted;
            sh->lock();
            if (!sh->by_path.count(key)) {
                sh->by_path[key] = Ent{fd, true, st->st_dev, st->st_ino, 0};
                fd_shard_[fd].store((uint8_t)(sh - shards_ + 1), std::memory_order_release);
                if (sh->by_path.size() > shard_cap_) evict(*sh, evicted);
            }
            sh->unlock();
            for (int x : evicted) ::close(x);
        }
        return fd;
    }
    // Close of a descriptor the device got from open(): a kept one stays open.
    int close(int fd) {
        const uint32_t k = fd >= 0 && fd < kMaxFd ? fd_shard_[fd].load(std::memory_order_acquire) : 0;
        if (k) {
            Shard& sh = shards_[k - 1];
            sh.lock();
            for (auto& kv : sh.by_path)
                if (kv.second.fd == fd && kv.second.busy) {
                    kv.second.busy = false;
                    kv.second.lru = ++sh.tick;
                    sh.unlock();
                    return 0;
                }
            sh.unlock();
        }
        return ::close(fd) == 0 ? 0 : -errno;
    }
    void clear() {
        for (auto& sh : shards_) {
            sh.lock();
            for (auto& kv : sh.by_path) {
                fd_shard_[kv.second.fd].store(0, std::memory_order_relaxed);
                ::close(kv.second.fd);
            }
            sh.by_path.clear();
            sh.unlock();
        }
    }
    size_t size() {
        size_t n = 0;
        for (auto& sh : shards_) {
            sh.lock();
            n += sh.by_path.size();
            sh.unlock();
        }
        return n;
    }

private:
    static constexpr uint32_t kShards = 64;
    static constexpr int kMaxFd = 1 << 16;
    struct Ent {
        int fd;
        bool busy;
        dev_t dev;
        ino_t ino;
        uint64_t lru;
    };
    struct Shard {
        std::atomic<bool> held{false};
        std::unordered_map<std::string, Ent> by_path;
        uint64_t tick = 0;
        void lock() {
            while (held.exchange(true, std::memory_order_acquire))
                while (held.load(std::memory_order_relaxed)) {
#if defined(__x86_64__)
                    __builtin_ia32_pause();
#endif
                }
        }
        void unlock() { held.store(false, std::memory_order_release); }
    };
    static std::string key_of(const char* name, int flags) {
        std::string k(1, (char)('0' + (flags & O_ACCMODE)));
        k += name;
        return k;
    }
    // A full shard (lock held): the least recently closed eighth of its idle
    // entries go (their fds into `out`).
    void evict(Shard& sh, std::vector<int>& out) {
        std::vector<std::pair<uint64_t, std::string>> idle;
        for (auto& kv : sh.by_path)
            if (!kv.second.busy) idle.emplace_back(kv.second.lru, kv.first);
        size_t n = shard_cap_ / 8 > 0 ? shard_cap_ / 8 : 1;
        if (n > idle.size()) n = idle.size();
        std::partial_sort(idle.begin(), idle.begin() + n, idle.end());
        for (size_t i = 0; i < n; ++i) {
            auto it = sh.by_path.find(idle[i].second);
            out.push_back(it->second.fd);
            fd_shard_[it->second.fd].store(0, std::memory_order_relaxed);
            sh.by_path.erase(it);
        }
    }
    Shard shards_[kShards];
    std::atomic<uint8_t> fd_shard_[kMaxFd];   // shard + 1 of a kept descriptor, or 0
    size_t shard_cap_ = 0;
};

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
static inline int64_t serve_slot(const RingView& rv, uint32_t s, ServiceCounters& ctr, WorkerBufs& wb,
                                 DescCache* dc = nullptr) {
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
        struct stat st;
        int64_t hfd;
        if (dc) {
            hfd = dc->open(name, flags, &st);
        } else {
            hfd = ::open(name, flags | O_CLOEXEC, 0644);
            if (hfd < 0) {
                hfd = -(int64_t)errno;
            } else if (::fstat((int)hfd, &st) != 0) {
                const int64_t e = -(int64_t)errno;
                ::close((int)hfd);
                hfd = e;
            }
        }
        if (hfd < 0) {
            res = hfd;
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
        res = dc ? dc->close(fd) : (::close(fd) == 0 ? 0 : -(int64_t)errno);
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

// The host side of the submission queue: `head` is the first ticket not
// taken yet, and claim[t & (sq_depth - 1)] == t + 1 once ticket t is taken.
// Workers take tickets out of order: each scans the window of kTakeWindow
// tickets from the head, in ticket order, and claims the first published one
// nobody has claimed with a CAS on that ticket's own word.  With one shared
// head word taken in order, every published request set off a CAS storm of
// all spinning workers on that word, and a ticket whose request the device
// had not published yet held up every later one: 64 opens posted within
// 9 us took the host ~30 us to pick up.  Scanning in ticket order keeps the
// property that a published request is taken before any request published
// after it, so a slot-less close is taken before the reopen of its entry (the
// submission queue's depth counts one outstanding close per entry).
static constexpr uint32_t kTakeWindow = 64;
struct TakeQueue {
    std::atomic<uint64_t> head{0};
    std::unique_ptr<std::atomic<uint64_t>[]> claim;
    uint64_t mask = 0;
    void init(uint32_t sq_depth) {
        if (claim && mask + 1 == sq_depth) return;
        claim.reset(new std::atomic<uint64_t>[sq_depth]);
        for (uint32_t i = 0; i < sq_depth; ++i) claim[i].store(0, std::memory_order_relaxed);
        mask = sq_depth - 1;
    }
};

// Takes a published request that no worker has taken.  Returns true and sets
// *slot (and, for a close posted without a slot, *close_fd: read before the
// claim, while the ticket cannot be reused, since the device never has more
// than sq_depth tickets outstanding and this one is not taken).
static inline bool try_take(const RingView& rv, TakeQueue& q, uint32_t* slot, int64_t* close_fd) {
    const uint64_t h = q.head.load(std::memory_order_acquire);
    const uint32_t win = kTakeWindow < rv.sq_depth ? kTakeWindow : rv.sq_depth;
    for (uint32_t k = 0; k < win; ++k) {
        const uint64_t t = h + k;
        const uint64_t e = __atomic_load_n(&rv.sq[t & q.mask], __ATOMIC_ACQUIRE);
        if ((e >> 16) != (t & 0xFFFFFFFFFFFFull)) continue;   // not published (yet)
#ifdef GFS_HOST_TAKE_HOOK
        GFS_HOST_TAKE_HOOK();   // (tests: a worker delayed between the two loads)
#endif
        std::atomic<uint64_t>& c = q.claim[t & q.mask];
        uint64_t cv = c.load(std::memory_order_relaxed);
        // Taken: by this lap's claim (t + 1), or, when this worker was delayed
        // since it read the entry, by a later lap's (the ticket was served, its
        // slot reused and the queue wrapped): a claim word only grows.  Claiming
        // it again would serve whatever request its slot holds now a second time.
        if (cv >= t + 1) continue;
        const uint32_t s = (uint32_t)(e & 0xFFFF);
        const int64_t fd = s == GFS_SQ_NOSLOT ? __atomic_load_n(&rv.closeq[t & q.mask], __ATOMIC_RELAXED) : -1;
        if (!c.compare_exchange_strong(cv, t + 1, std::memory_order_acq_rel)) continue;
        // move the head past the tickets taken in a row
        uint64_t hh = q.head.load(std::memory_order_relaxed);
        while (q.claim[hh & q.mask].load(std::memory_order_acquire) == hh + 1) {
            if (q.head.compare_exchange_weak(hh, hh + 1, std::memory_order_acq_rel)) ++hh;
        }
        *slot = s;
        *close_fd = fd;
        return true;
    }
    return false;
}

// A close posted without a slot (the device does not wait for it).
static inline void serve_slotless_close(int64_t fd, ServiceCounters& ctr, DescCache* dc = nullptr) {
    if (fd < 0 || (dc ? dc->close((int)fd) : ::close((int)fd)) != 0) ctr.errors.fetch_add(1, std::memory_order_relaxed);
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

// True when one of the first tickets from the head holds a published request
// that no worker has taken.
static inline bool backlog(const RingView& rv, const TakeQueue& q) {
    const uint64_t h = q.head.load(std::memory_order_relaxed);
    for (uint64_t t = h; t < h + 4; ++t) {
        const uint64_t e = __atomic_load_n(&rv.sq[t & q.mask], __ATOMIC_ACQUIRE);
        if ((e >> 16) == (t & 0xFFFFFFFFFFFFull) && q.claim[t & q.mask].load(std::memory_order_relaxed) != t + 1)
            return true;
    }
    return false;
}

// Worker `index` of the service.  Worker 0 polls the queue without sleeping
// while the service runs, so a lone request (an open between two launches)
// is picked up within a poll; the others spin for a while after their last
// request or a launch's kick, then sleep on the wake futex, woken by a worker
// that finds requests queued behind the one it took, or by their timeout
// (50 us for the first millisecond of sleep, then 1 ms).  A sleep is a few
// microseconds of CPU: idle workers sleeping in 2 us steps each kept a CPU
// busy, and a process under a CPU quota (a container's cgroup) that spends
// it all is stopped until the next period, which showed as multi-millisecond
// stalls of the serving loop.
static inline void worker_loop(RingView rv, TakeQueue* head, std::atomic<bool>* stop,
                               ServiceCounters* ctr, WakeState* wk, DescCache* dc, unsigned index) {
    prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
    WorkerBufs wb;
    uint64_t idle = 0;
    uint32_t naps = 0;   // futex sleeps since the last request or kick
    uint32_t kicks = wk->kicks.load(std::memory_order_relaxed);
    for (;;) {
        uint32_t s;
        int64_t close_fd;
        if (try_take(rv, *head, &s, &close_fd)) {
            if (wk->sleepers.load(std::memory_order_relaxed) && backlog(rv, *head)) wake_all(wk);
            const uint64_t t0 = mono_ns();
            if (s == GFS_SQ_NOSLOT) serve_slotless_close(close_fd, *ctr, dc);
            else serve_slot(rv, s, *ctr, wb, dc);
            ctr->busy_ns.fetch_add(mono_ns() - t0, std::memory_order_relaxed);
            idle = 0;
            naps = 0;
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
            naps = 0;
        }
        if (index == 0 || idle < 20000) {
#if defined(__x86_64__)
            __builtin_ia32_pause();
#endif
        } else {
            const uint32_t seen = wk->word.load(std::memory_order_acquire);
            wk->sleepers.fetch_add(1, std::memory_order_acq_rel);
            if (!backlog(rv, *head)) {
                struct timespec ts = {0, naps < 20 ? 50000 : 1000000};
                syscall(SYS_futex, (uint32_t*)&wk->word, FUTEX_WAIT_PRIVATE, seen, &ts, nullptr, 0);
                ++naps;
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
        head_.init(rv.sq_depth);
        for (unsigned i = 0; i < nthreads; ++i)
            threads_.emplace_back(worker_loop, rv, &head_, &stop_, &ctr_, &wake_, &dcache_, i);
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
    uint64_t head() const { return head_.head.load(); }
    ~Service() { stop(); }

private:
    std::vector<std::thread> threads_;
    TakeQueue head_;   // persists across stop/start: the device's tickets continue
    DescCache dcache_;   // persists across stop/start: reopens come after a restart
    std::atomic<bool> stop_{false};
    ServiceCounters ctr_;
    WakeState wake_;
    bool running_ = false;
};

} // namespace gfs_host
