// CPU model of the file-state protocol (gpufs_amd/csrc/fstate.h, the same
// code the device API runs): threads play workgroups that open, use and close
// one file-table entry concurrently, with host opens that sometimes fail and
// version bumps (truncate) while the file is held.  Checks:
//   * at most one thread is at the host (opening or closing) at a time, and a
//     host open only starts from a closed host file;
//   * every thread that holds an open sees the host file open;
//   * the thread that closes sees no holder left;
//   * at the end the entry is CLOSED with zero opens, and host opens that
//     succeeded equal host closes.
// Built with and without ThreadSanitizer.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>
#include <sched.h>
#include "../../gpufs_amd/csrc/fstate.h"

static std::atomic<int> failures{0};
#define CHECK(c, ...)                                                              \
    do {                                                                           \
        if (!(c)) {                                                                \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                   \
            fprintf(stderr, __VA_ARGS__);                                          \
            fprintf(stderr, "\n");                                                 \
            failures.fetch_add(1);                                                 \
        }                                                                          \
    } while (0)

struct HostPolicy {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    uint64_t fadd(uint64_t* w, uint64_t v) { return __atomic_fetch_add(w, v, __ATOMIC_ACQ_REL); }
    uint64_t load(uint64_t* w) { return __atomic_load_n(w, __ATOMIC_ACQUIRE); }
    uint64_t load_rlx(uint64_t* w) { return __atomic_load_n(w, __ATOMIC_RELAXED); }
    bool cas(uint64_t* w, uint64_t& exp, uint64_t v) {
        return __atomic_compare_exchange_n(w, &exp, v, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
    }
    bool expired() { return std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60); }
    void pause() { sched_yield(); }
};

enum { H_CLOSED = 0, H_OPENING = 1, H_OPEN = 2, H_CLOSING = 3 };

static uint64_t word;
static std::atomic<int> host{H_CLOSED};
static std::atomic<int> holders{0};
static std::atomic<long> host_opens{0}, host_open_fail{0}, host_closes{0}, joins{0}, bumps{0}, peeks{0};

static int n_threads = 16;

// Every count in the word belongs to a thread, so opens never exceed the thread count.
static void check_bound(const char* where) {
    const uint64_t w = __atomic_load_n(&word, __ATOMIC_ACQUIRE);
    CHECK((int)fst_opens(w) <= n_threads, "%s: opens %u > %d threads", where, fst_opens(w), n_threads);
    CHECK(fst_state(w) >= GFS_FS_OPENING && fst_state(w) <= GFS_FS_CLOSED, "%s: state %u", where, fst_state(w));
}

static void worker(int id, int iters) {
    std::mt19937 rng(1234 + id);
    for (int it = 0; it < iters; ++it) {
        HostPolicy P;
        check_bound("before open");
        // odd workers sometimes peek: a host open in progress returns at once,
        // counting nothing (the caller reads cached data and confirms later)
        const bool peek = (id & 1) && rng() % 2 == 0;
        int r = fst_open(&word, P, peek);
        check_bound("after open");
        if (r == FST_PEEK) {
            CHECK(peek, "FST_PEEK without peek");
            peeks.fetch_add(1);
            continue;
        }
        CHECK(r == FST_JOINED || r == FST_MUST_OPEN, "fst_open returned %d", r);
        if (r == FST_MUST_OPEN) {
            int exp = H_CLOSED;
            CHECK(host.compare_exchange_strong(exp, H_OPENING), "host open while host state %d", exp);
            for (int k = rng() % 8; k > 0; --k) std::this_thread::yield();   // a slow host open
            if (rng() % 4 == 0) {   // the host open fails
                host.store(H_CLOSED);
                host_open_fail.fetch_add(1);
                fst_open_done(&word, P, false, false);
                continue;
            }
            host.store(H_OPEN);
            host_opens.fetch_add(1);
            fst_open_done(&word, P, true, rng() % 8 == 0);
        } else {
            joins.fetch_add(1);
        }
        holders.fetch_add(1);
        CHECK(host.load() == H_OPEN, "holder sees host state %d", host.load());
        const uint64_t st = P.load(&word);
        CHECK(fst_state(st) == GFS_FS_OPEN && fst_opens(st) >= 1, "holder sees word state %u opens %u",
              fst_state(st), fst_opens(st));
        if (rng() % 32 == 0) {   // truncate: bump the version while holding the file
            uint64_t exp = P.load(&word);
            while (!P.cas(&word, exp, exp + (1ull << 32))) {
            }
            bumps.fetch_add(1);
        }
        if (rng() % 2 == 0) std::this_thread::yield();
        CHECK(host.load() == H_OPEN, "holder sees host state %d before close", host.load());
        holders.fetch_sub(1);
        int c = fst_close(&word, P);
        check_bound("after close");
        CHECK(c == 0 || c == 1, "fst_close returned %d", c);
        if (c == 1) {
            CHECK(holders.load() == 0, "last closer sees %d holders", holders.load());
            int exp = H_OPEN;
            CHECK(host.compare_exchange_strong(exp, H_CLOSING), "host close while host state %d", exp);
            if (rng() % 4 == 0) std::this_thread::yield();
            host.store(H_CLOSED);
            host_closes.fetch_add(1);
            fst_close_done(&word, P);
        }
    }
}

int main(int argc, char** argv) {
    const int threads = argc > 1 ? atoi(argv[1]) : 16;
    n_threads = threads;
    const int iters = argc > 2 ? atoi(argv[2]) : 20000;
    word = fst_make(GFS_FS_CLOSED, 0, 0);
    {
        HostPolicy P;
        CHECK(fst_close(&word, P) == -EBADF, "close of a closed entry");
    }
    std::vector<std::thread> ts;
    for (int i = 0; i < threads; ++i) ts.emplace_back(worker, i, iters);
    for (auto& t : ts) t.join();
    const uint64_t w = word;
    CHECK(fst_state(w) == GFS_FS_CLOSED && fst_opens(w) == 0, "final state %u opens %u", fst_state(w), fst_opens(w));
    CHECK(host.load() == H_CLOSED, "final host state %d", host.load());
    CHECK(host_opens.load() == host_closes.load(), "host opens %ld closes %ld", host_opens.load(), host_closes.load());
    CHECK(host_opens.load() + joins.load() + host_open_fail.load() + peeks.load() == (long)threads * iters,
          "open outcomes do not add up");
    CHECK(threads < 2 || peeks.load() > 0, "no peek met a host open in progress");
    printf("opens %ld (failed %ld) joins %ld peeks %ld closes %ld version bumps %ld final version %u\n",
           host_opens.load(), host_open_fail.load(), joins.load(), peeks.load(), host_closes.load(), bumps.load(),
           fst_version(w));
    if (failures.load()) {
        printf("fstate model FAILED (%d)\n", failures.load());
        return 1;
    }
    printf("fstate model OK\n");
    return 0;
}
