// CPU model of the GPU->host RPC ring: client threads play the device side of
// the protocol (claim a slot, write the request, publish (ticket << 16 | slot)
// in the submission queue, wait for the request's bit in a completion mask,
// read the result) against the real host service (host_service.h).  Batched
// clients post up to 64 reads sharing one completion word, take each page as
// its bit appears and return its slot at once, keeping the word's slot to the
// end, as a workgroup's page fill does.  Also models a wait that timed out:
// the client keeps (quarantines) the slots of its unanswered requests, the
// host answers them late, and fresh requests through the other slots stay
// exact.  Checks every result, the data in staging, slot conservation, and
// runs under ThreadSanitizer.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>
#include <fcntl.h>
#include <unistd.h>
// Now and then a worker stalls between reading a submission entry and
// claiming its ticket, long enough for the queue to wrap: it must not take
// the ticket again once a later lap has claimed the entry.
static std::atomic<uint32_t> take_hook_calls{0};
#define GFS_HOST_TAKE_HOOK()                                                                  \
    do {                                                                                      \
        if (take_hook_calls.fetch_add(1, std::memory_order_relaxed) % 61 == 0)                \
            std::this_thread::sleep_for(std::chrono::microseconds(400));                      \
    } while (0)
#include "../../gpufs_amd/csrc/host_service.h"

static int failures = 0;
#define CHECK(c, ...) do { if (!(c)) { fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); fprintf(stderr, __VA_ARGS__); fprintf(stderr, "\n"); __atomic_fetch_add(&failures, 1, __ATOMIC_RELAXED); } } while (0)

struct Ring {
    gfs_host::RingView rv;
    std::vector<gfs_rpc_slot> slots;
    std::vector<uint64_t> sq;
    std::vector<int64_t> closeq;
    std::vector<uint64_t> cq;
    std::vector<uint8_t> staging;
    std::mutex mu;
    std::vector<uint32_t> free_slots;
    std::atomic<uint64_t> tail{0};
    // the submission queue holds a request per slot plus up to 64 slot-less closes
    Ring(uint32_t depth, uint32_t log_page) : slots(depth), sq(depth + 64, ~0ull), cq(depth, 0),
        staging(((size_t)depth << log_page) + 64) {
        uint32_t q = 1; while (q < sq.size()) q <<= 1; sq.assign(q, ~0ull); closeq.assign(q, -1);
        rv.slots = slots.data(); rv.sq = sq.data(); rv.cq = cq.data(); rv.staging = staging.data();
        rv.closeq = closeq.data();
        rv.sq_depth = q; rv.ring_depth = depth; rv.log_page = log_page;
        for (uint32_t s = 0; s < depth; ++s) free_slots.push_back(s);
    }
    int claim() { std::lock_guard<std::mutex> g(mu); if (free_slots.empty()) return -1; int s = free_slots.back(); free_slots.pop_back(); return s; }
    void release(int s) { std::lock_guard<std::mutex> g(mu); free_slots.push_back(s); }
    // device side: publish one request and wait for it (cq index = slot)
    int64_t call(int s, uint32_t op, int fd, int flags, uint64_t off, uint64_t len, const char* name = nullptr) {
        gfs_rpc_slot* r = &slots[s];
        r->op = op; r->host_fd = fd; r->flags = flags; r->cq = (uint32_t)s; r->offset = off; r->length = len;
        r->result = INT64_MIN;
        if (name) { memset(r->name, 0, GFS_NAME_MAX); strncpy(r->name, name, GFS_NAME_MAX - 1); }
        __atomic_store_n(&cq[s], 0ull, __ATOMIC_RELAXED);
        publish(s);
        while ((__atomic_load_n(&cq[s], __ATOMIC_ACQUIRE) & 1ull) == 0) std::this_thread::yield();
        return __atomic_load_n(&r->result, __ATOMIC_RELAXED);
    }
    void publish(int s) {
        uint64_t t = tail.fetch_add(1);
        __atomic_store_n(&sq[t & (rv.sq_depth - 1)], (t << 16) | (uint64_t)s, __ATOMIC_RELEASE);
    }
    // device side of gclose: the fd under the ticket, no slot, no wait
    void post_close(int64_t fd) {
        uint64_t t = tail.fetch_add(1);
        __atomic_store_n(&closeq[t & (rv.sq_depth - 1)], fd, __ATOMIC_RELAXED);
        __atomic_store_n(&sq[t & (rv.sq_depth - 1)], (t << 16) | (uint64_t)GFS_SQ_NOSLOT, __ATOMIC_RELEASE);
    }
    // device side of a page fill: n <= 64 reads share slot ss[0]'s completion
    // word; each page is checked as soon as its bit is set and its slot (other
    // than ss[0]) goes back at once.  Returns the number of pages checked.
    int batch(const std::vector<int>& ss, int fd, const std::vector<uint64_t>& offs, uint64_t len,
              const std::vector<uint8_t>& data) {
        const int n = (int)ss.size();
        const uint32_t cqi = (uint32_t)ss[0];
        __atomic_store_n(&cq[cqi], 0ull, __ATOMIC_RELAXED);
        for (int i = 0; i < n; ++i) {
            gfs_rpc_slot* r = &slots[ss[i]];
            r->op = GFS_OP_READ; r->host_fd = fd; r->flags = 0; r->cq = cqi | ((uint32_t)i << 16);
            r->offset = offs[i]; r->length = len; r->result = INT64_MIN;
            publish(ss[i]);
        }
        const uint64_t all = n == 64 ? ~0ull : ((1ull << n) - 1);
        uint64_t done = 0;
        int checked = 0;
        while (done != all) {
            uint64_t m = __atomic_load_n(&cq[cqi], __ATOMIC_ACQUIRE) & all & ~done;
            if (!m) { std::this_thread::yield(); continue; }
            for (uint64_t b = m; b; b &= b - 1) {
                const int i = __builtin_ctzll(b);
                const int64_t got = __atomic_load_n(&slots[ss[i]].result, __ATOMIC_ACQUIRE);
                const uint64_t off = offs[i];
                const uint64_t exp = off >= data.size() ? 0 : std::min<uint64_t>(len, data.size() - off);
                CHECK(got == (int64_t)exp, "batch read %d off=%llu got=%lld", i, (unsigned long long)off, (long long)got);
                if (got > 0) CHECK(memcmp(&staging[(size_t)ss[i] << rv.log_page], &data[off], (size_t)got) == 0, "batch bytes");
                if (i != 0) release(ss[i]);
                ++checked;
            }
            done |= m;
        }
        release(ss[0]);
        return checked;
    }
};

int main(int argc, char** argv) {
    const char* dir = argc > 1 ? argv[1] : "/tmp";
    const uint32_t log_page = 12, depth = 8;
    std::string path = std::string(dir) + "/ring_model.bin";
    std::vector<uint8_t> data(1 << 20);
    std::mt19937_64 rng(7);
    for (auto& b : data) b = (uint8_t)rng();
    { int fd = open(path.c_str(), O_CREAT | O_TRUNC | O_WRONLY, 0644); CHECK(write(fd, data.data(), data.size()) == (ssize_t)data.size(), "write"); close(fd); }
    Ring ring(depth, log_page);
    gfs_host::Service svc;
    svc.start(ring.rv, 3);
    // OPEN through the ring
    int s0 = ring.claim();
    int64_t hfd = ring.call(s0, GFS_OP_OPEN, -1, O_RDONLY, 0, 0, path.c_str());
    CHECK(hfd >= 0, "open %lld", (long long)hfd);
    CHECK(ring.slots[s0].size == data.size(), "size");
    ring.release(s0);
    int s1 = ring.claim();
    int64_t bad = ring.call(s1, GFS_OP_OPEN, -1, O_RDONLY, 0, 0, (std::string(dir) + "/does_not_exist").c_str());
    CHECK(bad == -ENOENT, "missing open %lld", (long long)bad);
    ring.release(s1);
    // many concurrent reads from 12 client threads through 8 slots
    std::vector<std::thread> cl;
    std::atomic<int> done{0};
    for (int c = 0; c < 12; ++c) {
        cl.emplace_back([&, c]() {
            std::mt19937_64 r(100 + c);
            for (int k = 0; k < 300; ++k) {
                int s;
                while ((s = ring.claim()) < 0) std::this_thread::yield();
                uint64_t off = r() % (data.size() + 100);
                uint64_t len = 1 + r() % 4096;
                int64_t got = ring.call(s, GFS_OP_READ, (int)hfd, 0, off, len);
                uint64_t exp = off >= data.size() ? 0 : std::min<uint64_t>(len, data.size() - off);
                CHECK(got == (int64_t)exp, "read off=%llu len=%llu got=%lld", (unsigned long long)off, (unsigned long long)len, (long long)got);
                if (got > 0) CHECK(memcmp(&ring.staging[(size_t)s << log_page], &data[off], (size_t)got) == 0, "bytes");
                ring.release(s);
                done++;
            }
        });
    }
    for (auto& t : cl) t.join();
    CHECK(done == 12 * 300, "done");
    CHECK(ring.free_slots.size() == depth, "slots conserved");
    // batched fills: 4 clients, batches of 1..6 pages through the 8 slots
    std::atomic<int> pages{0};
    std::vector<std::thread> bl;
    for (int c = 0; c < 4; ++c) {
        bl.emplace_back([&, c]() {
            std::mt19937_64 r(900 + c);
            for (int k = 0; k < 100; ++k) {
                const int n = 1 + (int)(r() % 6);
                std::vector<int> ss;
                while ((int)ss.size() < n) {   // all-or-nothing claim, as a slot-limited batch
                    int s = ring.claim();
                    if (s >= 0) { ss.push_back(s); continue; }
                    for (int x : ss) ring.release(x);
                    ss.clear();
                    std::this_thread::yield();
                }
                std::vector<uint64_t> offs(n);
                for (auto& o : offs) o = (r() % (data.size() / 4096 + 1)) * 4096;
                pages += ring.batch(ss, (int)hfd, offs, 4096, data);
            }
        });
    }
    for (auto& t : bl) t.join();
    CHECK(ring.free_slots.size() == depth, "slots conserved after batches");
    const int batch_pages = pages.load();
    // write + OR-merge + truncate on a second file
    std::string wpath = std::string(dir) + "/ring_model_w.bin";
    { int fd = open(wpath.c_str(), O_CREAT | O_TRUNC | O_WRONLY, 0644); uint8_t b[8] = {1, 2, 4, 8, 0, 0, 0, 0}; CHECK(write(fd, b, 8) == 8, "w"); close(fd); }
    int s = ring.claim();
    int64_t wfd = ring.call(s, GFS_OP_OPEN, -1, O_RDWR, 0, 0, wpath.c_str());
    CHECK(wfd >= 0, "open w");
    uint8_t* st = &ring.staging[(size_t)s << log_page];
    uint8_t nb[12] = {16, 16, 16, 16, 16, 16, 16, 16, 7, 7, 7, 7};
    memcpy(st, nb, 12);
    CHECK(ring.call(s, GFS_OP_DIFF, (int)wfd, 0, 0, 12) == 12, "diff");
    uint8_t back[12];
    { int fd = open(wpath.c_str(), O_RDONLY); CHECK(read(fd, back, 12) == 12, "rb"); close(fd); }
    uint8_t exp[12] = {17, 18, 20, 24, 16, 16, 16, 16, 7, 7, 7, 7};
    CHECK(memcmp(back, exp, 12) == 0, "or-merge");
    memcpy(st, "abc", 3);
    CHECK(ring.call(s, GFS_OP_WRITE, (int)wfd, 0, 20, 3) == 3, "write");
    CHECK(ring.call(s, GFS_OP_TRUNC, (int)wfd, 0, 5, 0) == 0, "trunc");
    CHECK(ring.call(s, GFS_OP_CLOSE, (int)wfd, 0, 0, 0) == 0, "close");
    CHECK(ring.call(s, GFS_OP_READ, 12345, 0, 0, 10) == -EBADF, "bad fd");
    CHECK(ring.call(s, 99, -1, 0, 0, 0) == -EINVAL, "bad op");
    ring.release(s);
    { struct stat sb; stat(wpath.c_str(), &sb); CHECK(sb.st_size == 5, "size after trunc"); }
    // slot-less closes: 40 opens through the ring, each closed by a request
    // that carries no slot and is not waited for, from 4 client threads at
    // once; stopping the service drains them, so every fd is closed after it
    std::vector<int64_t> lfds;
    for (int k = 0; k < 40; ++k) {
        int so = ring.claim();
        lfds.push_back(ring.call(so, GFS_OP_OPEN, -1, O_RDONLY, 0, 0, path.c_str()));
        CHECK(lfds.back() >= 0, "open for a slot-less close");
        ring.release(so);
    }
    {
        std::vector<std::thread> pc;
        for (int c = 0; c < 4; ++c)
            pc.emplace_back([&, c]() { for (int k = c; k < 40; k += 4) ring.post_close(lfds[k]); });
        for (auto& t : pc) t.join();
    }
    svc.stop();
    for (int64_t f : lfds) CHECK(fcntl((int)f, F_GETFD) == -1 && errno == EBADF, "fd %lld still open", (long long)f);
    CHECK(svc.counters().requests.load() == (uint64_t)(2 + 3600 + batch_pages + 7 + 80), "request count %llu",
          (unsigned long long)svc.counters().requests.load());
    // 256 KiB pages: reads longer than the service's 64 KiB bounce buffer,
    // at any offset, ending at or past end of file
    {
        const uint32_t big_log = 18, big_depth = 4;
        Ring big(big_depth, big_log);
        gfs_host::Service bsvc;
        bsvc.start(big.rv, 2);
        int so = big.claim();
        const int64_t bfd = big.call(so, GFS_OP_OPEN, -1, O_RDONLY, 0, 0, path.c_str());
        CHECK(bfd >= 0, "open big");
        big.release(so);
        std::vector<std::thread> bc;
        for (int c = 0; c < 3; ++c) {
            bc.emplace_back([&, c]() {
                std::mt19937_64 r(500 + c);
                for (int k = 0; k < 60; ++k) {
                    int s;
                    while ((s = big.claim()) < 0) std::this_thread::yield();
                    const uint64_t off = k % 5 == 0 ? data.size() - (r() % 70000) : r() % (data.size() + 100);
                    const uint64_t len = 1 + r() % (1u << big_log);
                    const int64_t got = big.call(s, GFS_OP_READ, (int)bfd, 0, off, len);
                    const uint64_t exp = off >= data.size() ? 0 : std::min<uint64_t>(len, data.size() - off);
                    CHECK(got == (int64_t)exp, "big read off=%llu len=%llu got=%lld", (unsigned long long)off,
                          (unsigned long long)len, (long long)got);
                    if (got > 0) CHECK(memcmp(&big.staging[(size_t)s << big_log], &data[off], (size_t)got) == 0, "big bytes");
                    big.release(s);
                }
            });
        }
        for (auto& t : bc) t.join();
        bsvc.stop();
        CHECK(big.free_slots.size() == big_depth, "big slots conserved");
    }
    // Late completions: a client publishes a batch of three reads while no
    // host service runs, its bounded wait ends, and it keeps the three slots
    // (the device quarantines a slot whose request timed out: the host still
    // holds its ticket).  When the service starts it answers those requests
    // late, while fresh clients read through the other slots: every fresh read
    // is exact, and the late answers land only in the quarantined slots'
    // staging and completion word.
    {
        Ring lr(8, log_page);
        const int dfd = open(path.c_str(), O_RDONLY);
        CHECK(dfd >= 0, "open for late reads");
        std::vector<int> q = {lr.claim(), lr.claim(), lr.claim()};
        const uint32_t cqi = (uint32_t)q[0];
        __atomic_store_n(&lr.cq[cqi], 0ull, __ATOMIC_RELAXED);
        const uint64_t loff[3] = {4096, 40960, 8192 + 7};
        for (int i = 0; i < 3; ++i) {
            gfs_rpc_slot* r = &lr.slots[q[i]];
            r->op = GFS_OP_READ; r->host_fd = dfd; r->flags = 0; r->cq = cqi | ((uint32_t)i << 16);
            r->offset = loff[i]; r->length = 4096; r->result = INT64_MIN;
            lr.publish(q[i]);
        }
        const auto t0 = std::chrono::steady_clock::now();
        while ((__atomic_load_n(&lr.cq[cqi], __ATOMIC_ACQUIRE) & 7ull) != 7ull &&
               std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(20))
            std::this_thread::yield();
        CHECK(__atomic_load_n(&lr.cq[cqi], __ATOMIC_ACQUIRE) == 0, "no service yet: nothing answered");
        // timed out: q's slots are not released
        gfs_host::Service lsvc;
        lsvc.start(lr.rv, 2);
        std::vector<std::thread> fc;
        std::atomic<int> fresh{0};
        for (int c = 0; c < 3; ++c) {
            fc.emplace_back([&, c]() {
                std::mt19937_64 r(700 + c);
                for (int k = 0; k < 200; ++k) {
                    int s2;
                    while ((s2 = lr.claim()) < 0) std::this_thread::yield();
                    CHECK(s2 != q[0] && s2 != q[1] && s2 != q[2], "a quarantined slot was handed out");
                    const uint64_t off = r() % data.size();
                    const uint64_t len = 1 + r() % 4096;
                    const int64_t got = lr.call(s2, GFS_OP_READ, dfd, 0, off, len);
                    const uint64_t exp = std::min<uint64_t>(len, data.size() - off);
                    CHECK(got == (int64_t)exp, "fresh read beside late answers");
                    if (got > 0) CHECK(memcmp(&lr.staging[(size_t)s2 << log_page], &data[off], (size_t)got) == 0,
                                       "fresh bytes beside late answers");
                    lr.release(s2);
                    fresh++;
                }
            });
        }
        for (auto& t : fc) t.join();
        while ((__atomic_load_n(&lr.cq[cqi], __ATOMIC_ACQUIRE) & 7ull) != 7ull) std::this_thread::yield();
        lsvc.stop();
        CHECK(fresh == 600, "fresh reads");
        CHECK(lr.free_slots.size() == 5, "5 slots in the ring, 3 quarantined");
        for (int i = 0; i < 3; ++i) {
            CHECK(lr.slots[q[i]].result == 4096, "late answer %d", i);
            CHECK(memcmp(&lr.staging[(size_t)q[i] << log_page], &data[loff[i]], 4096) == 0, "late bytes %d", i);
        }
        close(dfd);
    }
    unlink(path.c_str()); unlink(wpath.c_str());
    if (failures) { fprintf(stderr, "%d failures\n", failures); return 1; }
    printf("ring model OK\n");
    return 0;
}
