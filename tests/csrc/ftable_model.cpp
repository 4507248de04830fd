// CPU model of the file table (gpufs_amd/csrc/ftable.h, the code the device
// API runs, with the open/close protocol of fstate.h): threads play
// workgroups that open names drawn from a pool three times larger than the
// table, hold one or two of them, and close them, so the table fills and
// reclaim passes take closed entries back while other threads look names up,
// insert, join and close.  The name hash is deliberately weak (few distinct
// home slots), so the index holds long probe runs and the backward-shift
// deletion moves entries.  Checks:
//   * an opener holds the entry of the name it asked for (a reclaim never
//     takes back an entry that an opener is about to join);
//   * the first opener after a close sees the host file closed, and the
//     entry's radix tree (built by the host opener, tagged with the entry)
//     stays intact while the file is held;
//   * a node popped from the free list is zeroed (it is in no live tree);
//   * at the end: gate clear, every entry in use is in the index once and
//     found from its home slot, free entries are in none, the free lists hold
//     no duplicates and no node of a live tree.
// Built with and without ThreadSanitizer.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>
#include <sched.h>
#include "../../gpufs_amd/csrc/ftable.h"

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

template <class T> struct ident { typedef T type; };
struct HostFt {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    template <class T> T load(T* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
    template <class T> T load_rlx(T* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
    template <class T> void store(T* p, typename ident<T>::type v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
    template <class T> void store_rlx(T* p, typename ident<T>::type v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
    template <class T> T fadd(T* p, typename ident<T>::type v) { return __atomic_fetch_add(p, v, __ATOMIC_ACQ_REL); }
    template <class T> bool cas(T* p, T& e, typename ident<T>::type v) {
        return __atomic_compare_exchange_n(p, &e, v, false, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE);
    }
    uint32_t fand(uint32_t* p, uint32_t v) { return __atomic_fetch_and(p, v, __ATOMIC_RELEASE); }
    bool expired() { return std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60); }
    void pause() { sched_yield(); }
};
// fstate.h policy
struct HostFst {
    HostFt f;
    uint64_t fadd(uint64_t* w, uint64_t v) { return f.fadd(w, v); }
    uint64_t load(uint64_t* w) { return f.load(w); }
    uint64_t load_rlx(uint64_t* w) { return f.load_rlx(w); }
    bool cas(uint64_t* w, uint64_t& exp, uint64_t v) { return f.cas(w, exp, v); }
    bool expired() { return f.expired(); }
    void pause() { sched_yield(); }
};

static gfs_dev G;
static int n_threads = 8, pool_names = 72;
static std::vector<std::atomic<int>> host_open;   // per entry: the simulated host file is open
static std::atomic<long> opens{0}, joins{0}, inserts{0}, host_fail{0}, full{0}, node_allocs{0};

static void make_name(int id, ft_name& n) {
    memset(&n, 0, sizeof(n));
    snprintf((char*)n.w, sizeof(n.w), "/data/file_%05d.txt", id);
    n.len = (int)strlen((const char*)n.w);
    n.hash = (uint32_t)(id % 11) * 5u;   // 11 home slots: long probe runs
}

static uint64_t node_alloc(HostFt& p) {
    uint64_t n = 0;
    if (p.load_rlx(&G.node_next) < G.node_cap) {
        const unsigned long long b = p.fadd(&G.node_next, 1ull);
        if (b < G.node_cap) n = b;
    }
    if (!n) n = ft_node_pop(&G, p);
    if (n) {
        for (uint32_t k = 0; k < GFS_FANOUT; ++k)
            CHECK(p.load_rlx(&G.nodes[n * GFS_FANOUT + k]) == 0, "allocated node %lu word %u not zero",
                  (unsigned long)n, k);
        node_allocs.fetch_add(1);
    }
    return n;
}

// A two-level tree for entry m: root r with children c1 (slot 0) and c2
// (slot 5); the leaves carry the tag m + 1 in their words 1 and 63.
static void build_tree(uint32_t m, HostFt& p) {
    const uint64_t r = node_alloc(p), c1 = node_alloc(p), c2 = node_alloc(p);
    CHECK(r && c1 && c2, "node pool exhausted");
    if (!(r && c1 && c2)) return;
    for (uint64_t c : {c1, c2}) {
        p.store_rlx(&G.nodes[c * GFS_FANOUT + 1], (uint64_t)m + 1);
        p.store_rlx(&G.nodes[c * GFS_FANOUT + 63], (uint64_t)m + 1);
    }
    p.store_rlx(&G.nodes[r * GFS_FANOUT + 0], c1);
    p.store_rlx(&G.nodes[r * GFS_FANOUT + 5], c2);
    p.store(&G.files[m].root, r | (2ull << 40));
}
static void check_tree(uint32_t m, HostFt& p, const char* where) {
    const uint64_t rw = p.load(&G.files[m].root);
    if (!rw) return;   // the host open failed before a tree was built
    const uint64_t r = rw & FT_NODE_MASK;
    CHECK((rw >> 40) == 2, "%s: entry %u root height %lu", where, m, (unsigned long)(rw >> 40));
    const uint64_t c1 = p.load_rlx(&G.nodes[r * GFS_FANOUT + 0]), c2 = p.load_rlx(&G.nodes[r * GFS_FANOUT + 5]);
    CHECK(c1 && c2 && c1 < G.node_cap && c2 < G.node_cap, "%s: entry %u children %lu %lu", where, m,
          (unsigned long)c1, (unsigned long)c2);
    if (!(c1 && c2 && c1 < G.node_cap && c2 < G.node_cap)) return;
    for (uint64_t c : {c1, c2})
        CHECK(p.load_rlx(&G.nodes[c * GFS_FANOUT + 1]) == (uint64_t)m + 1 &&
                  p.load_rlx(&G.nodes[c * GFS_FANOUT + 63]) == (uint64_t)m + 1,
              "%s: entry %u leaf %lu lost its tag", where, m, (unsigned long)c);
}

// gopen: returns the entry, or a negative code.
static int model_open(const ft_name& nm, std::mt19937& rng) {
    HostFt F;
    for (;;) {
        int r = ft_gate_enter(&G, F);
        if (r) return r;
        uint32_t idx = 0;
        r = ft_find_or_insert(&G, nm, F, &idx);
        int must = 0;
        if (r == FT_NEW) {
            ft_gate_leave(&G, F);
            inserts.fetch_add(1);
            must = 1;
        } else if (r == FT_FOUND) {
            HostFst P;
            const int o = fst_open(&G.files[idx].st, P);
            ft_gate_leave(&G, F);
            if (o < 0) return o;
            must = o == FST_MUST_OPEN;
            if (!must) joins.fetch_add(1);
        } else {
            ft_gate_leave(&G, F);
            if (r != E_FSTABLE_FULL) return r;
            const int c = ft_reclaim(&G, F);
            if (c < 0) return c;
            if (c == 0) {
                full.fetch_add(1);
                return E_FSTABLE_FULL;
            }
            continue;
        }
        gfs_file* e = &G.files[idx];
        CHECK(ft_name_eq(e, nm), "opener of %s holds entry %u named %s", (const char*)nm.w, idx, e->name);
        if (must) {
            // the host open: only this thread is at entry idx
            int exp = 0;
            CHECK(host_open[idx].compare_exchange_strong(exp, 1), "host open of entry %u while open", idx);
            HostFst P;
            if (rng() % 16 == 0) {
                host_open[idx].store(0);
                host_fail.fetch_add(1);
                fst_open_done(&e->st, P, false, false);
                return -EIO;
            }
            bool changed = false;
            const uint64_t rw = F.load_rlx(&e->root);
            if (rw && rng() % 8 == 0) {   // the file changed on the host: drop the old tree
                F.store_rlx(&e->root, 0ull);
                ft_tree_free(&G, rw, F);
                changed = true;
            }
            if (!F.load_rlx(&e->root)) build_tree(idx, F);
            opens.fetch_add(1);
            fst_open_done(&e->st, P, true, changed);
        }
        return (int)idx;
    }
}

static void model_close(uint32_t idx, std::mt19937& rng) {
    HostFst P;
    const int c = fst_close(&G.files[idx].st, P);
    CHECK(c == 0 || c == 1, "fst_close of entry %u returned %d", idx, c);
    if (c == 1) {
        if (rng() % 4 == 0) std::this_thread::yield();
        int exp = 1;
        CHECK(host_open[idx].compare_exchange_strong(exp, 0), "host close of entry %u while closed", idx);
        fst_close_done(&G.files[idx].st, P);
    }
}

static void worker(int id, int iters) {
    std::mt19937 rng(777 + id);
    for (int it = 0; it < iters; ++it) {
        const int k = rng() % 3 == 0 ? 2 : 1;   // sometimes hold two files at once
        int held[2] = {-1, -1}, ids[2] = {-1, -1};
        for (int j = 0; j < k; ++j) {
            // a hot set shared by every thread, and a cold tail
            ids[j] = rng() % 2 ? (int)(rng() % 6) : (int)(rng() % pool_names);
            if (j == 1 && ids[1] == ids[0]) ids[1] = (ids[0] + 1) % pool_names;
            ft_name nm;
            make_name(ids[j], nm);
            held[j] = model_open(nm, rng);
            CHECK(held[j] >= 0 || held[j] == -EIO || held[j] == E_FSTABLE_FULL, "open returned %d", held[j]);
        }
        HostFt F;
        for (int j = 0; j < k; ++j) {
            if (held[j] < 0) continue;
            CHECK(host_open[held[j]].load() == 1, "holder of entry %d sees the host file closed", held[j]);
            ft_name nm;
            make_name(ids[j], nm);
            CHECK(ft_name_eq(&G.files[held[j]], nm), "entry %d renamed while held", held[j]);
            check_tree((uint32_t)held[j], F, "held");
        }
        if (rng() % 2) std::this_thread::yield();
        for (int j = k - 1; j >= 0; --j)
            if (held[j] >= 0) {
                check_tree((uint32_t)held[j], F, "before close");
                model_close((uint32_t)held[j], rng);
            }
    }
}

// The same audit as gpufs_check_invariants, on the model's arrays.
static void audit() {
    CHECK(G.open_gate == 0, "gate %x at the end", G.open_gate);
    const uint32_t mask = G.name_cap - 1;
    const uint32_t nf = (uint32_t)std::min<unsigned long long>(G.files_next, G.max_files);
    std::vector<uint8_t> nseen(G.node_cap, 0);
    uint64_t in_index = 0, live = 0;
    for (uint32_t i = 0; i < G.name_cap; ++i) {
        const uint32_t v = G.name_index[i];
        if (!v) continue;
        CHECK(v != GFS_NAME_RESERVED && v <= nf, "index slot %u holds %u", i, v);
        if (v == GFS_NAME_RESERVED || v > nf) continue;
        CHECK(fst_state(G.files[v - 1].st) != GFS_FS_FREE, "index slot %u names free entry %u", i, v - 1);
        ++in_index;
    }
    for (uint32_t m = 0; m < nf; ++m) {
        const gfs_file* e = &G.files[m];
        const uint32_t st = fst_state(e->st);
        CHECK(st == GFS_FS_CLOSED || st == GFS_FS_FREE, "entry %u state %u opens %u at the end", m, st,
              fst_opens(e->st));
        CHECK(host_open[m].load() == 0, "entry %u: host file left open", m);
        if (st == GFS_FS_FREE) {
            CHECK(e->root == 0, "free entry %u keeps a tree", m);
            continue;
        }
        ++live;
        uint32_t pos = e->name_hash & mask;
        bool found = false;
        for (uint32_t probe = 0; probe <= mask; ++probe, pos = (pos + 1) & mask) {
            if (G.name_index[pos] == m + 1) {
                found = true;
                break;
            }
            if (G.name_index[pos] == 0) break;
        }
        CHECK(found, "entry %u (%s) not found from its home slot", m, e->name);
        for (uint32_t o = m + 1; o < nf; ++o)
            if (fst_state(G.files[o].st) != GFS_FS_FREE)
                CHECK(strncmp(e->name, G.files[o].name, GFS_NAME_MAX) != 0, "entries %u and %u share %s", m, o,
                      e->name);
        if (e->root) {
            const uint64_t r = e->root & FT_NODE_MASK;
            const uint64_t ns[3] = {r, G.nodes[r * GFS_FANOUT], G.nodes[r * GFS_FANOUT + 5]};
            for (uint64_t n : ns) {
                CHECK(n && n < G.node_cap && !nseen[n], "entry %u: node %lu bad or shared", m, (unsigned long)n);
                if (n && n < G.node_cap) nseen[n] = 1;
            }
        }
    }
    CHECK(live == in_index, "%lu entries in use, %lu in the index", (unsigned long)live, (unsigned long)in_index);
    std::vector<uint8_t> listed(nf, 0);
    for (uint64_t i = std::min<uint64_t>(G.free_taken, G.free_n); i < G.free_n; ++i) {
        const uint32_t m = G.file_free[i];
        CHECK(m < nf && !listed[m] && fst_state(G.files[m].st) == GFS_FS_FREE, "free list item %u", m);
        if (m < nf) listed[m] = 1;
    }
    uint64_t n = G.node_free & FT_NODE_MASK, steps = 0;
    for (; n; ++steps) {
        CHECK(n < G.node_cap && steps < G.node_cap && !nseen[n], "free node %lu bad, repeated or in a tree",
              (unsigned long)n);
        if (!(n < G.node_cap && steps < G.node_cap && !nseen[n])) break;
        nseen[n] = 1;
        for (uint32_t k = 1; k < GFS_FANOUT; ++k) CHECK(G.nodes[n * GFS_FANOUT + k] == 0, "free node not zeroed");
        n = G.nodes[n * GFS_FANOUT] & FT_NODE_MASK;
    }
    const uint64_t bumped = std::min<unsigned long long>(G.node_next, G.node_cap) - 1;
    uint64_t accounted = 0;
    for (uint64_t k = 1; k < G.node_cap; ++k) accounted += nseen[k];
    CHECK(accounted == bumped, "nodes: %lu handed out, %lu in trees or free", (unsigned long)bumped,
          (unsigned long)accounted);
}

int main(int argc, char** argv) {
    n_threads = argc > 1 ? atoi(argv[1]) : 8;
    const int iters = argc > 2 ? atoi(argv[2]) : 20000;
    const uint32_t max_files = argc > 3 ? (uint32_t)atoi(argv[3]) : 24;
    pool_names = 3 * (int)max_files;
    memset(&G, 0, sizeof(G));
    G.max_files = max_files;
    uint32_t cap = 1;
    while (cap < 2 * max_files) cap <<= 1;
    G.name_cap = cap;
    G.node_cap = 1 + 3 * max_files + 8;
    G.node_next = 1;
    std::vector<gfs_file> files(max_files);
    std::vector<uint32_t> index(cap, 0), ffree(max_files, 0);
    std::vector<uint64_t> nodes((size_t)G.node_cap * GFS_FANOUT, 0);
    memset(files.data(), 0, sizeof(gfs_file) * max_files);
    G.files = files.data();
    G.name_index = index.data();
    G.file_free = ffree.data();
    G.nodes = nodes.data();
    host_open = std::vector<std::atomic<int>>(max_files);
    for (auto& h : host_open) h.store(0);
    std::vector<std::thread> ts;
    for (int i = 0; i < n_threads; ++i) ts.emplace_back(worker, i, iters);
    for (auto& t : ts) t.join();
    audit();
    printf("inserts %ld host opens %ld (failed %ld) joins %ld full %ld reclaimed %llu nodes allocated %ld recycled "
           "%llu\n",
           inserts.load(), opens.load(), host_fail.load(), joins.load(), full.load(), G.ctr.files_reclaimed,
           node_allocs.load(), G.ctr.nodes_recycled);
    CHECK(G.ctr.files_reclaimed > 0, "no reclaim pass ran");
    if (failures.load()) {
        printf("ftable model FAILED (%d)\n", failures.load());
        return 1;
    }
    printf("ftable model OK\n");
    return 0;
}
