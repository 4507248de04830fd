// ftable.h -- the file table's name index, entry allocation, reclamation of
// closed entries and the radix-node free list.
//
// Shared by the device API (gpufs.h, agent-scope GPU atomics, one thread of a
// workgroup at a time) and the threaded CPU model in
// tests/csrc/ftable_model.cpp (host atomics); the policy type P supplies
//   T    P::load(T*)  / P::load_rlx(T*)          acquire / relaxed loads
//   void P::store(T*, T) / P::store_rlx(T*, T)   release / relaxed stores
//   T    P::fadd(T*, T)                          fetch-add, acquire-release
//   bool P::cas(T*, T& exp, T)                   acquire-release, updates exp on failure
//   uint32_t P::fand(uint32_t*, uint32_t)        fetch-and, release
//   bool P::expired(), void P::pause()           the caller's wait limit, a brief back-off
//
// The reference keeps a fixed table whose lookup reuses an entry left CLOSED
// when no EMPTY one is left (libgpufs/fs_structures.cu:335-378, esp.
// 364-369), under one global lock.  Here lookups and inserts are lock-free
// (open addressing, linear probing, value = entry + 1) and new entries come
// from a bump counter, then from a free list.  When both are exhausted, one
// opener runs a reclaim pass: it closes the "open gate" (every lookup, insert
// and fst_open join runs inside the gate, counted in open_gate), waits for the
// workgroups inside to leave, and then, alone on the index, takes back up to
// max_files / 8 CLOSED entries in clock order: each is removed from the index
// (backward-shift deletion, so probes need no tombstones), its radix tree goes
// to the node free list, its version is bumped (cached page references such as
// the read memo die with it) and it is put on the free list.
//
// An entry is taken back only when CLOSED with no open, and only while no
// opener is inside the gate, so no workgroup holds or is about to join it:
// workgroups join an entry only inside the gate, and a workgroup outside the
// gate holds an entry only while it is OPENING, OPEN or CLOSING.
#pragma once
#include <errno.h>
#include "fstate.h"
#include "gfs_layout.h"

#define FT_EXCL 0x80000000u   // open_gate: a reclaim pass holds the index
#define FT_FOUND 0            // ft_find_or_insert: the name is in the index
#define FT_NEW 1              // ft_find_or_insert: inserted, entry OPENING with this caller's open
#define FT_NODE_MASK ((1ull << 40) - 1)

struct ft_name {
    uint64_t w[GFS_NAME_MAX / 8];   // zero-padded
    int len;
    uint32_t hash;
};

GFS_HD inline bool ft_name_eq(const gfs_file* e, const ft_name& n) {
    const uint64_t* w = (const uint64_t*)e->name;
    uint64_t d = 0;
    for (int i = 0; i < GFS_NAME_MAX / 8; ++i) d |= w[i] ^ n.w[i];
    return d == 0;
}

// ---- the open gate
template <class P> GFS_HD int ft_gate_enter(gfs_dev* g, P& p) {
    for (;;) {
        const uint32_t v = p.fadd(&g->open_gate, 1u);
        if (!(v & FT_EXCL)) return 0;
        p.fadd(&g->open_gate, 0xFFFFFFFFu);   // a reclaim pass runs: step out and wait
        while (p.load(&g->open_gate) & FT_EXCL) {
            if (p.expired()) return -ETIMEDOUT;
            p.pause();
        }
    }
}
template <class P> GFS_HD void ft_gate_leave(gfs_dev* g, P& p) { p.fadd(&g->open_gate, 0xFFFFFFFFu); }

// ---- radix-node free list (a stack with a tag against ABA; the link of a
// free node is its word 0, its other words are zero)
template <class P> GFS_HD void ft_nodes_push(gfs_dev* g, uint64_t head, uint64_t tail, P& p) {
    unsigned long long old = p.load(&g->node_free);
    for (;;) {
        p.store_rlx(&g->nodes[tail * GFS_FANOUT], old & FT_NODE_MASK);
        const uint64_t nw = ((((old >> 40) + 1) & 0xFFFFFF) << 40) | head;
        if (p.cas(&g->node_free, old, nw)) return;
    }
}
// A zeroed node from the free list, or 0 when it is empty.
template <class P> GFS_HD uint64_t ft_node_pop(gfs_dev* g, P& p) {
    unsigned long long old = p.load(&g->node_free);
    for (;;) {
        const uint64_t n = old & FT_NODE_MASK;
        if (n == 0) return 0;
        // may read a node another thread has just popped and reused: the tag
        // then differs and the exchange fails
        const uint64_t next = p.load_rlx(&g->nodes[n * GFS_FANOUT]) & FT_NODE_MASK;
        const uint64_t nw = ((((old >> 40) + 1) & 0xFFFFFF) << 40) | next;
        if (p.cas(&g->node_free, old, nw)) {
            p.store_rlx(&g->nodes[n * GFS_FANOUT], 0ull);
            return n;
        }
    }
}
// Returns every node of the tree with root word `rw` to the free list (one
// thread; nobody else reaches the tree).  Returns the number of nodes.
template <class P> GFS_HD uint64_t ft_tree_free(gfs_dev* g, uint64_t rw, P& p) {
    if (rw == 0) return 0;
    const uint32_t h = (uint32_t)(rw >> 40);
    if (h == 0 || h > GFS_MAX_HEIGHT) return 0;
    uint64_t stk[GFS_MAX_HEIGHT];
    uint32_t at[GFS_MAX_HEIGHT];
    int d = 0;
    stk[0] = rw & FT_NODE_MASK;
    at[0] = 0;
    uint64_t head = 0, tail = 0, cnt = 0;
    while (d >= 0) {
        const uint64_t n = stk[d];
        // nodes above the leaf level hold child nodes; leaves hold frame references
        if (d + 1 < (int)h && at[d] < GFS_FANOUT) {
            const uint64_t c = p.load_rlx(&g->nodes[n * GFS_FANOUT + at[d]++]);
            if (c && c < g->node_cap) {
                ++d;
                stk[d] = c;
                at[d] = 0;
            }
            continue;
        }
        for (uint32_t k = 1; k < GFS_FANOUT; ++k) p.store_rlx(&g->nodes[n * GFS_FANOUT + k], 0ull);
        p.store_rlx(&g->nodes[n * GFS_FANOUT], head);
        if (!tail) tail = n;
        head = n;
        ++cnt;
        --d;
    }
    ft_nodes_push(g, head, tail, p);
    p.fadd(&g->ctr.nodes_recycled, (unsigned long long)cnt);
    return cnt;
}

// ---- entries (inside the gate)
template <class P> GFS_HD int64_t ft_take_entry(gfs_dev* g, P& p) {
    if (p.load_rlx(&g->files_next) < g->max_files) {
        const unsigned long long m = p.fadd(&g->files_next, 1ull);
        if (m < g->max_files) return (int64_t)m;
    }
    const uint32_t n = p.load(&g->free_n);
    if (p.load_rlx(&g->free_taken) < n) {
        const unsigned long long i = p.fadd(&g->free_taken, 1ull);
        if (i < n) return (int64_t)p.load_rlx(&g->file_free[i]);
    }
    return -1;
}

// Looks `nm` up, inserting it when absent (inside the gate).  FT_FOUND or
// FT_NEW with the entry in *idx, E_FSTABLE_FULL, or -ETIMEDOUT.  A new entry
// is OPENING with one open (the caller's), and its version continues from the
// entry's earlier lives.
template <class P> GFS_HD int ft_find_or_insert(gfs_dev* g, const ft_name& nm, P& p, uint32_t* idx) {
    const uint32_t mask = g->name_cap - 1;
    uint32_t pos = nm.hash & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe, pos = (pos + 1) & mask) {
        uint32_t v = p.load(&g->name_index[pos]);
        if (v == 0) {
            // reserve the slot first, so only the opener that publishes the
            // name takes an entry
            uint32_t exp = 0;
            if (p.cas(&g->name_index[pos], exp, GFS_NAME_RESERVED)) {
                const int64_t m = ft_take_entry(g, p);
                if (m < 0) {
                    p.store(&g->name_index[pos], 0u);
                    return E_FSTABLE_FULL;
                }
                gfs_file* e = &g->files[m];
                uint64_t* en = (uint64_t*)e->name;
                for (int i = 0; i < GFS_NAME_MAX / 8; ++i) en[i] = nm.w[i];
                e->name_hash = nm.hash;
                e->host_fd = -1;
                p.store_rlx(&e->root, 0ull);
                const uint64_t old = p.load_rlx(&e->st);
                p.store_rlx(&e->st, fst_make(GFS_FS_OPENING, 1, fst_version(old)));
                p.store(&g->name_index[pos], (uint32_t)m + 1u);   // publish the entry
                *idx = (uint32_t)m;
                return FT_NEW;
            }
            v = exp;
        }
        while (v == GFS_NAME_RESERVED) {   // another opener is publishing this slot
            if (p.expired()) return -ETIMEDOUT;
            p.pause();
            v = p.load(&g->name_index[pos]);
        }
        if (v == 0) {   // the reservation was given up (table full): look at the slot again
            --probe;
            pos = (pos - 1) & mask;
            continue;
        }
        const gfs_file* e = &g->files[v - 1];
        if (e->name_hash == nm.hash && ft_name_eq(e, nm)) {
            *idx = v - 1;
            return FT_FOUND;
        }
    }
    return E_FSTABLE_FULL;
}

// Backward-shift deletion of entry m's slot (exclusive holder of the index).
template <class P> GFS_HD void ft_index_remove(gfs_dev* g, uint32_t m, P& p) {
    const uint32_t mask = g->name_cap - 1;
    uint32_t i = g->files[m].name_hash & mask;
    for (uint32_t probe = 0;; ++probe, i = (i + 1) & mask) {
        if (probe > mask) return;
        const uint32_t v = p.load_rlx(&g->name_index[i]);
        if (v == m + 1) break;
        if (v == 0) return;   // not in the index
    }
    uint32_t j = i;
    for (;;) {
        j = (j + 1) & mask;
        const uint32_t v = p.load_rlx(&g->name_index[j]);
        if (v == 0) break;
        const uint32_t k = g->files[v - 1].name_hash & mask;
        // the entry at j stays when its home slot k lies cyclically in (i, j]
        const bool stays = i <= j ? (i < k && k <= j) : (i < k || k <= j);
        if (stays) continue;
        p.store_rlx(&g->name_index[i], v);
        i = j;
    }
    p.store_rlx(&g->name_index[i], 0u);
}

// One reclaim pass (called outside the gate after ft_find_or_insert found
// the table full).  Returns 1 when the caller should look again (entries were
// freed, here or by a pass that ran meanwhile), 0 when no entry could be
// taken back (every entry is in use), or -ETIMEDOUT.
template <class P> GFS_HD int ft_reclaim(gfs_dev* g, P& p) {
    uint32_t v = p.load(&g->open_gate);
    for (;;) {
        if (v & FT_EXCL) {
            while (p.load(&g->open_gate) & FT_EXCL) {
                if (p.expired()) return -ETIMEDOUT;
                p.pause();
            }
            return 1;
        }
        if (p.cas(&g->open_gate, v, v | FT_EXCL)) break;
    }
    while ((p.load(&g->open_gate) & ~FT_EXCL) != 0) {
        if (p.expired()) {
            p.fand(&g->open_gate, ~FT_EXCL);
            return -ETIMEDOUT;
        }
        p.pause();
    }
    // alone on the index, the entry allocator and every entry that is CLOSED
    const unsigned long long fnext = p.load_rlx(&g->files_next);
    const uint32_t n = fnext < g->max_files ? (uint32_t)fnext : g->max_files;
    const uint32_t quota = g->max_files / 8 ? g->max_files / 8 : 1u;
    const uint32_t fn = p.load_rlx(&g->free_n);
    const unsigned long long taken = p.load_rlx(&g->free_taken);
    uint32_t cnt = 0, found = 0;
    for (unsigned long long i = taken; i < fn; ++i) g->file_free[cnt++] = g->file_free[i];
    uint32_t s = 0;
    const uint32_t hand = n ? g->reclaim_hand % n : 0;
    for (; s < n && found < quota; ++s) {
        const uint32_t m = (hand + s) % n;
        gfs_file* e = &g->files[m];
        uint64_t st = p.load(&e->st);
        if (fst_state(st) != GFS_FS_CLOSED || fst_opens(st) != 0) continue;
        if (!p.cas(&e->st, st, fst_make(GFS_FS_FREE, 0, fst_version(st) + 1))) continue;
        ft_index_remove(g, m, p);
        const uint64_t rw = p.load_rlx(&e->root);
        p.store_rlx(&e->root, 0ull);
        ft_tree_free(g, rw, p);
        g->file_free[cnt++] = m;
        ++found;
    }
    g->reclaim_hand = n ? (hand + s) % n : 0;
    p.store_rlx(&g->free_taken, 0ull);
    p.store_rlx(&g->free_n, cnt);
    p.fadd(&g->ctr.files_reclaimed, (unsigned long long)found);
    p.fand(&g->open_gate, ~FT_EXCL);   // release: openers see the new index and free list
    return cnt ? 1 : 0;
}
