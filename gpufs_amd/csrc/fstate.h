// fstate.h -- open/close protocol on a file-table entry's state word.
//
// Shared by the device API (gpufs.h, with agent-scope GPU atomics) and the
// threaded CPU model in tests/csrc/fstate_model.cpp (with host atomics); the
// policy type P supplies the operations:
//   uint64_t P::fadd(uint64_t* w, uint64_t v)    fetch-add, acquire-release
//   uint64_t P::load(uint64_t* w)                acquire load
//   uint64_t P::load_rlx(uint64_t* w)            relaxed load (polls: the read-modify-write that
//                                                follows a change acquires)
//   bool     P::cas(uint64_t* w, uint64_t& exp, uint64_t v)   updates exp on failure
//   bool     P::expired()                        the caller's wait limit passed
//   void     P::pause()                          brief back-off
//
// Word: state [0,8) | opens [8,32) | version [32,64).  Once the entry is
// published in the name index every change is an atomic add of a delta (or a
// CAS from the value read to that value plus a delta), so a count added by an
// opener that is still deciding is never overwritten:
//   * an opener looks at the word first: OPEN -> it adds one open and is
//     joined if the state was still OPEN (else it takes the count back and
//     looks again); CLOSED -> it moves the state to OPENING with its open by
//     CAS and does the host open; OPENING / CLOSING -> it waits on loads.
//   * a closer subtracts its open; the one that leaves OPEN with zero opens
//     moves it to CLOSING by CAS, which fails when an opener joined meanwhile.
// When every workgroup of a launch opens the same file, an open or close of
// an open file is then one atomic, not a CAS retry loop over a shared word.
// (Reference semantics: libgpufs/fs_calls.cu:152-225 single_thread_open and
// the close path of gclose.)
#pragma once
#include <errno.h>
#include "gfs_layout.h"

#define FST_ONE (1ull << 8)
#define FST_NEG ((uint64_t)0 - FST_ONE)

GFS_HD inline uint32_t fst_state(uint64_t s) { return (uint32_t)(s & 0xFF); }
GFS_HD inline uint32_t fst_opens(uint64_t s) { return (uint32_t)((s >> 8) & 0xFFFFFF); }
GFS_HD inline uint32_t fst_version(uint64_t s) { return (uint32_t)(s >> 32); }
GFS_HD inline uint64_t fst_make(uint32_t state, uint32_t opens, uint32_t version) {
    return ((uint64_t)version << 32) | ((uint64_t)(opens & 0xFFFFFF) << 8) | state;
}
// Added to a word whose state is `from`, moves the state byte to `to`
// (arithmetic modulo 2^64; the state byte never borrows since it holds `from`).
GFS_HD inline uint64_t fst_delta(uint32_t from, uint32_t to) { return (uint64_t)to - (uint64_t)from; }

#define FST_JOINED 0     // the file is OPEN and this open is counted
#define FST_MUST_OPEN 1  // state is OPENING with this open counted: do the host open, then fst_open_done
#define FST_PEEK 2       // (peek) another caller's host open is in progress; nothing was counted

// With `peek`, a caller that finds another caller's host open in progress
// (OPENING) returns FST_PEEK at once instead of waiting to join.
template <class P> GFS_HD int fst_open(uint64_t* w, P& p, bool peek = false) {
    // Look before adding: when a launch's workgroups all open the same closed
    // file, one of them wins the CAS to OPENING and the others wait on plain
    // loads, then join with one add each (every atomic on the word queues
    // behind the others at one L2 channel).  The look is relaxed: the add or
    // CAS that acts on it acquires (an acquire load at agent scope would also
    // invalidate the opener's L2).
    uint64_t cur = p.load_rlx(w);
    for (;;) {
        const uint32_t s = fst_state(cur);
        if (s == GFS_FS_OPEN) {
            const uint64_t old = p.fadd(w, FST_ONE);
            if (fst_state(old) == GFS_FS_OPEN) return FST_JOINED;
            // a closer moved it to CLOSING first: give the count back, look again
            cur = p.fadd(w, FST_NEG) + FST_NEG;
            continue;
        }
        if (s == GFS_FS_CLOSED) {
            if (p.cas(w, cur, cur + fst_delta(GFS_FS_CLOSED, GFS_FS_OPENING) + FST_ONE)) return FST_MUST_OPEN;
            continue;   // cur holds the current word
        }
        // OPENING or CLOSING: another workgroup is at the host.  Relaxed
        // polls: an acquire load at agent scope invalidates the L2 of the
        // poller's XCD each time; the fadd or cas that follows acquires.
        if (peek && s == GFS_FS_OPENING) return FST_PEEK;
        if (p.expired()) return -ETIMEDOUT;
        p.pause();
        cur = p.load_rlx(w);
    }
}

// Outcome of the host open begun after FST_MUST_OPEN: OPEN keeping every
// count (a new cache version when `bump_version`), or CLOSED without the
// caller's count.
template <class P> GFS_HD void fst_open_done(uint64_t* w, P& p, bool ok, bool bump_version) {
    if (ok) p.fadd(w, fst_delta(GFS_FS_OPENING, GFS_FS_OPEN) + (bump_version ? (1ull << 32) : 0ull));
    else p.fadd(w, fst_delta(GFS_FS_OPENING, GFS_FS_CLOSED) + FST_NEG);
}

// Drops one open.  Returns -EBADF (the file is not open), 0 (other opens
// remain, or an opener joined), 1 (last: the state is CLOSING; close on the
// host, then fst_close_done).
template <class P> GFS_HD int fst_close(uint64_t* w, P& p) {
    const uint64_t st = p.load_rlx(w);   // a check only: the add below orders the close
    if (fst_state(st) != GFS_FS_OPEN || fst_opens(st) == 0) return -EBADF;
    uint64_t cur = p.fadd(w, FST_NEG) + FST_NEG;
    while (fst_state(cur) == GFS_FS_OPEN && fst_opens(cur) == 0) {
        if (p.cas(w, cur, cur + fst_delta(GFS_FS_OPEN, GFS_FS_CLOSING))) return 1;
        // cur holds the current word: a version bump retries, a joined opener ends it
    }
    return 0;
}

template <class P> GFS_HD void fst_close_done(uint64_t* w, P& p) {
    p.fadd(w, fst_delta(GFS_FS_CLOSING, GFS_FS_CLOSED));
}
