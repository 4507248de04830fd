// gpufs.h -- device-side file API for MI355X (gfx950), header-only HIP.
//
// Same calls, signatures and flag values as the reference's device API
// (libgpufs/fs_calls.cu.h:41-65): gopen, gclose, gread, gwrite, gmmap, gmunmap,
// gfsync, gftruncate, gmsync, fstat.  Every call is workgroup-collective: all
// threads of the workgroup call it with the same arguments (the contract of
// libgpufs/fs_calls.cu:685-687); calling it from a divergent subset of threads
// is undefined.  fstat is an ordinary per-thread read.
//
// Design (not a translation of the reference):
//  * page-cache index: one lock-free radix tree per open file, fan-out 64.  A
//    batch of up to 64 consecutive pages is resolved by wave 0 in one step:
//    lane i owns page i, the 64 leaf slots of a node are one coalesced 512-byte
//    load, and hits / misses / in-flight fills are separated with ballots.
//  * frames: a 64-bit state word per frame (pin count, state, dirty flag,
//    generation).  Leaf slots hold (frame, generation); a recycled frame has a
//    new generation, so a stale leaf is detected by the pinning atomic itself
//    (no ABA, no lock held across a fill).
//  * allocation: never-used frames first, then a wave-wide CLOCK sweep (64
//    frames per atomic on the hand) that evicts clean, unpinned frames with a
//    second-chance reference bit.
//  * host RPC: slots claimed from a device bitmap, requests published into a
//    submission queue in coherent pinned host memory with system-scope release
//    stores; the host service pread()s into pinned staging and sets the
//    request's bit in its batch's 64-bit completion mask; the workgroup copies
//    each page into its HBM frame with 16-byte loads and stores as soon as it
//    has arrived and returns its slot.
//  * file table: open/close by atomic adds on a per-file state word (fstate.h),
//    no lock.
//
// Workgroups must have at least 64 threads (one full wave64): wave 0 resolves
// up to 64 pages at a time, one page per lane.
//
// Errors: calls returning int return a negative code; gread/gwrite keep the
// reference's size_t return type and report errors as (size_t)(-errno), so
// `(ssize_t)ret < 0` tests for failure.  gread returns the bytes actually read
// (short at end of file); the reference always returns `size`
// (libgpufs/fs_calls.cu:720-736).
#pragma once
#include <hip/hip_runtime.h>
#include <errno.h>
#include <stdint.h>
#include "gfs_layout.h"
#include "fstate.h"
#include "ftable.h"
#include "../../include/gpufs_amd.h"

typedef unsigned char uchar;

// Device state of the bound context, one copy per translation unit (no
// relocatable device code is needed); gpufs_bind() sets it from the host.
static __device__ gfs_dev* gfs_g;

namespace gfs {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4u gv4u;
typedef __attribute__((address_space(3))) v4u lv4u;

#define GFS_NOFRAME 0xFFFFFFFFu
#define GFS_BUSYFRAME 0xFFFFFFFEu   // take_dirty: look at the page again later
#define GFS_K_NONE 0
#define GFS_K_HIT 1
#define GFS_K_OWN 2
#define GFS_K_WAIT 3
#define GFS_K_MISS 4
#define GFS_K_ERR 5
#define GFS_K_DONE 6

// Per-workgroup control block shared by all API calls (calls are collective
// and sequential within a workgroup, so one block suffices).
struct wg_ctl {
    uint32_t frame[64];
    uint32_t slot[64];
    uint8_t  kind[64];
    int64_t  ret;
    int32_t  err;
    uint64_t own;            // batch pages with a host read posted (bit = page index in the batch)
    uint64_t rel;            // of those, pages whose ring slot was already returned
    uint32_t cq;
    uint32_t flag;
    uint32_t tmo;            // wait-limit decision of thread 0, broadcast to the workgroup
    uint64_t a, b;
    uint32_t rot;
    uint32_t hits, misses;   // running page-lookup counts of this workgroup (wraps; callers take differences)
    uint32_t ra;             // gread: extra pages fetched after the request (read-ahead)
    uint32_t gen[64];        // resolve_batch: generation of each HIT / OWN frame
    uint32_t memo_f;         // gread: frame pinned through the page memo at entry, or GFS_NOFRAME
    uint32_t memo_f1;        // gread: the next page's frame when the read runs into it, or GFS_NOFRAME
    // gread's entry results (thread 0 writes them before gread's first barrier,
    // so no other call reads them after its own last barrier)
    int64_t gr_ret;
    uint64_t gr_root;
    uint32_t gr_ra;
    uint32_t pinf[64];       // pin_pages: frame of each pinned page
};

// Per-workgroup read history for sequential read-ahead.  It is advisory: it
// only decides how many pages beyond a read are fetched together with it.
// LDS keeps whatever the previous workgroup on the compute unit left there, so
// gopen clears the history; a read of a descriptor this workgroup did not open
// may meet a stale entry, which can cost one unneeded prefetch (or skip one),
// never a wrong byte.  (A tag built from the dispatch and kernel-argument
// pointers kept four more SGPRs live through every caller's loops and pushed
// the grep scan's window loop into scratch spills.)
// The pages of the workgroup's last gread batch, read-ahead pages included:
// their frames and generations stay in gfs_ctl.frame / gen (kind HIT or OWN)
// until another batch reuses the arrays.  A following read inside one of
// those pages pins that frame in gread's first round trip and skips the radix
// walk (a grep-style loop reads a 256 KiB page in four 64 KiB windows, and the
// pages its read-ahead brought in).  The pin only counts if the frame still
// carries the recorded generation, is VALID and is owned by this file and
// page, and the file still has `version`: a frame that was evicted or
// reassigned has a new generation, and truncation or a reopen after the file
// changed bumps the version.  Arrays overwritten by another call's batch can
// only name a frame that fails those checks or truly holds the page.  gopen
// clears the memo; one left in LDS by an earlier launch or by other code is
// bounded to the arrays and checked the same way.
struct wg_memo {
    int32_t fd;         // -1: none
    uint32_t cnt;       // pages of the batch (at most 64)
    uint32_t version;
    uint32_t pad;
    uint64_t pb;        // first page of the batch
};
struct wg_hist {
    int32_t fd;         // -1: no history
    uint32_t streak;    // forward-sequential reads in a row
    uint64_t first;     // pages of the last read
    uint64_t last;
    uint64_t fetched;   // pages below this were read or fetched ahead
};

__device__ __forceinline__ uint32_t tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ uint32_t nthreads() { return blockDim.x * blockDim.y * blockDim.z; }
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave() { return tid() >> 6; }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t popc(uint64_t m) { return __popcll(m); }

// Values broadcast through LDS are the same in every lane, but the compiler
// cannot prove it; branching on them directly makes it lower loops that hold
// barriers as divergent loops with lane-mask bookkeeping, which can keep some
// lanes iterating while others leave.  Every control-flow value read from a
// broadcast goes through v_readfirstlane so the branch is scalar (SGPR).
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int32_t uni(int32_t v) { return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v); }
// __builtin_amdgcn_readfirstlane returns int: widen through uint32_t, never
// straight from int (that would sign-extend bit 31 into the high word).
__device__ __forceinline__ uint64_t uni(uint64_t v) {
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | (uint64_t)lo;
}
__device__ __forceinline__ int64_t uni(int64_t v) { return (int64_t)uni((uint64_t)v); }
__device__ __forceinline__ unsigned long long uni(unsigned long long v) { return (unsigned long long)uni((uint64_t)v); }
__device__ __forceinline__ uint32_t first_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
__device__ __forceinline__ uint32_t nth_set(uint64_t m, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) m &= m - 1;
    return first_lane(m);
}

// ---- scoped atomics: agent scope for HBM state, system scope for host memory
// (the value parameter is a non-deduced context so literals convert to T)
template <class T> struct ident { typedef T type; };
template <class T> __device__ __forceinline__ T ld_acq(T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T ld_rlx(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_rel(T* p, typename ident<T>::type v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_rlx(T* p, typename ident<T>::type v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T fadd_rlx(T* p, typename ident<T>::type v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T fadd(T* p, typename ident<T>::type v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ bool cas(T* p, T& expected, typename ident<T>::type desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                                __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T sys_ld(T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T> __device__ __forceinline__ T sys_ld_acq(T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T> __device__ __forceinline__ void sys_st(T* p, typename ident<T>::type v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T> __device__ __forceinline__ void sys_st_rel(T* p, typename ident<T>::type v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void fence_acq_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_rel_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq_sys() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, ""); }
__device__ __forceinline__ void fence_rel_sys() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, ""); }
__device__ __forceinline__ void backoff() { __builtin_amdgcn_s_sleep(16); }
__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

// ftable.h policy: agent-scope atomics, the context's wait limit.
struct ft_dev {
    uint64_t t0, limit;
    template <class T> __device__ T load(T* x) { return ld_acq(x); }
    template <class T> __device__ T load_rlx(T* x) { return ld_rlx(x); }
    template <class T> __device__ void store(T* x, typename ident<T>::type v) { st_rel(x, v); }
    template <class T> __device__ void store_rlx(T* x, typename ident<T>::type v) { st_rlx(x, v); }
    template <class T> __device__ T fadd(T* x, typename ident<T>::type v) { return gfs::fadd(x, v); }
    template <class T> __device__ bool cas(T* x, T& e, typename ident<T>::type v) { return gfs::cas(x, e, v); }
    __device__ uint32_t fand(uint32_t* x, uint32_t v) {
        return __hip_atomic_fetch_and(x, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ bool expired() { return now_ticks() - t0 > limit; }
    __device__ void pause() { __builtin_amdgcn_s_sleep(16); }
};
// Phase hook for timing studies: expands to nothing unless a build defines it.
#ifndef GFS_TRACE
#define GFS_TRACE(id) ((void)0)
#endif

__device__ __forceinline__ uint32_t meta_state(uint64_t m) { return (uint32_t)((m >> GFS_ST_SHIFT) & 0xF); }
__device__ __forceinline__ uint32_t meta_gen(uint64_t m) { return (uint32_t)(m >> GFS_GEN_SHIFT); }
__device__ __forceinline__ uint32_t meta_ref(uint64_t m) { return (uint32_t)(m & GFS_REF_MASK); }
__device__ __forceinline__ uint64_t leaf_pack(uint32_t f, uint32_t gen) { return ((uint64_t)gen << 32) | (uint64_t)(f + 1u); }
__device__ __forceinline__ uint32_t leaf_frame(uint64_t v) { return (uint32_t)(v & 0xFFFFFFFFu) - 1u; }
__device__ __forceinline__ uint32_t leaf_gen(uint64_t v) { return (uint32_t)(v >> 32); }


__device__ __forceinline__ void count(unsigned long long* c, unsigned long long v) {
    fadd_rlx(c, v);
}
// Wave-aggregated counter bump: one atomic per wave (call with the whole wave).
__device__ __forceinline__ void wave_count(unsigned long long* c, bool pred) {
    uint64_t m = __ballot(pred);
    if (m && lane() == first_lane(m)) fadd_rlx(c, (unsigned long long)popc(m));
}
// The count of dirty frames (gfs_dev::dirty_frames) follows every change of a
// frame's DIRTY bit: the wave's lanes whose frame changed (pred) are added
// (up) or taken off with one atomic.
__device__ __forceinline__ void dirty_tally(gfs_dev* g, bool pred, bool up) {
    const uint64_t m = __ballot(pred);
    if (m && lane() == first_lane(m)) {
        const unsigned long long k = (unsigned long long)popc(m);
        fadd_rlx(&g->dirty_frames, up ? k : (unsigned long long)0 - k);
    }
}
// Sets frame f's DIRTY bit; true when the frame was clean.
__device__ __forceinline__ bool set_dirty(gfs_dev* g, uint32_t f) {
    return !(__hip_atomic_fetch_or(&g->meta[f], GFS_FL_DIRTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
             GFS_FL_DIRTY);
}

// ======================================================================
// Copy engine: workgroup-cooperative byte copies with 16-byte vector bodies.
// `j` is the thread's rank in [0,T).  Handles any alignment: equal 16-byte
// phase -> dwordx4; otherwise dst is aligned to 16 and each 16-byte unit is
// assembled from five aligned source dwords with v_alignbyte.  Every source
// dword loaded holds at least one byte of the segment.
// ======================================================================
// True when a generic pointer points into LDS.
__device__ __forceinline__ bool is_lds(const void* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_is_shared(p);
#else
    (void)p;
    return false;
#endif
}
__device__ __forceinline__ v4u load16(const uint8_t* p, bool src_lds, bool nt) {
    if (src_lds) return *(const lv4u*)(p);
    if (nt) return __builtin_nontemporal_load((const gv4u*)p);
    return *(const gv4u*)p;
}
__device__ __forceinline__ void store16(uint8_t* p, v4u v, bool dst_lds, bool nt) {
    if (dst_lds) { *(lv4u*)p = v; return; }
    if (nt) { __builtin_nontemporal_store(v, (gv4u*)p); return; }
    *(gv4u*)p = v;
}

// Address space of a copy operand, when the caller knows it.
enum buf_space { SPACE_ANY = 0, SPACE_LDS = 1, SPACE_GLOBAL = 2 };

// The non-temporal choices are template constants: as runtime flags the
// compiler merges the two load forms and the nontemporal hint is lost.  The
// operand spaces are template constants too: with SPACE_ANY every path below
// is compiled into the caller, and the global->global body's data registers
// then count against the caller's loops even where it never runs.  `depth` is
// the number of 16-byte loads each lane keeps in flight on the global->global
// path (eight for HBM-to-HBM streams, 4 where a copy sits inside
// loops that need the registers).
template <bool nt_load, bool nt_store, buf_space dst_space = SPACE_ANY, buf_space src_space = SPACE_ANY,
          int depth = 8>
__device__ inline void copy_seg(uint8_t* dst, const uint8_t* src, uint64_t len, uint32_t j, uint32_t T) {
    static_assert(depth == 4 || depth == 8, "copy depth");
    if (len == 0) return;
    const bool dl = dst_space == SPACE_ANY ? is_lds(dst) : dst_space == SPACE_LDS;
    const bool sl = src_space == SPACE_ANY ? is_lds(src) : src_space == SPACE_LDS;
    const uintptr_t d = (uintptr_t)dst, s = (uintptr_t)src;
    uint64_t head = (16 - (d & 15)) & 15;
    if (head > len) head = len;
    for (uint64_t k = j; k < head; k += T) dst[k] = src[k];
    const uint64_t rest = len - head;
    const uint64_t units = rest >> 4;
    const uint64_t tail = rest & 15;
    uint8_t* D = dst + head;
    const uint8_t* S = src + head;
    for (uint64_t k = j; k < tail; k += T) D[units * 16 + k] = S[units * 16 + k];
    if (units == 0) return;
    if (((s ^ d) & 15) == 0) {
        if (!sl && !dl) {
            // global -> global: `depth` 16-byte loads in flight per lane
            const gv4u* S4 = (const gv4u*)S;
            gv4u* D4 = (gv4u*)D;
            const uint64_t st = T;
            uint64_t u = j;
            for (; u + (depth - 1) * st < units; u += depth * st) {
                v4u r[depth];
#pragma unroll
                for (int k = 0; k < depth; ++k)
                    r[k] = nt_load ? __builtin_nontemporal_load(S4 + u + k * st) : S4[u + k * st];
#pragma unroll
                for (int k = 0; k < depth; ++k) {
                    if (nt_store) __builtin_nontemporal_store(r[k], D4 + u + k * st);
                    else D4[u + k * st] = r[k];
                }
            }
            for (; u < units; u += st) {
                v4u r0 = nt_load ? __builtin_nontemporal_load(S4 + u) : S4[u];
                if (nt_store) __builtin_nontemporal_store(r0, D4 + u);
                else D4[u] = r0;
            }
            return;
        }
        uint64_t u = j;
        const uint64_t st = T;
        if (dst_space == SPACE_LDS && !sl) {
            // into LDS (a window of at most 64 KiB plus a look-ahead): five
            // loads per lane in flight, the ones past the end switched off, so
            // a window a little over 4 x 16 x T bytes still takes one round trip
            for (; u < units; u += 5 * st) {
                v4u r[5];
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    if (u + k * st < units) r[k] = load16(S + 16 * (u + k * st), false, nt_load);
#pragma unroll
                for (int k = 0; k < 5; ++k)
                    if (u + k * st < units) store16(D + 16 * (u + k * st), r[k], true, nt_store);
            }
            return;
        }
        for (; u + 3 * st < units; u += 4 * st) {
            v4u a = load16(S + 16 * u, sl, nt_load);
            v4u b = load16(S + 16 * (u + st), sl, nt_load);
            v4u c = load16(S + 16 * (u + 2 * st), sl, nt_load);
            v4u e = load16(S + 16 * (u + 3 * st), sl, nt_load);
            store16(D + 16 * u, a, dl, nt_store);
            store16(D + 16 * (u + st), b, dl, nt_store);
            store16(D + 16 * (u + 2 * st), c, dl, nt_store);
            store16(D + 16 * (u + 3 * st), e, dl, nt_store);
        }
        for (; u < units; u += st) store16(D + 16 * u, load16(S + 16 * u, sl, nt_load), dl, nt_store);
        return;
    }
    // misaligned source: S is shifted by sh bytes from a dword boundary
    const uint32_t sh = (uint32_t)((uintptr_t)S & 3);
    const uint32_t* W = (const uint32_t*)((uintptr_t)S - sh);
    for (uint64_t u = j; u < units; u += T) {
        const uint32_t* w = W + 4 * u;
        v4u o;
        if (sh == 0) {
            o.x = w[0]; o.y = w[1]; o.z = w[2]; o.w = w[3];
        } else {
            uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
            o.x = __builtin_amdgcn_alignbyte(w1, w0, sh);
            o.y = __builtin_amdgcn_alignbyte(w2, w1, sh);
            o.z = __builtin_amdgcn_alignbyte(w3, w2, sh);
            o.w = __builtin_amdgcn_alignbyte(w4, w3, sh);
        }
        store16(D + 16 * u, o, dl, nt_store);
    }
}

// Zeroes bytes [from, to) of a frame (`to` a multiple of 16) with thread j of
// T: the head up to a 16-byte boundary byte by byte, then 16-byte stores.
__device__ inline void zero_tail(uint8_t* fp, uint64_t from, uint64_t to, uint32_t j, uint32_t T) {
    const uint64_t a = (from + 15) & ~15ull;
    if (j < a - from && from + j < to) fp[from + j] = 0;
    for (uint64_t k = (a >> 4) + j; k < (to >> 4); k += T) *(v4u*)(fp + 16 * k) = v4u{0, 0, 0, 0};
}

// Whole-workgroup copy of one buffer.
__device__ inline void wg_copy(uint8_t* dst, const uint8_t* src, uint64_t len) {
    copy_seg<false, false>(dst, src, len, tid(), nthreads());
}

// ======================================================================
// Radix tree (per file, fan-out 64).  root word = node | height << 40.
// Node n occupies g->nodes[n*64 .. n*64+63]; node 0 is the null node.
// ======================================================================
__device__ __forceinline__ uint64_t tree_cap(uint32_t h) {
    return h >= GFS_MAX_HEIGHT ? ~0ull : (1ull << (GFS_FANOUT_LOG * h));
}
// A zeroed node: from the never-used range, else from the free list that
// reclaimed files feed (ftable.h; out of line, it is rare).  0 when neither
// has one.
__device__ __noinline__ uint64_t node_pop_free(gfs_dev* g) {
    ft_dev P{0, 0};
    return ft_node_pop(g, P);
}
__device__ inline uint64_t node_alloc(gfs_dev* g) {
    if (ld_rlx(&g->node_next) < g->node_cap) {
        unsigned long long n = fadd_rlx(&g->node_next, 1ull);
        if (n < g->node_cap) return n;
    }
    return node_pop_free(g);
}

// Returns the node index of the leaf covering pages [leaf*64, leaf*64+64), creating
// interior nodes on the way when `create`.  0 = absent (or -ENOMEM in *err).
// `root_hint` (nonzero) is a root word read earlier by the caller: a tree only
// grows by pushing the old root down as child 0, so an older root still holds
// every page it covered.  One too small for `leaf` is not grown from: the
// current root is read first (a caller that retries with the same old hint
// would otherwise allocate, and lose, a growth node on every retry, and the
// node pool is a bump allocator).  A growth node lost to a racing workgroup
// is kept for this call's next growth step.
__device__ inline uint64_t tree_leaf(gfs_dev* g, gfs_file* f, uint64_t leaf, bool create, int* err,
                                     uint64_t root_hint = 0) {
    const uint64_t page = leaf << GFS_FANOUT_LOG;
    // Relaxed loads throughout: agent-scope atomic loads are coherent (they
    // do not read a stale line of the XCD's L2), each node word is read
    // through an address taken from the word above it, and a node is zeroed
    // before the CAS (a release) that links it.  An acquire load at agent
    // scope would also invalidate the XCD's L2, once per level.
    uint64_t rw = root_hint ? root_hint : ld_rlx(&f->root);
    bool fresh = root_hint == 0;
    uint64_t spare = 0;
    for (int guard = 0; guard < 16; ++guard) {
        uint32_t h = (uint32_t)(rw >> 40);
        if (rw != 0 && page < tree_cap(h)) break;
        if (!fresh) {
            rw = ld_rlx(&f->root);
            fresh = true;
            continue;
        }
        if (!create) return 0;
        uint64_t n = spare ? spare : node_alloc(g);
        if (!n) { *err = -ENOMEM; return 0; }
        uint64_t nw;
        if (rw == 0) {
            uint32_t nh = 1;
            while (page >= tree_cap(nh)) ++nh;
            st_rlx(&g->nodes[n * GFS_FANOUT], 0ull);   // a spare may hold an earlier attempt's child
            nw = n | ((uint64_t)nh << 40);
        } else {
            st_rlx(&g->nodes[n * GFS_FANOUT], rw & ((1ull << 40) - 1));
            nw = n | ((uint64_t)(h + 1) << 40);
        }
        uint64_t exp = rw;
        if (cas(&f->root, exp, nw)) {
            rw = nw;
            spare = 0;
        } else {
            rw = exp;   // lost the race: keep the node for the next step, if one is needed
            spare = n;
        }
    }
    uint32_t h = (uint32_t)(rw >> 40);
    uint64_t node = rw & ((1ull << 40) - 1);
    for (int lvl = (int)h - 1; lvl >= 1; --lvl) {
        uint64_t* sp = &g->nodes[node * GFS_FANOUT + ((page >> (GFS_FANOUT_LOG * lvl)) & (GFS_FANOUT - 1))];
        uint64_t c = ld_rlx(sp);
        if (!c) {
            if (!create) return 0;
            uint64_t n = node_alloc(g);
            if (!n) { *err = -ENOMEM; return 0; }
            uint64_t exp = 0;
            c = cas(sp, exp, n) ? n : exp;
        }
        node = c;
    }
    return node;
}

// ======================================================================
// Frame allocation (wave-wide).  Lanes with `need` receive a frame whose meta
// word is (new generation | FILLING | pin 1), or GFS_NOFRAME.
// ======================================================================
__device__ inline uint32_t frame_alloc_wave(gfs_dev* g, bool need, uint32_t sweeps) {
    uint32_t my = GFS_NOFRAME;
    uint64_t needm = __ballot(need);
    if (!needm) return my;
    const uint32_t ln = lane();
    const uint32_t N = g->n_frames;
    // never-used frames
    {
        uint32_t k = popc(needm);
        uint32_t leader = first_lane(needm);
        unsigned long long base = 0;
        if (ln == leader) base = fadd_rlx(&g->fresh_next, (unsigned long long)k);
        base = __shfl(base, leader);
        if (need) {
            unsigned long long f = base + mbcnt(needm);
            if (f < N) {
                my = (uint32_t)f;
                st_rlx(&g->meta[my], (GFS_F_FILLING << GFS_ST_SHIFT) | 1ull);
            }
        }
        needm = __ballot(need && my == GFS_NOFRAME);
    }
    // CLOCK sweep with second chance
    const uint32_t rounds = sweeps * ((N + 63) / 64) + 1;
    for (uint32_t r = 0; needm && r < rounds; ++r) {
        unsigned long long h = 0;
        if (ln == first_lane(needm)) h = fadd_rlx(&g->clock_hand, 64ull);
        h = __shfl(h, first_lane(needm));
        uint32_t f = (uint32_t)((h + ln) % N);
        uint64_t m = ld_rlx(&g->meta[f]);
        uint32_t st = meta_state(m);
        bool cand = meta_ref(m) == 0 && !(m & GFS_FL_DIRTY) && (st == GFS_F_FREE || st == GFS_F_VALID);
        if (cand && st == GFS_F_VALID && g->fref[f]) { g->fref[f] = 0; cand = false; }
        uint64_t candm = __ballot(cand);
        cand = cand && mbcnt(candm) < popc(needm);
        bool got = false;
        if (cand) {
            uint64_t nm = ((uint64_t)(meta_gen(m) + 1u) << GFS_GEN_SHIFT) | (GFS_F_FILLING << GFS_ST_SHIFT) | 1ull;
            uint64_t exp = m;
            got = cas(&g->meta[f], exp, nm);
            if (got && st == GFS_F_VALID) count(&g->ctr.evictions, 1);
        }
        uint64_t gotm = __ballot(got);
        uint32_t rank = mbcnt(needm);
        uint32_t src = 0;
        bool take = need && my == GFS_NOFRAME && rank < popc(gotm);
        if (take) src = nth_set(gotm, rank);
        uint32_t fv = __shfl(f, src);
        if (take) my = fv;
        needm = __ballot(need && my == GFS_NOFRAME);
        if (needm && r > 0 && (r % ((N + 63) / 64)) == 0) backoff();
    }
    if (need && my == GFS_NOFRAME) count(&g->ctr.alloc_fail, 1);
    return my;
}

// Give back a frame that was never published (or whose fill failed): bump the
// generation, state -> FREE, drop the owner's pin.  Transient pins of racing
// readers survive the add and are undone by them.
__device__ __forceinline__ void frame_release(gfs_dev* g, uint32_t f) {
    const uint64_t delta = (1ull << GFS_GEN_SHIFT) - (GFS_F_FILLING << GFS_ST_SHIFT) - 1ull;
    fadd(&g->meta[f], delta);
}

// ======================================================================
// RPC client
// ======================================================================
__device__ __forceinline__ uint64_t* slot_word(gfs_dev* g, uint32_t w) {
    return &g->slot_free[(size_t)w * GFS_SLOT_WORD_STRIDE];
}
// One ring slot for one thread, or -1 when none is free.
__device__ inline int rpc_claim(gfs_dev* g, bool need, uint32_t hint) {
    if (!need) return -1;
    const uint32_t W = g->slot_words;
    for (uint32_t a = 0; a < W; ++a) {
        uint32_t w = (hint + a) % W;
        uint64_t* wp = slot_word(g, w);
        uint64_t x = ld_rlx(wp);
        while (x) {
            uint32_t b = first_lane(x);
            uint64_t bit = 1ull << b;
            uint64_t old = __hip_atomic_fetch_and(wp, ~bit, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old & bit) return (int)(w * 64 + b);
            x = old & ~bit;
        }
    }
    return -1;
}
// The lowest k set bits of x.
__device__ __forceinline__ uint64_t low_bits(uint64_t x, uint32_t k) {
    if (popc(x) <= k) return x;
    uint64_t m = 0;
    for (uint32_t i = 0; i < k; ++i) {
        const uint64_t b = x & (0ull - x);
        m |= b;
        x ^= b;
    }
    return m;
}
// Wave-collective claim (call with all 64 lanes): every lane with `need` gets
// a slot, or -1 once none is left; slots go to the needing lanes in lane
// order.  Lane 0 scans the bitmap and takes all the bits the wave wants from
// a word with one atomic, so a wave costs one load and about one atomic per
// word it visits instead of one per lane and word.
__device__ inline int rpc_claim_wave(gfs_dev* g, bool need, uint32_t hint) {
    const uint64_t needm = __ballot(need);
    if (!needm) return -1;
    const uint32_t ln = lane();
    const uint32_t rank = mbcnt(needm);
    const uint32_t want = popc(needm);
    const uint32_t W = g->slot_words;
    uint32_t base = 0;
    int mine = -1;
    for (uint32_t a = 0; a < W && base < want; ++a) {
        const uint32_t w = (hint + a) % W;
        uint64_t* wp = slot_word(g, w);
        uint64_t x = 0;
        if (ln == 0) x = ld_rlx(wp);
        x = __shfl(x, 0);
        while (x && base < want) {
            const uint64_t m = low_bits(x, want - base);
            uint64_t old = 0;
            if (ln == 0) old = __hip_atomic_fetch_and(wp, ~m, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            old = __shfl(old, 0);
            const uint64_t got = old & m;
            const uint32_t c = popc(got);
            if (need && rank >= base && rank < base + c) {
                uint64_t t = got;
                for (uint32_t i = base; i < rank; ++i) t &= t - 1;
                mine = (int)(w * 64 + first_lane(t));
            }
            base += c;
            x = old & ~m;   // bits still free that this pass did not ask for
        }
    }
    return mine;
}
// A request whose wait timed out can still be served by the host later: it
// would pread into the slot's staging page, store the slot's result and set a
// bit in the slot's completion word.  Such a slot is never given back to the
// ring (it stays claimed until the context is destroyed), so no later request
// can meet those late writes, and every ticket outstanding at the host still
// holds a slot (the submission queue's depth relies on that).
__device__ __forceinline__ void rpc_quarantine(gfs_dev* g, uint32_t n) {
    if (n) fadd_rlx(&g->ctr.slots_lost, (unsigned long long)n);
}
__device__ __forceinline__ void rpc_release(gfs_dev* g, int s) {
    __hip_atomic_fetch_or(slot_word(g, (uint32_t)s >> 6), 1ull << (s & 63), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Wave-collective release (call with all 64 lanes): the slots of the lanes
// with `give` are returned with one atomic per bitmap word.
__device__ inline void rpc_release_wave(gfs_dev* g, bool give, int s) {
    uint64_t pend = __ballot(give);
    while (pend) {
        const uint32_t w = (uint32_t)__shfl(s, (int)first_lane(pend)) >> 6;
        const bool mine = ((pend >> lane()) & 1ull) && ((uint32_t)s >> 6) == w;
        uint64_t v = mine ? (1ull << (s & 63)) : 0ull;
        v |= __shfl_xor(v, 1);
        v |= __shfl_xor(v, 2);
        v |= __shfl_xor(v, 4);
        v |= __shfl_xor(v, 8);
        v |= __shfl_xor(v, 16);
        v |= __shfl_xor(v, 32);
        if (lane() == first_lane(pend))
            __hip_atomic_fetch_or(slot_word(g, w), v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        pend &= ~__ballot(mine);
    }
}
// Fill the request fields of slot s and publish it to the host.  `cq` is the
// completion word index | tag << 16: when done the host sets bit `tag` of it.
__device__ inline void rpc_post(gfs_dev* g, int s, uint32_t op, int32_t host_fd, int32_t flags,
                                uint64_t off, uint64_t len, uint32_t cq) {
    gfs_rpc_slot* r = &g->slots[s];
    sys_st(&r->op, op);
    sys_st(&r->host_fd, host_fd);
    sys_st(&r->flags, flags);
    sys_st(&r->cq, cq);
    sys_st(&r->offset, off);
    sys_st(&r->length, len);
    sys_st(&r->result, (int64_t)INT64_MIN);
    fence_rel_sys();
    unsigned long long t = fadd_rlx(&g->sq_tail, 1ull);
    sys_st_rel(&g->sq[t & (g->sq_depth - 1)], (uint64_t)((t << 16) | (uint64_t)s));
}
// Close of a host fd, posted without a slot and not waited for: the fd goes
// into closeq under the request's ticket.  The submission queue has room for a
// request from every slot plus one such close per file-table entry (an entry
// is reopened only through a later ticket, which the host takes after this
// one), so a ticket never lands on an entry the host has not taken yet.
__device__ inline void rpc_post_close(gfs_dev* g, int32_t host_fd) {
    const unsigned long long t = fadd_rlx(&g->sq_tail, 1ull);
    sys_st(&g->closeq[t & (g->sq_depth - 1)], (int64_t)host_fd);
    fence_rel_sys();
    sys_st_rel(&g->sq[t & (g->sq_depth - 1)], (uint64_t)((t << 16) | (uint64_t)GFS_SQ_NOSLOT));
}
// Wait until every bit of `want` is set in completion mask `cq`.  Returns 0
// or -ETIMEDOUT.
// Polls with relaxed loads and acquires once at the end: a system-scope
// acquire load also invalidates the caches on every poll.
__device__ inline int rpc_wait(gfs_dev* g, uint32_t cq, uint64_t want) {
    uint64_t t0 = now_ticks();
    uint32_t spins = 0;
    while ((sys_ld(&g->cq[cq]) & want) != want) {
        if ((++spins & 255) == 0 && now_ticks() - t0 > g->wait_ticks) return -ETIMEDOUT;
        backoff();
    }
    fence_acq_sys();
    return 0;
}
// One request from a single thread, posted: claims a slot, writes the name
// (if any) and posts.  Returns the slot, whose answer rpc_finish_single
// waits for, or -ETIMEDOUT (no slot within the wait limit).
__device__ inline int rpc_post_single(gfs_dev* g, uint32_t op, int32_t host_fd, int32_t flags, uint64_t off,
                                      uint64_t len, const char* name) {
    int s = -1;
    GFS_TRACE(23);
    uint64_t t0 = now_ticks();
    uint32_t hint = blockIdx.x * 7u + blockIdx.y * 13u;
    while ((s = rpc_claim(g, true, hint)) < 0) {
        if (now_ticks() - t0 > g->wait_ticks) return -ETIMEDOUT;
        backoff();
    }
    if (name) {
        // One word at a time, each store waited for before the next: with
        // every store of the name issued at once, 64 workgroups opening
        // different files together took 15 us longer per launch.
        gfs_rpc_slot* r = &g->slots[s];
        for (int i = 0; i < GFS_NAME_MAX; i += 8) {
            uint64_t w = 0;
            for (int b = 0; b < 8; ++b) w |= (uint64_t)(uint8_t)name[i + b] << (8 * b);
            sys_st((uint64_t*)(r->name + i), w);
        }
    }
    sys_st(&g->cq[s], 0u);
    GFS_TRACE(20);
    rpc_post(g, s, op, host_fd, flags, off, len, (uint32_t)s);
    GFS_TRACE(21);
    return s;
}
// Waits for the answer to slot s of rpc_post_single; returns the request's
// result.  *slot_out: the slot, left claimed for the caller to read extra
// results from and release, or -1 (released, or quarantined on a timeout).
__device__ inline int64_t rpc_finish_single(gfs_dev* g, int s, int* slot_out) {
    int w = rpc_wait(g, (uint32_t)s, 1ull);
    GFS_TRACE(22);
    if (w) {
        // the host still owns the slot: it is not returned, and no result is read from it
        rpc_quarantine(g, 1);
        if (slot_out) *slot_out = -1;
        return (int64_t)w;
    }
    const int64_t res = sys_ld(&g->slots[s].result);
    if (slot_out) *slot_out = s;   // caller reads extra results, then releases
    else rpc_release(g, s);
    return res;
}
// One synchronous request from a single thread (open / close / truncate / write).
__device__ inline int64_t rpc_call_single(gfs_dev* g, uint32_t op, int32_t host_fd, int32_t flags,
                                          uint64_t off, uint64_t len, const char* name, int* slot_out) {
    const int s = rpc_post_single(g, op, host_fd, flags, off, len, name);
    if (s < 0) return s;
    return rpc_finish_single(g, s, slot_out);
}

// ======================================================================
// File table
// ======================================================================
__device__ __forceinline__ bool fd_valid(gfs_dev* g, int fd) {
    return fd >= 0 && (unsigned long long)fd < ld_rlx(&g->files_next) && (uint32_t)fd < g->max_files;
}
// A path as GFS_NAME_MAX/8 little-endian words, zero from the terminator on.
// (len: strlen, or GFS_NAME_MAX when no terminator in the first GFS_NAME_MAX bytes)
typedef ft_name gname;
__device__ __forceinline__ uint64_t zero_bytes(uint64_t v) {
    // high bit of each zero byte; the lowest flagged byte is always a zero byte
    return (v - 0x0101010101010101ull) & ~v & 0x8080808080808080ull;
}
// Reads the NUL-terminated path `s` with aligned 8-byte loads issued together
// instead of one dependent load per byte.  First round: the words inside the
// 4 KiB page of `s`.  Second round, only when the first found no terminator at
// or after `s`: the remaining words, in the next page, which the string then
// reaches.  So no load touches a page that holds no byte of the string.
__device__ inline void name_load(const char* s, gname& n) {
    constexpr int NW = GFS_NAME_MAX / 8 + 1;   // covers GFS_NAME_MAX bytes at any alignment
    const uintptr_t a = (uintptr_t)s;
    const uint64_t* W = (const uint64_t*)(a & ~(uintptr_t)7);
    const uint32_t sh = (uint32_t)(a & 7) * 8;
    const uintptr_t page_end = (a | 4095) + 1;
    uint64_t raw[NW];
    bool term = false;
#pragma unroll
    for (int k = 0; k < NW; ++k) raw[k] = ((uintptr_t)(W + k) < page_end) ? W[k] : 0ull;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        if ((uintptr_t)(W + k) < page_end) {
            uint64_t v = raw[k];
            if (k == 0 && sh) v |= (1ull << sh) - 1;   // bytes before s are not part of the string
            term |= zero_bytes(v) != 0;
        }
    }
    if (!term) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
            if ((uintptr_t)(W + k) >= page_end) raw[k] = W[k];
    }
    int len = GFS_NAME_MAX;
#pragma unroll
    for (int i = GFS_NAME_MAX / 8 - 1; i >= 0; --i) {
        const uint64_t v = sh ? ((raw[i] >> sh) | (raw[i + 1] << (64 - sh))) : raw[i];
        n.w[i] = v;
        const uint64_t z = zero_bytes(v);
        if (z) len = 8 * i + (int)((__ffsll((unsigned long long)z) - 1) >> 3);
    }
    uint64_t h = 0xCBF29CE484222325ull ^ (uint64_t)len;
#pragma unroll
    for (int i = 0; i < GFS_NAME_MAX / 8; ++i) {
        const int lo = 8 * i;
        if (len <= lo) n.w[i] = 0;
        else if (len < lo + 8) n.w[i] &= (1ull << (8 * (len - lo))) - 1;
        h = (h ^ n.w[i]) * 0x100000001B3ull;
    }
    n.len = len;
    n.hash = (uint32_t)(h ^ (h >> 32));
}

// fstate.h policy: agent-scope atomics, the context's wait limit.
struct fst_dev {
    uint64_t t0, limit;
    uint32_t spins = 0;
    __device__ uint64_t fadd(uint64_t* w, uint64_t v) { return gfs::fadd(w, v); }
    __device__ uint64_t load(uint64_t* w) { return ld_acq(w); }
    __device__ uint64_t load_rlx(uint64_t* w) { return ld_rlx(w); }
    __device__ bool cas(uint64_t* w, uint64_t& exp, uint64_t v) { return gfs::cas(w, exp, v); }
    __device__ bool expired() { return now_ticks() - t0 > limit; }
    // Waits for another workgroup's host open or close: after a few short
    // pauses, longer ones (about 1.5 us), so that hundreds of waiting
    // workgroups do not crowd the L2 channel that holds the opener's word.
    __device__ void pause() {
        if (spins < 4) {
            ++spins;
            backoff();
        } else {
            __builtin_amdgcn_s_sleep(48);
        }
    }
};

// The rare paths of the file table, kept out of line so that the open path
// inlined into every kernel stays small.
__device__ __noinline__ void tree_drop(gfs_dev* g, uint64_t rw) {
    ft_dev F{0, 0};
    ft_tree_free(g, rw, F);
}
__device__ __noinline__ int reclaim_files(gfs_dev* g, uint64_t t0) {
    ft_dev F{t0, g->wait_ticks};
    return ft_reclaim(g, F);
}

// Host open of entry e (new or reopened), whose state is OPENING, in two
// halves: open_host_post posts the request (returns its slot or a negative
// code), open_host_finish waits for the answer, fills the host fields and
// moves the state to OPEN, or to CLOSED dropping this caller's open on
// failure.  Between the two the entry stays OPENING: joiners wait, and only
// this caller may change its fields.
__device__ inline int open_host_post(gfs_dev* g, gfs_file* e, int flags) {
    return rpc_post_single(g, GFS_OP_OPEN, -1, flags, 0, 0, e->name);
}
__device__ inline int open_host_finish(gfs_dev* g, gfs_file* e, int flags, bool reopen, int posted) {
    int s = -1;
    int64_t r = posted < 0 ? (int64_t)posted : rpc_finish_single(g, posted, &s);
    count(&g->ctr.rpc_opens, 1);
    uint64_t size = 0, ct = 0, mt = 0, ino = 0;
    if (s >= 0) {
        gfs_rpc_slot* sl = &g->slots[s];
        size = sys_ld(&sl->size), ct = sys_ld(&sl->ctime_ns), mt = sys_ld(&sl->mtime_ns), ino = sys_ld(&sl->ino);
        rpc_release(g, s);
    }
    GFS_TRACE(24);
    fst_dev P{0, 0};
    if (s < 0 || r < 0) {
        fst_open_done(&e->st, P, false, false);
        return (int)r;
    }
    bool changed = false;
    if (reopen && (e->ctime_ns != ct || e->mtime_ns != mt || e->size != size || e->ino != ino)) {
        // the file changed on the host: drop every cached page of the old
        // version (the entry was CLOSED, so nobody walks its old tree)
        changed = true;
        const uint64_t rw = ld_rlx(&e->root);
        st_rlx(&e->root, 0ull);
        tree_drop(g, rw);
    }
    e->host_fd = (int32_t)r;
    e->flags = flags;
    e->size = size;
    e->ctime_ns = ct;
    e->mtime_ns = mt;
    e->ino = ino;
    e->dirty = 0;
    fst_open_done(&e->st, P, true, changed);
    // after the state: a reader that saw OPENING has not seen this increment
    __hip_atomic_fetch_add(&e->opened, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
}
__device__ inline int open_host(gfs_dev* g, gfs_file* e, int flags, bool reopen) {
    return open_host_finish(g, e, flags, reopen, open_host_post(g, e, flags));
}

// A reopen of a closed entry whose host open was posted and not waited for
// (open_single with a gfs_open_pend): open_finish completes it.
struct gfs_open_pend {
    int idx, slot, flags;
    uint32_t pending;   // 1: this caller's reopen, open_post / open_finish owe it; 2: a peek (below)
    uint32_t peek_ok;   // in: a caller that finds another caller's reopen at the host may peek
    uint32_t opened0;   // peek: the entry's `opened` count when its state was OPENING
};

// Single-thread open (reference: single_thread_open, libgpufs/fs_calls.cu:152-225).
// The name is looked up (or inserted) and joined inside the open gate, so no
// reclaim pass (ftable.h) can take the entry back in between; a full table
// runs one and looks again.  With `pend`, a reopen of a closed entry returns
// the entry at once with pend->pending = 1, its host open not yet posted
// (open_post, then open_finish; the entry stays OPENING until then); every
// other case completes here.
__device__ __noinline__ int open_single(const char* name, int flags, gfs_open_pend* pend = nullptr) {
    gfs_dev* g = gfs_g;
    gname nm;
    name_load(name, nm);
    if (nm.len == 0) return -ENOENT;
    if (nm.len >= GFS_NAME_MAX) return -ENAMETOOLONG;
    const uint64_t t0 = now_ticks();
    ft_dev F{t0, g->wait_ticks};
    for (;;) {
        int r = ft_gate_enter(g, F);
        if (r) return r;
        uint32_t idx = 0;
        r = ft_find_or_insert(g, nm, F, &idx);
        if (r == FT_NEW) {
            ft_gate_leave(g, F);
            return open_host(g, &g->files[idx], flags, false) < 0 ? E_IPC_OPEN_ERROR : (int)idx;
        }
        if (r == FT_FOUND) {
            fst_dev P{t0, g->wait_ticks};
            GFS_TRACE(25);
            int o;
            if (pend && pend->peek_ok) {
                // Another caller's reopen at the host: take no open and read
                // the cached pages meanwhile (the caller confirms later that
                // an open succeeded after this look and the file did not
                // change).  `opened` is read first, with an acquire, so that a
                // state still OPENING means the count is from before the open.
                gfs_file* e = &g->files[idx];
                pend->opened0 = __hip_atomic_load(&e->opened, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                o = fst_open(&e->st, P, true);
            } else {
                o = fst_open(&g->files[idx].st, P);
            }
            GFS_TRACE(26);
            ft_gate_leave(g, F);
            if (o == FST_JOINED) return (int)idx;
            if (o < 0) return o;
            if (pend && o == FST_PEEK) {
                pend->idx = (int)idx;
                pend->pending = 2;
                return (int)idx;
            }
            if (pend) {
                pend->idx = (int)idx;
                pend->flags = flags;
                pend->slot = -ETIMEDOUT;   // until open_post
                pend->pending = 1;
                return (int)idx;
            }
            return open_host(g, &g->files[idx], flags, true) < 0 ? E_IPC_OPEN_ERROR : (int)idx;
        }
        ft_gate_leave(g, F);
        if (r != E_FSTABLE_FULL) return r;
        const int c = reclaim_files(g, t0);
        if (c < 0) return c;
        if (c == 0) return E_FSTABLE_FULL;
        if (F.expired()) return -ETIMEDOUT;
    }
}

// Posts the host open of a pending reopen (one thread).
__device__ inline void open_post(gfs_open_pend* pend) {
    pend->slot = open_host_post(gfs_g, &gfs_g->files[pend->idx], pend->flags);
}
// Completes a posted reopen (one thread): the entry's index, or
// E_IPC_OPEN_ERROR / a negative code when the host open failed.
__device__ __noinline__ int open_finish(gfs_open_pend* pend) {
    gfs_dev* g = gfs_g;
    pend->pending = 0;
    return open_host_finish(g, &g->files[pend->idx], pend->flags, true, pend->slot) < 0 ? E_IPC_OPEN_ERROR
                                                                                          : pend->idx;
}

} // namespace gfs

static __shared__ gfs::wg_ctl gfs_ctl;
static __shared__ gfs::wg_hist gfs_hist;
static __shared__ gfs::wg_memo gfs_memo;

namespace gfs {

// Workgroup-uniform wait-limit check: thread 0 compares its own clock reading
// with its own start time and broadcasts the verdict, so every thread takes
// the same branch around the barriers that follow.  Call with all threads.
__device__ inline bool wg_timed_out(gfs_dev* g, uint64_t t0) {
    __syncthreads();
    if (tid() == 0) {
        const bool late = now_ticks() - t0 > g->wait_ticks;
        gfs_ctl.tmo = late ? 1u : 0u;
        if (late) count(&g->ctr.timeouts, 1);
    }
    __syncthreads();
    return uni(gfs_ctl.tmo) != 0;
}

// Workgroup-uniform: more than three quarters of the frames are dirty.  A
// writer that sees it writes its file back before it dirties more pages, so
// readers keep finding clean frames to evict (the eviction sweep takes only
// clean ones).  Call with all threads.
__device__ inline bool wg_dirty_high(gfs_dev* g) {
    __syncthreads();
    if (tid() == 0) {
        const long long d = (long long)ld_rlx(&g->dirty_frames);
        gfs_ctl.tmo = d > (long long)(g->n_frames - (g->n_frames >> 2)) ? 1u : 0u;
    }
    __syncthreads();
    return uni(gfs_ctl.tmo) != 0;
}

// ======================================================================
// Page resolution for a batch of up to 64 consecutive pages (wave 0 only).
// On return gfs_ctl.kind[i] is HIT (pinned VALID frame), OWN (pinned FILLING
// frame this workgroup must fill; ring slot in slot[i]), WAIT (another
// workgroup is filling it, or no slot/frame was free) or DONE (not pending).
// `zero_fill_ok`: pages that will be fully overwritten or created zeroed do not
// need host data; they come back OWN with slot[i] = 0xFFFFFFFF.
// ======================================================================
__device__ inline void resolve_batch(gfs_dev* g, int fd, uint64_t pb, uint32_t cnt, uint64_t pending,
                                     uint64_t nohost_mask, bool for_write, uint64_t root_hint = 0) {
    const uint32_t ln = lane();
    gfs_file* file = &g->files[fd];
    const bool mine = ((pending >> ln) & 1ull) != 0 && ln < cnt;
    const uint64_t page = pb + ln;
    // leaf nodes: at most two distinct leaves in a 64-page window
    uint64_t leaf0 = pb >> GFS_FANOUT_LOG, leaf1 = (pb + (cnt ? cnt - 1 : 0)) >> GFS_FANOUT_LOG;
    uint64_t n0 = 0, n1 = 0;
    int err = 0;
    if (ln == 0) {
        n0 = tree_leaf(g, file, leaf0, true, &err, root_hint);
        n1 = (leaf1 == leaf0) ? n0 : tree_leaf(g, file, leaf1, true, &err, root_hint);
    }
    n0 = __shfl(n0, 0);
    n1 = __shfl(n1, 0);
    err = __shfl(err, 0);
    if (err) {
        if (ln == 0) gfs_ctl.err = err;
        if (ln < 64) gfs_ctl.kind[ln] = GFS_K_DONE;
        return;
    }
    uint64_t node = ((page >> GFS_FANOUT_LOG) == leaf0) ? n0 : n1;
    uint64_t* sp = &g->nodes[node * GFS_FANOUT + (page & (GFS_FANOUT - 1))];
    uint32_t kind = GFS_K_DONE, f = GFS_NOFRAME, fgen = 0;
    uint64_t v = 0;
    bool newly_dirty = false;
    if (mine) {
        v = ld_rlx(sp);
        kind = GFS_K_MISS;
        if (v == GFS_LEAF_RESERVED) {
            kind = GFS_K_WAIT;   // another workgroup is claiming a frame for it
        } else if (v) {
            uint32_t vf = leaf_frame(v), vg = leaf_gen(v);
            uint64_t old = fadd_rlx(&g->meta[vf], 1ull);
            if (meta_gen(old) == vg && meta_state(old) == GFS_F_VALID) {
                kind = GFS_K_HIT;
                f = vf;
                fgen = vg;
                g->fref[vf] = 1;   // CLOCK reference bit: a store, no load to wait for
                if (for_write) newly_dirty = set_dirty(g, vf);
            } else {
                fadd_rlx(&g->meta[vf], ~0ull);
                if (meta_gen(old) == vg && meta_state(old) == GFS_F_FILLING) kind = GFS_K_WAIT;
            }
        }
    }
    // A miss that needs host data claims a ring slot; it then reserves the
    // leaf slot, so of several workgroups missing the same page only one
    // claims a frame (frames claimed by the losers would churn the pool and
    // push useful pages out); a loser returns its ring slot and waits as for
    // a fill in flight.  A miss without a ring slot leaves the leaf alone.
    const bool nohost = ((nohost_mask >> ln) & 1ull) != 0;
    int s = rpc_claim_wave(g, kind == GFS_K_MISS && !nohost, blockIdx.x * 5u);
    bool resv = false;
    if (kind == GFS_K_MISS && (nohost || s >= 0)) {
        uint64_t exp = v;
        resv = cas(sp, exp, GFS_LEAF_RESERVED);
    }
    if (kind == GFS_K_MISS && !resv) {
        if (s >= 0) rpc_release(g, s);
        s = -1;
        kind = GFS_K_WAIT;
    }
    uint32_t nf = frame_alloc_wave(g, resv, 2);
    if (resv) {
        if (nf == GFS_NOFRAME) {
            if (s >= 0) rpc_release(g, s);
            s = -1;
            st_rel(sp, v);   // give the leaf back as it was
            kind = GFS_K_WAIT;
        } else {
            g->fkey[nf] = gfs_key((uint32_t)fd, page);
            g->fbytes[nf] = 0;
            const uint32_t gen = meta_gen(ld_rlx(&g->meta[nf]));
            st_rel(sp, leaf_pack(nf, gen));   // the reservation is ours: publish the frame
            kind = GFS_K_OWN;
            f = nf;
            fgen = gen;
            g->fref[nf] = 1;
        }
    }
    wave_count(&g->ctr.page_waits, kind == GFS_K_WAIT && mine);
    wave_count(&g->ctr.page_hits, kind == GFS_K_HIT);
    wave_count(&g->ctr.page_misses, kind == GFS_K_OWN);
    dirty_tally(g, newly_dirty, true);
    {
        const uint64_t hm = __ballot(kind == GFS_K_HIT), om = __ballot(kind == GFS_K_OWN);
        if (ln == 0) { gfs_ctl.hits += popc(hm); gfs_ctl.misses += popc(om); }
    }
    fence_acq_agent();
    if (ln < 64) {
        gfs_ctl.kind[ln] = (uint8_t)kind;
        gfs_ctl.frame[ln] = f;
        gfs_ctl.gen[ln] = fgen;
        gfs_ctl.slot[ln] = (kind == GFS_K_OWN && s >= 0) ? (uint32_t)s : 0xFFFFFFFFu;
    }
    // publish read requests for OWN pages that need host data
    bool post = kind == GFS_K_OWN && s >= 0;
    uint64_t postm = __ballot(post);
    uint32_t cq = 0;
    if (postm) {
        uint32_t leader = first_lane(postm);
        cq = (uint32_t)__shfl(s, leader);
        if (ln == leader) sys_st(&g->cq[cq], 0u);
        if (post) {
            // A reader asks for the bytes below the file size.  A writer asks
            // for the whole page: other writers of the file may have written
            // back bytes past the size it knows, and a page it reads in is
            // written back whole.  Past the host's end of file the read is
            // short and the frame is zero there.
            uint64_t poff = page << g->log_page;
            uint64_t plen = g->page_size;
            if (!for_write && poff + plen > file->size) plen = file->size > poff ? file->size - poff : 0;
            rpc_post(g, s, GFS_OP_READ, file->host_fd, 0, poff, plen, cq | (ln << 16));
        }
        wave_count(&g->ctr.rpc_reads, post);
    }
    if (ln == 0) { gfs_ctl.own = postm; gfs_ctl.rel = 0; gfs_ctl.cq = cq; }
}

// Drop the pins taken by resolve_batch; OWN frames are published VALID first
// (or released when their fill failed).
__device__ inline void finish_batch(gfs_dev* g, uint32_t cnt) {
    const uint32_t ln = lane();
    if (ln >= cnt) return;
    uint32_t k = gfs_ctl.kind[ln];
    uint32_t f = gfs_ctl.frame[ln];
    if (k == GFS_K_HIT) {
        fadd(&g->meta[f], ~0ull);
    } else if (k == GFS_K_OWN) {
        // FILLING -> VALID and drop the owner's pin in one release add
        fadd(&g->meta[f], (GFS_F_VALID - GFS_F_FILLING) * (1ull << GFS_ST_SHIFT) - 1ull);
    }
    if (gfs_ctl.slot[ln] != 0xFFFFFFFFu && !((gfs_ctl.rel >> ln) & 1ull)) rpc_release(g, (int)gfs_ctl.slot[ln]);
}

// Undo OWN frames whose fill failed: clear the leaf and free the frame.
__device__ inline void fail_own(gfs_dev* g, int fd, uint64_t pb, uint32_t cnt) {
    const uint32_t ln = lane();
    if (ln >= cnt || gfs_ctl.kind[ln] != GFS_K_ERR) return;
    uint32_t f = gfs_ctl.frame[ln];
    int err = 0;
    uint64_t page = pb + ln;
    uint64_t node = tree_leaf(g, &g->files[fd], page >> GFS_FANOUT_LOG, false, &err);
    if (node) {
        uint64_t* sp = &g->nodes[node * GFS_FANOUT + (page & (GFS_FANOUT - 1))];
        uint64_t exp = leaf_pack(f, meta_gen(ld_rlx(&g->meta[f])));
        cas(sp, exp, 0ull);
    }
    frame_release(g, f);
    if (gfs_ctl.slot[ln] != 0xFFFFFFFFu && !((gfs_ctl.rel >> ln) & 1ull)) rpc_release(g, (int)gfs_ctl.slot[ln]);
    gfs_ctl.slot[ln] = 0xFFFFFFFFu;
}

__device__ __forceinline__ uint8_t* frame_ptr(gfs_dev* g, uint32_t f) {
    return g->pool + ((uint64_t)f << g->log_page);
}
__device__ __forceinline__ uint8_t* staging_ptr(gfs_dev* g, uint32_t s) {
    return g->staging + ((uint64_t)s << g->log_page);
}

// Host reads of the batch's OWN pages, streamed (all threads): thread 0 waits
// for any read not yet taken, then the workgroup copies every page that has
// arrived from staging into its frame and returns those slots to the ring,
// while the host works on the rest.  The slot whose completion word the batch
// uses is kept to the end (finish_batch returns it).  Pages whose read failed
// or came up short are marked ERR and gfs_ctl.err is set.
__device__ inline void fill_from_host(gfs_dev* g, uint32_t cnt, uint64_t pb) {
    (void)cnt;
    (void)pb;
    const uint64_t own = uni(gfs_ctl.own);
    if (own == 0) return;
    const uint32_t cq = uni(gfs_ctl.cq);
    const uint32_t T = nthreads(), t = tid();
    uint64_t done = 0, filled = 0;
    uint32_t rot = 0;
    while (done != own) {
        if (t == 0) {
            const uint64_t t0 = now_ticks();
            uint32_t spins = 0;
            uint64_t m;
            gfs_ctl.tmo = 0u;
            // relaxed polls, one acquire once reads have arrived (an acquire
            // load at system scope would invalidate the caches on every poll)
            while (((m = sys_ld(&g->cq[cq])) & own & ~done) == 0) {
                if ((++spins & 255) == 0 && now_ticks() - t0 > g->wait_ticks) {
                    gfs_ctl.tmo = 1u;
                    count(&g->ctr.timeouts, 1);
                    m = own;   // everything not yet taken fails
                    break;
                }
                backoff();
            }
            fence_acq_sys();
            gfs_ctl.a = m & own & ~done;
        }
        __syncthreads();
        GFS_TRACE(13);
        const uint64_t got = uni(gfs_ctl.a);
        const bool tmo = uni(gfs_ctl.tmo) != 0u;
        if (wave() == 0) {
            const uint32_t ln = lane();
            if ((got >> ln) & 1ull) {
                // the host's bytes of the page: short at the host's end of
                // file (the rest of the frame is zeroed below)
                int64_t r = tmo ? (int64_t)-ETIMEDOUT : sys_ld(&g->slots[gfs_ctl.slot[ln]].result);
                if (r < 0 || (uint64_t)r > g->page_size) {
                    gfs_ctl.kind[ln] = GFS_K_ERR;
                    atomicCAS(&gfs_ctl.err, 0, r < 0 ? (int)r : -EIO);
                } else {
                    g->fbytes[gfs_ctl.frame[ln]] = (uint32_t)r;
                }
            }
            if (tmo) {
                // Requests not answered yet may still be served: their slots and
                // the slot holding the batch's completion word are quarantined
                // (fail_own / finish_batch skip a slot index of all ones).
                const bool lost = ((got >> ln) & 1ull) ||
                                  (((own >> ln) & 1ull) && ln < 64 && gfs_ctl.slot[ln] == cq);
                if (lost) gfs_ctl.slot[ln] = 0xFFFFFFFFu;
                const uint64_t lm = __ballot(lost);
                if (ln == 0) rpc_quarantine(g, popc(lm));
            }
        }
        __syncthreads();
        for (uint64_t m = got; m; m &= m - 1) {
            const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1;
            if (uni((uint32_t)gfs_ctl.kind[i]) != GFS_K_OWN) continue;
            const uint32_t f = gfs_ctl.frame[i];
            const uint64_t bytes = g->fbytes[f];
            // four loads in flight per lane: this copy is inlined into every
            // read loop, and eight would cost the warm path register spills
            copy_seg<true, false, SPACE_GLOBAL, SPACE_GLOBAL, 4>(frame_ptr(g, f), staging_ptr(g, gfs_ctl.slot[i]),
                                                                 bytes, (t + T - rot) % T, T);
            if (bytes < g->page_size) zero_tail(frame_ptr(g, f), bytes, g->page_size, t, T);
            rot = (rot + (uint32_t)(bytes >> 4)) % T;
            filled += bytes;
        }
        // every staging load of the round has returned (its data has been stored)
        __syncthreads();
        GFS_TRACE(12);
        if (wave() == 0) {
            const uint32_t ln = lane();
            const bool give = ((got >> ln) & 1ull) && gfs_ctl.kind[ln] == GFS_K_OWN && gfs_ctl.slot[ln] != cq;
            rpc_release_wave(g, give, give ? (int)gfs_ctl.slot[ln] : 0);
            const uint64_t gm = __ballot(give);
            if (ln == 0) gfs_ctl.rel |= gm;
        }
        done |= got;
        if (tmo) break;
    }
    if (t == 0 && filled) count(&g->ctr.bytes_filled, filled);
    __threadfence();
    __syncthreads();
}

} // namespace gfs

// ======================================================================
// Public device API (reference names and signatures)
// ======================================================================

// libgpufs/fs_calls.cu:227-236
__device__ inline int gopen(const char* filename, int flags) {
    __syncthreads();
    if (gfs::tid() == 0) {
        gfs_hist.fd = -1;
        gfs_memo.fd = -1;
        gfs_ctl.ret = gfs::open_single(filename, flags);
    }
    __syncthreads();
    return (int)gfs::uni(gfs_ctl.ret);
}

// libgpufs/fs_calls.cu:851-854 (per-thread, not collective)
__device__ inline size_t fstat(int fd) {
    gfs_dev* g = gfs_g;
    if (!gfs::fd_valid(g, fd)) return (size_t)(-EBADF);
    return (size_t)g->files[fd].size;
}

// libgpufs/fs_calls.cu:689-737.  Reads [offset, offset+size) clipped at end of
// file into `buffer` (global or LDS); returns the bytes read.  A caller that
// knows where `buffer` lives may say so (gread<gfs::SPACE_LDS>(...)): only that
// copy path is then compiled into it.
template <gfs::buf_space buffer_space = gfs::SPACE_ANY>
__device__ inline size_t gread(int fd, size_t offset, size_t size, uchar* buffer) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    // No barrier before thread 0's round trip: it writes only its own entry
    // fields, which no wave reads after the barrier that ends the previous
    // call, and it overlaps the other waves' last work before this call; the
    // barrier after it also keeps `buffer` until every wave is done with it.
    GFS_TRACE(5);
    uint32_t version = 0;   // thread 0: the file's version at entry
    if (tid() == 0) {
        // state, size and radix root in one round trip (relaxed loads, one
        // acquire fence); the caller's gopen already acquired what the opener
        // published.  A read inside the page of the workgroup's previous read
        // also pins that page's frame in the same round trip (wg_memo).
        int64_t r = 0;
        uint64_t fs = 0, root = 0;
        uint32_t mf = GFS_NOFRAME, mf1 = GFS_NOFRAME;
        if (!fd_valid(g, fd)) r = -EBADF;
        else {
            gfs_file* e = &g->files[fd];
            const uint64_t pin = g->page_size - (offset & (g->page_size - 1));
            // A read of one page, or running into the next, whose pages are
            // both in the memo batch.  LDS is not cleared between launches: a
            // memo this launch did not write is bounded to the batch arrays,
            // used only if its frame indexes are in range, and each frame's
            // owner key must name its page.
            const uint64_t p0 = offset >> g->log_page;
            const uint64_t mi = p0 - gfs_memo.pb;   // the first page's lane in the memo batch
            const bool two = size > pin;
            bool try_memo = gfs_memo.fd == fd && size > 0 && (!two || size - pin <= g->page_size) && mi < 63 &&
                            mi + (two ? 2 : 1) <= gfs_memo.cnt;
            uint32_t mfr0 = GFS_NOFRAME, mfr1 = GFS_NOFRAME, mgen0 = 0, mgen1 = 0;
            if (try_memo) {
                const uint32_t k0 = gfs_ctl.kind[mi];
                mfr0 = gfs_ctl.frame[mi];
                mgen0 = gfs_ctl.gen[mi];
                try_memo = (k0 == GFS_K_HIT || k0 == GFS_K_OWN) && mfr0 < g->n_frames;
                if (two) {
                    const uint32_t k1 = gfs_ctl.kind[mi + 1];
                    mfr1 = gfs_ctl.frame[mi + 1];
                    mgen1 = gfs_ctl.gen[mi + 1];
                    try_memo = try_memo && (k1 == GFS_K_HIT || k1 == GFS_K_OWN) && mfr1 < g->n_frames;
                }
            }
            const uint64_t st = ld_rlx(&e->st);
            fs = e->size;
            root = ld_rlx(&e->root);
            uint64_t mm0 = 0, mk0 = 0, mm1 = 0, mk1 = 0;
            if (try_memo) {
                mm0 = fadd_rlx(&g->meta[mfr0], 1ull);
                mk0 = g->fkey[mfr0];
                if (two) {
                    mm1 = fadd_rlx(&g->meta[mfr1], 1ull);
                    mk1 = g->fkey[mfr1];
                }
            }
            fence_acq_agent();   // the nodes under `root` (and the pinned frames) are read after it
            version = fst_version(st);
            if (fst_state(st) != GFS_FS_OPEN) r = -EBADF;
            else r = offset >= fs ? 0 : (int64_t)(size < fs - offset ? size : fs - offset);
            if (try_memo) {
                bool ok = r > 0 && version == gfs_memo.version && meta_gen(mm0) == mgen0 &&
                          meta_state(mm0) == GFS_F_VALID && mk0 == gfs_key((uint32_t)fd, p0);
                if (two)
                    ok = ok && meta_gen(mm1) == mgen1 && meta_state(mm1) == GFS_F_VALID &&
                         mk1 == gfs_key((uint32_t)fd, p0 + 1);
                if (ok) {
                    mf = mfr0;
                    mf1 = two ? mfr1 : GFS_NOFRAME;
                } else {
                    fadd_rlx(&g->meta[mfr0], ~0ull);
                    if (two) fadd_rlx(&g->meta[mfr1], ~0ull);
                }
            }
        }
        // read-ahead: the third and later forward-sequential reads of a file
        // that reach a page not fetched yet take up to GFS_READAHEAD_BYTES of
        // the following pages into their last batch (same round trip to the
        // host; only the requested bytes are copied out)
        uint32_t ext = 0;
        if (r > 0) {
            const uint32_t Lg = g->log_page;
            const uint64_t pf = offset >> Lg, pl = (offset + (uint64_t)r - 1) >> Lg;
            const bool known = gfs_hist.fd == fd;
            const bool seq = known && pf >= gfs_hist.first && pf <= gfs_hist.last + 1;
            const uint32_t streak = seq ? gfs_hist.streak + 1 : 0;
            uint64_t fetched = known ? gfs_hist.fetched : 0;
            if (mf == GFS_NOFRAME && streak >= 2 && pl >= fetched && !(g->tune_flags & GFS_TUNE_NO_READAHEAD)) {
                const uint64_t npages = (fs + g->page_size - 1) >> Lg;
                uint64_t want = ((uint64_t)GFS_READAHEAD_BYTES) >> Lg;
                if (want < 1) want = 1;
                const uint64_t room = 64 - (((pl - pf) % 64) + 1);   // free lanes of the last batch
                if (want > room) want = room;
                if (want > npages - 1 - pl) want = npages - 1 - pl;
                ext = (uint32_t)want;
            }
            if (pl + 1 + ext > fetched) fetched = pl + 1 + ext;
            gfs_hist.fd = fd;
            gfs_hist.streak = streak;
            gfs_hist.first = pf;
            gfs_hist.last = pl;
            gfs_hist.fetched = fetched;
        }
        gfs_ctl.gr_ret = r;
        gfs_ctl.gr_root = r > 0 ? root : 0;
        gfs_ctl.gr_ra = ext;
        gfs_ctl.memo_f = mf;
        gfs_ctl.memo_f1 = mf1;
    }
    __syncthreads();
    const int64_t n = uni(gfs_ctl.gr_ret);
    const uint64_t root_hint = uni(gfs_ctl.gr_root);
    const uint32_t ra_ext = uni(gfs_ctl.gr_ra);
    const uint32_t memo_f = uni(gfs_ctl.memo_f);   // n > 0 whenever a frame was pinned
    if (n <= 0) {
        // every wave has read the entry fields before thread 0 may rewrite them
        // in a next gread (which writes them before its first barrier)
        __syncthreads();
        return (size_t)n;
    }
    const uint32_t L = g->log_page;
    const uint64_t PS = g->page_size;
    if (memo_f != GFS_NOFRAME) {
        // One page, or two, pinned and checked at entry (thread 0's acquire,
        // then the barrier): copy, meet once, and thread 0 drops the pins.
        const uint32_t memo_f1 = uni(gfs_ctl.memo_f1);
        const uint64_t in0 = PS - (offset & (PS - 1));
        const uint64_t n0 = (uint64_t)n < in0 ? (uint64_t)n : in0;
        copy_seg<true, true, buffer_space, SPACE_GLOBAL>(buffer, frame_ptr(g, memo_f) + (offset & (PS - 1)), n0,
                                                         tid(), nthreads());
        if ((uint64_t)n > n0)   // the rest lies in the next page (memo_f1 is set whenever n > n0)
            copy_seg<true, true, buffer_space, SPACE_GLOBAL>(buffer + n0, frame_ptr(g, memo_f1), (uint64_t)n - n0,
                                                             tid(), nthreads());
        __syncthreads();
        if (tid() == 0) {
            // every load of the frames has returned (the barrier waited for the
            // copy's stores, which used them): a relaxed add that nothing
            // waits for drops the pins, and thread 0 is not held up
            fadd_rlx(&g->meta[memo_f], ~0ull);
            g->fref[memo_f] = 1;
            uint32_t pages = 1;
            if (memo_f1 != GFS_NOFRAME) {
                fadd_rlx(&g->meta[memo_f1], ~0ull);
                if ((uint64_t)n > n0) {
                    g->fref[memo_f1] = 1;
                    pages = 2;
                }
            }
            gfs_ctl.hits += pages;
            count(&g->ctr.page_hits, pages);
        }
        GFS_TRACE(6);
        return (size_t)n;
    }
    const bool page_sync = (uni(g->tune_flags) & GFS_TUNE_PAGE_SYNC) != 0;
    const uint64_t p_first = offset >> L, p_last = (offset + n - 1) >> L;
    const uint32_t bmax = 64;
    int err = 0;
    for (uint64_t pb = p_first; pb <= p_last && !err; pb += bmax) {
        // requested pages of this batch, then the read-ahead pages after the last one
        const uint32_t creq = (uint32_t)((p_last - pb + 1) < bmax ? (p_last - pb + 1) : bmax);
        const uint32_t cnt = creq + (pb + creq - 1 == p_last ? ra_ext : 0u);
        const uint64_t reqm = creq == 64 ? ~0ull : ((1ull << creq) - 1);
        uint64_t pending = cnt == 64 ? ~0ull : ((1ull << cnt) - 1);
        uint32_t idle = 0;
        uint64_t t0 = now_ticks();
        while (pending && !err) {
            __syncthreads();
            if (wave() == 0) {
                if (lane() == 0) gfs_ctl.err = 0;
                resolve_batch(g, fd, pb, cnt, pending, 0ull, false, root_hint);
            }
            __syncthreads();
            if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
            GFS_TRACE(1);
            fill_from_host(g, cnt, pb);
            GFS_TRACE(2);
            // copy every requested pinned page (hit or freshly filled) into the caller's buffer
            const uint32_t T = nthreads(), t = tid();
            uint32_t rot = 0;
            for (uint32_t i = 0; i < creq; ++i) {
                uint32_t k = uni((uint32_t)gfs_ctl.kind[i]);
                if (k != GFS_K_HIT && k != GFS_K_OWN) continue;
                uint64_t poff = (pb + i) << L;
                uint64_t lo = poff > offset ? poff : offset;
                uint64_t hi = poff + PS < offset + n ? poff + PS : offset + n;
                copy_seg<true, true, buffer_space, SPACE_GLOBAL>(buffer + (lo - offset),
                                                                 frame_ptr(g, gfs_ctl.frame[i]) + (lo - poff), hi - lo,
                                                                 (t + T - rot) % T, T);
                rot = (rot + (uint32_t)((hi - lo) >> 4)) % T;
                // keep the workgroup's waves on the same page (streams the pool in order);
                // thread 0's agent-scope acquire before the barrier measured ~1 % faster
                if (page_sync) {
                    if (tid() == 0) fence_acq_agent();
                    __syncthreads();
                }
            }
            __syncthreads();
            GFS_TRACE(3);
            if (tid() == 0 && pb + creq - 1 == p_last) {
                // this read's last batch (with its read-ahead pages) becomes the
                // memo for the next reads; its lanes keep kind / frame / gen
                gfs_memo.fd = fd;
                gfs_memo.pb = pb;
                gfs_memo.cnt = cnt;
                gfs_memo.version = version;
            }
            if (wave() == 0) {
                const uint32_t ln = lane();
                if (gfs_ctl.err) {
                    fail_own(g, fd, pb, cnt);
                    // a read-ahead page that failed does not fail the read
                    const uint64_t errm = __ballot(ln < creq && gfs_ctl.kind[ln] == GFS_K_ERR);
                    if (ln == 0 && errm == 0) gfs_ctl.err = 0;
                }
                finish_batch(g, cnt);
                uint64_t waitm = __ballot(ln < cnt && gfs_ctl.kind[ln] == GFS_K_WAIT);
                if (ln == 0) gfs_ctl.a = waitm;
            }
            __syncthreads();
            GFS_TRACE(4);
            if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
            // read-ahead pages are never waited for
            uint64_t nxt = uni(gfs_ctl.a) & reqm;
            if (nxt == pending) {
                if (++idle > 4) backoff();
                if ((idle & 63) == 0 && wg_timed_out(g, t0)) { err = -ETIMEDOUT; break; }
            } else {
                idle = 0;
            }
            pending = nxt;
        }
    }
    __syncthreads();
    GFS_TRACE(6);
    return err ? (size_t)(int64_t)err : (size_t)n;
}

namespace gfs {
// Pins pages [pb, pb+cnt) of open file `fd` (cnt <= 64, all below the end of
// the file) for the caller, reading the missing ones from the host as gread
// does, without copying anything out (workgroup-collective).  On success
// gfs_ctl.pinf[i] holds page pb+i's frame, VALID and pinned, until
// unpin_pages; the caller reads the frames' bytes directly, with
// agent-coherent loads (ld16_coh): a frame may have been filled on another
// XCD, whose L2 is not this one's, and this L2 may still hold lines of the
// frame's previous page.  (An acquire fence instead would invalidate this
// XCD's whole L2 on every pin: measured 25-30 us per pin with 256 workgroups
// pinning at once.)  Returns 0 or a negative code (nothing stays pinned then).
__device__ __noinline__ int pin_pages(gfs_dev* g, int fd, uint64_t pb, uint32_t cnt) {
    const uint64_t all = cnt >= 64 ? ~0ull : ((1ull << cnt) - 1);
    uint64_t pending = all, pinned = 0;
    __syncthreads();
    if (tid() == 0) {
        gfs_memo.fd = -1;   // the batch arrays are reused below
        // relaxed: every tree node and leaf word below it is read atomically
        gfs_ctl.gr_root = ld_rlx(&g->files[fd].root);
    }
    __syncthreads();
    const uint64_t root_hint = uni(gfs_ctl.gr_root);
    const uint64_t t0 = now_ticks();
    uint32_t idle = 0;
    int err = 0;
    while (pending && !err) {
        __syncthreads();
        if (wave() == 0) {
            if (lane() == 0) gfs_ctl.err = 0;
            resolve_batch(g, fd, pb, cnt, pending, 0ull, false, root_hint);
        }
        __syncthreads();
        if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
        fill_from_host(g, cnt, pb);
        if (wave() == 0) {
            const uint32_t ln = lane();
            if (gfs_ctl.err) fail_own(g, fd, pb, cnt);
            const uint32_t k = ln < cnt ? gfs_ctl.kind[ln] : GFS_K_DONE;
            const uint32_t f = gfs_ctl.frame[ln];
            if (k == GFS_K_OWN) {
                // published VALID, the allocation's pin is kept
                fadd(&g->meta[f], (GFS_F_VALID - GFS_F_FILLING) * (1ull << GFS_ST_SHIFT));
                if (gfs_ctl.slot[ln] != 0xFFFFFFFFu && !((gfs_ctl.rel >> ln) & 1ull)) rpc_release(g, (int)gfs_ctl.slot[ln]);
            }
            if (k == GFS_K_HIT || k == GFS_K_OWN) gfs_ctl.pinf[ln] = f;
            const uint64_t got = __ballot(k == GFS_K_HIT || k == GFS_K_OWN);
            const uint64_t waitm = __ballot(k == GFS_K_WAIT);
            if (ln == 0) {
                gfs_ctl.a = waitm;
                gfs_ctl.b = got;
            }
        }
        __syncthreads();
        pinned |= uni(gfs_ctl.b);
        if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
        const uint64_t nxt = uni(gfs_ctl.a);
        if (nxt == pending) {
            if (++idle > 4) backoff();
            if ((idle & 63) == 0 && wg_timed_out(g, t0)) { err = -ETIMEDOUT; break; }
        } else {
            idle = 0;
        }
        pending = nxt;
    }
    if (err) {
        if (wave() == 0) {
            const uint32_t ln = lane();
            if ((pinned >> ln) & 1ull) fadd(&g->meta[gfs_ctl.pinf[ln]], ~0ull);
        }
        __syncthreads();
        return err;
    }
    __syncthreads();
    return 0;
}
// Pins pages [pb, pb+cnt) (cnt <= 64) of entry `fd` only if every one of them
// is cached (VALID in a frame): no frame is claimed, filled or created and no
// host descriptor is used, so the entry need not be open (workgroup-collective).
// For reads that run while the file's reopen is still at the host: the entry
// is then OPENING, its tree is neither grown nor dropped except by the opener
// when the host finds the file changed, which also bumps the file's version;
// the caller compares the version after the open and discards what it read if
// it moved.  Node and frame indexes are bounds-checked, since nodes of a
// dropped tree may be reused under the walk.  Returns 0 (every page pinned,
// frames in gfs_ctl.pinf, released by unpin_pages) or 1 (nothing pinned).
__device__ __noinline__ int pin_cached(gfs_dev* g, int fd, uint64_t pb, uint32_t cnt) {
    __syncthreads();
    if (wave() == 0) {
        const uint32_t ln = lane();
        bool ok = true;
        uint32_t f = GFS_NOFRAME;
        if (ln < cnt) {
            const uint64_t page = pb + ln;
            const uint64_t rw = ld_rlx(&g->files[fd].root);
            const uint32_t h = (uint32_t)(rw >> 40);
            uint64_t node = rw & ((1ull << 40) - 1);
            ok = rw != 0 && h >= 1 && h <= GFS_MAX_HEIGHT && page < tree_cap(h) && node != 0 && node < g->node_cap;
            for (int lvl = (int)h - 1; ok && lvl >= 1; --lvl) {
                const uint64_t c =
                    ld_rlx(&g->nodes[node * GFS_FANOUT + ((page >> (GFS_FANOUT_LOG * lvl)) & (GFS_FANOUT - 1))]);
                ok = c != 0 && c < g->node_cap;
                node = c;
            }
            if (ok) {
                const uint64_t v = ld_rlx(&g->nodes[node * GFS_FANOUT + (page & (GFS_FANOUT - 1))]);
                ok = v != 0 && v != GFS_LEAF_RESERVED && leaf_frame(v) < g->n_frames;
                if (ok) {
                    f = leaf_frame(v);
                    const uint64_t old = fadd_rlx(&g->meta[f], 1ull);
                    ok = meta_gen(old) == leaf_gen(v) && meta_state(old) == GFS_F_VALID;
                    if (!ok) fadd_rlx(&g->meta[f], ~0ull);
                }
            }
        }
        const uint64_t bad = __ballot(ln < cnt && !ok);
        if (ln < cnt) {
            if (bad) {
                if (ok) fadd_rlx(&g->meta[f], ~0ull);   // give back the pins taken
            } else {
                gfs_ctl.pinf[ln] = f;
                g->fref[f] = 1;
            }
        }
        wave_count(&g->ctr.page_hits, ln < cnt && !bad);
        if (ln == 0) {
            gfs_ctl.a = bad;
            gfs_memo.fd = -1;
        }
    }
    __syncthreads();
    return uni(gfs_ctl.a) ? 1 : 0;
}
// Whether page `page` of open file `fd` is cached (VALID in a frame), looked
// up without pinning or creating anything (one thread).  Advisory: the answer
// can change right after.
__device__ inline bool page_cached(gfs_dev* g, int fd, uint64_t page) {
    int e = 0;
    const uint64_t node = tree_leaf(g, &g->files[fd], page >> GFS_FANOUT_LOG, false, &e);
    if (!node) return false;
    const uint64_t v = ld_rlx(&g->nodes[node * GFS_FANOUT + (page & (GFS_FANOUT - 1))]);
    if (v == 0 || v == GFS_LEAF_RESERVED) return false;
    const uint64_t m = ld_rlx(&g->meta[leaf_frame(v)]);
    return meta_gen(m) == leaf_gen(v) && meta_state(m) == GFS_F_VALID;
}
// 16 bytes of a pinned frame, read coherently at agent scope (two relaxed
// agent-scope atomic loads: global_load_dwordx2 sc1, no cache invalidation).
__device__ __forceinline__ v4u ld16_coh(const uint8_t* p) {
    const uint64_t a = __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load((uint64_t*)(p + 8), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v4u{(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
}
// Drops the pins of pin_pages (workgroup-collective); the caller's reads of
// the frames have all returned.
__device__ inline void unpin_pages(gfs_dev* g, uint32_t cnt) {
    __syncthreads();
    if (wave() == 0 && lane() < cnt) {
        const uint32_t f = gfs_ctl.pinf[lane()];
        g->fref[f] = 1;
        fadd_rlx(&g->meta[f], ~0ull);
    }
    __syncthreads();
}
// Copies bytes [off, off+n) of entry `fd` into `dst` from cached pages only
// (pin_cached, at most 64 pages at a time; the entry need not be open), the
// way gread copies pinned pages.  Returns 0, or 1 when a page was not cached
// (the bytes before it are copied).  Workgroup-collective.
__device__ inline int copy_cached(gfs_dev* g, int fd, uint64_t off, uint64_t n, uint8_t* dst) {
    if (n == 0) return 0;
    const uint32_t L = g->log_page;
    const uint64_t PS = g->page_size;
    const uint64_t p_first = off >> L, p_last = (off + n - 1) >> L;
    const uint32_t T = nthreads(), t = tid();
    for (uint64_t pb = p_first; pb <= p_last; pb += 64) {
        const uint32_t cnt = (uint32_t)(p_last - pb + 1 < 64 ? p_last - pb + 1 : 64);
        if (pin_cached(g, fd, pb, cnt)) return 1;
        // the frames may have been filled on another XCD
        if (t == 0) fence_acq_agent();
        __syncthreads();
        uint32_t rot = 0;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint64_t poff = (pb + i) << L;
            const uint64_t lo = poff > off ? poff : off;
            const uint64_t hi = poff + PS < off + n ? poff + PS : off + n;
            copy_seg<true, true, SPACE_GLOBAL, SPACE_GLOBAL>(dst + (lo - off), frame_ptr(g, gfs_ctl.pinf[i]) + (lo - poff),
                                                             hi - lo, (t + T - rot) % T, T);
            rot = (rot + (uint32_t)((hi - lo) >> 4)) % T;
            // as gread: the waves stay on one page
            if (t == 0) fence_acq_agent();
            __syncthreads();
        }
        unpin_pages(g, cnt);
    }
    return 0;
}
} // namespace gfs

namespace gfs {
__device__ inline int flush_file(gfs_dev* g, int fd, uint64_t min_size);
// flush_file out of line, for gwrite's write-back of its own file when the
// cache runs short of clean frames (a rare path: gwrite's loop stays small).
__device__ __noinline__ int flush_file_stalled(gfs_dev* g, int fd, uint64_t upto) {
    return flush_file(g, fd, upto);
}
} // namespace gfs

// Writes [offset, offset+size) of `buffer` through the page cache
// (libgpufs/fs_calls.cu:739-802).  Pages are fetched from the host first
// unless the file is O_GWRONCE (zero-filled) or the page is fully overwritten.
// When more than three quarters of the frames are dirty, or no frame is left
// for a page, the file's dirty pages are written back first.
__device__ inline size_t gwrite(int fd, size_t offset, size_t size, uchar* buffer) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    __syncthreads();
    if (tid() == 0) {
        int64_t r = (int64_t)size;
        if (!fd_valid(g, fd) || fst_state(ld_acq(&g->files[fd].st)) != GFS_FS_OPEN) r = -EBADF;
        else if ((g->files[fd].flags & O_ACCMODE) == O_RDONLY) r = -EBADF;
        gfs_ctl.ret = r;
        gfs_ctl.flag = r > 0 ? (uint32_t)g->files[fd].flags : 0;
    }
    __syncthreads();
    const int64_t n = uni(gfs_ctl.ret);
    const bool wronce = (uni(gfs_ctl.flag) & (O_CREAT | O_ACCMODE)) == (uint32_t)O_GWRONCE;
    if (n <= 0) return (size_t)n;
    const uint32_t L = g->log_page;
    const uint64_t PS = g->page_size;
    const uint64_t p_first = offset >> L, p_last = (offset + n - 1) >> L;
    const uint32_t bmax = 64;
    int err = 0;
    for (uint64_t pb = p_first; pb <= p_last && !err; pb += bmax) {
        const uint32_t cnt = (uint32_t)((p_last - pb + 1) < bmax ? (p_last - pb + 1) : bmax);
        uint64_t pending = cnt == 64 ? ~0ull : ((1ull << cnt) - 1);
        if (wg_dirty_high(g)) {
            err = flush_file_stalled(g, fd, offset + n);
            if (err) break;
        }
        // Pages that need no host data: O_GWRONCE (OR-merged on disk) or
        // fully overwritten.  A page past the end of file is read too: other
        // writers may have written it back and had it evicted since this
        // call read the size, and the page is written back whole.
        uint64_t nohost = 0;
        for (uint32_t i = 0; i < cnt; ++i) {
            uint64_t poff = (pb + i) << L;
            bool full = offset <= poff && offset + n >= poff + PS;
            if (wronce || full) nohost |= 1ull << i;
        }
        uint32_t idle = 0;
        uint64_t t0 = now_ticks();
        while (pending && !err) {
            __syncthreads();
            if (wave() == 0) {
                if (lane() == 0) gfs_ctl.err = 0;
                resolve_batch(g, fd, pb, cnt, pending, nohost, true);
            }
            __syncthreads();
            if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
            fill_from_host(g, cnt, pb);
            const uint32_t T = nthreads(), t = tid();
            // zero the never-filled part of host-less OWN pages, then copy the caller's bytes
            const uint64_t posted = uni(gfs_ctl.own);
            for (uint32_t i = 0; i < cnt; ++i) {
                if (gfs_ctl.kind[i] != GFS_K_OWN || ((posted >> i) & 1ull)) continue;
                uint8_t* fp = frame_ptr(g, gfs_ctl.frame[i]);
                for (uint64_t k = t; k < (PS >> 4); k += T) *(v4u*)(fp + 16 * k) = v4u{0, 0, 0, 0};
            }
            __syncthreads();
            for (uint32_t i = 0; i < cnt; ++i) {
                uint32_t k = gfs_ctl.kind[i];
                if (k != GFS_K_HIT && k != GFS_K_OWN) continue;
                uint64_t poff = (pb + i) << L;
                uint64_t lo = poff > offset ? poff : offset;
                uint64_t hi = poff + PS < offset + n ? poff + PS : offset + n;
                copy_seg<false, false, SPACE_GLOBAL>(frame_ptr(g, gfs_ctl.frame[i]) + (lo - poff), buffer + (lo - offset),
                                                     hi - lo, t, T);
                if (t == 0) atomicMax(&g->fbytes[gfs_ctl.frame[i]], (uint32_t)(hi - poff));
            }
            __threadfence();
            __syncthreads();
            if (wave() == 0) {
                const uint32_t ln = lane();
                bool nd = false;
                if (ln < cnt) {
                    uint32_t k = gfs_ctl.kind[ln];
                    if (k == GFS_K_OWN || k == GFS_K_HIT) nd = set_dirty(g, gfs_ctl.frame[ln]);
                }
                dirty_tally(g, nd, true);
                if (gfs_ctl.err) fail_own(g, fd, pb, cnt);
                finish_batch(g, cnt);
                uint64_t waitm = __ballot(ln < cnt && gfs_ctl.kind[ln] == GFS_K_WAIT);
                if (ln == 0) gfs_ctl.a = waitm;
            }
            __syncthreads();
            if (uni(gfs_ctl.err)) { err = uni(gfs_ctl.err); break; }
            uint64_t nxt = uni(gfs_ctl.a);
            if (nxt == pending) {
                if (++idle > 4) backoff();
                // Frames run out when the cache is full of dirty pages, which
                // the eviction sweep never takes (the reference's allocator
                // does not either, mallocfree.cu:81-226), and the check at the
                // batch's start cannot rule that out while other writers are
                // busy.  This workgroup writes back the file's dirty pages so
                // far, its own among them, which makes their frames evictable.
                if ((idle & 31) == 16) {
                    const int fr = flush_file_stalled(g, fd, offset + n);
                    if (fr) { err = fr; break; }
                }
                if ((idle & 63) == 0 && wg_timed_out(g, t0)) { err = -ETIMEDOUT; break; }
            } else {
                idle = 0;
            }
            pending = nxt;
        }
    }
    if (tid() == 0 && !err) {
        g->files[fd].dirty = 1;
        atomicMax((unsigned long long*)&g->files[fd].size, (unsigned long long)(offset + n));
    }
    __syncthreads();
    return err ? (size_t)(int64_t)err : (size_t)n;
}

namespace gfs {
// Takes the page of leaf word `v` for write-back: when the frame holds the
// live, VALID page, dirty and with no write-back of it in flight, one CAS pins
// it, clears DIRTY and sets WB; returns the frame.  GFS_NOFRAME: nothing to
// write (clean, or not cached).  GFS_BUSYFRAME: the page is dirty but its CAS
// kept failing, or another write-back of it is in flight; look again later.
// One write-back of a page at a time: two WRITEs of one page in flight at once
// could reach the disk in either order, the older bytes last.  DIRTY is
// cleared before the page is copied out: a gwrite after that sets it again,
// so its bytes are written back by a later flush.
__device__ inline uint32_t take_dirty(gfs_dev* g, uint64_t v) {
    uint32_t got = GFS_NOFRAME;
    if (v != 0 && v != GFS_LEAF_RESERVED) {
        const uint32_t f = leaf_frame(v);
        // a plain look first: clean pages cost no atomic
        uint64_t m = ld_rlx(&g->meta[f]);
        for (int tries = 0; tries < 8; ++tries) {
            if (meta_gen(m) != leaf_gen(v) || meta_state(m) != GFS_F_VALID) break;
            if (m & GFS_FL_WB) {
                got = GFS_BUSYFRAME;
                break;
            }
            if (!(m & GFS_FL_DIRTY)) break;
            uint64_t exp = m;
            if (cas(&g->meta[f], exp, ((m | GFS_FL_WB) & ~GFS_FL_DIRTY) + 1ull)) {
                got = f;
                break;
            }
            m = exp;
            got = GFS_BUSYFRAME;   // if the tries run out
        }
    }
    dirty_tally(g, got < GFS_BUSYFRAME, false);
    return got;
}
// Ends the write-back of frame f begun by take_dirty: WB off and its pin
// dropped; a page whose write did not happen (!ok) is marked dirty again, so
// it is neither evicted nor forgotten and a later gfsync / gclose retries it.
__device__ __forceinline__ void wb_end(gfs_dev* g, uint32_t f, bool ok) {
    bool nd = false;
    if (!ok) nd = !(__hip_atomic_fetch_or(&g->meta[f], GFS_FL_DIRTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                    GFS_FL_DIRTY);
    dirty_tally(g, nd, true);
    fadd(&g->meta[f], (uint64_t)0 - GFS_FL_WB - 1ull);
}

// Write back every dirty page of `fd` (workgroup-collective).  Pages of
// O_GWRONCE files are OR-merged with the bytes on disk by the host
// (libgpufs/host_loop.h:76-137).  The dirty pages of one radix leaf (up to 64)
// travel together: wave 0 pins them and claims a ring slot for each, the
// workgroup copies them into staging, every page's WRITE is posted at once on
// one shared completion mask, and thread 0 waits for that mask once.  A dirty
// page that found no free slot, or whose write-back by another workgroup is
// still in flight, is taken in a later round (bounded wait), so that when the
// call returns every page dirtied before it is on disk.  Returns 0 or the
// first error; pages whose write failed stay dirty.
__device__ inline int flush_file(gfs_dev* g, int fd, uint64_t min_size = 0) {
    gfs_file* file = &g->files[fd];
    const bool wronce = (file->flags & (O_CREAT | O_ACCMODE)) == O_GWRONCE;
    const uint64_t size = file->size > min_size ? file->size : min_size;
    const int32_t host_fd = file->host_fd;
    const uint64_t npages = (size + g->page_size - 1) >> g->log_page;
    const uint32_t T = nthreads(), t = tid();
    if (t == 0) gfs_memo.fd = -1;   // the batch arrays are reused below
    int err = 0;
    // Workgroups of one grid that write back the same file start at leaves
    // spread in proportion to their index (where a grid that writes its share
    // of a file by index has its own pages), then wrap round.
    const uint64_t nleaves = (npages + GFS_FANOUT - 1) >> GFS_FANOUT_LOG;
    const uint64_t first = nleaves ? (uint64_t)blockIdx.x * nleaves / gridDim.x : 0;
    for (uint64_t li = 0; li < nleaves && !err; ++li) {
        const uint64_t leaf = first + li < nleaves ? first + li : first + li - nleaves;
        const uint64_t base = leaf << GFS_FANOUT_LOG;
        const uint64_t in_leaf = npages - base < 64 ? npages - base : 64;
        uint64_t pending = in_leaf == 64 ? ~0ull : ((1ull << in_leaf) - 1);
        const uint64_t t0 = now_ticks();
        uint32_t idle = 0;
        while (pending && !err) {
            __syncthreads();
            if (wave() == 0) {
                const uint32_t ln = lane();
                int e = 0;
                uint64_t node = 0;
                if (ln == 0) node = tree_leaf(g, file, leaf, false, &e);
                node = __shfl(node, 0);
                const bool mine = node && ((pending >> ln) & 1ull);
                uint32_t f = mine ? take_dirty(g, ld_rlx(&g->nodes[node * GFS_FANOUT + ln])) : GFS_NOFRAME;
                const bool busy = f == GFS_BUSYFRAME;
                if (busy) f = GFS_NOFRAME;
                const int s = rpc_claim_wave(g, f != GFS_NOFRAME, blockIdx.x * 3u);
                const bool noslot = f != GFS_NOFRAME && s < 0;
                if (noslot) {   // stays dirty and pending
                    wb_end(g, f, false);
                    f = GFS_NOFRAME;
                }
                const bool post = f != GFS_NOFRAME;
                const uint64_t postm = __ballot(post);
                uint32_t bytes = 0;
                // The frame's extent, not clipped at the file size: the size
                // grows only when a gwrite ends, so another writer's bytes in
                // the page may lie past it, and the dirty flag is already off.
                if (post) bytes = g->fbytes[f];
                fence_acq_agent();
                gfs_ctl.kind[ln] = post ? GFS_K_OWN : GFS_K_DONE;
                gfs_ctl.frame[ln] = f;
                gfs_ctl.slot[ln] = post ? (uint32_t)s : 0xFFFFFFFFu;
                gfs_ctl.gen[ln] = bytes;   // bytes to write (the arrays' generation field is free here)
                const uint32_t cq = postm ? (uint32_t)__shfl(s, (int)first_lane(postm)) : 0u;
                if (postm && ln == first_lane(postm)) sys_st(&g->cq[cq], 0ull);
                const uint64_t again = __ballot(noslot || busy);
                if (ln == 0) {
                    gfs_ctl.own = postm;
                    gfs_ctl.cq = cq;
                    gfs_ctl.a = again;
                }
            }
            __syncthreads();
            const uint64_t postm = uni(gfs_ctl.own);
            for (uint64_t m = postm; m; m &= m - 1) {
                const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1;
                copy_seg<false, false, SPACE_GLOBAL, SPACE_GLOBAL, 4>(staging_ptr(g, gfs_ctl.slot[i]),
                                                                     frame_ptr(g, gfs_ctl.frame[i]), gfs_ctl.gen[i],
                                                                     t, T);
            }
            __threadfence_system();   // staging bytes before the requests that point at them
            __syncthreads();
            if (wave() == 0) {
                const uint32_t ln = lane();
                int perr = 0;
                if (postm) {
                    const uint32_t cq = uni(gfs_ctl.cq);
                    const bool post = (postm >> ln) & 1ull;
                    if (post)
                        rpc_post(g, (int)gfs_ctl.slot[ln], wronce ? GFS_OP_DIFF : GFS_OP_WRITE, host_fd, 0,
                                 (base + ln) << g->log_page, gfs_ctl.gen[ln], cq | (ln << 16));
                    wave_count(&g->ctr.rpc_writes, post);
                    int w = 0;
                    if (ln == 0) w = rpc_wait(g, cq, postm);
                    w = uni(w);
                    // after a timeout: the requests answered before the limit
                    const uint64_t answered = w ? uni((uint64_t)sys_ld(&g->cq[cq])) & postm : postm;
                    fence_acq_sys();
                    if (post) {
                        const uint32_t f = gfs_ctl.frame[ln];
                        const bool got = (answered >> ln) & 1ull;
                        const int64_t r = got ? sys_ld(&g->slots[gfs_ctl.slot[ln]].result) : (int64_t)-ETIMEDOUT;
                        const bool ok = r >= 0 && (uint64_t)r == gfs_ctl.gen[ln];
                        if (!ok) perr = r < 0 ? (int)r : -EIO;
                        else count(&g->ctr.bytes_written, (unsigned long long)r);
                        wb_end(g, f, ok);
                    }
                    // Slots go back to the ring, except after a timeout those the
                    // host may still answer: the unanswered ones and the one
                    // holding the completion word.
                    const bool lost = post && w && (!((answered >> ln) & 1ull) || gfs_ctl.slot[ln] == cq);
                    rpc_release_wave(g, post && !lost, post ? (int)gfs_ctl.slot[ln] : 0);
                    const uint64_t lm = __ballot(lost);
                    if (ln == 0) rpc_quarantine(g, popc(lm));
                }
                const uint64_t em = __ballot(perr != 0);
                const int first = em ? __shfl(perr, (int)first_lane(em)) : 0;
                if (ln == 0) gfs_ctl.ret = first;
            }
            __syncthreads();
            err = (int)uni(gfs_ctl.ret);
            const uint64_t nxt = uni(gfs_ctl.a);
            if (!err && nxt) {
                // dirty pages found no free slot, or their write-back by
                // another workgroup is in flight: wait, bounded
                if (postm == 0) {
                    if (++idle > 4) backoff();
                    if ((idle & 63) == 0 && wg_timed_out(g, t0)) err = -ETIMEDOUT;
                } else {
                    idle = 0;
                }
            }
            pending = nxt;
        }
    }
    __syncthreads();
    return err;
}

// Single-thread write-back of every dirty page of `fd` (one thread, the rest
// of its wave may be inactive): page by page through one ring slot.
__device__ inline int flush_file_single(gfs_dev* g, int fd) {
    gfs_file* file = &g->files[fd];
    const bool wronce = (file->flags & (O_CREAT | O_ACCMODE)) == O_GWRONCE;
    const uint64_t size = file->size;
    const uint64_t npages = (size + g->page_size - 1) >> g->log_page;
    int err = 0;
    for (uint64_t leaf = 0; (leaf << GFS_FANOUT_LOG) < npages; ++leaf) {
        int e = 0;
        const uint64_t node = tree_leaf(g, file, leaf, false, &e);
        if (!node) continue;
        for (uint32_t i = 0; i < GFS_FANOUT && (leaf << GFS_FANOUT_LOG) + i < npages; ++i) {
            uint32_t f;
            const uint64_t tb = now_ticks();
            while ((f = take_dirty(g, ld_rlx(&g->nodes[node * GFS_FANOUT + i]))) == GFS_BUSYFRAME &&
                   now_ticks() - tb <= g->wait_ticks)
                backoff();
            if (f == GFS_BUSYFRAME) {
                if (!err) err = -ETIMEDOUT;
                continue;
            }
            if (f == GFS_NOFRAME) continue;
            const uint64_t poff = ((leaf << GFS_FANOUT_LOG) + i) << g->log_page;
            const uint64_t bytes = g->fbytes[f];   // the frame's extent (see flush_file)
            int s = -1;
            const uint64_t t0 = now_ticks();
            while ((s = rpc_claim(g, true, blockIdx.x)) < 0 && now_ticks() - t0 <= g->wait_ticks) backoff();
            int64_t r = -ETIMEDOUT;
            if (s >= 0) {
                fence_acq_agent();
                copy_seg<false, false, SPACE_GLOBAL, SPACE_GLOBAL, 4>(staging_ptr(g, (uint32_t)s), frame_ptr(g, f),
                                                                     bytes, 0, 1);
                __threadfence_system();
                sys_st(&g->cq[s], 0ull);
                rpc_post(g, s, wronce ? GFS_OP_DIFF : GFS_OP_WRITE, file->host_fd, 0, poff, bytes, (uint32_t)s);
                count(&g->ctr.rpc_writes, 1);
                const int w = rpc_wait(g, (uint32_t)s, 1ull);
                if (w) {
                    rpc_quarantine(g, 1);
                    r = w;
                } else {
                    r = sys_ld(&g->slots[s].result);
                    rpc_release(g, s);
                }
            }
            const bool ok = r >= 0 && (uint64_t)r == bytes;
            if (!ok && !err) err = r < 0 ? (int)r : -EIO;
            if (ok) count(&g->ctr.bytes_written, bytes);
            wb_end(g, f, ok);
        }
    }
    return err;
}

// Forgets every cached page of `fd` (one thread): dirty flags are cleared, the
// radix tree is dropped and the file gets a new version, so no reader can meet
// the old pages (their frames are reclaimed by the CLOCK sweep).  Used by
// truncation, and by a last close whose write-back failed (the bytes that
// could not be written are not kept as cached data the disk does not hold).
__device__ inline void discard_pages(gfs_dev* g, int fd) {
    gfs_file* e = &g->files[fd];
    const uint64_t npages = (e->size + g->page_size - 1) >> g->log_page;
    for (uint64_t leaf = 0; (leaf << GFS_FANOUT_LOG) < npages; ++leaf) {
        int er = 0;
        const uint64_t node = tree_leaf(g, e, leaf, false, &er);
        if (!node) continue;
        for (uint32_t i = 0; i < GFS_FANOUT; ++i) {
            // the dirty flag of the page's live frame goes (a CAS that checks
            // the generation: the frame may be holding another page by now)
            const uint64_t v = ld_rlx(&g->nodes[node * GFS_FANOUT + i]);
            if (v == 0 || v == GFS_LEAF_RESERVED) continue;
            const uint32_t f = leaf_frame(v);
            uint64_t m = ld_rlx(&g->meta[f]);
            bool cleared = false;
            while (meta_gen(m) == leaf_gen(v) && (m & GFS_FL_DIRTY) && !cleared) {
                uint64_t exp = m;
                cleared = cas(&g->meta[f], exp, m & ~GFS_FL_DIRTY);
                m = exp;
            }
            dirty_tally(g, cleared, false);
        }
    }
    st_rlx(&e->root, 0ull);
    fadd(&e->st, 1ull << 32);   // version + 1 (fstate.h: every change is an added delta)
}

// Host truncation of open file `fd` to `size` (one thread), then its cached
// pages are dropped as in the reference (libgpufs/fs_calls.cu:127-140).  The
// caller writes dirty pages below `size` back first.
__device__ inline int truncate_single(gfs_dev* g, int fd, int64_t size) {
    gfs_file* e = &g->files[fd];
    const int64_t c = rpc_call_single(g, GFS_OP_TRUNC, e->host_fd, 0, (uint64_t)size, 0, nullptr, nullptr);
    if (c < 0) return (int)c;
    discard_pages(g, fd);
    e->size = (uint64_t)size;
    return 0;
}

// The last close of `fd` after its write-back (result `r`): the host close is
// posted, not waited for; on a failed write-back the pages are discarded.
__device__ inline void close_last(gfs_dev* g, int fd, int r) {
    gfs_file* e = &g->files[fd];
    if (r) discard_pages(g, fd);
    // the host closes its fd after the caller has moved on: nothing is
    // written through it any more (the write-back waited for its writes)
    rpc_post_close(g, e->host_fd);
    count(&g->ctr.rpc_closes, 1);
    e->host_fd = -1;
    e->dirty = 0;
    fst_dev P{0, 0};
    fst_close_done(&e->st, P);
}

__device__ __forceinline__ bool fd_open(gfs_dev* g, int fd) {
    return fd_valid(g, fd) && fst_state(ld_acq(&g->files[fd].st)) == GFS_FS_OPEN;
}
} // namespace gfs

// libgpufs/fs_calls.cu:107-125
__device__ inline int gfsync(int fd) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    __syncthreads();
    if (tid() == 0) gfs_ctl.ret = fd_open(g, fd) ? 0 : -EBADF;
    __syncthreads();
    if (uni(gfs_ctl.ret)) return (int)uni(gfs_ctl.ret);
    int r = flush_file(g, fd);
    __syncthreads();
    if (tid() == 0 && r == 0) g->files[fd].dirty = 0;
    __syncthreads();
    return r;
}

// libgpufs/fs_calls.cu:35-105.  The last closer writes back dirty pages and
// closes the host descriptor; cached clean pages stay valid for a reopen.
__device__ inline int gclose(int fd) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    __syncthreads();
    if (tid() == 0) {
        int r = 0;
        uint32_t last = 0;
        if (!fd_valid(g, fd)) r = -EBADF;
        else {
            fst_dev P{0, 0};
            const int c = fst_close(&g->files[fd].st, P);
            if (c < 0) r = c;
            else last = (uint32_t)c;
        }
        gfs_ctl.ret = r;
        gfs_ctl.flag = last;
    }
    __syncthreads();
    if (uni(gfs_ctl.ret) || !uni(gfs_ctl.flag)) return (int)uni(gfs_ctl.ret);
    gfs_file* e = &g->files[fd];
    if (tid() == 0) gfs_ctl.flag = e->dirty;
    __syncthreads();
    int r = uni(gfs_ctl.flag) ? flush_file(g, fd) : 0;
    __syncthreads();
    if (tid() == 0) {
        close_last(g, fd, r);
        gfs_ctl.ret = r;
    }
    __syncthreads();
    return (int)uni(gfs_ctl.ret);
}

// libgpufs/fs_calls.cu:127-149.  The reference supports only size 0; any size
// is accepted here.  Dirty pages are written back first when the new size
// keeps some of them, then the file is truncated on the host and its cached
// pages are dropped (new version) as in the reference.
__device__ inline int gftruncate(int fd, int size) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    __syncthreads();
    if (tid() == 0) gfs_ctl.ret = (fd_open(g, fd) && size >= 0) ? 0 : -EBADF;
    __syncthreads();
    int r = (int)uni(gfs_ctl.ret);
    if (r) return r;
    if (size > 0) r = flush_file(g, fd);
    __syncthreads();
    if (tid() == 0) gfs_ctl.ret = r ? r : truncate_single(g, fd, size);
    __syncthreads();
    return (int)uni(gfs_ctl.ret);
}

// ---- single-thread calls (libgpufs/fs_calls.cu.h:34-36): callable by one
// thread while the rest of its wave and workgroup do something else (the
// reference's samples call them inside their BEGIN_SINGLE_THREAD sections,
// samples/grep_text/grep_text_gpu.cu:258-268).  They touch no LDS state of the
// workgroup-collective calls, and move pages with that thread alone.

// libgpufs/fs_calls.cu:152-225.  Same results as gopen.
__device__ inline int single_thread_open(const char* filename, int flags) {
    return gfs::open_single(filename, flags);
}

// libgpufs/fs_calls.cu:127-140.  0 on success or a negative code (the
// reference returns truncate_cpu()==0, 1 on success, and only then skips its
// cache drop).  Any size is accepted; dirty pages below it are written first.
__device__ inline int single_thread_ftruncate(int fd, int size) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    if (!fd_open(g, fd) || size < 0) return -EBADF;
    if (size > 0) {
        const int r = flush_file_single(g, fd);
        if (r) return r;
    }
    return truncate_single(g, fd, size);
}

// Declared at libgpufs/fs_calls.cu.h:34 (and called by
// samples/grep_text/grep_text_gpu_mmap.cu:427) but never defined there: writes
// back the file's dirty pages, like gfsync, from one thread.
__device__ inline int single_thread_fsync(int fd) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    if (!fd_open(g, fd)) return -EBADF;
    const int r = flush_file_single(g, fd);
    if (r == 0) g->files[fd].dirty = 0;
    return r;
}

// Called by samples/grep_text/grep_text_gpu_mmap.cu:259-260,413-429 (not
// declared in the reference's headers): gclose from one thread.
__device__ inline int single_thread_close(int fd) {
    using namespace gfs;
    gfs_dev* g = gfs_g;
    if (!fd_valid(g, fd)) return -EBADF;
    fst_dev P{0, 0};
    const int c = fst_close(&g->files[fd].st, P);
    if (c <= 0) return c;
    const int r = g->files[fd].dirty ? flush_file_single(g, fd) : 0;
    close_last(g, fd, r);
    return r;
}

// libgpufs/fs_calls.cu:845-850: not implemented in the reference either (it
// asserts and returns 0); here it reports -ENOSYS.
__device__ inline uint gunlink(char* filename) {
    (void)filename;
    return (uint)(-ENOSYS);
}

// libgpufs/fs_calls.cu:422-481.  Pins the page holding `offset` and returns a
// pointer to `offset` inside it; [offset, offset+size) must not cross a page.
// Files opened for writing mark the page dirty.
__device__ inline volatile void* gmmap(void* addr, size_t size, int prot, int flags, int fd, off_t offset) {
    using namespace gfs;
    (void)addr; (void)prot; (void)flags;
    gfs_dev* g = gfs_g;
    __syncthreads();
    if (tid() == 0) {
        int64_t r = 0;
        if (!fd_valid(g, fd) || fst_state(ld_acq(&g->files[fd].st)) != GFS_FS_OPEN || offset < 0) r = -EBADF;
        else if (((uint64_t)offset & (g->page_size - 1)) + size > g->page_size) r = -EINVAL;
        gfs_ctl.ret = r;
        gfs_ctl.a = g->files[fd].size;
        gfs_ctl.flag = (uint32_t)g->files[fd].flags;
    }
    __syncthreads();
    if (uni(gfs_ctl.ret)) return GMAP_FAILED;
    const uint64_t fsize = uni(gfs_ctl.a);
    const int oflags = (int)uni(gfs_ctl.flag);
    const bool writable = (oflags & O_ACCMODE) != O_RDONLY;
    const bool wronce = (oflags & (O_CREAT | O_ACCMODE)) == O_GWRONCE;
    const uint64_t page = (uint64_t)offset >> g->log_page;
    // a writable mapping reads a page past the end of file too (see gwrite)
    const bool nohost = wronce || (!writable && (page << g->log_page) >= fsize);
    uint64_t t0 = now_ticks();
    for (;;) {
        __syncthreads();
        if (wave() == 0) {
            if (lane() == 0) gfs_ctl.err = 0;
            resolve_batch(g, fd, page, 1, 1ull, nohost ? 1ull : 0ull, writable);
        }
        __syncthreads();
        if (uni(gfs_ctl.err)) return GMAP_FAILED;
        fill_from_host(g, 1, page);
        if (uni((uint32_t)gfs_ctl.kind[0]) == GFS_K_OWN && (uni(gfs_ctl.own) & 1ull) == 0) {
            uint8_t* fp = frame_ptr(g, gfs_ctl.frame[0]);
            for (uint64_t k = tid(); k < (g->page_size >> 4); k += nthreads()) *(v4u*)(fp + 16 * k) = v4u{0, 0, 0, 0};
            __threadfence();
        }
        __syncthreads();
        if (tid() == 0) {
            const uint32_t k = gfs_ctl.kind[0];
            const uint32_t f = gfs_ctl.frame[0];
            if (k == GFS_K_ERR) {
                fail_own(g, fd, page, 1);
                gfs_ctl.ret = -EIO;
            } else if (k == GFS_K_OWN) {
                if (gfs_ctl.slot[0] != 0xFFFFFFFFu) rpc_release(g, (int)gfs_ctl.slot[0]);
                // publish VALID but keep the pin: it is dropped by gmunmap
                fadd(&g->meta[f], (GFS_F_VALID - GFS_F_FILLING) * (1ull << GFS_ST_SHIFT));
                gfs_ctl.ret = 1;
            } else if (k == GFS_K_HIT) {
                gfs_ctl.ret = 1;
            } else {
                gfs_ctl.ret = 0;
            }
            if (gfs_ctl.ret == 1 && writable) {
                dirty_tally(g, set_dirty(g, f), true);
                uint32_t end = (uint32_t)(((uint64_t)offset & (g->page_size - 1)) + size);
                atomicMax(&g->fbytes[f], end);
                g->files[fd].dirty = 1;
                atomicMax((unsigned long long*)&g->files[fd].size, (unsigned long long)(offset + size));
            }
        }
        __syncthreads();
        if (uni(gfs_ctl.ret) < 0) return GMAP_FAILED;
        if (uni(gfs_ctl.ret) == 1) break;
        if (wg_timed_out(g, t0)) return GMAP_FAILED;
        backoff();
    }
    return (volatile void*)uni((uint64_t)(uintptr_t)(frame_ptr(g, gfs_ctl.frame[0]) + ((uint64_t)offset & (g->page_size - 1))));
}

// libgpufs/fs_calls.cu:385-401.  Unpins the frame holding `addr`; -1 if
// `addr` is not inside the page pool.
__device__ inline int gmunmap(volatile void* addr, size_t length) {
    using namespace gfs;
    (void)length;
    gfs_dev* g = gfs_g;
    const uint8_t* a = (const uint8_t*)(uintptr_t)addr;
    uint64_t f = (uint64_t)(a - g->pool) >> g->log_page;
    if (a < g->pool || f >= g->n_frames) return -1;
    __threadfence();
    __syncthreads();
    if (tid() == 0) fadd(&g->meta[f], ~0ull);
    __syncthreads();
    return 0;
}


// Wave-collective mapping (reference gmmap_warp / getRwLockedPage_warp,
// libgpufs/fs_calls.cu:483-683): all 64 lanes of one wave call with the same
// arguments, and different waves of a workgroup may map different pages at
// the same time, so nothing here touches the workgroup's shared control block
// or waits at a workgroup barrier.  Lane 0 looks the page up and pins it
// (`ref` pins, dropped by gmunmap_warp), or claims a frame and a ring slot,
// reads the page from the host and publishes it; values reach the other lanes
// through v_readfirstlane; the page copy (or zero fill) uses all 64 lanes.
// Wait limits are decided by lane 0, so the whole wave leaves together.
__device__ inline volatile void* gmmap_warp(void* addr, size_t size, int prot, int flags, int fd, off_t offset,
                                           int ref = 1) {
    using namespace gfs;
    (void)addr; (void)prot; (void)flags;
    gfs_dev* g = gfs_g;
    const uint32_t ln = lane();
    if (ref < 1) return GMAP_FAILED;
    int32_t bad = 0, oflags = 0, host_fd = -1;
    uint64_t fsize = 0;
    if (ln == 0) {
        if (!fd_valid(g, fd) || fst_state(ld_acq(&g->files[fd].st)) != GFS_FS_OPEN || offset < 0) bad = -EBADF;
        else if (((uint64_t)offset & (g->page_size - 1)) + size > g->page_size) bad = -EINVAL;
        else {
            fsize = g->files[fd].size;
            oflags = g->files[fd].flags;
            host_fd = g->files[fd].host_fd;
        }
    }
    if (uni(bad)) return GMAP_FAILED;
    fsize = uni(fsize);
    oflags = uni(oflags);
    host_fd = uni(host_fd);
    const uint32_t L = g->log_page;
    const bool writable = (oflags & O_ACCMODE) != O_RDONLY;
    const bool wronce = (oflags & (O_CREAT | O_ACCMODE)) == O_GWRONCE;
    const uint64_t page = (uint64_t)offset >> L;
    // a writable mapping reads a page past the end of file too (see gwrite)
    const bool nohost = wronce || (!writable && (page << L) >= fsize);
    const uint64_t t0 = now_ticks();
    uint32_t f = GFS_NOFRAME;
    for (;;) {
        // lane 0: leaf slot, pin
        uint32_t kind = GFS_K_MISS;
        int32_t err = 0;
        uint64_t v = 0;
        uint64_t* sp = nullptr;
        if (ln == 0) {
            int e = 0;
            const uint64_t node = tree_leaf(g, &g->files[fd], page >> GFS_FANOUT_LOG, true, &e);
            err = e;
            if (!e) {
                sp = &g->nodes[node * GFS_FANOUT + (page & (GFS_FANOUT - 1))];
                v = ld_rlx(sp);
                if (v == GFS_LEAF_RESERVED) {
                    kind = GFS_K_WAIT;
                } else if (v) {
                    const uint32_t vf = leaf_frame(v), vg = leaf_gen(v);
                    const uint64_t old = fadd_rlx(&g->meta[vf], (uint64_t)ref);
                    if (meta_gen(old) == vg && meta_state(old) == GFS_F_VALID) {
                        kind = GFS_K_HIT;
                        f = vf;
                        g->fref[vf] = 1;
                    } else {
                        fadd_rlx(&g->meta[vf], (uint64_t)0 - (uint64_t)ref);
                        if (meta_gen(old) == vg && meta_state(old) == GFS_F_FILLING) kind = GFS_K_WAIT;
                    }
                }
            }
        }
        if (uni(err)) return GMAP_FAILED;
        kind = uni(kind);
        if (kind == GFS_K_HIT) {
            fence_acq_agent();
            if (ln == 0) count(&g->ctr.page_hits, 1);
            f = uni(f);
            break;
        }
        uint32_t late = 0;
        if (kind == GFS_K_MISS) {
            // reserve the leaf slot (see resolve_batch), then a ring slot for
            // the host read, then a frame
            uint32_t resv = 0;
            if (ln == 0) {
                uint64_t exp = v;
                resv = cas(sp, exp, GFS_LEAF_RESERVED) ? 1u : 0u;
            }
            resv = uni(resv);
            int32_t s = -1;
            if (resv && ln == 0 && !nohost) s = rpc_claim(g, true, blockIdx.x * 5u + wave());
            s = uni(s);
            if (resv && !(nohost || s >= 0) && ln == 0) st_rel(sp, v);   // no slot: give the leaf back
            if (resv && (nohost || s >= 0)) {
                uint32_t nf = uni(frame_alloc_wave(g, ln == 0, 2));
                uint32_t own = 0;
                if (ln == 0) {
                    if (nf != GFS_NOFRAME) {
                        g->fkey[nf] = gfs_key((uint32_t)fd, page);
                        g->fbytes[nf] = 0;
                        const uint32_t gen = meta_gen(ld_rlx(&g->meta[nf]));
                        st_rel(sp, leaf_pack(nf, gen));
                        g->fref[nf] = 1;
                        own = 1u;
                    } else {
                        st_rel(sp, v);
                    }
                }
                own = uni(own);
                if (own) {
                    int64_t r = 0;
                    if (!nohost) {
                        const uint64_t poff = page << L;
                        uint64_t want = fsize > poff ? fsize - poff : 0;
                        if (writable || want > g->page_size) want = g->page_size;   // as in resolve_batch
                        int w = 0;
                        if (ln == 0) {
                            sys_st(&g->cq[s], 0ull);
                            rpc_post(g, s, GFS_OP_READ, host_fd, 0, poff, want, (uint32_t)s);
                            count(&g->ctr.rpc_reads, 1);
                            w = rpc_wait(g, (uint32_t)s, 1ull);
                            r = w ? (int64_t)w : sys_ld(&g->slots[s].result);
                            if (r >= 0 && (uint64_t)r > want) r = -EIO;
                        }
                        r = uni(r);
                        if (r >= 0) {
                            copy_seg<true, false>(frame_ptr(g, nf), staging_ptr(g, (uint32_t)s), (uint64_t)r, ln, 64);
                            if ((uint64_t)r < g->page_size) zero_tail(frame_ptr(g, nf), (uint64_t)r, g->page_size, ln, 64);
                        }
                        // the wave's staging loads have all returned (their data was stored);
                        // a request that timed out keeps its slot (rpc_quarantine)
                        if (ln == 0) {
                            if (w) rpc_quarantine(g, 1);
                            else rpc_release(g, s);
                        }
                    } else {
                        uint8_t* fp = frame_ptr(g, nf);
                        for (uint64_t k = ln; k < (g->page_size >> 4); k += 64) *(v4u*)(fp + 16 * k) = v4u{0, 0, 0, 0};
                    }
                    if (r < 0) {
                        if (ln == 0) {   // unpublish and free the frame
                            uint64_t exp = leaf_pack(nf, meta_gen(ld_rlx(&g->meta[nf])));
                            cas(sp, exp, 0ull);
                            frame_release(g, nf);
                        }
                        return GMAP_FAILED;
                    }
                    __threadfence();   // frame bytes before VALID
                    if (ln == 0) {
                        g->fbytes[nf] = (uint32_t)r;
                        count(&g->ctr.page_misses, 1);
                        count(&g->ctr.bytes_filled, (unsigned long long)r);
                        // FILLING -> VALID, the allocation's pin becomes `ref` pins
                        fadd(&g->meta[nf], (GFS_F_VALID - GFS_F_FILLING) * (1ull << GFS_ST_SHIFT) + (uint64_t)(ref - 1));
                    }
                    f = nf;
                    break;
                }
                if (ln == 0 && s >= 0) rpc_release(g, s);
            }
        } else if (ln == 0) {
            count(&g->ctr.page_waits, 1);
        }
        if (ln == 0) late = (now_ticks() - t0 > g->wait_ticks) ? 1u : 0u;
        if (uni(late)) {
            if (ln == 0) count(&g->ctr.timeouts, 1);
            return GMAP_FAILED;
        }
        backoff();
    }
    if (writable && ln == 0) {
        dirty_tally(g, set_dirty(g, f), true);
        atomicMax(&g->fbytes[f], (uint32_t)(((uint64_t)offset & (g->page_size - 1)) + size));
        g->files[fd].dirty = 1;
        atomicMax((unsigned long long*)&g->files[fd].size, (unsigned long long)(offset + size));
    }
    return (volatile void*)uni((uint64_t)(uintptr_t)(frame_ptr(g, f) + ((uint64_t)offset & (g->page_size - 1))));
}

// Wave-collective unmap (reference gmunmap_warp, libgpufs/fs_calls.cu.h:45):
// drops `ref` pins of the frame holding `addr` (any address inside the mapped
// page); -1 if `addr` is not inside the page pool.
__device__ inline int gmunmap_warp(volatile void* addr, size_t length, int ref = 1) {
    using namespace gfs;
    (void)length;
    gfs_dev* g = gfs_g;
    const uint8_t* a = (const uint8_t*)(uintptr_t)uni((uint64_t)(uintptr_t)addr);
    const uint64_t f = (uint64_t)(a - g->pool) >> g->log_page;
    if (a < g->pool || f >= g->n_frames || ref < 1) return -1;
    __threadfence();   // the wave's accesses to the page come before the unpin
    if (lane() == 0) fadd(&g->meta[f], (uint64_t)0 - (uint64_t)ref);
    return 0;
}

// libgpufs/fs_calls.cu:804-843.  Writes back the page holding `addr` now.
__device__ inline int gmsync(volatile void* addr, size_t length, int flags) {
    using namespace gfs;
    (void)length; (void)flags;
    gfs_dev* g = gfs_g;
    const uint8_t* a = (const uint8_t*)(uintptr_t)addr;
    uint64_t f = (uint64_t)(a - g->pool) >> g->log_page;
    if (a < g->pool || f >= g->n_frames) return -1;
    __threadfence();
    const uint64_t key = g->fkey[f];
    const int fd = (int)(key >> 40);
    const uint64_t page = key & 0xFFFFFFFFFFull;
    gfs_file* file = &g->files[fd];
    const bool wronce = (file->flags & (O_CREAT | O_ACCMODE)) == O_GWRONCE;
    uint64_t poff = page << g->log_page;
    __syncthreads();
    if (tid() == 0) {
        int s = -1;
        uint64_t t0 = now_ticks();
        while ((s = rpc_claim(g, true, blockIdx.x)) < 0) {
            if (now_ticks() - t0 > g->wait_ticks) break;
            backoff();
        }
        if (s >= 0) {
            // The page is taken for write-back as take_dirty does (WB set,
            // DIRTY cleared, a pin), once no other write-back of it is in
            // flight (bounded wait; the mapping keeps the frame this page's).
            uint64_t m = ld_rlx(&g->meta[f]);
            bool took = false;
            while (!took) {
                if (m & GFS_FL_WB) {
                    if (now_ticks() - t0 > g->wait_ticks) break;
                    backoff();
                    m = ld_rlx(&g->meta[f]);
                    continue;
                }
                uint64_t exp = m;
                took = cas(&g->meta[f], exp, ((m | GFS_FL_WB) & ~GFS_FL_DIRTY) + 1ull);
                if (took) dirty_tally(g, (m & GFS_FL_DIRTY) != 0, false);
                else m = exp;
            }
            if (!took) {
                rpc_release(g, s);
                s = -1;
            }
        }
        gfs_ctl.slot[0] = s < 0 ? 0xFFFFFFFFu : (uint32_t)s;
        // the extent is read after the flag is cleared: a writer that grew it
        // before setting the flag is seen, a later one re-dirties
        if (s >= 0) gfs_ctl.ret = (int64_t)g->fbytes[f];
    }
    __syncthreads();
    const uint32_t s = uni(gfs_ctl.slot[0]);
    if (s == 0xFFFFFFFFu) return -ETIMEDOUT;
    const uint64_t bytes = (uint64_t)uni(gfs_ctl.ret);
    wg_copy(staging_ptr(g, s), frame_ptr(g, (uint32_t)f), bytes);
    __threadfence_system();
    __syncthreads();
    if (tid() == 0) {
        sys_st(&g->cq[s], 0u);
        rpc_post(g, (int)s, wronce ? GFS_OP_DIFF : GFS_OP_WRITE, file->host_fd, 0, poff, bytes, s);
        count(&g->ctr.rpc_writes, 1);
        const int w = rpc_wait(g, s, 1ull);
        int64_t r;
        if (w) {   // the host may still answer: the slot is not returned
            rpc_quarantine(g, 1);
            r = w;
        } else {
            r = sys_ld(&g->slots[s].result);
            rpc_release(g, (int)s);
        }
        const bool ok = r >= 0 && (uint64_t)r == bytes;
        if (ok) count(&g->ctr.bytes_written, bytes);
        gfs_ctl.ret = ok ? 0 : (r < 0 ? r : -EIO);
        wb_end(g, (uint32_t)f, ok);   // not on disk: the page stays dirty
    }
    __syncthreads();
    return (int)uni(gfs_ctl.ret);
}

// Host helper: store the context's device state in this translation unit.
// Call once per context (and again after switching contexts) before launching
// kernels that use the API above.
static inline int gpufs_bind(gpufs_ctx* ctx) {
    static void* bound = nullptr;
    static int bound_dev = -1;
    (void)gpufs_host_kick(ctx);   // a launch follows: have the host service awake for it
    void* p = gpufs_device_state(ctx);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -ENODEV;
    if (p == bound && dev == bound_dev) return 0;
    if (hipMemcpyToSymbol(HIP_SYMBOL(gfs_g), &p, sizeof(p), 0, hipMemcpyHostToDevice) != hipSuccess) return -EIO;
    bound = p;
    bound_dev = dev;
    return 0;
}
