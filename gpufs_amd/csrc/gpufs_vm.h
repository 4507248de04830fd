// gpufs_vm.h -- ActivePointers: file-backed pointers with a per-workgroup
// software TLB (reference: libgpufs/fat_pointer.cu.h, gvmmap in
// libgpufs/fs_calls.cu.h:54-61, used by samples/mosaic/mosaic_gpufs.cu:841,1020).
//
// A FatPointer<T, N> holds a file offset.  Dereferencing it resolves the page
// through a TLB<N> in shared memory: lanes of a wave that need the same page
// are served together; a TLB line holding that page is shared by reference
// count; a missing page takes over a line nobody holds and is mapped with
// gmmap_warp; the last holder to leave a line returns its page, so the pages
// pinned through lines never outnumber the pages pointers hold (a line that
// kept its page while idle could pin more pages, over all workgroups, than
// the cache has frames, and starve every other mapping).  While the line of a page is
// held for another page, the lanes map their page directly, each lane with a
// pin of its own, instead of waiting for the line (a wave whose own lanes
// hold the line would otherwise wait for itself).  N == 0 maps every page
// directly.  Moving a pointer within its page keeps it resolved; leaving the
// page, moveTo, copying, and destruction drop what it holds.
//
// Differences by design: every operation is wave-collective (all 64 lanes of
// the wave call it, as the reference's warp-level calls need their whole
// warp); HIP runs no constructor or destructor for __shared__ objects, so a
// workgroup calls tlb.init() before the first gvmmap and tlb.release() after
// its last pointer is gone; a page that cannot be mapped (I/O error, wait
// limit) makes the pointer refer to a scratch area beyond the page pool and
// counts a vm error in the context statistics.
#pragma once
#include "gpufs.h"

#define GFS_VM_NOFILE (-1)
#define GFS_VM_DIRECT (-2)   // holder: pins of the lanes themselves, no TLB line
#define GFS_VM_NONE (-1)     // holder: nothing held

template <int N>
struct TLB {
    static_assert(N >= 0 && (N & (N - 1)) == 0, "TLB size must be 0 or a power of two");
    struct line_t {
        int32_t fid;
        uint32_t frame;
        uint64_t vpage;
    };
    line_t lines[N > 0 ? N : 1];
    int32_t ref[N > 0 ? N : 1];   // holders; INT32_MIN added while a wave replaces the line

    // Workgroup-collective.
    __device__ void init() {
        for (uint32_t i = gfs::tid(); i < (uint32_t)N; i += gfs::nthreads()) {
            lines[i].fid = GFS_VM_NOFILE;
            lines[i].frame = 0;
            lines[i].vpage = ~0ull;
            ref[i] = 0;
        }
        __syncthreads();
    }
    // Workgroup-collective: drops the pin of every page the lines hold.  Call
    // after every FatPointer of the workgroup has been destroyed or moved away.
    __device__ void release() {
        __syncthreads();
        gfs_dev* g = gfs_g;
        for (uint32_t i = gfs::tid(); i < (uint32_t)N; i += gfs::nthreads()) {
            if (lines[i].fid != GFS_VM_NOFILE) {
                gfs::fadd(&g->meta[lines[i].frame], ~0ull);
                lines[i].fid = GFS_VM_NOFILE;
            }
            ref[i] = 0;
        }
        __syncthreads();
    }
};

namespace gfs {
// A few bytes past the last frame: the pool is allocated with room to spare
// there, so a pointer to it reads and writes harmlessly.
__device__ __forceinline__ uint8_t* vm_error_area(gfs_dev* g) {
    return g->pool + ((uint64_t)g->n_frames << g->log_page);
}
}   // namespace gfs

template <class T, int N>
class FatPointer {
public:
    __device__ FatPointer(int fd, uint64_t start, uint64_t size, int flags, TLB<N>* tlb)
        : m_fd(fd), m_flags(flags), m_start(start), m_end(start + size), m_off(start), m_tlb(tlb) {}

    // copies hold nothing (reference: "copies are invalid by definition")
    __device__ FatPointer(const FatPointer& o)
        : m_fd(o.m_fd), m_flags(o.m_flags), m_start(o.m_start), m_end(o.m_end), m_off(o.m_off), m_tlb(o.m_tlb) {}
    __device__ FatPointer& operator=(const FatPointer& o) {
        drop(true);
        m_fd = o.m_fd;
        m_flags = o.m_flags;
        m_start = o.m_start;
        m_end = o.m_end;
        m_off = o.m_off;
        m_tlb = o.m_tlb;
        return *this;
    }
    __device__ ~FatPointer() { drop(true); }

    // Offset from the start of the mapping, in bytes.
    __device__ FatPointer& moveTo(uint64_t offset) {
        drop(true);
        m_off = m_start + offset;
        return *this;
    }
    // Move by `n` elements; stays resolved while it stays in its page.
    __device__ FatPointer& move(int64_t n) {
        const uint64_t noff = m_off + (uint64_t)(n * (int64_t)sizeof(T));
        const uint32_t L = gfs_g->log_page;
        const bool stay = m_holder != GFS_VM_NONE && (noff >> L) == (m_off >> L);
        drop(!stay);   // every lane takes part, staying or not
        m_off = noff;
        return *this;
    }
    __device__ FatPointer& operator+=(int64_t n) { return move(n); }
    __device__ FatPointer& operator-=(int64_t n) { return move(-n); }

    __device__ T& operator*() {
        resolve();
        gfs_dev* g = gfs_g;
        if (m_holder == GFS_VM_NONE) return *(T*)gfs::vm_error_area(g);
        return *(T*)(g->pool + ((uint64_t)m_frame << g->log_page) + (m_off & (g->page_size - 1)));
    }

    __device__ uint64_t offset() const { return m_off - m_start; }

private:
    // Wave-collective: the lanes with `leave` set give up what they hold;
    // lanes holding the same thing release it together.
    __device__ void drop(bool leave) {
        gfs_dev* g = gfs_g;
        bool need = leave && m_holder != GFS_VM_NONE;
        for (;;) {
            const uint64_t m = __ballot(need);
            if (!m) break;
            const uint32_t leader = gfs::first_lane(m);
            const int32_t h = __shfl(m_holder, leader);
            const uint32_t f = __shfl(m_frame, leader);
            const bool same = need && m_holder == h && m_frame == f;
            const uint32_t n = gfs::popc(__ballot(same));
            if (gfs::lane() == leader) {
                if (h >= 0) {
                    // the last holder returns the line's page (a line keeps no
                    // page without holders, so the pages TLB lines pin never
                    // outnumber the pages pointers hold), unless a holder
                    // joined meanwhile: then the lock below fails
                    const int32_t old = atomicSub(&m_tlb->ref[h], (int32_t)n);
                    if (old == (int32_t)n && atomicCAS(&m_tlb->ref[h], 0, INT32_MIN) == 0) {
                        typename TLB<N>::line_t& ln = m_tlb->lines[h];
                        if (ln.fid != GFS_VM_NOFILE) {
                            gfs::fadd(&g->meta[ln.frame], ~0ull);
                            ln.fid = GFS_VM_NOFILE;
                        }
                        __threadfence_block();
                        atomicAdd(&m_tlb->ref[h], (int32_t)0x80000000u);   // unlock: back to 0
                    }
                } else {
                    gfs::fadd(&g->meta[f], (uint64_t)0 - (uint64_t)n);
                }
            }
            if (same) need = false;
        }
        if (leave) {
            m_holder = GFS_VM_NONE;
            m_failed = false;
        }
    }

    // Wave-collective: every lane ends up resolved (or on the error area).
    __device__ void resolve() {
        using namespace gfs;
        gfs_dev* g = gfs_g;
        const uint32_t L = g->log_page;
        bool need = m_holder == GFS_VM_NONE && !m_failed;
        if (!__ballot(need)) return;
        const uint64_t t0 = now_ticks();
        for (;;) {
            const uint64_t inv = __ballot(need);
            if (!inv) break;
            const uint32_t leader = first_lane(inv);
            const uint64_t vp = uni_from(m_off >> L, leader);
            const int32_t fd = (int32_t)uni_from((uint64_t)(uint32_t)m_fd, leader);
            const bool want = need && (m_off >> L) == vp && m_fd == fd;
            const uint32_t n = popc(__ballot(want));
            int32_t holder = GFS_VM_NONE;
            uint32_t frame = 0;
            if (N > 0) {
                const uint32_t h = (uint32_t)((vp + (uint64_t)fd * 0x9E3779B1ull) & (uint64_t)(N - 1));
                typename TLB<N>::line_t& ln = m_tlb->lines[h];
                int32_t old = 0;
                if (lane() == 0) old = atomicAdd(&m_tlb->ref[h], (int32_t)n);
                old = (int32_t)uni((uint32_t)old);
                if (old >= 0 && ln.fid == fd && ln.vpage == vp) {
                    holder = (int32_t)h;   // joined the line
                    frame = uni(ln.frame);
                } else {
                    if (lane() == 0) atomicSub(&m_tlb->ref[h], (int32_t)n);
                    int32_t c = 0;
                    if (lane() == 0) c = atomicCAS(&m_tlb->ref[h], 0, INT32_MIN);
                    c = (int32_t)uni((uint32_t)c);
                    if (c == 0) {
                        // the line is ours to refill: drop its page, map ours
                        if (lane() == 0 && ln.fid != GFS_VM_NOFILE) fadd(&g->meta[ln.frame], ~0ull);
                        volatile void* p = gmmap_warp(nullptr, g->page_size, 0, m_flags, fd, (off_t)(vp << L), 1);
                        if (p != GMAP_FAILED) {
                            frame = (uint32_t)(((const uint8_t*)(uintptr_t)p - g->pool) >> L);
                            if (lane() == 0) { ln.fid = fd; ln.vpage = vp; ln.frame = frame; }
                            holder = (int32_t)h;
                        } else if (lane() == 0) {
                            ln.fid = GFS_VM_NOFILE;
                        }
                        __threadfence_block();
                        // unlock: remove INT32_MIN and add our n holders (unsigned wrap)
                        if (lane() == 0)
                            atomicAdd(&m_tlb->ref[h], (int32_t)((holder >= 0 ? n : 0u) + 0x80000000u));
                    } else if (c < 0) {
                        // another wave is refilling the line: look again shortly
                        holder = GFS_VM_NONE;
                        uint32_t late = 0;
                        if (lane() == 0) late = now_ticks() - t0 > g->wait_ticks ? 1u : 0u;
                        if (!uni(late)) {
                            backoff();
                            continue;
                        }
                    }
                    if (holder == GFS_VM_NONE && c != 0) {
                        // held for another page: pins of our own
                        volatile void* p = gmmap_warp(nullptr, g->page_size, 0, m_flags, fd, (off_t)(vp << L), (int)n);
                        if (p != GMAP_FAILED) {
                            frame = (uint32_t)(((const uint8_t*)(uintptr_t)p - g->pool) >> L);
                            holder = GFS_VM_DIRECT;
                        }
                    }
                }
            } else {
                volatile void* p = gmmap_warp(nullptr, g->page_size, 0, m_flags, fd, (off_t)(vp << L), (int)n);
                if (p != GMAP_FAILED) {
                    frame = (uint32_t)(((const uint8_t*)(uintptr_t)p - g->pool) >> L);
                    holder = GFS_VM_DIRECT;
                }
            }
            if (want) {
                need = false;
                m_holder = holder;
                m_frame = frame;
                m_failed = holder == GFS_VM_NONE;
            }
            if (holder == GFS_VM_NONE && lane() == 0) count(&g->ctr.vm_errors, 1);
        }
    }

    // value of lane `src` (uniform: src comes from a ballot)
    __device__ static uint64_t uni_from(uint64_t v, uint32_t src) {
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src);
        return gfs::uni(((uint64_t)hi << 32) | lo);
    }

    int m_fd;
    int m_flags;
    uint64_t m_start, m_end, m_off;
    TLB<N>* m_tlb;
    int32_t m_holder = GFS_VM_NONE;
    uint32_t m_frame = 0;
    bool m_failed = false;
};

// Reference: gvmmap (libgpufs/fs_calls.cu.h:54-61).  `fd` is an open file;
// the pointer starts at `offset` and spans `size` bytes.
template <class T, int N>
__device__ inline FatPointer<T, N> gvmmap(void* addr, size_t size, int prot, int flags, int fd, off_t offset,
                                          TLB<N>* tlb) {
    (void)addr;
    (void)prot;
    return FatPointer<T, N>(fd, (uint64_t)offset, (uint64_t)size, flags, tlb);
}
