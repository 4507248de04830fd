// gfs_layout.h -- memory layout shared by the host runtime and the device API.
//
// Everything the device touches lives in one `gfs_dev` struct in HBM plus the
// arrays it points to.  The GPU->host RPC ring (slots, submission queue,
// completion counters, staging) lives in fine-grained coherent pinned host
// memory.  See the reference's equivalents:
//   PFrame          libgpufs/fs_structures.cu.h:24-60   -> frame meta word (gfs_dev::meta)
//   HashMap         libgpufs/hashMap.cu.h:26-38         -> per-file radix tree (gfs_dev::nodes)
//   PPool           libgpufs/mallocfree.cu.h:36-54      -> fresh counter + CLOCK hand
//   FTable_entry    libgpufs/fs_structures.cu.h:81-112  -> gfs_file
//   CPU_IPC_RW_Entry libgpufs/cpu_ipc.cu.h:75-95        -> gfs_rpc_slot
#pragma once
#include <stdint.h>
#ifdef __HIPCC__
#define GFS_HD __host__ __device__
#else
#define GFS_HD
#endif
#include <fcntl.h>
#include <sys/mman.h>

// ---- open flags and return codes kept from libgpufs/fs_constants.h:50-51,86-92
#define O_GRDONLY (O_RDONLY)
#define O_GWRONLY (O_WRONLY)
#define O_GCREAT  (O_CREAT)
#define O_GRDWR   (O_RDWR)
#define O_GWRONCE (O_GCREAT | O_GWRONLY)
#define GMAP_FAILED MAP_FAILED
#define E_FSTABLE_FULL (-1)
#define E_IPC_OPEN_ERROR (-2)

#define GFS_NAME_MAX 128                // bytes incl. NUL (reference: FILENAME_SIZE 64)
#define GFS_NAME_RESERVED 0xFFFFFFFFu   // name-index slot taken, entry not yet published
#define GFS_LEAF_RESERVED (~0ull)       // radix leaf slot claimed by a workgroup, frame not yet published
// Ring slot bitmap words sit 128 bytes apart: every claim and release is an
// atomic on them, and packed into one cache line they all queue at one L2
// channel (which also carries its share of every other stream).
#define GFS_SLOT_WORD_STRIDE 16u
#define GFS_FANOUT_LOG 6          // radix fan-out 64 = one wave64 lane per slot
#define GFS_FANOUT 64
#define GFS_MAX_HEIGHT 7          // 64^7 pages

// ---- frame meta word: [0,20) pin count | [20,24) state | [24,32) flags | [32,64) generation
#define GFS_REF_MASK   0xFFFFFull
#define GFS_ST_SHIFT   20
#define GFS_ST_MASK    (0xFull << GFS_ST_SHIFT)
#define GFS_FL_DIRTY   (1ull << 24)
#define GFS_FL_WB      (1ull << 25)   // a write-back of the frame's page is in flight
#define GFS_GEN_SHIFT  32
#define GFS_F_FREE     0ull
#define GFS_F_FILLING  1ull
#define GFS_F_VALID    2ull

// ---- file state word: [0,8) state | [8,32) open count | [32,64) version
#define GFS_FS_EMPTY   0u
#define GFS_FS_OPENING 1u
#define GFS_FS_OPEN    2u
#define GFS_FS_CLOSING 3u
#define GFS_FS_CLOSED  4u
#define GFS_FS_FREE    5u   // reclaimed: on the free list, in no name-index slot

// ---- tuning switches
#define GFS_TUNE_PAGE_SYNC 1u   // gread: workgroup barrier after each page copied to the caller
#define GFS_TUNE_NO_READAHEAD 2u   // gread: no sequential read-ahead
#define GFS_READAHEAD_BYTES (1u << 20)   // pages fetched ahead of a sequential reader (at most 63 pages)

// ---- RPC opcodes
#define GFS_SQ_NOSLOT 0xFFFFu   // submission-queue slot field of a close that carries no slot (ring_depth <= 65535)
#define GFS_OP_READ  1u
#define GFS_OP_WRITE 2u
#define GFS_OP_DIFF  3u     // write OR-merged with on-disk bytes (O_GWRONCE pages)
#define GFS_OP_OPEN  4u
#define GFS_OP_CLOSE 5u
#define GFS_OP_TRUNC 6u

typedef struct gfs_file {
    char     name[GFS_NAME_MAX];
    uint64_t st;        // state | opens | version (atomic)
    uint64_t root;      // radix root: node index [0,40) | height [40,48)  (atomic)
    uint64_t size;      // bytes
    uint64_t ctime_ns;  // host st_ctim of the last open: reopen validation
    uint64_t mtime_ns;
    uint64_t ino;
    int32_t  host_fd;
    int32_t  flags;     // open flags of the current open
    uint32_t dirty;     // set when a page was written since the last flush
    uint32_t name_hash;
    uint32_t opened;    // successful host opens of the entry (incremented after the state turns OPEN)
    uint32_t pad;
} gfs_file;

// One RPC slot; lives in coherent pinned host memory.  The GPU writes the
// request fields, the host writes `result` and the open results, then sets the
// request's bit in its completion mask.
typedef struct gfs_rpc_slot {
    uint32_t op;
    int32_t  host_fd;
    int32_t  flags;
    uint32_t cq;        // completion word index [0,16) | request tag [16,22): the host sets bit `tag` of cq[index]
    uint64_t offset;
    uint64_t length;
    int64_t  result;    // bytes moved, host fd for OPEN, or -errno
    uint64_t size;      // OPEN: st_size
    uint64_t ctime_ns;  // OPEN: st_ctim
    uint64_t mtime_ns;  // OPEN: st_mtim
    uint64_t ino;       // OPEN: st_ino
    uint64_t pad[7];
    char     name[GFS_NAME_MAX];
} gfs_rpc_slot;         // 256 bytes

typedef struct gfs_counters {
    unsigned long long page_hits, page_misses, page_waits, evictions, alloc_fail;
    unsigned long long rpc_reads, rpc_writes, rpc_opens, rpc_closes;
    unsigned long long bytes_filled, bytes_written;
    unsigned long long timeouts;
    unsigned long long vm_errors;
    unsigned long long slots_lost;   // ring slots quarantined after a timed-out request
    unsigned long long files_reclaimed;  // closed file-table entries taken back for other names
    unsigned long long nodes_recycled;   // radix nodes returned to the free list
    unsigned long long spec_parts;       // workgroups that read a cached file before its reopen completed
    unsigned long long spec_rescans;     // ... and read their range again (a page had left the cache)
    unsigned long long spec_changed;     // ... and found the file changed by the reopen
} gfs_counters;

typedef struct gfs_dev {
    // geometry
    uint32_t log_page;
    uint32_t n_frames;
    uint32_t max_files;
    uint32_t ring_depth;
    uint32_t sq_depth;      // power of two >= ring_depth + max_files (room for every slot's request and one close per file)
    uint32_t name_cap;      // power of two >= 2*max_files
    uint32_t node_cap;
    uint32_t slot_words;    // ceil(ring_depth / 64)
    uint64_t page_size;
    // page cache (HBM)
    uint8_t*  pool;         // n_frames << log_page bytes (+ padding)
    uint64_t* meta;         // per-frame state word
    uint64_t* fkey;         // per-frame owner key: file [40,64) | page [0,40)
    uint32_t* fbytes;       // per-frame valid bytes
    uint8_t*  fref;         // per-frame CLOCK reference bit
    uint64_t* nodes;        // radix nodes, GFS_FANOUT words each; node 0 is null
    gfs_file* files;
    uint32_t* name_index;   // open addressing, value = file index + 1
    uint64_t* slot_free;    // ring slot bitmap (1 = free); word w at [w * GFS_SLOT_WORD_STRIDE]
    // RPC ring (coherent pinned host memory, device-visible pointers)
    gfs_rpc_slot* slots;
    uint64_t* sq;           // submission queue: (ticket << 16) | slot
    uint64_t* cq;           // completion masks, one per slot (a batch of <= 64 requests shares its first slot's)
    uint8_t*  staging;      // ring_depth << log_page bytes (+ padding)
    int64_t*  closeq;       // sq_depth entries: host fd of the close posted with ticket t at [t & (sq_depth - 1)]
    // allocation cursors (device atomics)
    unsigned long long fresh_next;   // frames handed out from the never-used range
    unsigned long long clock_hand;
    unsigned long long node_next;
    unsigned long long files_next;
    unsigned long long sq_tail;
    uint64_t wait_ticks;             // bounded waits: wall-clock ticks before -ETIMEDOUT
    uint32_t tune_flags;             // GFS_TUNE_* (GPUFS_TUNE environment variable at init)
    uint32_t pad1;
    // reclamation (ftable.h)
    uint32_t* file_free;             // max_files entries: [free_taken, free_n) are FREE entries
    unsigned long long free_taken;   // entries taken from file_free (may pass free_n)
    uint32_t free_n;
    uint32_t reclaim_hand;           // next entry a reclaim pass looks at
    uint32_t open_gate;              // name-index users in flight [0,31) | exclusive bit 31
    uint32_t pad2;
    unsigned long long node_free;    // free radix nodes: stack head, tag [40,64) | node [0,40)
    long long dirty_frames;          // frames whose DIRTY bit is set (the writers' write-back throttle)
    gfs_counters ctr;
} gfs_dev;

GFS_HD static inline uint64_t gfs_key(uint32_t file, uint64_t page) {
    return ((uint64_t)file << 40) | (page & 0xFFFFFFFFFFull);
}
