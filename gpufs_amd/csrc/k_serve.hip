// k_serve.hip -- serving form: a resident request loop on the GPU fed by an
// open-loop host driver with continuous admission.
//
// Device: `workgroups` persistent workgroups each take the next request ticket
// from a device counter, wait for the host to publish that ticket in the
// request ring (coherent pinned host memory), run open -> read -> scan -> close
// on the requested file through the page cache (grep_one_file, shared with
// the plain and batched forms), write the per-word counts straight into the
// request's result row in pinned host memory and publish a completion word.
// Page cache and file table stay resident across requests (and across runs
// of the same server), and so do open files: the first request for a file
// keeps one open reference for the server (up to a cap), so later requests
// join the open file without a host round trip and their close is never the
// last one; the server's references are dropped when it stops.
//
// Host (gpufs_serve_run): releases requests at their arrival times, admits
// them FIFO while a request slot is free and the frames they may need fit the
// frame budget (admission by free cache frames), collects completions,
// releases slot and frames, and timestamps each request (arrival -> result).
// Requests arriving while others run are admitted as soon as slots and
// frames free up, so their work overlaps on the device.
#include "grep_scan.h"

#include <chrono>
#include <cstring>
#include <cstdlib>
#include <fcntl.h>
#include <sys/resource.h>
#include <unistd.h>
#include <deque>
#include <new>
#include <thread>
#include <vector>

#define SERVE_STOP 0xFFFFFFFFu
#define SERVE_HOLD_CLAIM 0xFFFFFFFFu   // held[f] while a workgroup opens f for the server

struct serve_req {
    uint64_t seq;       // ticket + 1 once published
    uint32_t file_id;   // SERVE_STOP ends the workgroup
    uint32_t slot;
    uint32_t parts;     // workgroups the request may be split over (1: one)
    uint32_t pad;
};

// An extra part of a split request (device memory), handed from the request's
// part 0 to a waiting workgroup: `ready` = ring position + 1 once published.
struct serve_part {
    uint64_t ready;
    uint64_t seq;       // the request's ticket + 1: its split epoch and completion word
    uint32_t file_id, slot, kf, part;
};
struct serve_part_ctl {
    uint32_t head, tail;   // ring positions taken / reserved
    uint32_t pad[62];
};
// Device-side state of the requests in flight, per request slot.
struct serve_dev {
    grep_split_state* st;   // split state
    uint32_t* row;          // counts, added into by the parts
    int32_t* status;        // written by the part that completes the request
    unsigned int* pages;    // page hits, misses of all its parts
    serve_part* ring;       // extra parts
    uint32_t ring_mask;
    serve_part_ctl* ctl;
    uint64_t min_part;      // bytes: a split request's parts are at least this long
};

#define SERVE_KIND_STOP 0u
#define SERVE_KIND_REQ 1u
#define SERVE_KIND_PART 2u

__global__ __launch_bounds__(1024) void k_serve(const char* names, uint32_t stride, uint32_t nfiles,
                                                const char* dict, uint32_t nw, uint32_t chunk,
                                                serve_req* ring, uint32_t ring_mask, uint32_t* results,
                                                int32_t* res_status, uint32_t* res_pages, uint64_t* done,
                                                unsigned long long* ticket, const uint32_t* halt,
                                                uint64_t claim_ticks, uint32_t* held, unsigned int* held_n,
                                                uint32_t held_cap, serve_dev D, uint32_t* res_parts) {
    __shared__ __attribute__((aligned(16))) grep_lds S;
    __shared__ uint32_t fid, slot, nparts, partj, kind, have_tk;
    __shared__ uint64_t myseq, mytk;
    __shared__ uint32_t h0, m0, hold;
    grep_build_dict(S, dict, nw);
    const uint32_t t = gfs::tid();
    if (t == 0) have_tk = 0;
    // Completion of a split request by the part that finished it: the row is
    // stored into the request's result slot, then its status, page counts and
    // completion word (workgroup-collective).
    auto complete_split = [&](uint32_t s, uint64_t sq, uint32_t kf) {
        for (uint32_t q = t; q < nw; q += gfs::nthreads()) results[(size_t)s * nw + q] = D.row[(size_t)s * nw + q];
        __threadfence_system();
        __syncthreads();
        if (t == 0) {
            gfs::sys_st(&res_pages[2 * (size_t)s], gfs::ld_rlx(&D.pages[2 * (size_t)s]));
            gfs::sys_st(&res_pages[2 * (size_t)s + 1], gfs::ld_rlx(&D.pages[2 * (size_t)s + 1]));
            gfs::sys_st(&res_status[s], D.status[s]);
            gfs::sys_st(&res_parts[s], kf);
            gfs::fence_rel_sys();
            gfs::sys_st_rel(&done[s], sq);
        }
        __syncthreads();
    };
    // the part's page lookups go to the request's totals before it arrives
    auto add_pages = [&](uint32_t s) {
        if (t == 0) {
            atomicAdd(&D.pages[2 * (size_t)s], gfs_ctl.hits - h0);
            atomicAdd(&D.pages[2 * (size_t)s + 1], gfs_ctl.misses - m0);
        }
    };
    for (;;) {
        if (t == 0) {
            if (!have_tk) {
                mytk = atomicAdd(ticket, 1ull);
                have_tk = 1;
            }
            const unsigned long long k = mytk;
            serve_req* e = &ring[k & ring_mask];
            uint32_t spins = 0;
            kind = SERVE_KIND_STOP;
            // Waits as long as the server runs, however long it idles: a ticket
            // once taken is served.  Meanwhile extra parts of split requests
            // are taken first (they belong to earlier requests).  The wait ends
            // when the request is published or when the host sets the halt word
            // (gpufs_serve_stop, after it has published a stop request for every
            // workgroup).  Relaxed polls (a system-scope acquire load
            // invalidates the caches each time), one acquire when something is
            // there.
            for (;;) {
                const uint32_t ph = gfs::ld_rlx(&D.ctl->head), pt = gfs::ld_rlx(&D.ctl->tail);
                if ((int32_t)(pt - ph) > 0) {
                    uint32_t hh = ph;
                    if (__hip_atomic_compare_exchange_strong(&D.ctl->head, &hh, ph + 1, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        // reserved by a running part 0, published right after it:
                        // bounded wait
                        serve_part* it = &D.ring[ph & D.ring_mask];
                        const uint64_t c0 = gfs::now_ticks();
                        while (gfs::ld_rlx(&it->ready) != (uint64_t)ph + 1 && gfs::now_ticks() - c0 < claim_ticks)
                            __builtin_amdgcn_s_sleep(2);
                        gfs::fence_acq_agent();
                        if (gfs::ld_rlx(&it->ready) == (uint64_t)ph + 1) {
                            kind = SERVE_KIND_PART;
                            fid = it->file_id;
                            slot = it->slot;
                            nparts = it->kf;
                            partj = it->part;
                            myseq = it->seq;
                            break;
                        }
                    }
                    continue;
                }
                if (gfs::sys_ld(&e->seq) == k + 1) {
                    gfs::fence_acq_sys();
                    kind = SERVE_KIND_REQ;
                    fid = gfs::sys_ld(&e->file_id);
                    slot = gfs::sys_ld(&e->slot);
                    nparts = gfs::sys_ld(&e->parts);
                    partj = 0;
                    myseq = k + 1;
                    have_tk = 0;
                    break;
                }
                if (++spins > 64) __builtin_amdgcn_s_sleep(127);
                else __builtin_amdgcn_s_sleep(8);
                if ((spins & 63) == 0 && gfs::sys_ld(halt) != 0u) break;
            }
            h0 = gfs_ctl.hits;
            m0 = gfs_ctl.misses;
        }
        __syncthreads();
        const uint32_t kd = gfs::uni(kind);
        const uint32_t f = kd == SERVE_KIND_STOP ? SERVE_STOP : gfs::uni(fid), s = gfs::uni(slot);
        const uint32_t np = gfs::uni(nparts), pj = gfs::uni(partj);
        const uint64_t sq = gfs::uni(myseq);
        __syncthreads();
        if (kd == SERVE_KIND_PART) {
            // an extra part of another workgroup's request
            const bool fin = grep_file_split(S, names + (size_t)f * stride, nw, chunk, D.row + (size_t)s * nw,
                                             D.status + s, pj, np, D.min_part, D.st + s, np, (uint32_t)sq, true,
                                             false, [](uint32_t) {}, [&] { add_pages(s); });
            if (fin) complete_split(s, sq, np);
            __syncthreads();
            continue;
        }
        if (f == SERVE_STOP) {
            // drop the server's open references of this workgroup's share of the files
            for (uint32_t x = blockIdx.x; x < nfiles; x += gridDim.x) {
                if (t == 0) {
                    // a claim resolves as soon as its workgroup's open returns (bounded wait)
                    const uint64_t c0 = gfs::now_ticks();
                    uint32_t hv;
                    while ((hv = __hip_atomic_load(&held[x], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) ==
                               SERVE_HOLD_CLAIM &&
                           gfs::now_ticks() - c0 < claim_ticks)
                        __builtin_amdgcn_s_sleep(64);
                    hold = hv;
                }
                __syncthreads();
                const uint32_t hv = gfs::uni(hold);
                __syncthreads();
                if (hv != 0 && hv != SERVE_HOLD_CLAIM) {
                    gclose((int)hv - 1);
                    if (t == 0) __hip_atomic_store(&held[x], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            break;
        }
        if (f < nfiles && held_cap) {
            // the first request for a file keeps an open reference for the server
            if (t == 0) {
                hold = 0;
                uint32_t exp = 0;
                if (__hip_atomic_load(&held[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                    __hip_atomic_load(held_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < held_cap &&
                    __hip_atomic_compare_exchange_strong(&held[f], &exp, SERVE_HOLD_CLAIM, __ATOMIC_ACQ_REL,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    if (atomicAdd(held_n, 1u) < held_cap) hold = 1;
                    else {
                        atomicSub(held_n, 1u);
                        __hip_atomic_store(&held[f], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
            __syncthreads();
            if (gfs::uni(hold)) {
                const int hfd = gopen(names + (size_t)f * stride, O_GRDONLY);
                if (t == 0) {
                    if (hfd < 0) atomicSub(held_n, 1u);
                    __hip_atomic_store(&held[f], hfd >= 0 ? (uint32_t)hfd + 1u : 0u, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (f < nfiles && np > 1) {
            // part 0 of a split request: it hands the file's other parts to
            // waiting workgroups through the part ring; whichever part
            // finishes last completes the request
            if (t == 0) {
                __hip_atomic_store(&D.pages[2 * (size_t)s], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&D.pages[2 * (size_t)s + 1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            uint32_t kfr = 1;
            const bool fin = grep_file_split(
                S, names + (size_t)f * stride, nw, chunk, D.row + (size_t)s * nw, D.status + s, 0, np, D.min_part,
                D.st + s, 0, (uint32_t)sq, true, false,
                [&](uint32_t kf) {
                    kfr = kf;
                    if (t != 0 || kf < 2) return;
                    const uint32_t r0 = atomicAdd(&D.ctl->tail, kf - 1);
                    for (uint32_t x = 1; x < kf; ++x) {
                        const uint32_t pos = r0 + x - 1;
                        serve_part* it = &D.ring[pos & D.ring_mask];
                        it->seq = sq;
                        it->file_id = f;
                        it->slot = s;
                        it->kf = kf;
                        it->part = x;
                        __hip_atomic_store(&it->ready, (uint64_t)pos + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    }
                },
                [&] { add_pages(s); });
            if (fin) complete_split(s, sq, kfr);
            __syncthreads();
            continue;
        }
        int st = f < nfiles ? grep_one_file(S, names + (size_t)f * stride, nw, chunk, results + (size_t)s * nw)
                            : -EINVAL;
        __threadfence_system();
        __syncthreads();
        if (t == 0) {
            gfs::sys_st(&res_pages[2 * (size_t)s], gfs_ctl.hits - h0);
            gfs::sys_st(&res_pages[2 * (size_t)s + 1], gfs_ctl.misses - m0);
            gfs::sys_st(&res_status[s], st);
            gfs::sys_st(&res_parts[s], 1u);
            gfs::fence_rel_sys();
            gfs::sys_st_rel(&done[s], sq);
        }
        // barrier between the two thread-0 regions of consecutive iterations
        // (see k_grep_batched)
        __syncthreads();
    }
}

struct gpufs_server {
    gpufs_ctx* ctx;
    hipStream_t stream;
    uint32_t workgroups, threads, slots, nw, chunk, nfiles, stride;
    uint64_t frame_budget, page_size;
    std::vector<uint64_t> sizes;
    serve_req* ring;     // host coherent
    uint32_t ring_mask;
    uint32_t* results;   // host coherent: slots x nw
    int32_t* rstatus;    // host coherent: slots
    uint32_t* rpages;    // host coherent: slots x 2 (page hits, misses of the request)
    uint64_t* done;      // host coherent: slots
    uint32_t* halt;      // host coherent: set by gpufs_serve_stop, ends every workgroup's request wait
    uint32_t* rparts;    // host coherent: slots (workgroups the request ran on)
    uint64_t hits, misses;   // summed over completed requests
    uint64_t split_requests, split_parts;   // requests run on more than one workgroup, and their parts
    uint32_t parts_max;
    uint64_t min_part;
    serve_dev D;         // device: per-slot split state, rows, page counts; the part ring
    unsigned long long* ticket;   // device
    uint32_t* held;               // device, per file: 0, SERVE_HOLD_CLAIM, or the server's open fd + 1
    unsigned int* held_n;         // device: files held open by the server
    uint64_t next_ticket;
    bool running;
};

// Files the server keeps open at most: about half the process's descriptor
// limit (the host service opens one descriptor per open file), at most 4096;
// GPUFS_SERVE_HOLD sets it explicitly.
// Grow the process's descriptor table once, up front, to cover `fds` more
// descriptors: the kernel grows it by doubling, and in a threaded process each
// growth waits for an RCU grace period (about 10-20 ms measured), during which
// every host open queues.  A server that keeps files open would otherwise hit
// those stalls in the middle of requests.
static void serve_reserve_fds(uint32_t fds) {
    if (fds == 0) return;
    struct rlimit rl;
    if (getrlimit(RLIMIT_NOFILE, &rl) != 0) return;
    const uint64_t soft = rl.rlim_cur == RLIM_INFINITY ? 1ull << 20 : (uint64_t)rl.rlim_cur;
    const uint64_t want = 2ull * fds + 1024;
    const int target = (int)(want < soft ? want : soft - 1);
    const int fd = ::open("/dev/null", O_RDONLY | O_CLOEXEC);
    if (fd < 0) return;
    if (target > fd && ::dup2(fd, target) == target) ::close(target);
    ::close(fd);
}

static uint32_t serve_hold_cap() {
    if (const char* e = getenv("GPUFS_SERVE_HOLD")) {   // explicit cap (0: the server holds nothing)
        const long v = strtol(e, nullptr, 10);
        if (v >= 0) return (uint32_t)(v < 4096 ? v : 4096);
    }
    struct rlimit rl;
    if (getrlimit(RLIMIT_NOFILE, &rl) != 0) return 0;
    const uint64_t soft = rl.rlim_cur == RLIM_INFINITY ? 1ull << 20 : (uint64_t)rl.rlim_cur;
    const uint64_t cap = soft > 256 ? soft / 2 - 64 : 0;
    return (uint32_t)(cap < 4096 ? cap : 4096);
}

void gpufs_resident_add(gpufs_ctx* c, int d);

static uint64_t serve_now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void serve_publish(gpufs_server* S, uint32_t file_id, uint32_t slot, uint32_t parts = 1) {
    uint64_t k = S->next_ticket++;
    serve_req* e = &S->ring[k & S->ring_mask];
    e->file_id = file_id;
    e->slot = slot;
    e->parts = parts;
    __atomic_store_n(&e->seq, k + 1, __ATOMIC_RELEASE);
}

static void serve_free(gpufs_server* S) {
    (void)hipHostFree(S->ring);
    (void)hipHostFree(S->results);
    (void)hipHostFree(S->rstatus);
    (void)hipHostFree(S->rpages);
    (void)hipHostFree(S->done);
    (void)hipHostFree(S->halt);
    (void)hipHostFree(S->rparts);
    (void)hipFree(S->D.st);
    (void)hipFree(S->D.row);
    (void)hipFree(S->D.status);
    (void)hipFree(S->D.pages);
    (void)hipFree(S->D.ring);
    (void)hipFree(S->D.ctl);
    (void)hipFree(S->ticket);
    (void)hipFree(S->held);
    (void)hipFree(S->held_n);
    delete S;
}

extern "C" int gpufs_serve_start(gpufs_ctx* ctx, const gpufs_serve_config* cfg, const char* names_dev,
                                 uint32_t name_stride, uint32_t n_files, const uint64_t* file_sizes,
                                 const char* dict_dev, gpufs_server** out) {
    if (!ctx || !cfg || !names_dev || !file_sizes || !dict_dev || !out || n_files == 0) return -EINVAL;
    if (cfg->workgroups == 0 || cfg->slots == 0 || cfg->slots > 65536) return -EINVAL;
    int v = grep_args_ok(n_files, cfg->n_words, cfg->chunk_bytes, cfg->threads);
    if (v) return v;
    if ((v = gpufs_bind(ctx))) return v;
    gpufs_config cc;
    gpufs_get_config(ctx, &cc);
    gpufs_server* S = new (std::nothrow) gpufs_server();
    if (!S) return -ENOMEM;
    S->ctx = ctx;
    S->stream = (hipStream_t)gpufs_stream(ctx);
    S->workgroups = cfg->workgroups;
    S->threads = cfg->threads;
    S->slots = cfg->slots;
    S->nw = cfg->n_words;
    S->chunk = cfg->chunk_bytes;
    S->nfiles = n_files;
    S->stride = name_stride;
    S->page_size = cc.page_size;
    const uint64_t frames = cc.cache_bytes / cc.page_size;
    S->frame_budget = cfg->frame_budget ? cfg->frame_budget : frames - frames / 10;
    S->sizes.assign(file_sizes, file_sizes + n_files);
    S->parts_max = cfg->parts_max ? cfg->parts_max : 8;
    S->min_part = (uint64_t)(cfg->min_part_kib ? cfg->min_part_kib : 512) << 10;
    if (S->parts_max > 64 || S->min_part < 4096) {
        delete S;
        return -EINVAL;
    }
    uint32_t rs = 1;
    while (rs < 4 * (cfg->slots + cfg->workgroups)) rs <<= 1;
    S->ring_mask = rs - 1;
    // the part ring holds every extra part of the requests in flight
    uint32_t ps = 1;
    while (ps < cfg->slots * (S->parts_max - 1) + 1) ps <<= 1;
    const unsigned hf = hipHostMallocMapped | hipHostMallocCoherent;
    bool ok = hipHostMalloc((void**)&S->ring, sizeof(serve_req) * rs, hf) == hipSuccess &&
              hipHostMalloc((void**)&S->results, sizeof(uint32_t) * (size_t)cfg->slots * cfg->n_words, hf) == hipSuccess &&
              hipHostMalloc((void**)&S->rstatus, sizeof(int32_t) * cfg->slots, hf) == hipSuccess &&
              hipHostMalloc((void**)&S->rpages, sizeof(uint32_t) * 2 * (size_t)cfg->slots, hf) == hipSuccess &&
              hipHostMalloc((void**)&S->done, sizeof(uint64_t) * cfg->slots, hf) == hipSuccess &&
              hipHostMalloc((void**)&S->halt, sizeof(uint32_t), hf) == hipSuccess &&
              hipHostMalloc((void**)&S->rparts, sizeof(uint32_t) * cfg->slots, hf) == hipSuccess &&
              hipMalloc((void**)&S->D.st, sizeof(grep_split_state) * cfg->slots) == hipSuccess &&
              hipMalloc((void**)&S->D.row, sizeof(uint32_t) * (size_t)cfg->slots * cfg->n_words) == hipSuccess &&
              hipMalloc((void**)&S->D.status, sizeof(int32_t) * cfg->slots) == hipSuccess &&
              hipMalloc((void**)&S->D.pages, sizeof(unsigned int) * 2 * (size_t)cfg->slots) == hipSuccess &&
              hipMalloc((void**)&S->D.ring, sizeof(serve_part) * ps) == hipSuccess &&
              hipMalloc((void**)&S->D.ctl, sizeof(serve_part_ctl)) == hipSuccess &&
              hipMalloc((void**)&S->ticket, sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc((void**)&S->held, sizeof(uint32_t) * (size_t)n_files) == hipSuccess &&
              hipMalloc((void**)&S->held_n, sizeof(unsigned int)) == hipSuccess;
    if (!ok) {
        serve_free(S);
        return -ENOMEM;
    }
    memset(S->ring, 0, sizeof(serve_req) * rs);
    memset(S->done, 0, sizeof(uint64_t) * cfg->slots);
    memset(S->rparts, 0, sizeof(uint32_t) * cfg->slots);
    *S->halt = 0;
    gfs_dev hd;
    S->D.ring_mask = ps - 1;
    S->D.min_part = S->min_part;
    if (hipMemsetAsync(S->D.st, 0, sizeof(grep_split_state) * cfg->slots, S->stream) != hipSuccess ||
        hipMemsetAsync(S->D.ring, 0, sizeof(serve_part) * ps, S->stream) != hipSuccess ||
        hipMemsetAsync(S->D.ctl, 0, sizeof(serve_part_ctl), S->stream) != hipSuccess ||
        hipMemsetAsync(S->ticket, 0, sizeof(unsigned long long), S->stream) != hipSuccess ||
        hipMemsetAsync(S->held, 0, sizeof(uint32_t) * (size_t)n_files, S->stream) != hipSuccess ||
        hipMemsetAsync(S->held_n, 0, sizeof(unsigned int), S->stream) != hipSuccess ||
        hipMemcpyAsync(&hd, gpufs_device_state(ctx), sizeof(hd), hipMemcpyDeviceToHost, S->stream) != hipSuccess ||
        hipStreamSynchronize(S->stream) != hipSuccess) {
        serve_free(S);
        return -EIO;
    }
    S->next_ticket = 0;
    gpufs_host_start(ctx);
    const uint32_t hold_cap = serve_hold_cap();
    serve_reserve_fds(hold_cap);
    // at stop, a workgroup waits at most the RPC wait limit for another one's
    // open of a held file to resolve
    hipLaunchKernelGGL(k_serve, dim3(S->workgroups), dim3(S->threads), 0, S->stream, names_dev, name_stride,
                       n_files, dict_dev, cfg->n_words, cfg->chunk_bytes, S->ring, S->ring_mask, S->results,
                       S->rstatus, S->rpages, S->done, S->ticket, (const uint32_t*)S->halt, hd.wait_ticks, S->held,
                       S->held_n, hold_cap, S->D, S->rparts);
    if (hipGetLastError() != hipSuccess) {
        serve_free(S);
        return -EIO;
    }
    S->running = true;
    gpufs_resident_add(ctx, 1);
    *out = S;
    return 0;
}

// Open-loop run: request i arrives arrival_s[i] seconds after the start
// (non-decreasing).  Fills counts_out (n x n_words), status_out, latency_s_out
// (arrival -> completion observed by the host); *elapsed_s_out = first
// arrival to last completion.
extern "C" int gpufs_serve_run(gpufs_server* S, const uint32_t* file_ids, const double* arrival_s,
                               uint32_t n, uint32_t* counts_out, int32_t* status_out, double* latency_s_out,
                               double* elapsed_s_out) {
    if (!S || !S->running || !file_ids || !arrival_s || !counts_out || !status_out || !latency_s_out) return -EINVAL;
    std::vector<int64_t> slot_req(S->slots, -1);
    std::vector<uint64_t> slot_frames(S->slots, 0), slot_seq(S->slots, 0);
    std::vector<uint32_t> free_slots;
    for (uint32_t s = S->slots; s-- > 0;) free_slots.push_back(s);
    std::deque<uint32_t> pending;
    uint64_t reserved = 0;
    uint32_t next_arrival = 0, completed = 0, inflight = 0;
    const uint64_t t0 = serve_now_ns();
    auto frames_for = [&](uint32_t i) -> uint64_t {
        uint32_t f = file_ids[i];
        uint64_t sz = f < S->nfiles ? S->sizes[f] : 0;
        uint64_t p = (sz + S->page_size - 1) / S->page_size;
        return p < 1 ? 1 : p;
    };
    // a file of at least two minimum parts is split over up to parts_max
    // workgroups (its part 0 decides from the size it opens: the same count
    // unless the file changed)
    auto parts_for = [&](uint32_t i) -> uint32_t {
        uint32_t f = file_ids[i];
        uint64_t k = (f < S->nfiles ? S->sizes[f] : 0) / S->min_part;
        return k < 2 ? 1u : (uint32_t)(k < S->parts_max ? k : S->parts_max);
    };
    uint64_t last_done = t0;
    uint64_t guard_t = t0;
    uint64_t polls = 0;
    while (completed < n) {
        const uint64_t now = serve_now_ns();
        while (next_arrival < n && (double)(now - t0) * 1e-9 >= arrival_s[next_arrival]) pending.push_back(next_arrival++);
        // completions
        for (uint32_t s = 0; s < S->slots; ++s) {
            if (slot_req[s] < 0) continue;
            if (__atomic_load_n(&S->done[s], __ATOMIC_ACQUIRE) != slot_seq[s]) continue;
            const uint32_t i = (uint32_t)slot_req[s];
            const uint64_t tc = serve_now_ns();
            memcpy(counts_out + (size_t)i * S->nw, S->results + (size_t)s * S->nw, sizeof(uint32_t) * S->nw);
            status_out[i] = S->rstatus[s];
            S->hits += S->rpages[2 * (size_t)s];
            S->misses += S->rpages[2 * (size_t)s + 1];
            if (S->rparts[s] > 1) {
                ++S->split_requests;
                S->split_parts += S->rparts[s];
            }
            latency_s_out[i] = (double)(tc - t0) * 1e-9 - arrival_s[i];
            last_done = tc;
            reserved -= slot_frames[s];
            slot_req[s] = -1;
            free_slots.push_back(s);
            ++completed;
            --inflight;
            guard_t = tc;
        }
        // admission: FIFO, a free slot and the frame budget
        while (!pending.empty() && !free_slots.empty()) {
            const uint32_t i = pending.front();
            const uint64_t need = frames_for(i);
            if (inflight > 0 && reserved + need > S->frame_budget) break;
            pending.pop_front();
            const uint32_t s = free_slots.back();
            free_slots.pop_back();
            slot_req[s] = i;
            slot_frames[s] = need;
            reserved += need;
            slot_seq[s] = S->next_ticket + 1;
            serve_publish(S, file_ids[i], s, parts_for(i));
            ++inflight;
        }
        if (inflight > 0 && serve_now_ns() - guard_t > 120ull * 1000000000ull) return -ETIMEDOUT;
        // the server's workgroups never leave while it runs: a finished (or
        // failed) kernel means no request can be served any more
        if (inflight > 0 && (++polls & 1023) == 0 && hipStreamQuery(S->stream) != hipErrorNotReady) return -EIO;
        if (inflight == 0) guard_t = serve_now_ns();
    }
    if (elapsed_s_out) *elapsed_s_out = (double)(last_done - t0) * 1e-9 - (n ? arrival_s[0] : 0.0);
    return 0;
}

extern "C" int gpufs_serve_stop(gpufs_server* S) {
    if (!S) return -EINVAL;
    int rc = 0;
    if (S->running) {
        for (uint32_t w = 0; w < S->workgroups; ++w) serve_publish(S, SERVE_STOP, 0);
        // every waiting workgroup also sees the halt word (it ends the wait of a
        // ticket no stop request was published for)
        __atomic_store_n(S->halt, 1u, __ATOMIC_RELEASE);
        // bounded wait for the workgroups to leave
        const uint64_t t0 = serve_now_ns();
        hipError_t e;
        while ((e = hipStreamQuery(S->stream)) == hipErrorNotReady) {
            if (serve_now_ns() - t0 > 60ull * 1000000000ull) return -ETIMEDOUT;
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        if (e != hipSuccess) rc = -EIO;
        S->running = false;
        gpufs_resident_add(S->ctx, -1);
    }
    serve_free(S);
    return rc;
}

extern "C" int gpufs_serve_pages(gpufs_server* S, uint64_t* page_hits, uint64_t* page_misses) {
    if (!S || !page_hits || !page_misses) return -EINVAL;
    *page_hits = S->hits;
    *page_misses = S->misses;
    return 0;
}

extern "C" int gpufs_serve_parts(gpufs_server* S, uint64_t* split_requests, uint64_t* parts) {
    if (!S || !split_requests || !parts) return -EINVAL;
    *split_requests = S->split_requests;
    *parts = S->split_parts;
    return 0;
}
