"""Autotuner for the batched form: page size, workgroup count / size and RPC
ring depth are chosen by measuring the real workload on a sample of its files.

Search: a sweep of a small grid (page size x threads per workgroup x
workgroups x ring depth x parts per file), each point run cold (page cache
emptied, host page cache warm) on the same file sample; the point with the
highest files/s wins.  Parts per file only change the run when the sample
holds files large enough to split, so that axis is swept on the best point of
the others.  Every point builds a fresh context (the page size is fixed per
context).
"""
import ctypes
import itertools
import time

import numpy as np

from . import _lib

DEFAULT_SPACE = {
    "page_size": [64 << 10, 256 << 10, 1 << 20],
    "threads": [512, 1024],
    "workgroups_per_cu": [1],          # the scan's LDS holds one workgroup per compute unit
    "ring_depth": [256, 1024],
    "parts_per_file": [1, 4, 16],      # workgroups a large file may be split over
}


def _device_cus():
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count


def batched_trial(cfg, names_d, n_files, dict_d, n_words, chunk, cache_bytes, host_threads, torch, runs=2):
    """Cold runs of the batched scan with configuration `cfg` (page cache
    emptied before each, the best kept: a single short run is noisy enough to
    mislead the choice); returns (files/s, seconds)."""
    L = _lib.lib()
    ctx = _lib.Context(cache_bytes=cache_bytes, page_size=cfg["page_size"], ring_depth=cfg["ring_depth"],
                       host_threads=host_threads, max_files=max(1024, 2 * n_files))
    try:
        counts = torch.zeros(n_files * n_words, dtype=torch.int32, device="cuda")
        status = torch.zeros(n_files, dtype=torch.int32, device="cuda")
        parts = cfg.get("parts_per_file", 8)
        wsb = int(L.gpufs_grep_batched_workspace_bytes(n_files, parts))
        ws = torch.zeros((wsb + 7) // 8, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        ctx.host_start()
        stream = torch.cuda.ExternalStream(ctx.stream)
        dt, ok = None, True
        for r in range(runs):
            if r:
                ctx.cache_drop()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            rc = L.gpufs_launch_grep_batched(ctx.h, ctypes.c_void_p(names_d.data_ptr()), 256, n_files,
                                             ctypes.c_void_p(dict_d.data_ptr()), n_words, chunk,
                                             ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                             ctypes.c_void_p(ws.data_ptr()), cfg["workgroups"], cfg["threads"], parts,
                                             0, None, None)
            _lib.check("gpufs_launch_grep_batched", rc)
            e1.record(stream)
            e1.synchronize()
            t = e0.elapsed_time(e1) / 1e3
            dt = t if dt is None else min(dt, t)
            ok = ok and bool((status.cpu().numpy() == 0).all())
        ctx.host_stop()
    finally:
        ctx.close()
    return (n_files / dt if ok else 0.0), dt


def tune_batched(names_d, n_files, dict_d, n_words, chunk, cache_bytes, host_threads=16, space=None, log=None):
    """Sweep `space` (see DEFAULT_SPACE) on the given files.  Returns
    (best_cfg, table) where table lists every point with its files/s."""
    import torch
    space = dict(DEFAULT_SPACE, **(space or {}))
    cus = _device_cus()
    table = []
    best = None
    def trial(cfg):
        t = time.time()
        fps, dt = batched_trial(cfg, names_d, n_files, dict_d, n_words, chunk, cache_bytes, host_threads, torch)
        row = dict(cfg, files_per_s=round(fps, 1), kernel_s=round(dt, 5))
        table.append(row)
        if log:
            log(f"autotune {row} ({time.time() - t:.1f}s)")
        return fps

    p0 = space["parts_per_file"][len(space["parts_per_file"]) // 2]
    for page, threads, wpc, ring in itertools.product(space["page_size"], space["threads"],
                                                      space["workgroups_per_cu"], space["ring_depth"]):
        if page * ring > (1 << 30):   # pinned staging kept at or below 1 GiB
            continue
        cfg = {"page_size": page, "threads": threads, "workgroups": int(wpc * cus), "ring_depth": ring,
               "parts_per_file": p0}
        fps = trial(cfg)
        if best is None or fps > best[1]:
            best = (cfg, fps)
    for parts in space["parts_per_file"]:
        if parts == p0:
            continue
        cfg = dict(best[0], parts_per_file=parts)
        fps = trial(cfg)
        if fps > best[1]:
            best = (cfg, fps)
    return best[0], table
