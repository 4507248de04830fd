"""Helpers for the GPU parity tests: device buffers (torch), launch wrappers
around the C ABI, and context construction.  Every call goes through
libgpufs_amd.so; a missing library or GPU is an error, not a skip."""
import ctypes
import os
import sys

import numpy as np

from gpufs_amd import _lib

STRIDE = 256


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def say(msg):
    print(f"[gpu-step] {msg}", flush=True)
    sys.stdout.flush()


def dev_bytes(data):
    torch = torch_cuda()
    a = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data)
    t = torch.from_numpy(a.copy()).to("cuda")
    torch.cuda.synchronize()
    return t


def dev_names(paths):
    torch = torch_cuda()
    a = np.zeros((len(paths), STRIDE), dtype=np.uint8)
    for i, p in enumerate(paths):
        b = os.fsencode(p)
        assert len(b) < 128
        a[i, :len(b)] = np.frombuffer(b, dtype=np.uint8)
    t = torch.from_numpy(a).to("cuda")
    torch.cuda.synchronize()
    return t


def dev_zeros(n, dtype=None):
    torch = torch_cuda()
    t = torch.zeros(n, dtype=dtype or torch.uint8, device="cuda")
    torch.cuda.synchronize()
    return t


# device waits give up after 10 s in the tests, so a stalled request surfaces as
# -ETIMEDOUT (and the stats' `timeouts` counter) instead of a long wait
os.environ.setdefault("GPUFS_WAIT_TIMEOUT_S", "10")


def make_ctx(**cfg):
    cfg.setdefault("cache_bytes", 64 << 20)
    cfg.setdefault("page_size", 64 << 10)
    cfg.setdefault("ring_depth", 128)
    cfg.setdefault("host_threads", 4)
    cfg.setdefault("max_files", 4096)
    return _lib.Context(**cfg)


def finish(ctx):
    """Serve RPCs until the context stream is idle (reference run_gpufs_handler)."""
    ctx.run_handler()
    torch_cuda().cuda.synchronize()


def read_file(ctx, path, off, length, flags=os.O_RDONLY, threads=256, dst_pad=0):
    """One workgroup: gopen, gread(off, length) into a device buffer at byte
    offset dst_pad, fstat, gclose.  Returns (ret, size, close_rc, bytes)."""
    torch = torch_cuda()
    name = dev_names([path])
    dst = dev_zeros(length + dst_pad + 16)
    res = dev_zeros(4, torch.int64)
    rc = _lib.lib().gpufs_launch_read(ctx.h, ctypes.c_void_p(name.data_ptr()), flags, off, length,
                                      ctypes.c_void_p(dst.data_ptr() + dst_pad), ctypes.c_void_p(res.data_ptr()),
                                      threads, None)
    _lib.check("gpufs_launch_read", rc)
    finish(ctx)
    r = res.cpu().numpy()
    got = int(r[0])
    data = dst.cpu().numpy()[dst_pad:dst_pad + max(got, 0)].tobytes()
    return got, int(r[1]), int(r[2]), data


def write_file(ctx, path, off, data, flags, fsync=False, threads=256):
    torch = torch_cuda()
    name = dev_names([path])
    src = dev_bytes(data)
    res = dev_zeros(4, torch.int64)
    rc = _lib.lib().gpufs_launch_write(ctx.h, ctypes.c_void_p(name.data_ptr()), flags, off, len(data),
                                       ctypes.c_void_p(src.data_ptr()), 1 if fsync else 0,
                                       ctypes.c_void_p(res.data_ptr()), threads, None)
    _lib.check("gpufs_launch_write", rc)
    finish(ctx)
    return [int(x) for x in res.cpu().numpy()]


def gread_ranges(ctx, path, n_ranges, range_bytes, threads=1024, dst=None):
    torch = torch_cuda()
    name = dev_names([path])
    if dst is None:
        dst = dev_zeros(n_ranges * range_bytes)
    out = dev_zeros(n_ranges, torch.int64)
    rc = _lib.lib().gpufs_launch_gread_ranges(ctx.h, ctypes.c_void_p(name.data_ptr()), n_ranges, range_bytes,
                                              ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                              threads, None)
    _lib.check("gpufs_launch_gread_ranges", rc)
    finish(ctx)
    return out.cpu().numpy(), dst


def gwrite_ranges(ctx, path, flags, n_ranges, range_bytes, src, fsync=True, threads=1024):
    """gpufs_launch_gwrite_ranges: workgroup w writes src[w*range:(w+1)*range]
    at w*range.  `src` is a device uint8 tensor.  Returns the per-workgroup
    results."""
    torch = torch_cuda()
    name = dev_names([path])
    out = dev_zeros(n_ranges, torch.int64)
    rc = _lib.lib().gpufs_launch_gwrite_ranges(ctx.h, ctypes.c_void_p(name.data_ptr()), flags, n_ranges,
                                               range_bytes, ctypes.c_void_p(src.data_ptr()), 1 if fsync else 0,
                                               ctypes.c_void_p(out.data_ptr()), threads, None)
    _lib.check("gpufs_launch_gwrite_ranges", rc)
    finish(ctx)
    return out.cpu().numpy()


def grep(ctx, paths, dict_words, chunk, threads=1024, parts=None, min_part=0):
    """gpufs_launch_grep, or gpufs_launch_grep_split when `parts` is given.
    The count rows start as garbage: the call writes them entirely."""
    torch = torch_cuda()
    names = dev_names(paths)
    d = dev_bytes(np.ascontiguousarray(dict_words, dtype=np.uint8).reshape(-1))
    nw = int(np.asarray(dict_words).shape[0])
    counts = torch.full((len(paths) * nw,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    status = torch.full((len(paths),), -12345, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    L = _lib.lib()
    common = (ctx.h, ctypes.c_void_p(names.data_ptr()), STRIDE, len(paths), ctypes.c_void_p(d.data_ptr()), nw, chunk,
              ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(status.data_ptr()))
    if parts is None:
        rc = L.gpufs_launch_grep(*common, threads, None)
    else:
        rc = L.gpufs_launch_grep_split(*common, parts, min_part, threads, None)
    _lib.check("gpufs_launch_grep", rc)
    finish(ctx)
    return counts.cpu().numpy().view(np.uint32).reshape(len(paths), nw), status.cpu().numpy()


def gemv_mmap(ctx, path, rows, cols, x, rows_per_wg, threads=1024):
    torch = torch_cuda()
    name = dev_names([path])
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()
    y = torch.zeros(rows, dtype=torch.float32, device="cuda")
    nwg = (rows + rows_per_wg - 1) // rows_per_wg
    status = dev_zeros(nwg, torch.int32)
    csum = dev_zeros(nwg, torch.int64)
    torch.cuda.synchronize()
    rc = _lib.lib().gpufs_launch_gemv_mmap(ctx.h, ctypes.c_void_p(name.data_ptr()), rows, cols,
                                           ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                           rows_per_wg, ctypes.c_void_p(status.data_ptr()),
                                           ctypes.c_void_p(csum.data_ptr()), threads, None)
    _lib.check("gpufs_launch_gemv_mmap", rc)
    finish(ctx)
    return y.cpu().numpy(), status.cpu().numpy(), csum.cpu().numpy().view(np.uint64)


def batched_workspace(nfiles, parts):
    torch = torch_cuda()
    n = int(_lib.lib().gpufs_grep_batched_workspace_bytes(nfiles, parts))
    return dev_zeros((n + 7) // 8, torch.int64)


def grep_batched(ctx, names_t, nfiles, dict_t, nw, chunk, counts, status, ws, workgroups, threads=1024,
                 graph=False, parts=8, min_part=0, times=None):
    """names_t/dict_t/counts/status/ws: device tensors (ws from batched_workspace).
    Returns a graph handle when graph=True (caller replays and destroys it)."""
    L = _lib.lib()
    args = (ctx.h, ctypes.c_void_p(names_t.data_ptr()), STRIDE, nfiles, ctypes.c_void_p(dict_t.data_ptr()), nw,
            chunk, ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(status.data_ptr()),
            ctypes.c_void_p(ws.data_ptr()), workgroups, threads, parts, min_part,
            ctypes.c_void_p(times.data_ptr()) if times is not None else None)
    if graph:
        h = ctypes.c_void_p()
        _lib.check("gpufs_grep_batched_graph", L.gpufs_grep_batched_graph(*args, ctypes.byref(h)))
        return h
    _lib.check("gpufs_launch_grep_batched", L.gpufs_launch_grep_batched(*args, None))
    return None


from gpufs_amd.serve import Server  # noqa: E402,F401
