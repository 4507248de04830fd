"""Host-side handle of the serving form (gpufs_serve_* in the C ABI).

Device buffers (file names, dictionary) are torch CUDA tensors; the request
loop itself (arrival schedule, admission, completion) runs in the native
library.
"""
import ctypes
import os

import numpy as np

from . import _lib

STRIDE = 256


def _dev(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t


class Server:
    """Resident request loop (gpufs_serve_*): start, run open-loop batches, stop."""

    def __init__(self, ctx, paths, sizes, dict_words, workgroups=8, slots=16, threads=1024, chunk=65536,
                 frame_budget=0, parts_max=0, min_part_kib=0):
        import torch
        names = np.zeros((len(paths), STRIDE), dtype=np.uint8)
        for i, p in enumerate(paths):
            b = os.fsencode(p)
            if len(b) >= 128:
                raise ValueError(f"path too long: {p}")
            names[i, :len(b)] = np.frombuffer(b, dtype=np.uint8)
        self.names = _dev(torch, names)
        self.dict = _dev(torch, np.ascontiguousarray(dict_words, dtype=np.uint8).reshape(-1))
        self.nw = int(np.asarray(dict_words).shape[0])
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        cfg = _lib.ServeConfig(workgroups, threads, slots, chunk, self.nw, frame_budget, parts_max, min_part_kib)
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        _lib.check("gpufs_serve_start", _lib.lib().gpufs_serve_start(
            ctx.h, ctypes.byref(cfg), ctypes.c_void_p(self.names.data_ptr()), STRIDE, len(paths),
            self.sizes.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.dict.data_ptr()), ctypes.byref(h)))
        self.h = h

    def run(self, file_ids, arrivals):
        ids = np.ascontiguousarray(file_ids, dtype=np.uint32)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        n = ids.size
        counts = np.zeros((n, self.nw), dtype=np.uint32)
        status = np.zeros(n, dtype=np.int32)
        lat = np.zeros(n, dtype=np.float64)
        el = ctypes.c_double()
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        _lib.check("gpufs_serve_run", _lib.lib().gpufs_serve_run(self.h, P(ids), P(arr), n, P(counts), P(status),
                                                                  P(lat), ctypes.byref(el)))
        return counts, status, lat, el.value

    def pages(self):
        """(page hits, page misses) of the requests completed since start."""
        h, m = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check("gpufs_serve_pages", _lib.lib().gpufs_serve_pages(self.h, ctypes.byref(h), ctypes.byref(m)))
        return int(h.value), int(m.value)

    def parts(self):
        """(requests split over several workgroups, their parts) since start."""
        r, p = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.check("gpufs_serve_parts", _lib.lib().gpufs_serve_parts(self.h, ctypes.byref(r), ctypes.byref(p)))
        return int(r.value), int(p.value)

    def stop(self):
        if self.h:
            rc = _lib.lib().gpufs_serve_stop(self.h)
            self.h = None
            _lib.check("gpufs_serve_stop", rc)
