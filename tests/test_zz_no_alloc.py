"""gpufs_launch_grep_noalloc (the split grep scan with its per-file state in
the library's fixed device array instead of an allocation) gives the same
counts and statuses as gpufs_launch_grep_split, on two calls in a row, the
second over fewer files, so it also starts from the state the first left."""
import ctypes
import errno

import numpy as np
import pytest

from gpufs_amd import _lib, datagen as dg
from tests.gpu_helpers import STRIDE, dev_bytes, dev_names, make_ctx, say, torch_cuda

pytestmark = pytest.mark.gpu


def _run(ctx, fn, paths, q, chunk, parts, min_part):
    torch = torch_cuda()
    names = dev_names(paths)
    d = dev_bytes(np.ascontiguousarray(q, dtype=np.uint8).reshape(-1))
    nw = q.shape[0]
    counts = torch.full((len(paths) * nw,), 0x3C3C3C3C, dtype=torch.int32, device="cuda")
    status = torch.full((len(paths),), -777, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(_lib.lib(), fn)(ctx.h, ctypes.c_void_p(names.data_ptr()), STRIDE, len(paths),
                                 ctypes.c_void_p(d.data_ptr()), nw, chunk, ctypes.c_void_p(counts.data_ptr()),
                                 ctypes.c_void_p(status.data_ptr()), parts, min_part, 512, None)
    _lib.check(fn, rc)
    ctx.run_handler()
    torch.cuda.synchronize()
    return counts.cpu().numpy().view(np.uint32).reshape(len(paths), nw), status.cpu().numpy()


def test_noalloc_matches_split_twice(tmp_path, orc):
    say("test_noalloc_matches_split_twice")
    voc = dg.vocab(3000, seed=91)
    q = np.ascontiguousarray(voc[dg.query(3000, 512, seed=92)])
    sizes = [300000, 1, 0, 70001, 200000, 5000]
    paths = [str(tmp_path / f"n{i}.txt") for i in range(len(sizes))]
    dg.text_files(paths, sizes, 93, voc, threads=4)
    cpu, cst = orc.scan_files(paths, q, 4096, threads=4)
    assert (cst == 0).all()
    with make_ctx(page_size=65536, cache_bytes=32 << 20) as ctx:
        for sel, parts in ((slice(None), 4), (slice(0, 3), 3)):   # the second call: fewer files
            p = paths[sel]
            a, sa = _run(ctx, "gpufs_launch_grep_split", p, q, 4096, parts, 4096)
            b, sb = _run(ctx, "gpufs_launch_grep_noalloc", p, q, 4096, parts, 4096)
            assert (sa == 0).all() and (sb == 0).all(), (sa, sb)
            assert (a == b).all()
            assert (b == cpu[sel]).all()
        # beyond the array's capacity: a clear error, nothing launched
        L = _lib.lib()
        rc = L.gpufs_launch_grep_noalloc(ctx.h, ctypes.c_void_p(1), STRIDE, 65537, ctypes.c_void_p(1), 512, 4096,
                                         ctypes.c_void_p(1), ctypes.c_void_p(1), 1, 0, 512, None)
        assert rc == -errno.E2BIG
        assert ctx.check_invariants() == 0
