"""Wave-level mapping (gmmap_warp / gmunmap_warp): every wave maps fixed-size
records, several records share a page and waves of one workgroup map
concurrently; the per-record sums of 8-byte words must equal the file's, with
a cache small enough to evict, and every pin must be returned."""
import ctypes

import numpy as np
import pytest

from gpufs_amd import _lib
from tests.gpu_helpers import dev_names, dev_zeros, finish, make_ctx, say, torch_cuda

pytestmark = pytest.mark.gpu


def _records(ctx, path, rec, nrec, workgroups, threads):
    torch = torch_cuda()
    name = dev_names([path])
    sums = dev_zeros(nrec, torch.int64)
    status = dev_zeros(nrec, torch.int32)
    torch.cuda.synchronize()
    _lib.check("gpufs_launch_mmap_warp_records", _lib.lib().gpufs_launch_mmap_warp_records(
        ctx.h, ctypes.c_void_p(name.data_ptr()), rec, nrec, ctypes.c_void_p(sums.data_ptr()),
        ctypes.c_void_p(status.data_ptr()), workgroups, threads, None))
    finish(ctx)
    return sums.cpu().numpy().view(np.uint64), status.cpu().numpy()


@pytest.mark.parametrize("page_kib,cache_kib,rec,workgroups,threads", [
    (16, 1024, 4096, 8, 256),     # 4 records per page, 64 frames, 32 waves
    (64, 512, 2048, 4, 1024),     # 32 records per page, 8 frames, 64 waves meet on pages
])
def test_mmap_warp_records(tmp_path, page_kib, cache_kib, rec, workgroups, threads):
    say(f"test_mmap_warp_records page={page_kib}K cache={cache_kib}K rec={rec} wg={workgroups}x{threads}")
    nrec = 1000
    data = np.random.default_rng(rec).integers(0, 256, nrec * rec, dtype=np.uint8)
    p = str(tmp_path / "records.bin")
    data.tofile(p)
    want = np.sum(data.view(np.uint64).reshape(nrec, rec // 8), axis=1, dtype=np.uint64)
    with make_ctx(page_size=page_kib << 10, cache_bytes=cache_kib << 10, ring_depth=64) as ctx:
        ctx.host_start()
        for _ in range(2):   # second pass: part hits, part refills
            sums, st = _records(ctx, p, rec, nrec, workgroups, threads)
            assert (st == 0).all(), st[st != 0][:8]
            assert (sums == want).all(), np.nonzero(sums != want)[0][:8]
        s = ctx.stats()
        assert s["page_misses"] > 0 and s["page_hits"] > 0
        ctx.host_stop()
        assert ctx.check_invariants() == 0   # no pinned or half-filled frame left
