"""GPU tests of the device calls outside the read path: the single-thread
calls (single_thread_open / _ftruncate / _fsync / _close), gftruncate
followed by gread, gmsync of a mapped page written in place, gfsync, and
gunlink (not implemented in the reference either).  Each case runs a script
of device calls in one workgroup (gpufs_launch_api_script) and checks the
bytes on disk with the host's own reads."""
import ctypes
import errno
import os

import numpy as np
import pytest

from gpufs_amd import O_GRDONLY, O_GRDWR, _lib
from tests.gpu_helpers import dev_bytes, dev_names, dev_zeros, make_ctx, say, torch_cuda

pytestmark = pytest.mark.gpu

GOPEN, GREAD, GWRITE, GFSYNC, GFTRUNCATE, GCLOSE, FSTAT = 1, 2, 3, 4, 5, 6, 7
ST_OPEN, ST_FTRUNCATE, ST_FSYNC, ST_CLOSE, GMMAP_FILL, GUNLINK = 8, 9, 10, 11, 12, 13


def script(ctx, path, ops, buf=None, threads=256):
    """Runs ops [(op, a, b, c), ...] in one workgroup; returns (results, buffer bytes)."""
    torch = torch_cuda()
    name = dev_names([path])
    o = np.zeros((len(ops), 4), dtype=np.int64)
    for k, rec in enumerate(ops):
        o[k, :len(rec)] = rec
    od = torch.from_numpy(o.reshape(-1)).cuda()
    b = dev_bytes(buf) if buf is not None else dev_zeros(1 << 20)
    res = dev_zeros(len(ops), torch.int64)
    torch.cuda.synchronize()
    _lib.check("gpufs_launch_api_script", _lib.lib().gpufs_launch_api_script(
        ctx.h, ctypes.c_void_p(name.data_ptr()), ctypes.c_void_p(od.data_ptr()), len(ops),
        ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(res.data_ptr()), threads, None))
    ctx.run_handler()
    torch.cuda.synchronize()
    return [int(x) for x in res.cpu().numpy()], b.cpu().numpy()


def rand_bytes(n, seed):
    return np.random.default_rng(seed).integers(0, 256, size=n, dtype=np.uint8)


def test_single_thread_ftruncate_like_grep_text_sample(tmp_path):
    """One thread truncates the output file to 0 (grep_text_gpu.cu:258-268),
    after which gread returns nothing and fstat says 0."""
    say("test_single_thread_ftruncate_like_grep_text_sample")
    p = tmp_path / "out.txt"
    rand_bytes(300000, 1).tofile(str(p))
    with make_ctx(page_size=64 << 10) as ctx:
        r, _ = script(ctx, str(p), [(ST_OPEN, O_GRDWR), (GREAD, 0, 200000, 0), (ST_FTRUNCATE, 0), (FSTAT,),
                                    (GREAD, 0, 4096, 0), (ST_CLOSE,)])
        assert r[0] >= 0 and r[1] == 200000 and r[2] == 0 and r[3] == 0 and r[4] == 0 and r[5] == 0, r
        assert p.stat().st_size == 0
        assert ctx.check_invariants() == 0


def test_gftruncate_then_gread(tmp_path):
    """gftruncate to 0 after the pages were cached: gread reads 0 bytes (the
    cached pages are dropped).  Then to a nonzero size: the prefix survives."""
    say("test_gftruncate_then_gread")
    p = tmp_path / "t.bin"
    data = rand_bytes(500000, 2)
    data.tofile(str(p))
    with make_ctx(page_size=64 << 10) as ctx:
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDWR), (GREAD, 0, 500000, 0), (GFTRUNCATE, 0), (GREAD, 0, 500000, 0),
                                    (FSTAT,), (GCLOSE,)])
        assert r[1] == 500000 and r[2] == 0 and r[3] == 0 and r[4] == 0 and r[5] == 0, r
        assert p.stat().st_size == 0
        data.tofile(str(p))
        # a write into the first page, then truncation inside the second page: the
        # written bytes below the new size reach the disk first
        patch = rand_bytes(1000, 3)
        buf = np.zeros(1 << 20, dtype=np.uint8)
        buf[:1000] = patch
        r, out = script(ctx, str(p), [(GOPEN, O_GRDWR), (GWRITE, 5000, 1000, 0), (GFTRUNCATE, 70001),
                                      (GREAD, 0, 500000, 4096), (FSTAT,), (GCLOSE,)], buf=buf)
        exp = data[:70001].copy()
        exp[5000:6000] = patch
        assert r[1] == 1000 and r[2] == 0 and r[3] == 70001 and r[4] == 70001 and r[5] == 0, r
        assert (out[4096:4096 + 70001] == exp).all()
        assert (np.fromfile(str(p), dtype=np.uint8) == exp).all()
        assert ctx.check_invariants() == 0


def test_gmsync_writes_mapped_page(tmp_path):
    """A page mapped with gmmap and written in place reaches the disk through
    gmsync while the file is still open (no close, no fsync)."""
    say("test_gmsync_writes_mapped_page")
    p = tmp_path / "m.bin"
    data = rand_bytes(200000, 4)
    data.tofile(str(p))
    with make_ctx(page_size=64 << 10) as ctx:
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDWR), (GMMAP_FILL, 70000, 100, 0xAB)])
        assert r == [r[0], 0] and r[0] >= 0, r
        exp = data.copy()
        exp[70000:70100] = 0xAB
        assert (np.fromfile(str(p), dtype=np.uint8) == exp).all()   # the file is still open
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDWR), (GCLOSE,), (GCLOSE,)])
        assert r[1:] == [0, 0], r
        assert (np.fromfile(str(p), dtype=np.uint8) == exp).all()
        assert ctx.check_invariants() == 0


def test_single_thread_fsync_and_gfsync(tmp_path):
    """gwrite then single_thread_fsync (one thread) puts the bytes on disk
    with the file still open; the same for gfsync over 70 pages (more than one
    radix leaf, one batch of writes per leaf)."""
    say("test_single_thread_fsync_and_gfsync")
    p = tmp_path / "f.bin"
    data = rand_bytes(5 << 20, 5)
    data.tofile(str(p))
    patch = rand_bytes(70 * 4096 + 123, 6)
    buf = np.zeros(1 << 20, dtype=np.uint8)
    buf[:patch.size] = patch
    with make_ctx(page_size=4096, cache_bytes=8 << 20, ring_depth=32) as ctx:
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDWR), (GWRITE, 777, 5000, 0), (ST_FSYNC,)], buf=buf)
        assert r[1:] == [5000, 0], r
        exp = data.copy()
        exp[777:5777] = patch[:5000]
        assert (np.fromfile(str(p), dtype=np.uint8) == exp).all()
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDWR), (GWRITE, 1 << 20, patch.size, 0), (GFSYNC,), (GCLOSE,)],
                      buf=buf)
        assert r[1:] == [patch.size, 0, 0], r
        exp[(1 << 20):(1 << 20) + patch.size] = patch
        assert (np.fromfile(str(p), dtype=np.uint8) == exp).all()
        r, _ = script(ctx, str(p), [(GOPEN, O_GRDONLY), (GCLOSE,), (GCLOSE,)])
        assert r[1:] == [0, 0], r
        s = ctx.stats()
        assert s["rpc_writes"] >= 71 and s["slots_lost"] == 0, s
        assert ctx.check_invariants() == 0


def test_single_thread_open_close_and_gunlink(tmp_path):
    say("test_single_thread_open_close_and_gunlink")
    p = tmp_path / "o.bin"
    data = rand_bytes(10000, 7)
    data.tofile(str(p))
    with make_ctx(page_size=4096) as ctx:
        r, out = script(ctx, str(p), [(ST_OPEN, O_GRDONLY), (GREAD, 0, 10000, 0), (GUNLINK,), (ST_CLOSE,),
                                      (ST_CLOSE,)])
        assert r[0] >= 0 and r[1] == 10000 and r[3] == 0, r
        assert np.uint32(r[2] & 0xFFFFFFFF).view(np.int32) == -errno.ENOSYS, r
        assert r[4] == -errno.EBADF, r   # the file is no longer open
        assert (out[:10000] == data).all()
        assert p.exists()
        missing = str(tmp_path / "absent.bin")
        r, _ = script(ctx, missing, [(ST_OPEN, O_GRDONLY)])
        assert r[0] == -2, r   # E_IPC_OPEN_ERROR, as gopen
        assert ctx.check_invariants() == 0
