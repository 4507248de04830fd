"""GPU parity of the write path (gwrite / gfsync / gclose write-back, O_GWRONCE
OR-merge): the bytes on disk after the kernel equal the expected bytes."""
import os

import numpy as np
import pytest

from gpufs_amd import O_GRDWR, O_GWRONCE, O_GWRONLY, O_GCREAT
from tests.gpu_helpers import dev_bytes, gwrite_ranges, make_ctx, read_file, say, write_file

pytestmark = pytest.mark.gpu


def test_write_new_file_then_read_back(tmp_path):
    say("test_write_new_file_then_read_back")
    p = tmp_path / "new.bin"
    data = np.random.default_rng(1).integers(0, 256, size=200000, dtype=np.uint8).tobytes()
    with make_ctx(page_size=65536) as ctx:
        r = write_file(ctx, str(p), 0, data, O_GWRONLY | O_GCREAT)
        assert r[0] == len(data) and r[2] == 0, r
        assert p.read_bytes() == data
        got, fsize, _, out = read_file(ctx, str(p), 0, len(data))
        assert got == len(data) and out == data
        assert ctx.check_invariants() == 0


def test_partial_page_read_modify_write(tmp_path):
    say("test_partial_page_read_modify_write")
    p = tmp_path / "rmw.bin"
    base = np.random.default_rng(2).integers(0, 256, size=300000, dtype=np.uint8)
    base.tofile(str(p))
    patch = np.random.default_rng(3).integers(0, 256, size=5000, dtype=np.uint8).tobytes()
    off = 65536 - 1000
    with make_ctx(page_size=65536) as ctx:
        r = write_file(ctx, str(p), off, patch, O_GRDWR, fsync=True)
        assert r[0] == len(patch) and r[3] == 0 and r[2] == 0, r
        exp = bytearray(base.tobytes())
        exp[off:off + len(patch)] = patch
        assert p.read_bytes() == bytes(exp)
        assert ctx.check_invariants() == 0


def test_write_extends_file(tmp_path):
    say("test_write_extends_file")
    p = tmp_path / "ext.bin"
    p.write_bytes(b"abc")
    with make_ctx(page_size=4096) as ctx:
        r = write_file(ctx, str(p), 10000, b"XYZ", O_GRDWR)
        assert r[0] == 3 and r[1] == 10003, r
        exp = b"abc" + b"\0" * (10000 - 3) + b"XYZ"
        assert p.read_bytes() == exp


def test_write_past_4gib_into_new_file(tmp_path):
    """gwrite at a 64-bit offset: 100,000 bytes at 5 GiB + 3 of a new file
    extend it there (a hole of zeros before them) and read back through the
    cache."""
    say("test_write_past_4gib_into_new_file")
    p = tmp_path / "sparse_w.bin"
    off = (5 << 30) + 3
    data = np.random.default_rng(4).integers(0, 256, size=100000, dtype=np.uint8).tobytes()
    with make_ctx(page_size=65536) as ctx:
        r = write_file(ctx, str(p), off, data, O_GWRONLY | O_GCREAT)
        assert r[0] == len(data) and r[1] == off + len(data) and r[2] == 0, r
        assert os.path.getsize(p) == off + len(data)
        with open(p, "rb") as f:
            f.seek(off - 10)
            assert f.read() == b"\0" * 10 + data
        got, fsize, _, out = read_file(ctx, str(p), off - 10, len(data) + 10)
        assert got == len(data) + 10 and fsize == off + len(data)
        assert out == b"\0" * 10 + data
        assert ctx.check_invariants() == 0


def test_wronce_or_merges_with_disk(tmp_path):
    say("test_wronce_or_merges_with_disk")
    p = tmp_path / "once.bin"
    p.write_bytes(bytes([0x01, 0x10, 0x00, 0xF0]) * 1000)
    new = bytes([0x02, 0x01, 0x80, 0x00]) * 500
    with make_ctx(page_size=4096) as ctx:
        r = write_file(ctx, str(p), 0, new, O_GWRONCE)
        assert r[0] == len(new) and r[2] == 0, r
    got = p.read_bytes()
    assert got[:2000] == bytes([0x03, 0x11, 0x80, 0xF0]) * 500
    assert got[2000:] == bytes([0x01, 0x10, 0x00, 0xF0]) * 500


def test_gwrite_larger_than_cache(tmp_path):
    """One gwrite of three times the page cache, no gfsync: the cache fills
    with dirty pages, which the eviction sweep never takes, so the writer
    writes its file back to free frames and the call completes."""
    say("test_gwrite_larger_than_cache")
    p = tmp_path / "big.bin"
    data = np.random.default_rng(5).integers(0, 256, size=(12 << 20) + 777, dtype=np.uint8).tobytes()
    with make_ctx(page_size=65536, cache_bytes=4 << 20) as ctx:
        r = write_file(ctx, str(p), 0, data, O_GWRONLY | O_GCREAT, threads=1024)
        assert r[0] == len(data) and r[1] == len(data) and r[2] == 0, r
        assert p.read_bytes() == data
        assert ctx.check_invariants() == 0
        assert ctx.stats()["bytes_written"] == len(data)


@pytest.mark.parametrize("fsync", [True, False])
def test_gwrite_ranges_new_file(tmp_path, fsync):
    """64 workgroups write a new file, one ragged range each (ranges share
    pages with their neighbours), with and without gfsync, through a cache
    large enough to hold the file."""
    say(f"test_gwrite_ranges_new_file fsync={fsync}")
    p = tmp_path / "ranges.bin"
    n, rng = 64, (256 << 10) + 4099
    data = np.random.default_rng(6).integers(0, 256, size=n * rng, dtype=np.uint8)
    with make_ctx(page_size=65536, cache_bytes=64 << 20) as ctx:
        out = gwrite_ranges(ctx, str(p), O_GWRONLY | O_GCREAT, n, rng, dev_bytes(data.tobytes()), fsync=fsync)
        assert (out == rng).all(), out
        assert p.read_bytes() == data.tobytes()
        assert ctx.check_invariants() == 0


@pytest.mark.parametrize("mode", ["wronly", "wronce"])
def test_gwrite_ranges_through_small_cache(tmp_path, mode):
    """48 workgroups write 24 MiB through an 8 MiB cache without gfsync:
    writers write back when three quarters of the frames are dirty, or when
    a page finds no frame.  O_GWRONCE OR-merges with the bytes on disk."""
    say(f"test_gwrite_ranges_through_small_cache {mode}")
    p = tmp_path / "small_cache.bin"
    n, rng = 48, (512 << 10) + 8191
    rs = np.random.default_rng(7)
    data = rs.integers(0, 256, size=n * rng, dtype=np.uint8)
    if mode == "wronce":
        base = rs.integers(0, 256, size=n * rng, dtype=np.uint8)
        base.tofile(str(p))
        flags, expect = O_GWRONCE, (base | data).tobytes()
    else:
        flags, expect = O_GWRONLY | O_GCREAT, data.tobytes()
    with make_ctx(page_size=65536, cache_bytes=8 << 20) as ctx:
        out = gwrite_ranges(ctx, str(p), flags, n, rng, dev_bytes(data.tobytes()), fsync=False)
        assert (out == rng).all(), out
        assert p.read_bytes() == expect
        assert ctx.check_invariants() == 0
        s = ctx.stats()
        assert s["bytes_written"] >= n * rng, s
