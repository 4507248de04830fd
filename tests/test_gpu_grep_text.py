"""The grep_text sample end to end: query records, database list and output
file are GPUfs files; the output file's lines, sorted, must equal the
oracle's (the reference writes them in an order set by timing)."""
import os
import shutil
import tempfile

import numpy as np
import pytest

from gpufs_amd import _lib, datagen as dg
from tests.gpu_helpers import dev_zeros, finish, make_ctx, say, torch_cuda

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture
def short_dir():
    d = tempfile.mkdtemp(prefix="gt", dir="/tmp")   # database names stay under 64 bytes
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _run(ctx, src, out, dbs, chunk, workgroups, threads=1024, records=False):
    torch = torch_cuda()
    st = dev_zeros(1, torch.int32)
    torch.cuda.synchronize()
    fn = _lib.lib().gpufs_grep_text_records if records else _lib.lib().gpufs_grep_text
    grid = fn(ctx.h, src.encode(), out.encode(), dbs.encode(), chunk, workgroups, threads, st.data_ptr(), None)
    _lib.check("gpufs_grep_text", grid)
    finish(ctx)
    return grid, int(st.cpu().item())


def _sorted_lines(b):
    return sorted(b.decode().splitlines())


@pytest.mark.parametrize("chunk,workgroups", [(65536, 4), (4096, 16)])
def test_grep_text_reference_fixture(orc, short_dir, chunk, workgroups):
    """The reference's shipped text and a subset of its dict.gpu, three databases."""
    say(f"test_grep_text_reference_fixture chunk={chunk} wg={workgroups}")
    words = np.fromfile(os.path.join(GOLDEN, "dict_subset.bin"), dtype=np.uint8).reshape(-1, 32)
    src = os.path.join(short_dir, "dict.gpu")
    words.tofile(src)
    text = open(os.path.join(GOLDEN, "pg100_small.txt"), "rb").read()
    dbs = []
    for i, body in enumerate([text, text[::-1], text * 5]):
        p = os.path.join(short_dir, f"db{i}.txt")
        open(p, "wb").write(body)
        dbs.append(p)
    lst = os.path.join(short_dir, "dbs.txt")
    open(lst, "w").write(dbs[0] + "\n" + dbs[1] + ",\t" + dbs[2] + "\n")
    out = os.path.join(short_dir, "out.txt")
    want = orc.grep_text_lines(words, open(lst, "rb").read(), chunk)
    # an older, longer output: the kernel truncates it (from one GPU thread)
    # before any line is written, so none of these bytes may survive
    open(out, "wb").write(b"\xff" * (2 * len(want) + 4096))
    with make_ctx(page_size=64 << 10, cache_bytes=64 << 20) as ctx:
        grid, st = _run(ctx, src, out, lst, chunk, workgroups)
        assert st == 0
        got = open(out, "rb").read()
        assert _sorted_lines(got) == _sorted_lines(want)
        assert len(got) == len(want)
        # a second run truncates and rewrites the output
        grid2, st2 = _run(ctx, src, out, lst, chunk, workgroups)
        assert st2 == 0 and _sorted_lines(open(out, "rb").read()) == _sorted_lines(want)
        assert ctx.check_invariants() == 0


def test_grep_text_more_records_than_one_workgroup_holds(orc, short_dir):
    """3000 query records (duplicates included) need at least three
    workgroups; the launcher widens the grid, the lines stay the same."""
    say("test_grep_text_more_records_than_one_workgroup_holds")
    voc = dg.vocab(5000, seed=31)
    q = np.ascontiguousarray(np.concatenate([voc[dg.query(5000, 2500, seed=32)], voc[:500]]))
    src = os.path.join(short_dir, "q.gpu")
    q.tofile(src)
    paths = [os.path.join(short_dir, f"t{i}.txt") for i in range(2)]
    dg.text_files(paths, [300000, 90001], 33, voc, threads=4)
    lst = os.path.join(short_dir, "l.txt")
    open(lst, "w").write("\n".join(paths) + "\n")
    out = os.path.join(short_dir, "o.txt")
    want = orc.grep_text_lines(q, open(lst, "rb").read(), 65536)
    with make_ctx(page_size=64 << 10, cache_bytes=64 << 20) as ctx:
        grid, st = _run(ctx, src, out, lst, 65536, 1)
        assert grid >= 3 and st == 0
        assert _sorted_lines(open(out, "rb").read()) == _sorted_lines(want)
        assert ctx.check_invariants() == 0


def test_grep_text_missing_database_fails_cleanly(short_dir):
    say("test_grep_text_missing_database_fails_cleanly")
    src = os.path.join(short_dir, "d.gpu")
    np.zeros((4, 32), dtype=np.uint8).tofile(src)
    lst = os.path.join(short_dir, "l.txt")
    open(lst, "w").write(os.path.join(short_dir, "nope.txt") + "\n")
    with make_ctx() as ctx:
        grid, st = _run(ctx, src, os.path.join(short_dir, "o.txt"), lst, 65536, 2)
        assert st < 0
        assert ctx.check_invariants() == 0


def test_grep_text_binary_records(orc, short_dir):
    """The same scan writing display.cpp's 100-byte records: printed with
    gpufs_amd.display they are the text form's lines (as a multiset), and the
    record fields match the oracle's lines."""
    say("test_grep_text_binary_records")
    from gpufs_amd import display
    words = np.fromfile(os.path.join(GOLDEN, "dict_subset.bin"), dtype=np.uint8).reshape(-1, 32)
    src = os.path.join(short_dir, "dict.gpu")
    words.tofile(src)
    text = open(os.path.join(GOLDEN, "pg100_small.txt"), "rb").read()
    dbs = []
    for i, body in enumerate([text, text[::-1]]):
        p = os.path.join(short_dir, f"db{i}.txt")
        open(p, "wb").write(body)
        dbs.append(p)
    lst = os.path.join(short_dir, "dbs.txt")
    open(lst, "w").write("\n".join(dbs) + "\n")
    out = os.path.join(short_dir, "out.bin")
    want = orc.grep_text_lines(words, open(lst, "rb").read(), 4096)
    open(out, "wb").write(b"\x7f" * 12345)   # an older, longer output: truncated first
    with make_ctx(page_size=64 << 10, cache_bytes=64 << 20) as ctx:
        grid, st = _run(ctx, src, out, lst, 4096, 8, records=True)
        assert st == 0
        data = open(out, "rb").read()
        assert len(data) % display.RECORD_BYTES == 0 and len(data) // display.RECORD_BYTES == len(want.splitlines())
        recs = display.records(data)
        got = sorted(f"{w.decode()} {c} {n.decode()}" for w, c, n in recs)
        assert got == sorted(want.decode().splitlines())
        for (w, c, n), line in zip(recs, display.lines(data)):
            assert c > 0 and n.decode() in dbs
            if len(w) < 32:   # a full 32-byte query prints on into its count, as display.cpp does
                assert line == w + b" " + str(c).encode() + b" " + n + b"\n"
        assert ctx.check_invariants() == 0
