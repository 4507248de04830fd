"""File-table reclamation: a context whose table holds 64 entries opens three
times as many distinct files, again after all of them were closed, and again
after some changed on the host.  Closed entries are taken back for new names
(the reference reuses CLOSED entries, libgpufs/fs_structures.cu:364-369), their
radix trees go to the node free list, and every count stays bit-exact against
the CPU path.  A launch that holds more files open at once than the table has
entries fails those opens with E_FSTABLE_FULL and leaves the others exact."""
import os

import numpy as np
import pytest

from gpufs_amd import datagen as dg
from tests.gpu_helpers import (batched_workspace, dev_bytes, dev_names, dev_zeros, grep, grep_batched, make_ctx, say,
                               torch_cuda)

pytestmark = pytest.mark.gpu

MAX_FILES = 64


def _corpus(tmp_path, n, seed):
    voc = dg.vocab(3000, seed=seed)
    q = np.ascontiguousarray(voc[dg.query(3000, 512, seed=seed + 1)])
    sizes = dg.lognormal_sizes(n, seed=seed + 2, median=30000.0, sigma=1.2, lo=1, hi=700000)
    paths = [str(tmp_path / f"r{i:04d}.txt") for i in range(n)]
    dg.text_files(paths, sizes, seed + 3, voc, threads=8)
    return paths, q, voc


def _batched(ctx, names, n, d, nw, counts, status, ws):
    status.fill_(-12345)
    torch_cuda().cuda.synchronize()
    ctx.host_start()
    grep_batched(ctx, names, n, d, nw, 65536, counts, status, ws, workgroups=16, parts=4, min_part=64 << 10)
    ctx.run_handler()
    ctx.host_stop()
    return counts.cpu().numpy().view(np.uint32).reshape(n, nw).copy(), status.cpu().numpy().copy()


def test_reclaim_three_times_the_table(tmp_path, orc):
    say("test_reclaim_three_times_the_table")
    torch = torch_cuda()
    n = 3 * MAX_FILES
    paths, q, voc = _corpus(tmp_path, n, 900)
    nw = q.shape[0]
    cpu, cst = orc.scan_files(paths, q, 65536, threads=8)
    assert (cst == 0).all()
    with make_ctx(max_files=MAX_FILES, page_size=64 << 10, cache_bytes=32 << 20, ring_depth=256) as ctx:
        names = dev_names(paths)
        d = dev_bytes(q.reshape(-1))
        counts = dev_zeros(n * nw, torch.int32)
        status = dev_zeros(n, torch.int32)
        ws = batched_workspace(n, 4)
        for rep in range(2):
            got, st = _batched(ctx, names, n, d, nw, counts, status, ws)
            bad = np.nonzero(st != 0)[0]
            assert bad.size == 0, (rep, bad[:20].tolist(), st[bad[:20]].tolist(), ctx.stats())
            assert (got == cpu).all(), rep
            s = ctx.stats()
            assert s["files_reclaimed"] >= n - MAX_FILES, s
            assert s["files_used"] <= MAX_FILES, s
            assert ctx.check_invariants() == 0, rep
        recycled = ctx.stats()["nodes_recycled"]
        assert recycled > 0
        # some files change on the host: entries still in the table reopen as
        # a new version (their old trees are freed), the others are inserted again
        changed = list(range(0, n, 7))
        dg.text_files([paths[i] for i in changed], [12345 + 17 * i for i in changed], 77, voc, threads=8)
        cpu2, cst2 = orc.scan_files(paths, q, 65536, threads=8)
        assert (cst2 == 0).all() and not (cpu2 == cpu).all()
        got, st = _batched(ctx, names, n, d, nw, counts, status, ws)
        assert (st == 0).all(), st[st != 0][:20]
        assert (got == cpu2).all()
        assert ctx.stats()["nodes_recycled"] > recycled
        assert ctx.check_invariants() == 0
        # one launch per file, one after another, on the same table
        plain, pst = grep(ctx, paths[:24], q, 65536, parts=1)
        assert (pst == 0).all() and (plain == cpu2[:24]).all()
        assert ctx.check_invariants() == 0


def test_more_files_open_than_entries(tmp_path, orc):
    """One workgroup per file for 192 files at once on a 64-entry table: an open
    that finds every entry in use fails with E_FSTABLE_FULL (-1); every other
    file's counts are exact, and the table is consistent afterwards."""
    say("test_more_files_open_than_entries")
    n = 3 * MAX_FILES
    paths, q, _ = _corpus(tmp_path, n, 950)
    cpu, cst = orc.scan_files(paths, q, 65536, threads=8)
    assert (cst == 0).all()
    with make_ctx(max_files=MAX_FILES, page_size=64 << 10, cache_bytes=32 << 20, ring_depth=256) as ctx:
        got, st = grep(ctx, paths, q, 65536, parts=1)
        ok = st == 0
        assert np.isin(st, [0, -1]).all(), np.unique(st)
        assert ok.sum() >= MAX_FILES, ok.sum()
        assert (got[ok] == cpu[ok]).all()
        assert ctx.check_invariants() == 0
        # the same files again, now that every entry is closed: all succeed
        got, st = grep(ctx, paths[:MAX_FILES], q, 65536, parts=1)
        assert (st == 0).all() and (got == cpu[:MAX_FILES]).all()
        assert ctx.check_invariants() == 0
        assert os.path.exists(paths[0])
