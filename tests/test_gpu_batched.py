"""Batched form of the scan (persistent workgroups taking files from a device
queue and the extra parts of large files from a device ring, also replayed from
a hipGraph): per-file counts bit-exact against the CPU path and equal to the
plain one-launch-per-file form."""
import numpy as np
import pytest

from gpufs_amd import _lib, datagen as dg
from tests.gpu_helpers import (batched_workspace, dev_bytes, dev_names, dev_zeros, grep, grep_batched, make_ctx, say,
                               torch_cuda)

pytestmark = pytest.mark.gpu


def _corpus(tmp_path, n, seed):
    voc = dg.vocab(4000, seed=seed)
    q = np.ascontiguousarray(voc[dg.query(4000, 1024, seed=seed + 1)])
    sizes = dg.lognormal_sizes(n, seed=seed + 2, median=16384.0, sigma=1.5, lo=1, hi=1 << 20)
    # empty file, one byte, a last window longer than the chunk, and files split into several parts
    sizes[:6] = [0, 1, 65536 + 40, 3 << 20, (2 << 20) + 12345, 600001]
    paths = [str(tmp_path / f"u{i:05d}.txt") for i in range(n)]
    dg.text_files(paths, sizes, seed + 3, voc, threads=8)
    return paths, q


def test_batched_and_graph_match_cpu_and_plain(tmp_path, orc):
    say("test_batched_and_graph_match_cpu_and_plain")
    torch = torch_cuda()
    n = 300
    paths, q = _corpus(tmp_path, n, 500)
    cpu, cst = orc.scan_files(paths, q, 65536, threads=8)
    assert (cst == 0).all()
    with make_ctx(page_size=64 << 10, cache_bytes=64 << 20, ring_depth=256) as ctx:
        names = dev_names(paths)
        d = dev_bytes(q.reshape(-1))
        counts = dev_zeros(n * 1024, torch.int32)
        status = dev_zeros(n, torch.int32)
        times = dev_zeros(2 * n, torch.int64)
        ws = batched_workspace(n, 4)
        ctx.host_start()
        # 4 parts per file at most, parts of at least 256 KiB: the 3 MiB file takes 4 workgroups
        grep_batched(ctx, names, n, d, 1024, 65536, counts, status, ws, workgroups=16, parts=4, min_part=256 << 10,
                     times=times)
        ctx.run_handler()
        got = counts.cpu().numpy().view(np.uint32).reshape(n, 1024)
        st = status.cpu().numpy()
        bad = np.nonzero(st != 0)[0]
        assert bad.size == 0, (bad[:20].tolist(), st[bad[:20]].tolist(), ctx.stats())
        assert (got == cpu).all()
        tk = times.cpu().numpy().reshape(n, 2)
        assert (tk[:, 1] >= tk[:, 0]).all() and (tk[:, 0] > 0).all()
        # graph replays, with eager work in between, each checked
        g = grep_batched(ctx, names, n, d, 1024, 65536, counts, status, ws, workgroups=32, graph=True, parts=4,
                         min_part=256 << 10)
        try:
            for rep in range(3):
                counts.fill_(7)
                status.fill_(-99)
                torch.cuda.synchronize()
                _lib.check("graph", _lib.lib().gpufs_graph_launch(g, ctx.stream))
                ctx.run_handler()
                assert (status.cpu().numpy() == 0).all(), rep
                assert (counts.cpu().numpy().view(np.uint32).reshape(n, 1024) == cpu).all(), rep
        finally:
            _lib.lib().gpufs_graph_destroy(g)
        ctx.host_stop()
        # plain form: one launch per file, same results
        plain, pst = grep(ctx, paths[:40], q, 65536, parts=1)
        assert (pst == 0).all() and (plain == cpu[:40]).all()
        assert ctx.check_invariants() == 0
