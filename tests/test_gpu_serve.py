"""Serving form: a resident request loop with continuous admission.  Every
request's counts equal the CPU path for its file, and are the same when the
request runs alone and when it runs in a Poisson mix with others in flight."""
import errno
import resource

import numpy as np
import pytest

from gpufs_amd import _lib, datagen as dg
from tests.gpu_helpers import Server, make_ctx, say

pytestmark = pytest.mark.gpu


def test_serve_alone_equals_mix_equals_cpu(tmp_path, orc):
    say("test_serve_alone_equals_mix_equals_cpu")
    n = 200
    voc = dg.vocab(4000, seed=71)
    q = np.ascontiguousarray(voc[dg.query(4000, 1024, seed=72)])
    sizes = dg.lognormal_sizes(n, seed=73, median=16384.0, sigma=1.5, lo=1, hi=1 << 20)
    paths = [str(tmp_path / f"s{i:04d}.txt") for i in range(n)]
    dg.text_files(paths, sizes, 74, voc, threads=8)
    cpu, cst = orc.scan_files(paths, q, 65536, threads=8)
    assert (cst == 0).all()
    ids = dg.zipf_sample(n, 600, seed=75, s=1.1)
    arr = np.cumsum(dg.poisson_gaps(20000.0, 600, seed=76))
    with make_ctx(page_size=64 << 10, cache_bytes=64 << 20, ring_depth=256) as ctx:
        srv = Server(ctx, paths, sizes, q, workgroups=8, slots=16)
        try:
            alone = {}
            for f in sorted(set(ids[:40].tolist())):
                c, st, lat, _ = srv.run([f], [0.0])
                assert st[0] == 0 and (c[0] == cpu[f]).all(), f
                alone[f] = c[0].copy()
            c, st, lat, el = srv.run(ids, arr)
            assert (st == 0).all(), st[:10]
            for i, f in enumerate(ids):
                assert (c[i] == cpu[f]).all(), (i, f)
                if f in alone:
                    assert (c[i] == alone[f]).all()
            assert (lat > 0).all() and el > 0
            hits, misses = srv.pages()
            assert hits > 0   # popular files are served from the resident cache
            assert misses > 0
            # device-copy calls refuse to run beside the resident workgroups
            with pytest.raises(_lib.GpufsError) as ei:
                ctx.stats()
            assert ei.value.code == -errno.EBUSY
        finally:
            srv.stop()
        s = ctx.stats()
        # the workgroups' own counts agree with the context counters
        assert s["page_hits"] == hits and s["page_misses"] == misses
        # the server keeps each file open from its first request on (while under its cap of half
        # the descriptor limit, at most 4096), so every distinct file is opened on the host once,
        # and it closes them all when it stops
        soft = resource.getrlimit(resource.RLIMIT_NOFILE)[0]
        cap = 4096 if soft == resource.RLIM_INFINITY else (min(4096, soft // 2 - 64) if soft > 256 else 0)
        distinct = len(set(ids.tolist()))
        if cap >= distinct:
            assert s["rpc_opens"] == distinct, (s, distinct)
        assert s["rpc_opens"] == s["rpc_closes"], s
        assert ctx.check_invariants() == 0


def test_serve_splits_large_files(tmp_path, orc):
    """Requests for large files run on several workgroups (parts handed to
    waiting workgroups through the device part ring; the last part to finish
    completes the request), cold and cached: every result equals the CPU
    path, alone and in a Poisson mix with small requests in flight."""
    say("test_serve_splits_large_files")
    n = 40
    voc = dg.vocab(4000, seed=91)
    q = np.ascontiguousarray(voc[dg.query(4000, 1024, seed=92)])
    sizes = dg.lognormal_sizes(n, seed=93, median=65536.0, sigma=2.0, lo=1, hi=6 << 20)
    sizes[:6] = [6 << 20, (4 << 20) + 12345, 3 << 20, (1 << 20) + 7, 700001, 300000]   # several parts each
    paths = [str(tmp_path / f"L{i:03d}.txt") for i in range(n)]
    dg.text_files(paths, sizes, 94, voc, threads=8)
    cpu, cst = orc.scan_files(paths, q, 65536, threads=8)
    assert (cst == 0).all()
    ids = np.concatenate([np.arange(6), dg.zipf_sample(n, 300, seed=95, s=1.1)]).astype(np.uint32)
    arr = np.cumsum(dg.poisson_gaps(5000.0, ids.size, seed=96))
    with make_ctx(page_size=64 << 10, cache_bytes=256 << 20, ring_depth=256) as ctx:
        srv = Server(ctx, paths, sizes, q, workgroups=16, slots=8, parts_max=8, min_part_kib=128)
        try:
            for f in range(6):   # cold, then cached, alone
                for rnd in ("cold", "warm"):
                    c, st, lat, _ = srv.run([f], [0.0])
                    assert st[0] == 0 and (c[0] == cpu[f]).all(), (f, rnd)
            split, parts = srv.parts()
            assert split >= 8 and parts > 2 * split, (split, parts)
            c, st, lat, el = srv.run(ids, arr)
            assert (st == 0).all(), st[:10]
            for i, f in enumerate(ids):
                assert (c[i] == cpu[f]).all(), (i, f)
        finally:
            srv.stop()
        assert ctx.check_invariants() == 0


def test_serve_survives_idle(tmp_path, orc, monkeypatch):
    """A server left idle for longer than the device wait limit (1 s here, so
    3 s of idling is three times it) still serves: its workgroups never give
    up a request ticket they have taken.  Both runs are bit-exact."""
    import time
    say("test_serve_survives_idle")
    monkeypatch.setenv("GPUFS_WAIT_TIMEOUT_S", "1")
    n = 24
    voc = dg.vocab(2000, seed=81)
    q = np.ascontiguousarray(voc[dg.query(2000, 256, seed=82)])
    sizes = dg.lognormal_sizes(n, seed=83, median=8192.0, sigma=1.0, lo=1, hi=1 << 18)
    paths = [str(tmp_path / f"i{i:03d}.txt") for i in range(n)]
    dg.text_files(paths, sizes, 84, voc, threads=4)
    cpu, cst = orc.scan_files(paths, q, 65536, threads=4)
    assert (cst == 0).all()
    with make_ctx(page_size=64 << 10, cache_bytes=32 << 20, ring_depth=64) as ctx:
        srv = Server(ctx, paths, sizes, q, workgroups=4, slots=8)
        try:
            ids = np.arange(n, dtype=np.uint32)
            c, st, _, _ = srv.run(ids[:12], np.zeros(12))
            assert (st == 0).all() and (c == cpu[:12]).all()
            time.sleep(3.0)
            c, st, _, _ = srv.run(ids, np.linspace(0.0, 0.01, n))
            assert (st == 0).all(), st
            assert (c == cpu).all()
        finally:
            srv.stop()
        assert ctx.check_invariants() == 0
