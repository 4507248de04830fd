"""GPU parity of the grep_text-style scan (gopen + gread into LDS + whole-word
counts) against the CPU oracle (literal restatement of match_string and the
chunk loop) and against the generator's tally."""
import os

import numpy as np
import pytest

from gpufs_amd import datagen as dg
from tests.gpu_helpers import grep, make_ctx, say

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_grep_reference_fixture(tmp_path, orc):
    say("test_grep_reference_fixture")
    import json
    g = os.path.join(HERE, "golden")
    text = open(os.path.join(g, "pg100_small.txt"), "rb").read()
    d = np.fromfile(os.path.join(g, "dict_subset.bin"), dtype=np.uint8).reshape(-1, 32)
    exp = json.load(open(os.path.join(g, "pg100_counts.json")))
    p = tmp_path / "pg100.txt"
    p.write_bytes(text)
    with make_ctx(page_size=4096, cache_bytes=4 << 20) as ctx:
        for chunk in (16384, 64):
            counts, status = grep(ctx, [str(p)], d, chunk)
            assert status[0] == 0
            assert (counts[0] == np.array(exp["counts"][str(chunk)], dtype=np.uint32)).all(), chunk
        assert ctx.check_invariants() == 0


def test_grep_synthetic_matches_tally_and_oracle(tmp_path, orc):
    say("test_grep_synthetic_matches_tally_and_oracle")
    voc = dg.vocab(5000, seed=31)
    qi = dg.query(5000, 1024, seed=32)
    q = voc[qi]
    paths = [str(tmp_path / f"t{i}.txt") for i in range(8)]
    sizes = [256 << 10] * 6 + [1000, (256 << 10) + 77]
    tally = dg.text_files(paths, sizes, 33, voc, with_tally=True, threads=4)
    cpu, st = orc.scan_files(paths, q, 65536, threads=4)
    assert (st == 0).all()
    with make_ctx(page_size=65536, cache_bytes=32 << 20) as ctx:
        counts, status = grep(ctx, paths, q, 65536)
        assert (status == 0).all(), status
        assert (counts == cpu).all()
        for i in range(6):
            assert (counts[i] == tally[i][qi]).all()
        # warm pass gives the same counts
        counts2, _ = grep(ctx, paths, q, 65536)
        assert (counts2 == counts).all()
        assert ctx.check_invariants() == 0


@pytest.mark.parametrize("chunk", [8, 64, 100, 4096])
def test_grep_adversarial_small_chunks(tmp_path, orc, chunk):
    say(f"test_grep_adversarial_small_chunks chunk={chunk}")
    rng = np.random.default_rng(chunk)
    alphabet = np.frombuffer(b"aab_A ' \n\0\xff-1", dtype=np.uint8)
    paths, datas = [], []
    for i in range(6):
        n = int(rng.integers(0, 20000))
        t = rng.choice(alphabet, size=n) if i % 2 else rng.choice(np.frombuffer(b"aab ", dtype=np.uint8), size=n)
        if i == 4:
            t = np.frombuffer(b"x" * 40 + b"yy z" + b"a" * 70 + b" aab", dtype=np.uint8)
        p = str(tmp_path / f"a{i}.bin")
        t.tofile(p)
        paths.append(p)
        datas.append(t)
    words = [b"", b"a", b"aa", b"aab", b"b", b"A", b"_", b"a_b", b"yy", b"z", b"x" * 32, b"a" * 32, b"aab", b"cross-bun"]
    for _ in range(40):
        words.append(bytes(rng.choice(np.frombuffer(b"ab_A", dtype=np.uint8), size=int(rng.integers(1, 8)))))
    d = np.zeros((len(words), 32), dtype=np.uint8)
    for i, w in enumerate(words):
        d[i, :len(w)] = np.frombuffer(w, dtype=np.uint8)
    with make_ctx(page_size=4096, cache_bytes=8 << 20) as ctx:
        counts, status = grep(ctx, paths, d, chunk, threads=256)
        assert (status == 0).all(), status
        for i, t in enumerate(datas):
            assert (counts[i] == orc.scan_naive(t, d, chunk)).all(), i
        assert ctx.check_invariants() == 0


def test_grep_many_workgroups_one_file(tmp_path, orc):
    """64 workgroups open and read the same file at once: concurrent opens,
    shared pages and fills waited on by other workgroups."""
    say("test_grep_many_workgroups_one_file")
    voc = dg.vocab(3000, seed=41)
    q = voc[dg.query(3000, 512, seed=42)]
    p = str(tmp_path / "same.txt")
    dg.text_files([p], [3 << 20], 43, voc)
    cpu, _ = orc.scan_files([p], q, 65536)
    with make_ctx(page_size=65536, cache_bytes=16 << 20, ring_depth=64) as ctx:
        counts, status = grep(ctx, [p] * 64, q, 65536)
        assert (status == 0).all(), status
        assert (counts == cpu[0][None, :]).all()
        s = ctx.stats()
        assert s["page_misses"] == 48
        assert s["rpc_opens"] >= 1
        assert ctx.check_invariants() == 0


def test_grep_sequential_read_ahead(tmp_path, orc):
    """Small windows over small pages: sequential reads fetch pages ahead in
    the same round trip; every page is still read from the host exactly once
    (nothing twice, nothing past the end) and the counts stay exact."""
    say("test_grep_sequential_read_ahead")
    voc = dg.vocab(2000, seed=51)
    q = voc[dg.query(2000, 256, seed=52)]
    size = (1 << 20) + 100
    paths = [str(tmp_path / f"ra{i}.txt") for i in range(4)]
    dg.text_files(paths, [size] * 4, 53, voc)
    cpu, _ = orc.scan_files(paths, q, 4096)
    with make_ctx(page_size=16 << 10, cache_bytes=32 << 20, ring_depth=128) as ctx:
        counts, status = grep(ctx, paths, q, 4096)
        assert (status == 0).all(), status
        assert (counts == cpu).all()
        s = ctx.stats()
        pages = 4 * ((size + (16 << 10) - 1) // (16 << 10))
        assert s["page_misses"] == pages and s["rpc_reads"] == pages, s
        assert ctx.check_invariants() == 0


def test_grep_page_memo_under_eviction(tmp_path, orc):
    """Reads inside the page of a workgroup's previous read reuse its frame
    without the radix walk, but only while the frame still holds that page:
    here the cache (2 MiB) is a third of the data and pages are evicted between
    a workgroup's windows on them, so both the reuse and its fallback run."""
    say("test_grep_page_memo_under_eviction")
    voc = dg.vocab(2000, seed=61)
    q = voc[dg.query(2000, 256, seed=62)]
    size = (768 << 10) + 100
    paths = [str(tmp_path / f"m{i}.txt") for i in range(8)]
    dg.text_files(paths, [size] * 8, 63, voc)
    cpu, _ = orc.scan_files(paths, q, 4096)
    with make_ctx(page_size=16 << 10, cache_bytes=2 << 20, ring_depth=128) as ctx:
        counts, status = grep(ctx, paths, q, 4096, parts=1)
        assert (status == 0).all(), status
        assert (counts == cpu).all()
        s = ctx.stats()
        assert s["evictions"] > 0, s
        assert ctx.check_invariants() == 0


@pytest.mark.parametrize("chunk", [8, 64, 100, 65536])
def test_grep_split_parts_match_reference(tmp_path, orc, chunk):
    """Several workgroups per file (2, 5 and 16 parts, parts of at least 256
    bytes): counts equal the reference loop's, including files whose last
    window repeats (tail re-scans), empty and one-byte files, and files with
    runs of 32 or more word bytes that make a part boundary or the reference's
    look-ahead cut a token (the recount path)."""
    say(f"test_grep_split_parts_match_reference chunk={chunk}")
    rng = np.random.default_rng(100 + chunk)
    letters = np.frombuffer(b"aab_Zz", dtype=np.uint8)
    seps = np.frombuffer(b" \n,.0\x00\xff", dtype=np.uint8)

    def text(n, long_runs):
        out, tot = [], 0
        while tot < n:
            k = int(rng.integers(1, 45 if long_runs else 9))
            out.append(letters[rng.integers(0, len(letters), k)])
            out.append(seps[rng.integers(0, len(seps), int(rng.integers(1, 3)))])
            tot += len(out[-1]) + len(out[-2])
        return np.concatenate(out)[:n].copy() if out else np.zeros(0, dtype=np.uint8)

    sizes = [0, 1, 63, 64, 65, 5000, 20000, 200003, 300000] if chunk == 65536 else [0, 1, 64, 700, 5000, 20000, 60001]
    paths, datas = [], []
    for i, n in enumerate(sizes):
        t = text(n, long_runs=(i % 3 == 2))
        p = str(tmp_path / f"s{i}.bin")
        t.tofile(p)
        paths.append(p)
        datas.append(t)
    words = [b"a", b"aa", b"aab", b"ab", b"b", b"Zz", b"_", b"a_b", b"zZ", b"ab_a", b"", b"a" * 32]
    for _ in range(30):
        words.append(bytes(rng.choice(letters, size=int(rng.integers(1, 7)))))
    d = np.zeros((len(words), 32), dtype=np.uint8)
    for i, w in enumerate(words):
        d[i, :len(w)] = np.frombuffer(w, dtype=np.uint8)
    ref = [orc.scan_fast(t.tobytes(), d, chunk) for t in datas]
    with make_ctx(page_size=4096 if chunk < 4096 else 65536, cache_bytes=16 << 20) as ctx:
        for parts in (2, 5, 16):
            counts, status = grep(ctx, paths, d, chunk, threads=512, parts=parts, min_part=256)
            assert (status == 0).all(), (parts, status)
            for i in range(len(paths)):
                assert (counts[i] == ref[i]).all(), (parts, i, sizes[i])
        assert ctx.check_invariants() == 0


def test_grep_split_under_eviction(tmp_path, orc):
    """A split scan whose files do not fit the cache: parts and the finishing
    workgroup's chain (which pins pages a batch at a time) evict each other's
    pages; counts still equal the reference's."""
    say("test_grep_split_under_eviction")
    voc = dg.vocab(3000, seed=41)
    q = voc[dg.query(3000, 1024, seed=42)]
    paths = [str(tmp_path / f"e{i}.txt") for i in range(6)]
    dg.text_files(paths, [3 << 20] * 6, 43, voc, threads=4)
    cpu, st = orc.scan_files(paths, q, 65536, threads=4)
    assert (st == 0).all()
    with make_ctx(page_size=65536, cache_bytes=12 << 20, ring_depth=128) as ctx:
        counts, status = grep(ctx, paths, q, 65536, parts=4, min_part=65536)
        assert (status == 0).all(), (status, ctx.stats())
        assert (counts == cpu).all()
        s = ctx.stats()
        assert s["evictions"] > 0, s
        assert ctx.check_invariants() == 0


def test_grep_split_sees_changed_file(tmp_path, orc):
    """A warm split scan reads the cached pages while the files' reopens are
    at the host; a file changed on the host in between (same size, larger,
    smaller) must be counted as it is now, through the reopened file."""
    say("test_grep_split_sees_changed_file")
    voc = dg.vocab(3000, seed=51)
    q = voc[dg.query(3000, 1024, seed=52)]
    paths = [str(tmp_path / f"c{i}.txt") for i in range(5)]
    dg.text_files(paths, [1 << 20] * 4 + [300000], 53, voc, threads=4)
    with make_ctx(page_size=65536, cache_bytes=64 << 20) as ctx:
        for _ in range(2):   # cold, then warm (every page cached, entries closed)
            counts, status = grep(ctx, paths, q, 65536, parts=4, min_part=65536)
            assert (status == 0).all(), status
        cpu, st = orc.scan_files(paths, q, 65536, threads=4)
        assert (st == 0).all()
        assert (counts == cpu).all()
        # rewrite three files in place (same inode): same size, larger, smaller
        repl = [str(tmp_path / f"n{i}.txt") for i in range(3)]
        dg.text_files(repl, [1 << 20, (1 << 20) + 70000, 500000], 54, voc, threads=4)
        for src, dst in zip(repl, paths[:3]):
            data = open(src, "rb").read()
            with open(dst, "r+b") as f:
                f.write(data)
                f.truncate(len(data))
            m = os.stat(dst).st_mtime_ns
            os.utime(dst, ns=(m + 2_000_000_000, m + 2_000_000_000))
        cpu2, st = orc.scan_files(paths, q, 65536, threads=4)
        assert (st == 0).all()
        assert not (cpu2[:3] == cpu[:3]).all()
        ctx.stats_reset()
        for k in range(2):
            counts, status = grep(ctx, paths, q, 65536, parts=4, min_part=65536)
            assert (status == 0).all(), (k, status)
            for i in range(len(paths)):
                assert (counts[i] == cpu2[i]).all(), (k, i)
            if k == 0:
                s = ctx.stats()
                # every part of the 5 cached files read before the reopen;
                # the 3 rewritten files were found changed by every part
                assert s["spec_parts"] == 20 and s["spec_changed"] == 12, s
        assert ctx.check_invariants() == 0


@pytest.mark.parametrize("nfiles,fsize,cache_mib", [(16, 512 << 10, 4), (160, 256 << 10, 8)])
def test_grep_split_overlapped_reopen_under_eviction(tmp_path, orc, nfiles, fsize, cache_mib):
    """Files that fit part 0's pin budget in a cache smaller than the data:
    in every pass some files are still wholly cached (their parts scan
    before the reopen completes), others are not.  With more workgroups than
    compute units (160 files x 4 parts), a file's other parts start after its
    part 0 has ended and other files' fills may have evicted its pages: such
    a part scans its range again through the open file.  Every pass equals
    the reference."""
    say(f"test_grep_split_overlapped_reopen_under_eviction nfiles={nfiles}")
    voc = dg.vocab(3000, seed=61)
    q = voc[dg.query(3000, 1024, seed=62)]
    paths = [str(tmp_path / f"v{i}.txt") for i in range(nfiles)]
    dg.text_files(paths, [fsize] * (nfiles - 4) + [300000, 200000, 65536, 1], 63, voc, threads=4)
    cpu, st = orc.scan_files(paths, q, 65536, threads=4)
    assert (st == 0).all()
    with make_ctx(page_size=65536, cache_bytes=cache_mib << 20, ring_depth=64, max_files=256) as ctx:
        for k in range(4):
            order = paths if k % 2 == 0 else paths[::-1]
            counts, status = grep(ctx, order, q, 65536, parts=4, min_part=65536)
            assert (status == 0).all(), (k, status, ctx.stats())
            want = cpu if k % 2 == 0 else cpu[::-1]
            bad = [i for i in range(len(order)) if not (counts[i] == want[i]).all()]
            assert not bad, (k, bad)
        s = ctx.stats()
        say(f"  spec_parts {s['spec_parts']} spec_rescans {s['spec_rescans']} evictions {s['evictions']}")
        assert s["evictions"] > 0, s
        if nfiles == 16:
            assert s["spec_parts"] > 0, s
        assert ctx.check_invariants() == 0
