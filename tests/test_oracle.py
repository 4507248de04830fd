"""Pins the CPU oracle: FNV specification vectors, hand-derived scan cases
(including the chunk-loop corner cases of grep_text_gpu.cu:298-318), the
literal restatement vs the fast one on adversarial inputs, the reference's own
text + dictionary fixtures, the generator's tally, and numpy for the GEMV."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def D(*words):
    a = np.zeros((len(words), 32), dtype=np.uint8)
    for i, w in enumerate(words):
        b = w if isinstance(w, bytes) else w.encode()
        a[i, :len(b)] = np.frombuffer(b[:32], dtype=np.uint8)
    return a


# ---------------------------------------------------------------- FNV-1a-64
def test_fnv_spec_vectors(orc):
    assert orc.fnv1a64(b"") == 0xcbf29ce484222325
    assert orc.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert orc.fnv1a64(b"foobar") == 0x85944171f73967e8


def test_fnv_differs_on_order_and_length(orc):
    assert orc.fnv1a64(b"ab") != orc.fnv1a64(b"ba")
    assert orc.fnv1a64(b"a\0") != orc.fnv1a64(b"a")


# ---------------------------------------------------------------- scan, by hand
CASES = [
    # (text, dict, chunk, expected)  -- each derived by hand from match_string
    (b"cat cats scat cat\n", ["cat"], 64, [2]),
    (b"cat", ["cat"], 64, [1]),                       # last window gets a NUL terminator
    (b"Cat cat CAT", ["cat", "Cat"], 64, [1, 1]),      # case-sensitive
    (b"snake_case snake case", ["snake", "case", "snake_case"], 64, [1, 1, 1]),  # '_' is a word byte
    (b"don't dont 4cat cat4", ["don", "t", "dont", "cat"], 64, [1, 1, 1, 2]),    # digits and ' split words
    (b"caf\xc3\xa9 cafe", ["caf", "cafe"], 64, [1, 1]),  # bytes >= 0x80 are not word bytes
    (b"cross-bun cross bun", ["cross-bun", "cross", ""], 64, [0, 2, 0]),  # non-word query / empty query
    (b"cat cat", ["cat", "cat"], 64, [2, 2]),          # duplicated query words count alike
    (b"aa\0aa", ["aa"], 64, [2]),                      # NUL inside a file separates words
    # chunk 8, 17 bytes: window [0,17)+NUL is the last window but the cursor only
    # advances by 8, so [8,17) = "bbbbb cc\n" is scanned again: "cc" twice and
    # the suffix "bbbbb" once.
    (b"aaaa bbbbbbbb cc\n", ["aaaa", "bbbbbbbb", "cc", "bbbbb"], 8, [1, 1, 2, 1]),
    # 40 word bytes then "yy z": the first window's look-ahead hits 32 word bytes,
    # so a NUL cuts the run at byte 40 and "yy" starts the next window as a word.
    (b"x" * 40 + b"yy z", ["yy", "z", "x" * 32], 8, [1, 1, 0]),
    # a 32-byte query matches a 32-byte word only
    (b"a" * 32 + b" " + b"a" * 33 + b" " + b"a" * 31, ["a" * 32], 128, [1]),
]


@pytest.mark.parametrize("text,words,chunk,expected", CASES)
def test_scan_known_answers(orc, text, words, chunk, expected):
    d = D(*words)
    assert list(orc.scan_naive(text, d, chunk)) == expected
    assert list(orc.scan_fast(text, d, chunk)) == expected


def test_scan_empty_input(orc):
    d = D("a", "b")
    assert list(orc.scan_naive(b"", d, 64)) == [0, 0]
    assert list(orc.scan_fast(b"", d, 64)) == [0, 0]


def test_scan_bytes_after_nul_in_query_are_ignored(orc):
    d = np.zeros((1, 32), dtype=np.uint8)
    d[0, :2] = np.frombuffer(b"ab", dtype=np.uint8)
    d[0, 3:5] = np.frombuffer(b"cd", dtype=np.uint8)
    assert list(orc.scan_naive(b"ab abcd", d, 64)) == [1]
    assert list(orc.scan_fast(b"ab abcd", d, 64)) == [1]


# ---------------------------------------------------------------- naive vs fast
ALPHABET = np.frombuffer(b"aab_A ' \n\0\xff-1", dtype=np.uint8)


@pytest.mark.parametrize("seed", range(40))
def test_scan_fast_equals_literal_on_adversarial_inputs(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 3000))
    if seed % 4 == 0:   # long word runs around window boundaries
        text = rng.choice(np.frombuffer(b"aab", dtype=np.uint8), size=n)
        text[rng.integers(0, max(n, 1), size=n // 50)] = 32
    else:
        text = rng.choice(ALPHABET, size=n)
    words = []
    for _ in range(int(rng.integers(1, 40))):
        L = int(rng.integers(0, 34))
        w = bytes(rng.choice(np.frombuffer(b"aab_A", dtype=np.uint8), size=L))
        words.append(w)
    if n > 10:   # also query words cut from the text itself
        for _ in range(10):
            a = int(rng.integers(0, n - 1))
            words.append(bytes(text[a:a + int(rng.integers(1, 12))]).split(b"\0")[0])
    words.append(words[0])   # a duplicate
    d = D(*words)
    chunk = int(rng.choice([8, 33, 64, 100, 1000]))
    assert (orc.scan_fast(text, d, chunk) == orc.scan_naive(text, d, chunk)).all()


# ---------------------------------------------------------------- reference data fixtures
def test_golden_pg100_counts(orc):
    g = os.path.join(HERE, "golden")
    text = open(os.path.join(g, "pg100_small.txt"), "rb").read()
    d = np.fromfile(os.path.join(g, "dict_subset.bin"), dtype=np.uint8).reshape(-1, 32)
    exp = json.load(open(os.path.join(g, "pg100_counts.json")))
    assert exp["n_words"] == d.shape[0]
    for chunk, counts in exp["counts"].items():
        counts = np.array(counts, dtype=np.uint32)
        assert (orc.scan_naive(text, d, int(chunk)) == counts).all()
        assert (orc.scan_fast(text, d, int(chunk)) == counts).all()
    # hand check: "the" occurs; the two duplicated trailing records agree
    words = [bytes(r).split(b"\0")[0] for r in d]
    i = [k for k, w in enumerate(words) if w == b"the"]
    assert len(i) >= 2 and len({exp["counts"]["16384"][k] for k in i}) == 1 and exp["counts"]["16384"][i[0]] > 0


# ---------------------------------------------------------------- generator tally
def test_scan_equals_generator_tally():
    from gpufs_amd import datagen as dg
    from oracle import oracle as orc
    voc = dg.vocab(5000, seed=11)
    text, tally = dg.text(12, voc, 1 << 20)
    counts = orc.scan_fast(text, voc, 65536)
    assert (counts == tally).all()
    q = dg.query(5000, 256, seed=13)
    assert (orc.scan_naive(text[:200000], voc[q[:16]], 65536) ==
            orc.scan_fast(text[:200000], voc[q[:16]], 65536)).all()


def test_scan_files_cpu_path_equals_in_memory(tmp_path):
    from gpufs_amd import datagen as dg
    from oracle import oracle as orc
    voc = dg.vocab(2000, seed=5)
    paths = [str(tmp_path / f"f{i}.txt") for i in range(6)]
    sizes = [0, 1, 4097, 70000, 131072 + 33, 300000]
    tally = dg.text_files(paths, sizes, 77, voc, with_tally=True, threads=3)
    q = voc[dg.query(2000, 300, seed=3)]
    for chunk in (65536, 1000):
        counts, status = orc.scan_files(paths, q, chunk, threads=3)
        assert (status == 0).all()
        for i, p in enumerate(paths):
            data = open(p, "rb").read()
            assert len(data) == sizes[i]
            assert (counts[i] == orc.scan_fast(data, q, chunk)).all()
    # no window corner case triggers at chunk 65536 for these sizes except when the
    # final window is longer than the chunk; files below 65568 bytes have one window
    qi = dg.query(2000, 300, seed=3)
    counts, _ = orc.scan_files(paths, q, 65536)
    for i in (0, 1, 2, 3):
        assert (counts[i] == tally[i][qi]).all()


# ---------------------------------------------------------------- FNV over ranges, GEMV, merge
def test_pread_ranges_fnv(tmp_path, orc):
    from gpufs_amd import datagen as dg
    p = str(tmp_path / "r.bin")
    dg.random_file(p, 3 << 20, seed=9, threads=2)
    data = np.fromfile(p, dtype=np.uint8)
    h, got = orc.pread_ranges(p, 6, 1 << 19, threads=3)
    for r in range(6):
        assert int(h[r]) == orc.fnv1a64(data[r << 19:(r + 1) << 19])
        assert got[r] == 1 << 19
    dst = np.zeros(3 << 20, dtype=np.uint8)
    orc.pread_ranges(p, 6, 1 << 19, dst=dst, fnv=False, threads=2)
    assert (dst == data).all()


def test_gemv_f64_against_numpy(orc):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((37, 129)).astype(np.float32)
    x = rng.standard_normal(129).astype(np.float32)
    y = orc.gemv_f64(A, x)
    ref = A.astype(np.float64) @ x.astype(np.float64)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    eye = np.eye(16, dtype=np.float32)
    assert (orc.gemv_f64(eye, x[:16]) == x[:16].astype(np.float64)).all()


def test_gemv_file_against_numpy(orc, tmp_path):
    """The file GEMV: y equals numpy's float64 product, the fp32 mode stays
    within float32 rounding, and each block's checksum is the wrapping sum of
    the block's 8-byte words (ragged last block included)."""
    rng = np.random.default_rng(1)
    rows, cols, rpb = 37, 96, 8
    A = rng.standard_normal((rows, cols)).astype(np.float32)
    x = rng.standard_normal(cols).astype(np.float32)
    p = str(tmp_path / "m.bin")
    A.tofile(p)
    y, cs = orc.gemv_file(p, rows, cols, x, rpb, f64=True, threads=3)
    ref = A.astype(np.float64) @ x.astype(np.float64)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    words = A.view(np.uint64).reshape(rows, cols // 2)
    for b in range(len(cs)):
        exp = np.sum(words[b * rpb:(b + 1) * rpb], dtype=np.uint64)
        assert cs[b] == exp, b
    y32, cs32 = orc.gemv_file(p, rows, cols, x, rpb, f64=False, threads=2)
    assert (cs32 == cs).all()
    scale = np.abs(A).astype(np.float64) @ np.abs(x).astype(np.float64)
    assert (np.abs(y32 - ref) <= 1e-5 * scale).all()
    # identity rows: y is x exactly
    eye = np.eye(cols, dtype=np.float32)
    eye.tofile(p)
    ye, _ = orc.gemv_file(p, cols, cols, x, 5, f64=True, threads=4)
    assert (ye == x.astype(np.float64)).all()


def test_or_merge_known_answer(orc):
    page = np.array([0x01, 0x10, 0x00, 0xF0], dtype=np.uint8)
    disk = np.array([0x02, 0x01, 0x80], dtype=np.uint8)
    assert list(orc.or_merge(page, disk)) == [0x03, 0x11, 0x80, 0xF0]


def _dict(words):
    d = np.zeros((len(words), 32), dtype=np.uint8)
    for i, w in enumerate(words):
        b = w.encode()
        d[i, :len(b)] = np.frombuffer(b, dtype=np.uint8)
    return d


@pytest.fixture
def short_dir():
    """Database names must stay under 64 bytes (the sample's name limit)."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="gt", dir="/tmp")
    yield __import__("pathlib").Path(d)
    shutil.rmtree(d, ignore_errors=True)


def test_grep_text_lines_known_answers(orc, short_dir):
    """grep_text output lines: format "<word> <count> <db>\\n", one line per
    query record (duplicates included) and window with a nonzero count, full
    decimal counts, get_next list parsing."""
    a = short_dir / "a.txt"
    a.write_bytes(b"cat dog cat\nbird")
    b = short_dir / "b.txt"
    b.write_bytes(b"ab " * 1234)
    d = _dict(["cat", "dog", "cow", "cat", "ab"])
    lst = f"  {a},{b}\r\n".encode()
    out = orc.grep_text_lines(d, lst, 65536).decode()
    assert out == (f"cat 2 {a}\ndog 1 {a}\ncat 2 {a}\n" f"ab 1234 {b}\n")
    # tabs / newlines as separators, an empty list, a 64-byte name ends the list
    assert orc.grep_text_lines(d, f"\t{b}\n\n{a}\n".encode(), 65536).decode() == (
        f"ab 1234 {b}\ncat 2 {a}\ndog 1 {a}\ncat 2 {a}\n")
    assert orc.grep_text_lines(d, b"\n , \n", 65536) == b""
    long_name = "/" + "x" * 63
    assert orc.grep_text_lines(d, f"{long_name}\n{a}\n".encode(), 65536) == b""


def test_grep_text_lines_sum_to_file_counts(orc, short_dir):
    """Per-window counts (several windows per file) add up to the file's
    counts from the pinned scan, once per query record holding the word."""
    from gpufs_amd import datagen as dg
    voc = dg.vocab(300, seed=5)
    q = np.ascontiguousarray(voc[dg.query(300, 40, seed=6)])
    paths = [short_dir / f"f{i}.txt" for i in range(3)]
    dg.text_files([str(p) for p in paths], [20000, 7777, 300], 9, voc, threads=2)
    chunk = 1024
    out = orc.grep_text_lines(q, ("\n".join(str(p) for p in paths) + "\n").encode(), chunk).decode()
    words = [bytes(r).rstrip(b"\0").decode() for r in q]
    for p in paths:
        total = {}
        for line in out.splitlines():
            w, c, name = line.split(" ")
            if name == str(p):
                total[w] = total.get(w, 0) + int(c)
        ref = orc.scan_fast(np.fromfile(p, dtype=np.uint8), q, chunk)
        for i, w in enumerate(words):
            assert total.get(w, 0) == ref[i] * words.count(w), (p, w)
    nlines = len(out.splitlines())
    assert nlines > 3 * len(set(words))   # several windows per file contribute lines
