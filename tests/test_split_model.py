"""CPU model of the split grep scan (grep_file_split in gpufs_amd/csrc/grep_scan.h):
a file is cut into parts at non-word bytes, each part counts the tokens that
start in it through windows that also end before non-word bytes, and the last
part adds the reference's re-scanned tail found from the reference's window
chain, or falls back to the reference loop when a boundary meets a run of 32
or more word bytes.  The model follows the device code step by step; the
counts must equal the oracle's restatement of the reference chunk loop
(grep_text_gpu.cu:298-318) on adversarial inputs, and the fast path (no
fallback) must be the one taken on ordinary text."""
import numpy as np
import pytest

ALPHA = np.zeros(256, dtype=bool)
for c in range(256):
    ALPHA[c] = (65 <= c <= 90) or (97 <= c <= 122) or c == 95


def run_at(data, x, avail):
    """Word bytes at data[x:], at most `avail` of them examined (device: lds_word_run)."""
    n = 0
    while n < avail and x + n < len(data) and ALPHA[data[x + n]]:
        n += 1
    return n


def window_tokens(data, a, e, qset, counts):
    """Tokens of the window [a, e) scanned from a word start at a, bytes from e on
    counting as separators (match_string over one window)."""
    i = a
    while i < e:
        if ALPHA[data[i]]:
            j = i
            while j < e and ALPHA[data[j]]:
                j += 1
            w = bytes(data[i:j])
            if w in qset:
                counts[w] = counts.get(w, 0) + 1
            i = j
        else:
            i += 1


def part_windows(data, chunk, pb, pe):
    size = len(data)
    cursor, first, lr, wins = pb, pb > 0, False, []
    while True:
        final = pe - cursor <= chunk
        rend = min(pe + 64 if final else cursor + chunk + 32, size)
        start = cursor
        if first:
            r = run_at(data, cursor, min(64, rend - cursor))
            if r >= 32:
                lr = True
            else:
                start = cursor + r
            first = False
        if final:
            e = size
            if pe < size:
                r = run_at(data, pe, min(64, rend - pe))
                if r >= 32:
                    lr = True
                else:
                    e = pe + r
            wins.append((start, e))
            return wins, lr
        ov = run_at(data, cursor + chunk, min(32, rend - cursor - chunk))
        if ov >= 32:
            lr = True
        wins.append((start, cursor + chunk + ov))
        cursor += chunk + ov
        if cursor >= pe:
            return wins, lr


def chain_last(data, chunk):
    size, c = len(data), 0
    while size - c >= chunk + 32:
        q = c + chunk
        r = run_at(data, q, 32)
        if r == 32:
            return c, True
        c = q + r
    return c, False


def split_counts(data, words, chunk, parts, min_part):
    """Returns (counts per word, took_fallback)."""
    size = len(data)
    assert min_part >= 256   # the launcher's bound: parts are then never empty
    kf = max(1, min(parts, size // min_part))
    cstr = [w.split(b"\0")[0] for w in words]   # a record is a C string (grep_text_gpu.cu:139-160)
    qset = {w for w in cstr if w and all(ALPHA[b] for b in w)}
    if kf == 1:
        return None, True   # one part: the reference loop itself
    counts, longrun = {}, False
    def split_point(j):   # part 0 takes 4/5 of a share (it also walks the chain)
        return 0 if j == 0 else (size if j >= kf else (size * (5 * j - 1) // (5 * kf - 1)) & ~63)
    for j in range(kf):
        pb, pe = split_point(j), split_point(j + 1)
        assert pe - pb >= 64 or kf == 1
        wins, lr = part_windows(data, chunk, pb, pe)
        longrun |= lr
        for a, e in wins:
            window_tokens(data, a, e, qset, counts)
    cl, cut = chain_last(data, chunk)
    if longrun or cut:
        return None, True
    # every later window is a last one too: the cursor moves by chunk each time
    x = cl + chunk
    while x < size:
        window_tokens(data, x, size, qset, counts)
        x += chunk
    return [counts.get(w, 0) if w in qset else 0 for w in cstr], False


def dict_rec(words):
    d = np.zeros((len(words), 32), dtype=np.uint8)
    for i, w in enumerate(words):
        d[i, :len(w)] = np.frombuffer(w, dtype=np.uint8)
    return d


def adversarial(rng, n, long_runs):
    letters = np.frombuffer(b"abcab_Zz", dtype=np.uint8)
    seps = np.frombuffer(b" \n,.0\x00\xff", dtype=np.uint8)
    out = []
    while sum(len(x) for x in out) < n:
        L = int(rng.integers(1, 40 if long_runs else 9))
        out.append(letters[rng.integers(0, len(letters), L)])
        out.append(seps[rng.integers(0, len(seps), int(rng.integers(1, 3)))])
    return np.concatenate(out)[:n].copy()


@pytest.mark.parametrize("long_runs", [False, True])
def test_split_equals_reference(orc, long_runs):
    rng = np.random.default_rng(11 if long_runs else 12)
    fast_path = 0
    for trial in range(60):
        n = int(rng.integers(64, 6000))
        data = adversarial(rng, n, long_runs)
        words = [b"ab", b"abc", b"cab", b"a", b"b_", b"Zz", b"ab_Z", b"cabca", b"x", b"ba",
                 bytes(data[10:14]) if n > 14 else b"q"]
        words = [w for w in words if w]
        chunk = int(rng.choice([8, 16, 64, 100, 256]))
        parts = int(rng.integers(2, 9))
        got, fell_back = split_counts(data, words, chunk, parts, 256)
        ref = list(orc.scan_fast(data.tobytes(), dict_rec(words), chunk))
        if fell_back:
            continue   # the device then runs the reference loop itself
        fast_path += 1
        assert got == ref, (trial, n, chunk, parts)
    if not long_runs:
        assert fast_path >= 50   # ordinary text takes the split path
    else:
        assert fast_path >= 5


def test_split_tail_rescan_counts_twice(orc):
    """A file whose reference last window is longer than the chunk: its final
    bytes are counted again by the reference, and the split path adds them.
    With a chunk below 32 bytes the last window repeats several times."""
    n = next(n for n in range(200, 400) if len(b"ab " * n) - chain_last(np.frombuffer(b"ab " * n, dtype=np.uint8), 64)[0] > 64)
    data = np.frombuffer(b"ab " * n + b"cab", dtype=np.uint8).copy()
    words = [b"ab", b"cab"]
    chunk = 64
    cl, cut = chain_last(data, chunk)
    assert not cut and len(data) - cl > chunk
    got, fb = split_counts(data, words, chunk, 2, 256)
    ref = list(orc.scan_fast(data.tobytes(), dict_rec(words), chunk))
    assert not fb and got == ref
    assert ref != [n, 1]   # the re-scanned tail is counted again
    data = np.frombuffer(b"ab " * 200, dtype=np.uint8).copy()
    got, fb = split_counts(data, [b"ab"], 8, 2, 256)
    ref = list(orc.scan_fast(data.tobytes(), dict_rec([b"ab"]), 8))
    assert not fb and got == ref and ref[0] > 201


def test_model_planted_mistake_is_caught(orc):
    """A plausible slip (forgetting the re-scanned tail) fails against the oracle."""
    rng = np.random.default_rng(5)
    bad = 0
    for _ in range(40):
        data = adversarial(rng, int(rng.integers(200, 3000)), False)
        words = [b"ab", b"abc", b"cab", b"a", b"Zz"]
        chunk = 64
        got, fb = split_counts(data, words, chunk, 3, 256)
        if fb:
            continue
        cl, _ = chain_last(data, chunk)
        if len(data) - cl > chunk:
            counts, qset = {}, set(words)
            x = cl + chunk
            while x < len(data):
                window_tokens(data, x, len(data), qset, counts)
                x += chunk
            no_tail = [g - counts.get(w, 0) for g, w in zip(got, words)]
            ref = list(orc.scan_fast(data.tobytes(), dict_rec(words), chunk))
            bad += no_tail != ref
    assert bad > 0
