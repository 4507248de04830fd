"""The synthetic workload generators are deterministic and follow their stated recipes."""
import numpy as np

from gpufs_amd import datagen as dg


def test_vocab_unique_lowercase_lengths():
    v = dg.vocab(5000, seed=1)
    words = [bytes(r).split(b"\0")[0] for r in v]
    assert len(set(words)) == 5000
    assert all(3 <= len(w) <= 12 and w.isalpha() and w.islower() for w in words)
    assert (dg.vocab(5000, seed=1) == v).all()
    assert not (dg.vocab(5000, seed=2) == v).all()


def test_query_distinct():
    q = dg.query(50000, 1024, seed=3)
    assert len(set(q.tolist())) == 1024 and q.max() < 50000


def test_text_format_and_size():
    v = dg.vocab(1000, seed=4)
    t, tally = dg.text(5, v, 100000)
    assert t.size == 100000
    s = bytes(t)
    assert b"  " not in s.rstrip(b" ")
    lines = s.rstrip(b" ").split(b"\n")
    for line in lines[:-1]:
        assert 8 <= len(line.split(b" ")) <= 16
    assert sum(len(l.split()) for l in lines) == tally.sum()
    # Zipf(1): rank 1 beats rank 100 by a wide margin
    assert tally[0] > 20 * max(1, tally[99])


def test_random_fill_is_chunk_addressable():
    full = dg.random_bytes(0, 3 << 20, seed=7)
    for off, ln in [(0, 10), (12345, 999), ((1 << 20) - 3, 10), ((2 << 20) + 5, (1 << 20) - 5)]:
        assert (dg.random_bytes(off, ln, seed=7) == full[off:off + ln]).all()


def test_random_file_matches_fill(tmp_path):
    p = str(tmp_path / "r.bin")
    dg.random_file(p, (5 << 20) + 77, seed=8, threads=3)
    data = np.fromfile(p, dtype=np.uint8)
    assert data.size == (5 << 20) + 77
    assert (data == dg.random_bytes(0, data.size, seed=8)).all()


def test_normal_and_matrix(tmp_path):
    x = dg.normal_f32(100000, seed=9)
    assert abs(float(x.mean())) < 0.02 and abs(float(x.std()) - 1) < 0.02
    p = str(tmp_path / "m.bin")
    dg.matrix_file(p, 10, 64, seed=10, threads=2)
    m = np.fromfile(p, dtype=np.float32).reshape(10, 64)
    assert (m[3] == dg.normal_f32(64, seed=10, stream=3)).all()


def test_lognormal_sizes():
    s = dg.lognormal_sizes(20000, seed=11)
    assert s.min() >= 4096 and s.max() <= 64 << 20
    med = float(np.median(s))
    assert 0.9 * 65536 < med < 1.1 * 65536


def test_zipf_and_poisson():
    z = dg.zipf_sample(1000, 50000, seed=12, s=1.1)
    c = np.bincount(z, minlength=1000)
    assert c[0] > c[10] > c[500]
    g = dg.poisson_gaps(1000.0, 20000, seed=13)
    assert abs(g.mean() - 1e-3) < 5e-5
