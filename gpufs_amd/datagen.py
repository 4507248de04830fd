"""Synthetic inputs for the read-path workloads (ctypes over libgpufs_datagen.so).

Every generator is seeded and deterministic; see csrc/datagen.cpp for the
exact recipes.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None

P, U32, U64, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int

# seeds of the benchmark inputs
SEED_VOCAB = 0x5EED0001
SEED_QUERY = 0x5EED0002
SEED_C0 = 0x5EED0100
SEED_C1 = 0x5EED0200
SEED_C2 = 0x5EED0300
SEED_C3 = 0x5EED0400
SEED_C4 = 0x5EED0500


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "libgpufs_datagen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run `python -m gpufs_amd.build`")
        L = ctypes.CDLL(path)
        L.dg_vocab.argtypes = [U64, U32, P]
        L.dg_query.argtypes = [U64, U32, U32, P]
        L.dg_text.argtypes = [U64, P, U32, ctypes.c_double, U64, P, P]
        L.dg_text_files.argtypes = [P, U32, U32, P, U64, P, U32, ctypes.c_double, P, I]
        L.dg_random_fill.argtypes = [P, U64, U64, U64]
        L.dg_random_file.argtypes = [ctypes.c_char_p, U64, U64, I]
        L.dg_normal_f32.argtypes = [P, U64, U64, U64]
        L.dg_matrix_file.argtypes = [ctypes.c_char_p, U64, U64, U64, I]
        L.dg_lognormal_sizes.argtypes = [U64, U32, ctypes.c_double, ctypes.c_double, U64, U64, P]
        L.dg_zipf_sample.argtypes = [U64, U32, ctypes.c_double, U32, P]
        L.dg_poisson_gaps.argtypes = [U64, ctypes.c_double, U32, P]
        for f in (L.dg_vocab, L.dg_query, L.dg_text, L.dg_text_files, L.dg_random_fill, L.dg_random_file,
                  L.dg_normal_f32, L.dg_matrix_file, L.dg_lognormal_sizes, L.dg_zipf_sample, L.dg_poisson_gaps):
            f.restype = I
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {rc}")


def vocab(n=50000, seed=SEED_VOCAB):
    out = np.zeros((n, 32), dtype=np.uint8)
    _ok(lib().dg_vocab(seed, n, _p(out)), "dg_vocab")
    return out


def query(nvocab, nq=1024, seed=SEED_QUERY):
    out = np.zeros(nq, dtype=np.uint32)
    _ok(lib().dg_query(seed, nvocab, nq, _p(out)), "dg_query")
    return out


def text(seed, voc, size, s=1.0):
    out = np.zeros(size, dtype=np.uint8)
    tally = np.zeros(voc.shape[0], dtype=np.uint32)
    _ok(lib().dg_text(seed, _p(voc), voc.shape[0], s, size, _p(out), _p(tally)), "dg_text")
    return out, tally


def names_array(paths, stride=256):
    a = np.zeros((len(paths), stride), dtype=np.uint8)
    for i, p in enumerate(paths):
        b = os.fsencode(p)
        if len(b) >= min(stride, 128):
            raise ValueError(f"path too long: {p}")
        a[i, :len(b)] = np.frombuffer(b, dtype=np.uint8)
    return a


def text_files(paths, sizes, seed_base, voc, s=1.0, with_tally=False, threads=8):
    names = names_array(paths)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    tally = np.zeros((len(paths), voc.shape[0]), dtype=np.uint32) if with_tally else None
    _ok(lib().dg_text_files(_p(names), names.shape[1], len(paths), _p(sizes), seed_base, _p(voc), voc.shape[0], s,
                            _p(tally) if tally is not None else None, threads), "dg_text_files")
    return tally


def random_bytes(off, length, seed):
    out = np.zeros(length, dtype=np.uint8)
    _ok(lib().dg_random_fill(_p(out), off, length, seed), "dg_random_fill")
    return out


def random_file(path, size, seed, threads=8):
    _ok(lib().dg_random_file(os.fsencode(path), size, seed, threads), "dg_random_file")


def normal_f32(n, seed, stream=0):
    out = np.zeros(n + (n & 1), dtype=np.float32)
    _ok(lib().dg_normal_f32(_p(out), n, seed, stream), "dg_normal_f32")
    return out[:n]


def matrix_file(path, rows, cols, seed, threads=8):
    _ok(lib().dg_matrix_file(os.fsencode(path), rows, cols, seed, threads), "dg_matrix_file")


def lognormal_sizes(n, seed, median=65536.0, sigma=1.5, lo=4096, hi=64 << 20):
    out = np.zeros(n, dtype=np.uint64)
    _ok(lib().dg_lognormal_sizes(seed, n, median, sigma, lo, hi, _p(out)), "dg_lognormal_sizes")
    return out


def zipf_sample(nitems, count, seed, s=1.1):
    out = np.zeros(count, dtype=np.uint32)
    _ok(lib().dg_zipf_sample(seed, nitems, s, count, _p(out)), "dg_zipf_sample")
    return out


def poisson_gaps(rate, count, seed):
    out = np.zeros(count, dtype=np.float64)
    _ok(lib().dg_poisson_gaps(seed, rate, count, _p(out)), "dg_poisson_gaps")
    return out
