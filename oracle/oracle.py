"""ctypes loader for the CPU oracle (oracle.c).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, never by the gpufs_amd package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.orc_scan_naive.argtypes = [P, ctypes.c_size_t, u32, P, u32, P]
        L.orc_scan_fast.argtypes = [P, ctypes.c_size_t, u32, P, u32, P]
        L.orc_scan_files.argtypes = [P, u32, u32, u32, P, u32, P, P, i32]
        L.orc_fnv1a64.argtypes = [P, ctypes.c_size_t]
        L.orc_fnv1a64.restype = u64
        L.orc_pread_ranges.argtypes = [ctypes.c_char_p, u32, u64, P, P, P, i32]
        L.orc_gemv_f64.argtypes = [P, u64, u64, P, P]
        L.orc_gemv_file.argtypes = [ctypes.c_char_p, u64, u64, P, u64, P, P, i32, i32]
        L.orc_gemv_file.restype = i32
        L.orc_grep_text_lines.argtypes = [P, u32, ctypes.c_char_p, u64, u32, P, u64]
        L.orc_grep_text_lines.restype = ctypes.c_int64
        L.orc_or_merge.argtypes = [P, P, u64]
        for f in (L.orc_scan_naive, L.orc_scan_fast, L.orc_scan_files, L.orc_pread_ranges):
            f.restype = i32
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dict_array(dict_words):
    d = np.asarray(dict_words, dtype=np.uint8)
    assert d.ndim == 2 and d.shape[1] == 32
    return np.ascontiguousarray(d)


def scan_naive(data, dict_words, chunk):
    data = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    d = _dict_array(dict_words)
    out = np.zeros(d.shape[0], dtype=np.uint32)
    rc = lib().orc_scan_naive(_ptr(data), data.size, chunk, _ptr(d), d.shape[0], _ptr(out))
    assert rc == 0
    return out


def scan_fast(data, dict_words, chunk):
    data = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    d = _dict_array(dict_words)
    out = np.zeros(d.shape[0], dtype=np.uint32)
    rc = lib().orc_scan_fast(_ptr(data), data.size, chunk, _ptr(d), d.shape[0], _ptr(out))
    assert rc == 0
    return out


def scan_files(paths, dict_words, chunk, threads=1, stride=256):
    d = _dict_array(dict_words)
    names = np.zeros((len(paths), stride), dtype=np.uint8)
    for i, p in enumerate(paths):
        b = os.fsencode(p)
        assert len(b) < stride
        names[i, : len(b)] = np.frombuffer(b, dtype=np.uint8)
    counts = np.zeros((len(paths), d.shape[0]), dtype=np.uint32)
    status = np.zeros(len(paths), dtype=np.int32)
    rc = lib().orc_scan_files(_ptr(names), stride, len(paths), chunk, _ptr(d), d.shape[0], _ptr(counts),
                              _ptr(status), threads)
    assert rc == 0
    return counts, status


def fnv1a64(data):
    a = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    return int(lib().orc_fnv1a64(_ptr(a), a.size))


def pread_ranges(path, nranges, range_bytes, dst=None, fnv=True, threads=1):
    h = np.zeros(nranges, dtype=np.uint64) if fnv else None
    got = np.zeros(nranges, dtype=np.int64)
    rc = lib().orc_pread_ranges(os.fsencode(path), nranges, range_bytes,
                                _ptr(dst) if dst is not None else None,
                                _ptr(h) if h is not None else None, _ptr(got), threads)
    assert rc == 0, rc
    return h, got


def gemv_f64(A, x):
    A = np.ascontiguousarray(A, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(A.shape[0], dtype=np.float64)
    lib().orc_gemv_f64(_ptr(A), A.shape[0], A.shape[1], _ptr(x), _ptr(y))
    return y


def gemv_file(path, rows, cols, x, rows_per_block, f64=True, threads=1):
    """(y float64[rows], per-block wrapping uint64 sums of the block's 8-byte words)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(rows, dtype=np.float64)
    nb = (rows + rows_per_block - 1) // rows_per_block
    cs = np.zeros(nb, dtype=np.uint64)
    rc = lib().orc_gemv_file(os.fsencode(path), rows, cols, _ptr(x), rows_per_block, _ptr(y), _ptr(cs),
                             1 if f64 else 0, threads)
    if rc != 0:
        raise OSError(-rc, f"orc_gemv_file({path})")
    return y, cs


def grep_text_lines(dict_words, dblist, chunk):
    """Output of the grep_text sample for a database list (bytes, deterministic order)."""
    d = _dict_array(dict_words)
    nw = d.size // 32
    lst = dblist if isinstance(dblist, bytes) else dblist.encode()
    n = lib().orc_grep_text_lines(_ptr(d), nw, lst, len(lst), chunk, None, 0)
    if n < 0:
        raise OSError(-n, "orc_grep_text_lines")
    out = np.zeros(max(1, n), dtype=np.uint8)
    m = lib().orc_grep_text_lines(_ptr(d), nw, lst, len(lst), chunk, _ptr(out), n)
    assert m == n
    return out[:n].tobytes()


def or_merge(page, disk):
    p = np.array(page, dtype=np.uint8, copy=True)
    d = np.ascontiguousarray(disk, dtype=np.uint8)
    lib().orc_or_merge(_ptr(p), _ptr(d), min(d.size, p.size))
    return p
