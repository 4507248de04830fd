"""The C ABI library loads and exports every entry point include/*.h declares
(no GPU work is done here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        for m in re.finditer(r"\b(gpufs_[A-Za-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    from gpufs_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    for n in names:
        getattr(L, n)
    # the Python binding types every declared entry point
    assert names == set(_lib.SIGNATURES), sorted(names ^ set(_lib.SIGNATURES))


def test_config_default_and_strerror():
    from gpufs_amd import _lib
    c = _lib.default_config()
    assert c.page_size == 256 << 10
    assert c.cache_bytes == 1 << 30
    assert c.ring_depth > 0 and c.host_threads > 0 and c.max_files > 0
    L = _lib.lib()
    assert L.gpufs_strerror(-1) == b"file table full"
    assert L.gpufs_strerror(-2) == b"host open failed"
    assert L.gpufs_strerror(-12).startswith(b"Cannot allocate")


def test_config_struct_layout_matches_header():
    from gpufs_amd import _lib
    assert ctypes.sizeof(_lib.Config) == 40
    assert ctypes.sizeof(_lib.Stats) == 8 * len(_lib.STATS_FIELDS)


def test_null_arguments_are_rejected_without_gpu():
    from gpufs_amd import _lib
    L = _lib.lib()
    assert L.gpufs_init(None, None) == -22
    assert L.gpufs_destroy(None) == -22
    assert L.gpufs_run_handler(None) == -22
    assert L.gpufs_launch_gread_ranges(None, None, 0, 0, None, None, 64, None) == -22


def test_compat_example_built():
    """build() also builds the reference-style driver used by the GPU test."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpufs_amd", "examples",
                       "compat_copy")
    assert os.access(exe, os.X_OK)
