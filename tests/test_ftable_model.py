"""The file table's name index, entry reclamation and radix-node free list
(gpufs_amd/csrc/ftable.h, the code the device API runs) exercised by host
threads on the CPU: a table of 16-64 entries serves names from a pool three
times larger, so reclaim passes run while other threads look up, insert, join
and close.  Built with ThreadSanitizer when the toolchain supports it."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path, sanitize):
    exe = str(tmp_path / ("ftable_model_tsan" if sanitize else "ftable_model"))
    cmd = ["g++", "-O1" if sanitize else "-O2", "-g", "-std=c++17", "-pthread",
           os.path.join(HERE, "csrc", "ftable_model.cpp"), "-o", exe]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.parametrize("sanitize,threads,iters,max_files",
                         [(False, 8, 20000, 24), (False, 32, 5000, 16), (False, 64, 2000, 64), (True, 8, 3000, 24)])
def test_ftable_reclaim(tmp_path, sanitize, threads, iters, max_files):
    exe = _build(tmp_path, sanitize)
    if exe is None:
        if sanitize:
            pytest.skip("ThreadSanitizer not available")
        raise AssertionError("ftable model failed to build")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(threads), str(iters), str(max_files)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ftable model OK" in r.stdout
