"""The file-state open/close protocol (gpufs_amd/csrc/fstate.h, the code the
device API runs) exercised by host threads on the CPU; built with
ThreadSanitizer when the toolchain supports it."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path, sanitize):
    exe = str(tmp_path / ("fstate_model_tsan" if sanitize else "fstate_model"))
    cmd = ["g++", "-O1" if sanitize else "-O2", "-g", "-std=c++17", "-pthread",
           os.path.join(HERE, "csrc", "fstate_model.cpp"), "-o", exe]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.parametrize("sanitize,threads,iters", [(False, 16, 20000), (False, 64, 2000), (True, 8, 3000)])
def test_fstate_protocol(tmp_path, sanitize, threads, iters):
    exe = _build(tmp_path, sanitize)
    if exe is None:
        if sanitize:
            pytest.skip("ThreadSanitizer not available")
        raise AssertionError("fstate model failed to build")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(threads), str(iters)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=180, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "fstate model OK" in r.stdout
