"""Host side of the RPC ring tested on the CPU: client threads play the
device protocol against the real host service; built with ThreadSanitizer
when the toolchain supports it."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path, sanitize):
    exe = str(tmp_path / ("ring_model_tsan" if sanitize else "ring_model"))
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", os.path.join(HERE, "csrc", "ring_model.cpp"), "-o", exe]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.parametrize("sanitize", [False, True])
def test_ring_protocol(tmp_path, sanitize):
    exe = _build(tmp_path, sanitize)
    if exe is None:
        if sanitize:
            pytest.skip("ThreadSanitizer not available")
        raise AssertionError("ring model failed to build")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ring model OK" in r.stdout
