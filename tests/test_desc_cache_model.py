"""The host service's descriptor cache (descriptors kept across a file's close
and reopen) tested on the CPU: reuse for an unchanged path, a real open for a
path renamed over or removed, nothing kept for O_CREAT, no descriptor handed
out twice under concurrent opens; built with ThreadSanitizer when available."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path, sanitize):
    exe = str(tmp_path / ("desc_cache_tsan" if sanitize else "desc_cache"))
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", os.path.join(HERE, "csrc", "desc_cache_model.cpp"), "-o", exe]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.parametrize("sanitize", [False, True])
def test_desc_cache(tmp_path, sanitize):
    exe = _build(tmp_path, sanitize)
    if exe is None:
        if sanitize:
            pytest.skip("ThreadSanitizer not available")
        raise AssertionError("descriptor cache model failed to build")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "desc cache model OK" in r.stdout
