"""A reference-style host driver (initializer / streamMgr->kernelStream /
run_gpufs_handler / delete, the shape of samples/memcopy/fs.cu) built against
gpufs_amd/csrc/gpufs_compat.h copies a file through gread; its bytes must equal
the file."""
import os
import subprocess

import numpy as np
import pytest

from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpufs_amd", "examples",
                   "compat_copy")


@pytest.mark.parametrize("size,nwg", [(3 * (1 << 20) + 17, 16), (1 << 20, 1)])
def test_reference_style_driver(tmp_path, size, nwg):
    say(f"test_reference_style_driver size={size} nwg={nwg}")
    assert os.access(EXE, os.X_OK), "gpufs_amd/examples/compat_copy is not built"
    p = tmp_path / "in.bin"
    np.random.default_rng(size).integers(0, 256, size, dtype=np.uint8).tofile(p)
    env = dict(os.environ, GPUFS_WAIT_TIMEOUT_S="10", GPUFS_COMPAT_CACHE_BYTES=str(64 << 20),
               GPUFS_COMPAT_PAGE_SIZE=str(64 << 10))
    r = subprocess.run([EXE, str(p), str(nwg)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "compat_copy OK" in r.stdout
