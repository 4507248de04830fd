"""GPU parity of the gmmap GEMV (mapped-matrix workload) with continuous
eviction: mapped bytes exact (64-bit word checksum per workgroup against the
file), y within 1e-5 of the FP64 oracle (max |y - y64| / max |y64|), and y
bit-identical between two runs (fixed reduction order)."""
import numpy as np
import pytest

from gpufs_amd import datagen as dg
from tests.gpu_helpers import gemv_mmap, make_ctx, say

pytestmark = pytest.mark.gpu


def _word_sums(A, rows_per_wg):
    u = A.view(np.uint64).reshape(A.shape[0], -1)
    per_row = u.sum(axis=1, dtype=np.uint64)
    n = (A.shape[0] + rows_per_wg - 1) // rows_per_wg
    return np.array([per_row[i * rows_per_wg:(i + 1) * rows_per_wg].sum(dtype=np.uint64) for i in range(n)],
                    dtype=np.uint64)


@pytest.mark.parametrize("page_kib,cache_mib", [(256, 16), (64, 4)])
def test_gemv_mmap_with_eviction(tmp_path, orc, page_kib, cache_mib):
    say(f"test_gemv_mmap_with_eviction page={page_kib}KiB cache={cache_mib}MiB")
    rows, cols, rpw = 512, 32768, 16            # 64 MiB matrix through a 16 / 4 MiB cache
    p = str(tmp_path / "A.bin")
    dg.matrix_file(p, rows, cols, seed=21, threads=4)
    A = np.fromfile(p, dtype=np.float32).reshape(rows, cols)
    x = dg.normal_f32(cols, seed=22, stream=1 << 40)
    y64 = orc.gemv_f64(A, x)
    with make_ctx(page_size=page_kib << 10, cache_bytes=cache_mib << 20, ring_depth=128) as ctx:
        y, st, cs = gemv_mmap(ctx, p, rows, cols, x, rpw)
        assert (st == 0).all(), st
        want = _word_sums(A, rpw)
        bad = np.nonzero(cs != want)[0]
        assert bad.size == 0, f"first pass: workgroups with wrong mapped bytes {bad.tolist()[:16]}, stats {ctx.stats()}"
        err = np.abs(y.astype(np.float64) - y64).max() / np.abs(y64).max()
        assert err <= 1e-5, err
        y2, st2, cs2 = gemv_mmap(ctx, p, rows, cols, x, rpw)
        bad_wg = np.nonzero(cs2 != _word_sums(A, rpw))[0]
        bad_rows = np.nonzero(y2 != y)[0]
        s = ctx.stats()
        assert (st2 == 0).all() and bad_wg.size == 0 and bad_rows.size == 0, (
            f"second pass: workgroups with wrong mapped bytes {bad_wg.tolist()[:16]}, "
            f"rows that differ {bad_rows.tolist()[:16]} of {bad_rows.size}, stats {s}")
        assert s["evictions"] > 0
        assert ctx.check_invariants() == 0
