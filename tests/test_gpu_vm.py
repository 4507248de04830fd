"""ActivePointers (FatPointer / gvmmap with a per-workgroup TLB): threads walk
records with moveTo and += across pages; every record sum must equal the
file's, no dereference may fail, and every page pin must be returned once the
pointers and the TLB are gone."""
import ctypes

import numpy as np
import pytest

from gpufs_amd import _lib
from tests.gpu_helpers import dev_names, dev_zeros, finish, make_ctx, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tlb,page_kib,cache_mib,rec_elems,workgroups,threads", [
    (8, 16, 4, 1024, 4, 256),      # 8 lines for 16 waves: refills and lanes mapping directly
    (32, 64, 8, 1024, 8, 512),     # the mosaic sample's TLB size, 16 records per page
    (0, 16, 4, 2048, 2, 1024),     # no TLB: every page mapped by the lanes themselves
])
def test_vm_records(tmp_path, tlb, page_kib, cache_mib, rec_elems, workgroups, threads):
    say(f"test_vm_records tlb={tlb} page={page_kib}K cache={cache_mib}M rec={rec_elems} wg={workgroups}x{threads}")
    torch = torch_cuda()
    nrec = 3000
    rng = np.random.default_rng(tlb * 7 + page_kib)
    data = rng.integers(0, 1 << 32, nrec * rec_elems, dtype=np.uint64).astype(np.uint32)
    p = str(tmp_path / "recs.bin")
    data.tofile(p)
    rec_sums = data.reshape(nrec, rec_elems).astype(np.uint64).sum(axis=1)
    per_wave = 40
    nslots = workgroups * (threads // 64) * per_wave
    order = rng.integers(0, nrec, nslots, dtype=np.uint32)
    with make_ctx(page_size=page_kib << 10, cache_bytes=cache_mib << 20, ring_depth=128) as ctx:
        ctx.host_start()
        name = dev_names([p])
        order_d = torch.from_numpy(order).cuda()
        for _ in range(2):
            sums = dev_zeros(nslots, torch.int64)
            st = dev_zeros(1, torch.int32)
            torch.cuda.synchronize()
            _lib.check("gpufs_launch_vm_records", _lib.lib().gpufs_launch_vm_records(
                ctx.h, ctypes.c_void_p(name.data_ptr()), rec_elems, ctypes.c_void_p(order_d.data_ptr()), per_wave,
                ctypes.c_void_p(sums.data_ptr()), ctypes.c_void_p(st.data_ptr()), tlb, workgroups, threads, None))
            finish(ctx)
            assert int(st.cpu().item()) == 0
            got = sums.cpu().numpy().view(np.uint64)
            bad = np.nonzero(got != rec_sums[order])[0]
            assert bad.size == 0, (bad[:8], got[bad[:4]], rec_sums[order][bad[:4]])
        s = ctx.stats()
        assert s["vm_errors"] == 0 and s["timeouts"] == 0
        ctx.host_stop()
        assert ctx.check_invariants() == 0   # every pin returned
