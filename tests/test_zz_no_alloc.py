"""ixg_accumulate_device_noalloc against ixg_accumulate_device, through the
C ABI: two calls in a row on the same GPU (the second with a smaller frame
and fewer samples, so stale scratch contents would show), every per-pixel
FP64 sum bit for bit; frames larger than one scratch chunk; a panorama in
the scratch slot; the size limits rejected with IXG_RESOURCE_LIMIT and the
caller's buffer untouched."""
import numpy as np
import pytest

from indicatrix_amd import abi, tracer, workloads
from indicatrix_amd._lib import IxgError

pytestmark = pytest.mark.gpu


def _sums(backend, scene, offset, spp, noalloc):
    import torch
    acc = torch.zeros(scene.width * scene.height * 3, dtype=torch.float64, device="cuda:0")
    acc += 0.25  # the call ADDs
    if noalloc:
        backend.accumulate_device_noalloc(scene, offset, spp, acc.data_ptr())
    else:
        backend.accumulate_device(scene, offset, spp, acc.data_ptr())
    torch.cuda.synchronize()
    return acc.cpu().numpy()


def test_noalloc_equals_accumulate_device_on_two_calls_in_a_row():
    be = tracer.GpuBackend(0)
    try:
        cases = [(workloads.film_c1_scene(128, 128), 64, 64),        # the C1 frame: one full scratch chunk
                 (workloads.film_c1_scene(40, 24), 7, 8),            # smaller and fewer samples next
                 (workloads.film_c2_scene(64, 48, 8), 3, 8),          # uniaxial, absorbing
                 (workloads.film_c1_scene(200, 160), 0, 40)]          # 1.28 M paths: two chunks
        for sc, offset, spp in cases:
            print(f"noalloc {sc.width}x{sc.height}x{spp} offset {offset}", flush=True)
            a = _sums(be, sc, offset, spp, noalloc=True)
            b = _sums(be, sc, offset, spp, noalloc=False)
            assert a.tobytes() == b.tobytes()
            assert a.sum() > 0.25 * a.size
    finally:
        be.close()


def test_noalloc_panorama_in_the_scratch_slot():
    be = tracer.GpuBackend(0)
    try:
        sc = workloads.film_c1_scene(32, 24)
        rng = np.random.default_rng(3)
        sc.set_hdr_map(rng.uniform(0.0, 3.0, (64, 128, 3)).astype(np.float32))
        print("noalloc panorama 32x24x8", flush=True)
        assert _sums(be, sc, 0, 8, True).tobytes() == _sums(be, sc, 0, 8, False).tobytes()
    finally:
        be.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    import torch
    be = tracer.GpuBackend(0)
    try:
        sc = workloads.film_c1_scene(1, 1)
        acc = torch.full((3,), 0.5, dtype=torch.float64, device="cuda:0")
        with pytest.raises(IxgError) as e:
            be.accumulate_device_noalloc(sc, 0, abi.NOALLOC_MAX_PATHS + 1, acc.data_ptr())
        assert e.value.status == abi.IXG_RESOURCE_LIMIT
        big = workloads.film_c1_scene(4, 4)
        big.set_hdr_map(np.ones((513, 1024, 3), np.float32))   # one row over the slot
        acc16 = torch.full((48,), 0.5, dtype=torch.float64, device="cuda:0")
        with pytest.raises(IxgError) as e:
            be.accumulate_device_noalloc(big, 0, 1, acc16.data_ptr())
        assert e.value.status == abi.IXG_RESOURCE_LIMIT
        torch.cuda.synchronize()
        assert torch.all(acc == 0.5) and torch.all(acc16 == 0.5)
        assert not be.is_lost()
        # exactly at the limits is accepted
        be.accumulate_device_noalloc(sc, 0, abi.NOALLOC_MAX_PATHS, acc.data_ptr())
        ok = workloads.film_c1_scene(4, 4)
        ok.set_hdr_map(np.ones((512, 1024, 3), np.float32))
        be.accumulate_device_noalloc(ok, 0, 1, acc16.data_ptr())
        torch.cuda.synchronize()
        assert torch.all(acc > 0.5)
    finally:
        be.close()
