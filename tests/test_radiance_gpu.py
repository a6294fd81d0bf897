"""Radiance payload kernels on the device against the oracle, bit for bit:
the byte shuffle and its inverse over every f32 bit pattern class, and the
guarded add (shuffled and raw) with its dropped-pixel count, on aligned
bodies, ragged tails and offset (unaligned) buffers."""
import sys
import time

import numpy as np
import pytest

from oracle import radiance_oracle as ro
from radiance_vectors import mixed_bits, mixed_bits_fast, radiance_like

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("n_values", [0, 1, 3, 4, 12, 997, 3 * 4096, 3 * 1920 * 1080, 3 * 1000003])
def test_shuffle_round_trip_matches_oracle(backend, n_values):
    import torch
    from indicatrix_amd import radiance
    print(f"shuffle n_values={n_values}", flush=True)
    bits = mixed_bits(n_values, 7) if n_values < 5000 else mixed_bits_fast(n_values, 7)
    src = _dev(bits.view(np.int32))
    sh = torch.empty_like(src)
    back = torch.empty_like(src)
    radiance.shuffle(backend, src, sh)
    radiance.unshuffle(backend, sh, back)
    raw = bits.astype("<u4").tobytes()
    assert sh.cpu().numpy().tobytes() == ro.shuffle(raw)
    assert back.cpu().numpy().tobytes() == raw


def test_shuffle_unaligned_buffers(backend):
    import torch
    from indicatrix_amd import radiance
    bits = mixed_bits_fast(3 * 4099, 3)
    buf = _dev(np.concatenate([[0], bits.view(np.int32)]).astype(np.int32))
    src = buf[1:]                      # 4-byte offset: the byte-wise path
    sh = torch.empty(src.numel() + 1, dtype=torch.int32, device="cuda:0")[1:]
    radiance.shuffle(backend, src, sh)
    assert sh.cpu().numpy().tobytes() == ro.shuffle(bits.astype("<u4").tobytes())


@pytest.mark.parametrize("pixels", [1, 5, 331, 4096, 4099, 512 * 512 + 3])
@pytest.mark.parametrize("shuffled", [True, False])
def test_add_valid_matches_oracle(backend, pixels, shuffled):
    import torch
    from indicatrix_amd import radiance
    print(f"add pixels={pixels} shuffled={shuffled}", flush=True)
    delta = radiance_like(pixels, 11 + pixels)
    raw = delta.astype("<f4").tobytes()
    payload = ro.shuffle(raw) if shuffled else raw
    start = radiance_like(pixels, 99, bad_every=pixels + 1)   # finite running sums
    start[::5] = 0.0
    acc = _dev(start)
    dropped = torch.full((1,), 12345, dtype=torch.int32, device="cuda:0")
    radiance.add_valid(backend, _dev(np.frombuffer(payload, np.uint8)), acc, shuffled, dropped)
    want, want_dropped = ro.add_valid(payload, start, shuffled)
    assert acc.cpu().numpy().view(np.uint32).tobytes() == want.view(np.uint32).tobytes()
    assert int(dropped.item()) == want_dropped


def test_add_valid_subnormal_and_signed_zero_bits(backend):
    import torch
    from indicatrix_amd import radiance
    bits = mixed_bits(3 * 1024, 21)
    payload = ro.shuffle(bits.astype("<u4").tobytes())
    start = mixed_bits(3 * 1024, 22).view(np.float32).copy()
    start[~np.isfinite(start)] = 1.0     # running sums are finite
    start = start.reshape(-1, 3)
    acc = _dev(start)
    dropped = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    radiance.add_valid(backend, _dev(np.frombuffer(payload, np.uint8)), acc, True, dropped)
    want, want_dropped = ro.add_valid(payload, start, True)
    assert acc.cpu().numpy().view(np.uint32).tobytes() == want.view(np.uint32).tobytes()
    assert int(dropped.item()) == want_dropped


def test_payload_throughput_4k_frame(backend):
    """4096 x 4096 pixels (201 MB payload): shuffle, unshuffle and the
    shuffled add, timed over 20 calls each (each call synchronises)."""
    import torch
    from indicatrix_amd import radiance
    px = 4096 * 4096
    src = torch.rand(px * 3, device="cuda:0")
    sh = torch.empty_like(src)
    back = torch.empty_like(src)
    acc = torch.zeros_like(src)
    dropped = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    nbytes = px * 12
    for name, fn, traffic in (
            ("shuffle", lambda: radiance.shuffle(backend, src, sh), 2 * nbytes),
            ("unshuffle", lambda: radiance.unshuffle(backend, sh, back), 2 * nbytes),
            ("add_shuffled", lambda: radiance.add_valid(backend, sh, acc, True, dropped), 3 * nbytes)):
        fn()
        t0 = time.perf_counter()
        for _ in range(20):
            fn()
        dt = (time.perf_counter() - t0) / 20
        print(f"{name}: {dt * 1e3:.3f} ms/call, {traffic / dt / 1e9:.0f} GB/s algorithmic", flush=True)
    assert torch.equal(back, src)
    assert int(dropped.item()) == 0


def test_radiance_calls_honour_a_lost_context():
    """A context marked lost (forced from the host by an injected device loss
    on the film path) declines the payload kernels with IXG_DEVICE_LOST and
    stays lost; after ixg_try_recover they run again."""
    import torch
    from indicatrix_amd import radiance, tracer, workloads
    from indicatrix_amd._lib import IxgError
    print("radiance on a lost context", flush=True)
    be = tracer.GpuBackend(0)
    try:
        src = torch.arange(12, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        radiance.shuffle(be, src, dst)
        be.inject_fault(0)
        acc = np.zeros((4 * 4, 3), np.float64)
        with pytest.raises(IxgError):
            be.try_accumulate(workloads.film_c1_scene(4, 4), 0, 1, acc)
        assert be.is_lost()
        for call in (lambda: radiance.shuffle(be, src, dst), lambda: radiance.unshuffle(be, src, dst),
                     lambda: radiance.add_valid(be, src, torch.zeros(3, dtype=torch.float32, device="cuda:0")
                                                .repeat(4), False,
                                                torch.zeros(1, dtype=torch.int32, device="cuda:0"))):
            with pytest.raises(IxgError) as e:
                call()
            assert e.value.status == 3  # IXG_DEVICE_LOST
            assert be.is_lost()
        assert be.try_recover()
        radiance.shuffle(be, src, dst)
        assert dst.cpu().numpy().tobytes() == ro.shuffle(src.cpu().numpy().astype("<u4").tobytes())
    finally:
        be.close()
