"""Radiance payloads on the CPU: the oracle against the reference's own test
vectors (indicatrix-net/src/radiance/shuffle.rs tests, client/accumulate
tests) and the library's header / frame admission checks."""
import struct

import numpy as np
import pytest

from oracle import radiance_oracle as ro
from radiance_vectors import mixed_bits
from indicatrix_amd import abi, radiance


def f32s(*v):
    return struct.pack("<%df" % len(v), *v)


def test_shuffle_groups_bytes_by_plane():
    # shuffle.rs shuffle_groups_bytes_by_plane
    raw = bytes([0x10, 0x11, 0x12, 0x13, 0x20, 0x21, 0x22, 0x23])
    assert ro.shuffle(raw) == bytes([0x10, 0x20, 0x11, 0x21, 0x12, 0x22, 0x13, 0x23])
    assert ro.unshuffle(ro.shuffle(raw)) == raw


@pytest.mark.parametrize("n,seed", [(0, 1), (1, 2), (3, 3), (12, 4), (997, 5), (3 * 4096, 6)])
def test_shuffle_round_trips_bit_exactly(n, seed):
    raw = mixed_bits(n, seed).astype("<u4").tobytes()
    sh = ro.shuffle(raw)
    assert ro.unshuffle(sh) == raw
    if n:
        # plane k holds byte k of every value
        b = np.frombuffer(raw, np.uint8)
        assert sh[:n] == b[0::4].tobytes() and sh[3 * n:] == b[3::4].tobytes()


def test_shuffle_rejects_a_partial_value():
    with pytest.raises(ValueError, match="whole number of f32"):
        ro.shuffle(b"\x01\x02\x03")


def test_add_valid_skips_and_counts_invalid_pixels():
    # shuffle.rs add_unshuffled_valid_skips_and_counts_invalid_pixels
    nan, inf = float("nan"), float("inf")
    px = [(1.0, 2.0, 3.0), (nan, 1.0, 1.0), (1.0, inf, 1.0), (1.0, 1.0, -0.5), (0.0, -0.0, 4.0)]
    raw = b"".join(f32s(*p) for p in px)
    for shuffled, payload in ((True, ro.shuffle(raw)), (False, raw)):
        acc, dropped = ro.add_valid(payload, np.ones((5, 3), np.float32), shuffled)
        assert dropped == 3
        assert acc.tolist() == [[2, 3, 4], [1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 5]]


def test_is_valid_sample_signed_zero_passes():
    assert ro.is_valid_sample([0.0, -0.0, 0.0]).all()
    assert not ro.is_valid_sample([1e-45, 1.0, -1e-45]).any()


def test_frame_within_range_vectors():
    # accumulate/tests.rs frame_within_range_is_inclusive_start_exclusive_end_and_overflow_safe
    for f in (ro.frame_within_range, radiance.frame_within_range):
        assert f(10, 5, 10, 15)
        assert f(12, 3, 10, 15)
        assert not f(9, 1, 10, 15)
        assert not f(14, 2, 10, 15)
        assert not f(2 ** 32 - 1, 2, 0, 2 ** 32 - 1)
        assert f(10, 0, 10, 15)


def test_check_frame_kinds_and_messages():
    # accumulate/tests.rs an_out_of_range_frame_is_rejected
    with pytest.raises(radiance.RadianceError) as e:
        radiance.check_frame(70, 4, (64, 72))
    assert e.value.kind == "FrameOutsideRange"
    assert str(e.value) == "FRAME [70, +4) lies outside the requested sample range [64, 72)"
    # a_cumulative_overrun_from_two_individually_in_range_frames_is_rejected
    radiance.check_frame(64, 5, (64, 72), samples_done=0)
    with pytest.raises(radiance.RadianceError) as e:
        radiance.check_frame(67, 5, (64, 72), samples_done=5)
    assert e.value.kind == "SampleCountOverrun"
    assert (e.value.fields["cumulative"], e.value.fields["range_len"]) == (10, 8)
    # a_zero_sample_frame_with_a_payload_is_rejected (checked before any range)
    with pytest.raises(radiance.RadianceError) as e:
        radiance.check_frame(0, 0)
    assert e.value.kind == "ZeroSampleFrame"
    radiance.check_frame(5, 3)              # no requested range: anything non-empty
    radiance.check_frame(64, 3, (64, 72), samples_done=5)
    with pytest.raises(radiance.RadianceError) as e:
        radiance.check_frame(10, 5, (0, 100), samples_done=2 ** 32 - 2)
    assert e.value.fields["cumulative"] == 2 ** 32 - 1   # saturating


def test_check_payload_order_and_messages():
    chk = radiance.check_payload
    chk(abi.PAYLOAD_RAW, 48, 48, 2, 2)
    chk(abi.PAYLOAD_RAW, 0, 0, 0, 0)
    with pytest.raises(radiance.RadianceError) as e:
        chk(abi.PAYLOAD_RAW, 48, 47, 2, 2)
    assert e.value.kind == "LengthMismatch"
    assert str(e.value) == "radiance payload for 2x2 must be 48 bytes, got 47"
    with pytest.raises(radiance.RadianceError) as e:
        chk(abi.PAYLOAD_RAW, 47, 48, 2, 2)
    assert e.value.kind == "RawLenMismatch"
    assert str(e.value) == "header raw_len 47 disagrees with 2x2 = 48 bytes"
    # TooLarge comes first, before raw_len is looked at
    with pytest.raises(radiance.RadianceError) as e:
        chk(abi.PAYLOAD_SHUFFLE_LZ4, 0, 0, 8192, 8192)
    assert e.value.kind == "TooLarge"
    assert str(e.value) == "a 8192x8192 radiance payload would exceed the 536870912 byte frame cap"
    # exactly at the cap is allowed: 44739242 pixels * 12 = 536870904
    chk(abi.PAYLOAD_RAW, 536870904, 536870904, 44739242, 1)
    # Display of UnsupportedEncoding prints the encoding's Debug form
    # (radiance/mod.rs:149-151), the zstd arm with its level
    for level in (1, 3, 22):
        with pytest.raises(radiance.RadianceError) as e:
            chk(abi.payload_zstd(level), 48, 20, 2, 2)
        assert e.value.kind == "UnsupportedEncoding"
        assert str(e.value) == f"payload encoding ShuffleZstd {{ level: {level} }} is not supported by this build"
        assert e.value.fields["encoding"] == abi.payload_zstd(level)
    with pytest.raises(radiance.RadianceError) as e:
        chk(abi.PAYLOAD_SHUFFLE_LZ4, 48, 20, 2, 2)
    assert str(e.value) == "payload encoding ShuffleLz4 is not supported by this build"
    chk(abi.PAYLOAD_SHUFFLE_LZ4, 48, 20, 2, 2, compression=True)
    chk(abi.payload_zstd(1), 48, 20, 2, 2, compression=True)
    # a level on an arm that has none, or beyond u8, is not an encoding
    for bad in (abi.PAYLOAD_RAW | (1 << 8), abi.PAYLOAD_SHUFFLE_LZ4 | (2 << 8), 1 << 16, -1, 3):
        with pytest.raises(ValueError):
            chk(bad, 48, 48, 2, 2)


def test_planted_mistake_in_the_oracle_is_caught():
    # a transposed shuffle (value-major instead of plane-major) fails the vector
    raw = bytes([0x10, 0x11, 0x12, 0x13, 0x20, 0x21, 0x22, 0x23])
    wrong = np.frombuffer(raw, np.uint8).reshape(4, -1).T.copy().tobytes()
    assert wrong != bytes([0x10, 0x20, 0x11, 0x21, 0x12, 0x22, 0x13, 0x23])
