"""CPU restatement of the radiance payload byte work (test infrastructure:
only the tests and the benchmark's CPU leg use it as the checker).

- shuffle / unshuffle: indicatrix-net/src/radiance/shuffle.rs:35,60
- is_valid_sample: shuffle.rs:25
- add_valid: add_unshuffled_valid (shuffle.rs:122) and the Raw arm of
  PayloadDecoder::decode_and_add (payload.rs:225-250)
- frame_within_range: client/accumulate/mod.rs:198

Pinned by the reference's own test vectors (shuffle.rs tests, accumulate
tests) in tests/test_radiance.py and by the byte-permutation identity.
"""
import numpy as np


def shuffle(raw: bytes) -> bytes:
    b = np.frombuffer(raw, dtype=np.uint8)
    if b.size % 4:
        raise ValueError("payload is not a whole number of f32 values")
    return b.reshape(-1, 4).T.copy().tobytes()


def unshuffle(planes: bytes) -> bytes:
    b = np.frombuffer(planes, dtype=np.uint8)
    if b.size % 4:
        raise ValueError("payload is not a whole number of f32 values")
    return b.reshape(4, -1).T.copy().tobytes()


def is_valid_sample(v) -> np.ndarray:
    v = np.asarray(v, dtype=np.float32).reshape(-1, 3)
    with np.errstate(invalid="ignore"):
        return np.all(np.isfinite(v), axis=1) & np.all(v >= 0.0, axis=1)


def add_valid(payload: bytes, acc: np.ndarray, shuffled: bool):
    """Returns (new acc, dropped) for acc f32 (pixels, 3)."""
    raw = unshuffle(payload) if shuffled else payload
    d = np.frombuffer(raw, dtype="<f4").reshape(-1, 3)
    ok = is_valid_sample(d)
    out = acc.astype(np.float32).copy()
    with np.errstate(invalid="ignore", over="ignore"):
        out[ok] = (out[ok] + d[ok]).astype(np.float32)
    return out, int((~ok).sum())


def frame_within_range(first_sample, samples, start, end) -> bool:
    frame_end = first_sample + samples
    if frame_end > 0xFFFFFFFF:
        return False
    return first_sample >= start and frame_end <= end
