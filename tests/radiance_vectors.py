"""Seeded f32 bit patterns for the radiance payload tests: an xorshift32
stream with every third value drawn from special patterns (NaN payloads,
signed zeros, subnormals, infinities, extremes)."""
import numpy as np

SPECIAL = [0x7FC00000, 0x7FC00001, 0xFFC00000, 0x7F800001, 0x00000000, 0x80000000,
           0x00000001, 0x807FFFFF, 0x7F800000, 0xFF800000, 0x7F7FFFFF, 0x3F800000]


def mixed_bits(n, seed):
    s = (seed | 1) & 0xFFFFFFFF
    out = np.empty(n, dtype=np.uint32)
    for i in range(n):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        out[i] = SPECIAL[(i // 3) % len(SPECIAL)] if i % 3 == 0 else s
    return out


def mixed_bits_fast(n, seed):
    """Large payloads: numpy random bits with the same special share."""
    rng = np.random.default_rng(seed)
    out = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    idx = np.arange(0, n, 3)
    out[idx] = np.array(SPECIAL, dtype=np.uint32)[(idx // 3) % len(SPECIAL)]
    return out


def radiance_like(pixels, seed, bad_every=7):
    """Non-negative finite XYZ sums with an invalid pixel every bad_every."""
    rng = np.random.default_rng(seed)
    v = (rng.random((pixels, 3), dtype=np.float32) * 4.0).astype(np.float32)
    bad = np.arange(0, pixels, bad_every)
    kinds = np.array([np.nan, np.inf, -0.5, -np.inf], dtype=np.float32)
    v[bad, bad % 3] = kinds[(bad // bad_every) % 4]
    return v
