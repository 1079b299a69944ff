"""Seeded, counter-based synthetic column generator (host twin of csrc/datagen.hip).

rand(seed, stream, i) = mix(key(seed, stream) + i), key = mix(mix(seed) ^ stream), where mix is
the SplitMix64 finaliser with its golden-ratio increment (Steele, Lea, Flood, OOPSLA 2014).
Uniform integers in [lo, hi] take lo + (((r >> 32) * (hi - lo + 1)) >> 32).  A code i receives a
single-bit flip iff rand(seed, s, i) < threshold, at bit ((rand(seed, s+1, i) >> 32) * k) >> 32.
The device kernels compute the same functions, so host and device columns are identical.
"""
from __future__ import annotations

import numpy as np

GOLDEN = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB
MASK64 = (1 << 64) - 1


def mix_int(x: int) -> int:
    z = (x + GOLDEN) & MASK64
    z = ((z ^ (z >> 30)) * M1) & MASK64
    z = ((z ^ (z >> 27)) * M2) & MASK64
    return z ^ (z >> 31)


def stream_key(seed: int, stream: int) -> int:
    return mix_int(mix_int(seed) ^ stream)


def mix(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64, copy=False)
    with np.errstate(over="ignore"):
        z = x + np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
    return z ^ (z >> np.uint64(31))


def rand_u64(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    idx = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix(idx + np.uint64(stream_key(seed, stream)))


def uniform(n: int, seed: int, stream: int, lo: int, hi: int) -> np.ndarray:
    """n values uniform in [lo, hi] (hi - lo < 2^32), as uint64."""
    rng = hi - lo + 1
    assert 0 < rng <= (1 << 32)
    r = rand_u64(seed, stream, n)
    with np.errstate(over="ignore"):
        return np.uint64(lo) + (((r >> np.uint64(32)) * np.uint64(rng)) >> np.uint64(32))


DTYPES = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def encode(values: np.ndarray, a: int, width: int) -> np.ndarray:
    """AN-encode: (values * a) mod 2^(8*width), stored in the width's unsigned dtype."""
    with np.errstate(over="ignore"):
        return (values.astype(np.uint64) * np.uint64(a)).astype(DTYPES[width])


def flip_threshold(rate: float) -> int:
    return min(int(rate * 2.0**64), MASK64)


def flip_positions(n: int, seed: int, stream: int, threshold: int) -> np.ndarray:
    if threshold == 0 or n == 0:
        return np.zeros(0, dtype=np.uint64)
    return np.nonzero(rand_u64(seed, stream, n) < np.uint64(threshold))[0].astype(np.uint64)


def inject_flips(codes: np.ndarray, seed: int, stream: int, threshold: int) -> np.ndarray:
    """Flips one bit in each selected code, in place; returns the sorted flipped positions."""
    pos = flip_positions(codes.size, seed, stream, threshold)
    if pos.size:
        k = codes.dtype.itemsize * 8
        r = mix(pos + np.uint64(stream_key(seed, stream + 1)))
        bits = ((r >> np.uint64(32)) * np.uint64(k)) >> np.uint64(32)
        one = np.ones(pos.size, dtype=codes.dtype)
        codes[pos.astype(np.int64)] ^= (one << bits.astype(codes.dtype))
    return pos
