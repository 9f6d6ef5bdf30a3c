"""Explicit, splittable random keys (the role JAX PRNG keys play in the
reference API, dequantizer/src/quantum_emulator/base_quantum_emulator.py:17-23).

A key is a pair of 32-bit words.  ``split`` derives two child keys with
NumPy's SeedSequence; ``generator`` turns a key into a NumPy Generator.  The
values differ from JAX threefry streams; the structure (explicit keys, split
before every draw) follows the reference.
"""
from typing import Tuple

import numpy as np

Key = Tuple[int, int]


def PRNGKey(seed: int) -> Key:  # noqa: N802 (reference spelling)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return (seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF


def as_key(key) -> Key:
    if isinstance(key, (int, np.integer)):
        return PRNGKey(int(key))
    words = [int(w) for w in np.asarray(key).reshape(-1)]
    if len(words) != 2:
        raise ValueError("a key is a pair of 32-bit words")
    return words[0] & 0xFFFFFFFF, words[1] & 0xFFFFFFFF


def split(key) -> Tuple[Key, Key]:
    words = np.random.SeedSequence(list(as_key(key))).generate_state(4, dtype=np.uint32)
    return (int(words[0]), int(words[1])), (int(words[2]), int(words[3]))


def generator(key) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence(list(as_key(key))))
