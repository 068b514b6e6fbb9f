"""Deterministic counter-based RNG shared by every workload.

SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): output i of a stream seeded with s is
mix(s + (i+1)*0x9E3779B97F4A7C15).  Uniforms take the top 53 bits;
normals use Box-Muller.  Being counter based, it vectorises in numpy and
gives bit-identical inputs on every host.
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def _mix(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & M64

    def next_u64(self) -> int:
        self.state = (self.state + GAMMA) & M64
        return _mix(self.state)

    def uniform(self) -> float:
        """Uniform double in [0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        span = hi - lo + 1
        return lo + min(int(self.uniform() * span), span - 1)


def uniforms(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+count-1 of the stream `seed` as uniforms in [0,1)."""
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M64) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal_state(seed: int, n: int) -> np.ndarray:
    """2^n amplitudes with re, im i.i.d. N(0,1) (Box-Muller), unit 2-norm."""
    N = 1 << n
    u = uniforms(seed, 4 * N).reshape(N, 4)
    r1 = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    r2 = np.sqrt(-2.0 * np.log1p(-u[:, 2]))
    re = r1 * np.cos(2 * math.pi * u[:, 1])
    im = r2 * np.cos(2 * math.pi * u[:, 3])
    psi = re + 1j * im
    return psi / np.linalg.norm(psi)
