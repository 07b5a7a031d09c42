"""NumPy restatement of the device start-vector generator (pc_rand_*), so the
oracle can be fed the same random blocks as the HIP path.

Element idx of stream ``seed`` is splitmix64's output for state
``seed + (idx+1)*0x9E3779B97F4A7C15``; uniform = (z >> 11) * 2^-53; column c of
an n x k block takes elements [c*n, (c+1)*n); normals use Box-Muller on
elements (2i, 2i+1).  Seeds per call follow ``eigensolvers.SeedStream``.
"""
import numpy as np

from parsec_python_amd.eigensolvers import SeedStream

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(z):
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, idx: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        st = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
    return (_mix(st) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def block(seed: int, n: int, k: int, normal: bool) -> np.ndarray:
    idx = (np.arange(k, dtype=np.uint64)[None, :] * np.uint64(n) + np.arange(n, dtype=np.uint64)[:, None])
    if not normal:
        return uniform(seed, idx)
    u1 = uniform(seed, np.uint64(2) * idx)
    u2 = uniform(seed, np.uint64(2) * idx + np.uint64(1))
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)


class MirrorRNG:
    """Drop-in for the oracle's ``rng`` argument (``random`` / ``standard_normal``)."""

    def __init__(self, seed: int):
        self.seeds = SeedStream(seed)

    def random(self, shape):
        n, k = shape
        return block(self.seeds.next(), n, k, False)

    def standard_normal(self, shape):
        n, k = shape
        return block(self.seeds.next(), n, k, True)
