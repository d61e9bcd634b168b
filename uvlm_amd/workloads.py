"""Seeded synthetic particle sets for the FMM vortex-particle workloads.

The generator is project-defined so that every consumer (HIP path, oracle, bench) sees
bit-identical inputs: splitmix64 -> uniform (x >> 11) * 2^-53 -> Box-Muller normals.
Positions, strengths and core sizes follow the workload definitions below; every array
is float64, positions/strengths (n, 3), sigma (n,).
"""
import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based splitmix64 stream; draws are vectorised but sequential in value."""

    def __init__(self, seed):
        self.state = np.uint64(seed)

    def _next_u64(self, k):
        with np.errstate(over="ignore"):
            idx = np.arange(1, k + 1, dtype=np.uint64)
            z = self.state + idx * _GAMMA
            self.state = self.state + np.uint64(k) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, k):
        return (self._next_u64(k) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, k):
        m = (k + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        out = np.empty(2 * m)
        out[0::2] = r * np.cos(2.0 * np.pi * u2)
        out[1::2] = r * np.sin(2.0 * np.pi * u2)
        return out[:k]


def vortex_ring(rng, stations, samples, radius=1.0, core=0.1, gamma=1.0, z0=0.0):
    """Thin ring in the z=z0 plane: Gaussian core samples around each azimuthal station,
    strength Gamma * (2 pi R / stations) / samples along the local tangent."""
    phi = 2.0 * np.pi * np.arange(stations) / stations
    xi = rng.normal(stations * samples).reshape(stations, samples) * core
    eta = rng.normal(stations * samples).reshape(stations, samples) * core
    c, s = np.cos(phi)[:, None], np.sin(phi)[:, None]
    X = np.empty((stations, samples, 3))
    X[..., 0] = (radius + xi) * c
    X[..., 1] = (radius + xi) * s
    X[..., 2] = z0 + eta
    a = gamma * (2.0 * np.pi * radius / stations) / samples
    A = np.empty((stations, samples, 3))
    A[..., 0] = -a * s
    A[..., 1] = a * c
    A[..., 2] = 0.0
    return X.reshape(-1, 3), A.reshape(-1, 3)


def ring(n_stations=64, samples=128, seed=1001):
    """One thin ring: R=1, core 0.1, Gamma=1, sigma=0.1 (8,192 particles by default)."""
    rng = SplitMix64(seed)
    X, A = vortex_ring(rng, n_stations, samples)
    return X, A, np.full(X.shape[0], 0.1)


def cube(n=1 << 20, seed=1002, n_sample=16384):
    """Uniform random cube [-1,1]^3; strengths N(0,1) * h^3 with h = 2/N^(1/3); sigma = 2h.
    Returns X, alpha, sigma and `n_sample` distinct target indices."""
    rng = SplitMix64(seed)
    X = (rng.uniform(3 * n).reshape(n, 3) * 2.0 - 1.0)
    h = 2.0 / n ** (1.0 / 3.0)
    A = rng.normal(3 * n).reshape(n, 3) * h ** 3
    sig = np.full(n, 2.0 * h)
    idx = sample_without_replacement(rng, n, min(n_sample, n))
    return X, A, sig, idx


def leapfrog(n_stations=16384, samples=128, seed=1003, sigma=0.08):
    """Two coaxial rings, R=1, core 0.1, Gamma=1 each, at z=0 and z=0.5 (4,194,304 by default)."""
    rng = SplitMix64(seed)
    X1, A1 = vortex_ring(rng, n_stations, samples, z0=0.0)
    X2, A2 = vortex_ring(rng, n_stations, samples, z0=0.5)
    X = np.concatenate([X1, X2])
    A = np.concatenate([A1, A2])
    return X, A, np.full(X.shape[0], sigma)


def blobs(n=1 << 20, seed=1004, n_probe=4096, n_check=64):
    """Three anisotropic Gaussian blobs (centres U[-1,1]^3, axis spreads U[0.05,0.3], random
    rotations), strengths N(0,1), sigma 0.05; probes from the same mixture, a unit probe
    vector v, and `n_check` particle indices for gradient checks."""
    rng = SplitMix64(seed)
    centres = rng.uniform(9).reshape(3, 3) * 2.0 - 1.0
    spreads = 0.05 + 0.25 * rng.uniform(9).reshape(3, 3)
    rots = []
    for _ in range(3):
        q, r = np.linalg.qr(rng.normal(9).reshape(3, 3))
        rots.append(q * np.sign(np.diag(r))[None, :])
    counts = [n // 3 + (1 if i < n % 3 else 0) for i in range(3)]

    def draw(ks):
        out = []
        for b, k in enumerate(ks):
            z = rng.normal(3 * k).reshape(k, 3) * spreads[b][None, :]
            out.append(centres[b][None, :] + z @ rots[b].T)
        return np.concatenate(out)

    X = draw(counts)
    A = rng.normal(3 * n).reshape(n, 3)
    sig = np.full(n, 0.05)
    pc = [n_probe // 3 + (1 if i < n_probe % 3 else 0) for i in range(3)]
    probes = draw(pc)
    v = rng.normal(3)
    v /= np.linalg.norm(v)
    check = sample_without_replacement(rng, n, min(n_check, n))
    return X, A, sig, probes, v, check


def sample_without_replacement(rng, n, k):
    """k distinct indices in [0, n) from a seeded partial Fisher-Yates shuffle."""
    u = rng.uniform(k)
    chosen = {}
    out = np.empty(k, dtype=np.int64)
    for i in range(k):
        j = i + int(u[i] * (n - i))
        out[i] = chosen.get(j, j)
        chosen[j] = chosen.get(i, i)
    return out


def scaled(name, n):
    """A workload of the same shape at a smaller particle count (for tests and CPU samples)."""
    if name == "ring":
        st = max(8, int(round(math.sqrt(n / 2))))
        return ring(n_stations=st, samples=max(1, n // st))
    if name == "cube":
        X, A, s, _ = cube(n=n, n_sample=0)
        return X, A, s
    if name == "leapfrog":
        st = max(8, n // 256)
        return leapfrog(n_stations=st, samples=max(1, n // (2 * st)))
    if name == "blobs":
        X, A, s, _, _, _ = blobs(n=n, n_probe=16, n_check=1)
        return X, A, s
    raise ValueError(name)
