"""Synthetic inputs of the benchmark workloads.

Every stream comes from SplitMix64 (u = (z >> 11) * 2^-53, value = 2u - 1),
so the same workload can be rebuilt anywhere from its seeds:

  grid_points(n1d)          n1d^3 tensor grid of uniform_nodes(-1, 1, n1d)
                            (myHeaders.hpp:111-120), dimension 0 outermost
                            (user_kernel.hpp:79-95);
  batch_specs(count, seed)  independent problems with N log-uniform in
                            [n_lo, n_hi] and NDIM drawn from {2, 3};
  problem_points(...)       their points, uniform iid in [-1, 1]^d;
  poisson_arrivals(...)     request arrival times of a serving stream.
"""
import math

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_uniform(seed, n):
    """n values of SplitMix64 stream `seed` mapped to [-1, 1)."""
    s = np.uint64(seed)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = s + idx * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def grid_points(n1d, a=-1.0, b=1.0, ndim=3):
    nodes = a + (b - a) / float(n1d - 1) * np.arange(n1d, dtype=np.float64)
    g = np.meshgrid(*([nodes] * ndim), indexing="ij")
    return np.ascontiguousarray(np.stack([c.ravel() for c in g], axis=1))


def log_uniform_int(u, lo, hi):
    """u in [-1, 1) -> integer log-uniform in [lo, hi]."""
    return int(round(math.exp(math.log(lo) + (u + 1.0) / 2.0 * (math.log(hi) - math.log(lo)))))


def batch_specs(count, seed=4, n_lo=2000, n_hi=64000, dims=(2, 3)):
    """[(n, ndim)] of `count` independent problems."""
    out = []
    for p in range(count):
        u = splitmix_uniform(seed * 100003 + p, 2)
        n = log_uniform_int(u[0], n_lo, n_hi)
        nd = dims[0] if len(dims) == 1 or u[1] < 0 else dims[1]
        out.append((n, nd))
    return out


def problem_points(index, n, ndim, seed=7):
    """Points of problem `index`: uniform iid in [-1, 1]^ndim."""
    return splitmix_uniform(seed * 1000003 + index, n * ndim).reshape(n, ndim)


def poisson_arrivals(rate, count, seed=5):
    """Arrival times (s) of `count` requests of a Poisson stream at `rate`/s."""
    u = (splitmix_uniform(seed, count) + 1.0) / 2.0
    gaps = -np.log1p(-u) / rate
    return np.cumsum(gaps)
