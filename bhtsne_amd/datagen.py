"""Seeded synthetic inputs for the benchmark workloads (Gaussian mixtures) and the seed
embedding.  numpy's PCG64 generator keyed by the stated seed; the same bytes feed the GPU run and
the CPU baseline."""
from __future__ import annotations

import hashlib

import numpy as np


def gmm(n, d, components, centre_var, seed, axis_std=None, weights="equal"):
    """n points in d dimensions from a Gaussian mixture: centres ~ N(0, centre_var I), per-axis
    standard deviation ``axis_std`` (scalar or length-d array), equal or Dirichlet(1) weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.standard_normal((components, d)) * np.sqrt(centre_var)
    if weights == "dirichlet":
        w = rng.dirichlet(np.ones(components))
    else:
        w = np.full(components, 1.0 / components)
    labels = rng.choice(components, size=n, p=w)
    std = 1.0 if axis_std is None else np.asarray(axis_std, np.float64)
    x = centres[labels] + rng.standard_normal((n, d)) * std
    return x.astype(np.float32), labels.astype(np.int32)


def config_c1(n=5000):
    """Barnes-Hut 2-D: N=5,000, D_in=50, 10 components, centres N(0, 25 I), sigma 1, seed 0."""
    return gmm(n, 50, 10, 25.0, 0)


def config_c2(n=70000):
    """FIt-SNE 2-D: N=70,000, D_in=20, 10 components, centres N(0, 16 I),
    per-axis std 1/sqrt(i+1), seed 1."""
    return gmm(n, 20, 10, 16.0, 1, axis_std=1.0 / np.sqrt(np.arange(20) + 1.0))


def config_c3(n=1_000_000):
    """FIt-SNE 2-D at scale: N=1,000,000, D_in=50, 50 components, centres N(0, 36 I), sigma 1,
    Dirichlet(1) weights, seed 2."""
    return gmm(n, 50, 50, 36.0, 2, weights="dirichlet")


def log_uniform_int(rng, lo, hi):
    return int(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def config_c4_job(j):
    """Job j of the batched workload (seeded by j): N log-uniform in [500, 50,000], D_in from
    {16, 32, 64}, 5-20 components, centres N(0, 25 I), sigma 1; even jobs FIt-SNE, odd jobs
    Barnes-Hut (theta 0.5).  Returns (x, method)."""
    rng = np.random.Generator(np.random.PCG64(10_000 + j))
    n = log_uniform_int(rng, 500, 50_000)
    d_in = int(rng.choice([16, 32, 64]))
    comps = int(rng.integers(5, 21))
    x, _ = gmm(n, d_in, comps, 25.0, 20_000 + j)
    return x, ("fitsne" if j % 2 == 0 else "bh")


def config_c5_request(r):
    """Request r of the serving workload (seeded by r): N log-uniform in [1,000, 200,000], D_in
    from {20, 50}, D_out from {1, 2}, 10 components, centres N(0, 25 I), sigma 1.
    Returns (x, d_out)."""
    rng = np.random.Generator(np.random.PCG64(30_000 + r))
    n = log_uniform_int(rng, 1_000, 200_000)
    d_in = int(rng.choice([20, 50]))
    d_out = int(rng.choice([1, 2]))
    x, _ = gmm(n, d_in, 10, 25.0, 40_000 + r)
    return x, d_out


def poisson_arrivals(count, rate, seed):
    """Arrival times (seconds) of a Poisson stream of `count` requests at `rate` per second."""
    rng = np.random.Generator(np.random.PCG64(50_000 + seed))
    return np.cumsum(rng.exponential(1.0 / rate, count))


def seed_embedding(n, d_out, seed):
    """Seed embedding ~ N(0, 1e-4) (the law of src/tsne/mod.rs:142-147), seeded."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    return (rng.standard_normal((n, d_out)) * 1e-4).astype(np.float32)


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
