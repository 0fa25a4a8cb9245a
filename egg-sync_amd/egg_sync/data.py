"""Seeded synthetic data for the EGG-SynC solve, plus the reference's input contract.

``min_max_normalize`` follows the reference's python/data.py:100-103: the solver
assumes every coordinate lies in [0, 1] (cell ids treat the data as a unit
cube, src/algorithms/EGG_SynC.cu:63-76,1250).

The generators replace the reference's unseeded sklearn ``make_blobs`` data
(python/generator.py:174-222) with fully seeded numpy draws.  Sampling order is
fixed: centres, then per-blob sigmas, then blob sizes, then the points of each
blob in blob order, then the uniform noise, then one permutation of all
points; the result is min-max normalised in float64 and cast to float32.
"""
from __future__ import annotations

import hashlib

import numpy as np


def min_max_normalize(X):
    """(X - min) / (max - min) per column (python/data.py:100-103)."""
    X = np.asarray(X, dtype=np.float64)
    lo = X.min(axis=0)
    hi = X.max(axis=0)
    span = np.where(hi > lo, hi - lo, 1.0)
    return (X - lo) / span


def _zipf_sizes(total: int, k: int, s: float) -> np.ndarray:
    w = np.arange(1, k + 1, dtype=np.float64) ** (-s)
    w /= w.sum()
    sizes = np.floor(w * total).astype(np.int64)
    sizes[: total - int(sizes.sum())] += 1  # remainder to the largest blobs
    return sizes


def gaussian_blobs(n: int, d: int, n_blobs: int, *, seed: int, center_lo: float, center_hi: float,
                   sigma=None, sigma_lo=None, sigma_hi=None, per_axis_sigma: bool = False,
                   sizes: str = "equal", zipf_s: float = 1.1, noise: float = 0.0) -> np.ndarray:
    """Gaussian blobs plus uniform noise in [0,1]^d, min-max normalised, float32 [n, d]."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(center_lo, center_hi, size=(n_blobs, d))
    if sigma is not None:
        sig = np.full((n_blobs, d), float(sigma))
    elif per_axis_sigma:
        sig = rng.uniform(sigma_lo, sigma_hi, size=(n_blobs, d))
    else:
        sig = np.repeat(rng.uniform(sigma_lo, sigma_hi, size=(n_blobs, 1)), d, axis=1)
    n_noise = int(round(noise * n))
    n_blob = n - n_noise
    if sizes == "equal":
        sz = np.full(n_blobs, n_blob // n_blobs, dtype=np.int64)
        sz[: n_blob - int(sz.sum())] += 1
    elif sizes == "zipf":
        sz = _zipf_sizes(n_blob, n_blobs, zipf_s)
    else:
        raise ValueError(f"unknown sizes mode {sizes!r}")
    parts = [centers[k] + sig[k] * rng.standard_normal((int(sz[k]), d)) for k in range(n_blobs)]
    parts.append(rng.uniform(0.0, 1.0, size=(n_noise, d)))
    X = np.concatenate(parts, axis=0)
    X = X[rng.permutation(n)]
    return min_max_normalize(X).astype(np.float32)


# The workloads the solver is measured and checked on.  Each returns (X, eps).

def config0(seed: int = 0, n: int = 2000):
    """Plain SynC, CPU-runnable: n=2,000, d=2; 4 equal blobs sigma=0.03, centres U[0.2,0.8]^2,
    5% uniform noise; eps=0.05."""
    X = gaussian_blobs(n, 2, 4, seed=seed, center_lo=0.2, center_hi=0.8, sigma=0.03,
                       sizes="equal", noise=0.05)
    return X, 0.05


def config1(seed: int = 1, n: int = 1_000_000):
    """Plain EGG-SynC on GPU: n=1,000,000, d=3; 20 blobs, centres U[0.1,0.9]^3, per-axis sigma
    U[0.005,0.02], Zipf(1.1) sizes, 2% noise; eps=0.02."""
    X = gaussian_blobs(n, 3, 20, seed=seed, center_lo=0.1, center_hi=0.9, sigma_lo=0.005,
                       sigma_hi=0.02, per_axis_sigma=True, sizes="zipf", zipf_s=1.1, noise=0.02)
    return X, 0.02


def config4(seed: int = 4, n: int = 4_000_000):
    """Learned-jump evaluation workload: the config-1 generator at n=4,000,000, d=3, eps=0.01."""
    X, _ = config1(seed=seed, n=n)
    return X, 0.01


def config2(seed: int = 2, n: int = 250_000):
    """Higher-dim sparse grid: n=250,000, d=8; 10 blobs sigma=0.03, centres U[0.2,0.8]^8,
    1% noise; eps=0.1."""
    X = gaussian_blobs(n, 8, 10, seed=seed, center_lo=0.2, center_hi=0.8, sigma=0.03,
                       sizes="equal", noise=0.01)
    return X, 0.1


def config3_instances(seed: int = 3, count: int = 64, n_lo: int = 1000, n_hi: int = 200_000):
    """Batched workload: `count` instances, n log-uniform in [n_lo, n_hi], d in {2,3,5,8},
    2-30 blobs with sigma U[0.01,0.04], 3% noise.  Each is swept over eps in
    {0.01, 0.02, 0.05, 0.1}.  Returns a list of float32 arrays."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        d = int(rng.choice([2, 3, 5, 8]))
        k = int(rng.integers(2, 31))
        s = int(rng.integers(0, 2**31 - 1))
        out.append(gaussian_blobs(n, d, k, seed=s, center_lo=0.1, center_hi=0.9,
                                  sigma_lo=0.01, sigma_hi=0.04, sizes="equal", noise=0.03))
    return out


CONFIG3_EPS = (0.01, 0.02, 0.05, 0.1)


def digest(X: np.ndarray) -> str:
    """SHA-256 of the float32 bytes (pins the generators)."""
    return hashlib.sha256(np.ascontiguousarray(X, dtype=np.float32).tobytes()).hexdigest()
