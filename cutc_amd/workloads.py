"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

`batch_specs` / `batch_problems`: many independent 3-way ALS problems of
uneven size — per-mode dims log-uniform in [200, 20000], nnz log-normal
(median 1.4e5, sigma 1.03) clipped to [1e4, 2e6] and to a quarter of the
tensor's cells, uniform distinct coordinates, rank drawn from {8, 16, 32, 64},
values from a rank-R/2 CP truth (entries U(0,1)) plus N(0, 0.1) noise, 10%
test split.
"""
from __future__ import annotations

import numpy as np

from . import VALUES_RAW, synth

BATCH_SEED = 0x5EED0004


def batch_specs(n=512, seed=BATCH_SEED, dim_range=(200, 20000), nnz_range=(10_000, 2_000_000)):
    rng = np.random.default_rng(seed)
    lo, hi = np.log(dim_range[0]), np.log(dim_range[1])
    mu = 0.5 * (np.log(nnz_range[0]) + np.log(nnz_range[1]))
    sigma = (np.log(nnz_range[1]) - mu) / 2.576  # the clip bounds sit at +-2.576 sigma
    specs = []
    for _ in range(n):
        dims = tuple(int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(3))
        nnz = float(np.exp(rng.normal(mu, sigma)))
        nnz = int(np.clip(nnz, nnz_range[0], nnz_range[1]))
        nnz = int(min(nnz, 0.25 * np.prod(np.array(dims, dtype=np.float64))))
        rank = int(rng.choice([8, 16, 32, 64]))
        specs.append((dims, nnz, rank))
    return specs


def batch_problems(n=512, seed=BATCH_SEED, **kw):
    """-> list of (train, test, rank) for the batch workload."""
    out = []
    for p, (dims, nnz, rank) in enumerate(batch_specs(n, seed, **kw)):
        tr, te = synth(dims, nnz, true_rank=max(1, rank // 2), value_map=VALUES_RAW, noise_std=0.1,
                       test_fraction=0.1, seed=seed + 7919 * (p + 1))
        out.append((tr, te, rank))
    return out


def init_factors(dims, rank, rng):
    """U(-1, 1) FP32 factors, one (dims[m], rank) matrix per mode."""
    return [rng.uniform(-1, 1, (d, rank)).astype(np.float32) for d in dims]
