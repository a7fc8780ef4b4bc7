"""Synthetic YCSB workloads of the benchmark configurations.

Every workload is deterministic in the seed: a stream of batches is generated
batch by batch with each batch's own Zipf skew, read ratio and length range,
and global transaction ids run across the stream (so a transaction's ops
depend only on (seed, its id, its batch's parameters)).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import SEED, Ops, generate


@dataclass
class Stream:
    ops: Ops
    batch_sizes: list[int]
    batch_params: list[dict] = field(default_factory=list)

    @property
    def ntx(self) -> int:
        return self.ops.ntx


def _rng(seed: int, tag: int) -> np.random.Generator:
    return np.random.default_rng([seed & 0xFFFFFFFF, seed >> 32, tag])


def concat(parts: list[Ops]) -> Ops:
    ops = np.concatenate([p.ops for p in parts]) if parts else np.zeros(0, np.uint64)
    offs = [np.zeros(1, np.uint64)]
    base = 0
    for p in parts:
        offs.append(p.tx_off[1:] + np.uint64(base))
        base += p.nops
    return Ops(ops, np.concatenate(offs))


def batched_stream(rows: int, nbatches: int, size_lo: int, size_hi: int, *, len_lo: int, len_hi: int,
                   theta_lo: float, theta_hi: float, read_ratios=(0.5, 0.9, 1.0), seed: int = SEED,
                   threads: int = 0) -> Stream:
    """Config-4 style stream: batch sizes log-uniform in [size_lo, size_hi], per
    batch theta ~ U[theta_lo, theta_hi] and read ratio drawn from read_ratios;
    transaction length uniform in [len_lo, len_hi]."""
    r = _rng(seed, 4)
    sizes = np.exp(r.uniform(math.log(size_lo), math.log(size_hi), nbatches)).astype(np.int64)
    sizes = np.clip(sizes, size_lo, size_hi)
    thetas = r.uniform(theta_lo, theta_hi, nbatches)
    rrs = r.choice(np.asarray(read_ratios, dtype=np.float64), nbatches)
    parts, params = [], []
    base = 0
    for b in range(nbatches):
        n = int(sizes[b])
        parts.append(generate(rows, n, theta=float(thetas[b]), read_ratio=float(rrs[b]), len_min=len_lo,
                              len_max=len_hi, seed=seed, tx_base=base, threads=threads))
        params.append({"size": n, "theta": float(thetas[b]), "read_ratio": float(rrs[b])})
        base += n
    return Stream(concat(parts), [int(x) for x in sizes], params)


def config4(seed: int = SEED, nbatches: int = 256, rows: int = 10485760, threads: int = 0) -> Stream:
    """256 batches, sizes log-uniform in [1024, 262144], length U[4, 32], theta
    U[0, 0.99], read ratio in {0.5, 0.9, 1.0}, over 10,485,760 rows."""
    return batched_stream(rows, nbatches, 1024, 262144, len_lo=4, len_hi=32, theta_lo=0.0, theta_hi=0.99,
                          seed=seed, threads=threads)


def uniform_batches(ntx: int, batch: int) -> list[int]:
    return [min(batch, ntx - s) for s in range(0, ntx, batch)]


def serving_stream(rows: int = 67108864, nbatches: int = 256, size_lo: int = 256, size_hi: int = 65536, *,
                   length: int = 16, theta: float = 0.9, read_ratio: float = 0.8, seed: int = SEED,
                   threads: int = 0, max_txs: int | None = None) -> Stream:
    """Serving requests: batch sizes log-uniform in [size_lo, size_hi], every
    transaction `length` ops with one Zipf skew and read ratio.  With max_txs,
    only the first batches holding at least max_txs transactions are kept."""
    r = _rng(seed, 5)
    sizes = np.exp(r.uniform(math.log(size_lo), math.log(size_hi), nbatches)).astype(np.int64)
    sizes = [int(x) for x in np.clip(sizes, size_lo, size_hi)]
    if max_txs is not None:
        keep = int(np.searchsorted(np.cumsum(sizes), max_txs)) + 1
        sizes = sizes[:keep]
    ops = generate(rows, int(sum(sizes)), theta=theta, read_ratio=read_ratio, len_min=length, len_max=length,
                   seed=seed, threads=threads)
    return Stream(ops, sizes, [{"size": n, "theta": theta, "read_ratio": read_ratio} for n in sizes])


def poisson_arrivals(batch_sizes, txps: float, seed: int = SEED) -> np.ndarray:
    """Arrival times (ms) of a Poisson process of batches whose mean rate
    carries `txps` transactions per second: exponential gaps with mean
    mean(batch size) / txps."""
    sizes = np.asarray(batch_sizes, dtype=np.float64)
    if sizes.size == 0:
        return np.zeros(0)
    mean_gap_ms = 1e3 * float(sizes.mean()) / txps
    gaps = _rng(seed, 6).exponential(mean_gap_ms, sizes.size)
    gaps[0] = 0.0
    return np.cumsum(gaps)
