"""Host logic of the synthetic streams (no GPU)."""
import numpy as np

import gpu_cc_tb_amd as g
from gpu_cc_tb_amd import workloads as wl


def test_batched_stream_shape():
    s = wl.batched_stream(4096, 20, 16, 400, len_lo=4, len_hi=32, theta_lo=0.0, theta_hi=0.99)
    assert len(s.batch_sizes) == 20 and sum(s.batch_sizes) == s.ntx
    assert all(16 <= n <= 400 for n in s.batch_sizes)
    lens = np.diff(s.ops.tx_off.astype(np.int64))
    assert lens.min() >= 4 and lens.max() <= 32
    # each batch was generated with its own parameters: transaction ids run
    # across the stream, so the same stream is generated twice identically
    t = wl.batched_stream(4096, 20, 16, 400, len_lo=4, len_hi=32, theta_lo=0.0, theta_hi=0.99)
    assert np.array_equal(s.ops.ops, t.ops.ops) and s.batch_sizes == t.batch_sizes


def test_serving_stream_truncation():
    s = wl.serving_stream(1 << 20, 4096, 256, 65536, max_txs=100000)
    assert s.ntx >= 100000 and s.ntx - s.batch_sizes[-1] < 100000
    lens = np.diff(s.ops.tx_off.astype(np.int64))
    assert (lens == 16).all()
    reads = float(np.mean(~g.op_is_write(s.ops.ops)))
    assert abs(reads - 0.8) < 0.01


def test_poisson_arrivals():
    sizes = [1000] * 20000
    arr = wl.poisson_arrivals(sizes, txps=1e6, seed=3)
    assert arr[0] == 0 and np.all(np.diff(arr) >= 0)
    gaps = np.diff(arr)
    # mean gap = mean batch / rate = 1 ms; exponential: std == mean
    assert abs(gaps.mean() - 1.0) < 0.03
    assert abs(gaps.std() - 1.0) < 0.05
    assert np.array_equal(arr, wl.poisson_arrivals(sizes, txps=1e6, seed=3))
