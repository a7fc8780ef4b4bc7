"""Serving form on the GPU: histories of concurrent launches against the
serial replay, mixed launches against batch-at-a-time runs, and the admission
limits."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc
from gpu_cc_tb_amd import workloads as wl

ROWS = 4096


def check_replay(w, res, rows, seed=g.SEED):
    init = orc.seed_table(seed, rows, 10)
    exp_t, exp_r, ties = orc.replay(rows, 10, seed, init, w.ops, w.tx_off, res["order"])
    assert ties == 0, "serialization keys must be distinct"
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr]), "reads differ from the serial replay"
    assert np.array_equal(res["table"], exp_t), "final table differs from the serial replay"


def serve(scheme, s, arrival=None, rows=ROWS, **kw):
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=1 << 20) as ctx:
        ctx.seed_table()
        ctx.load_ops(s.ops)
        st = ctx.serve(s.batch_sizes, arrival, latencies=True, **kw)
        res = ctx.fetch()
    return st, res


def small_stream(theta, nb=24, seed=g.SEED):
    return wl.serving_stream(ROWS, nb, 16, 600, length=16, theta=theta, read_ratio=0.8, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["mvcc", "silo", "tictoc", "to", "tpl_nw", "tpl_wd", "gacco"])
def test_serve_replay(scheme):
    """Peak load (every batch at once, coalesced into launches of at most 2000
    transactions): committed history equals its serial replay."""
    s = small_stream(0.9)
    print(f"[serve] scheme={scheme} ntx={s.ntx} batches={len(s.batch_sizes)}", flush=True)
    st, res = serve(scheme, s, max_launch_txs=2000)
    assert st["committed"] == s.ntx and st["error_flags"] == 0
    assert 1 < st["launches"] < len(s.batch_sizes)
    lat = st["latency_ms"]
    assert np.all(lat > 0) and st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    check_replay(s.ops, res, ROWS)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["mvcc", "silo", "gacco"])
def test_serve_mix_vs_alone(scheme):
    """The same requests mixed into one continuously fed launch and run one
    batch per launch: both histories replay exactly; where the outcome is
    fixed by the input it is bit-identical (gacco: transaction-id order in
    both; a conflict-free stream: every scheme)."""
    s = small_stream(0.9, seed=g.SEED + 1)
    print(f"[mix] scheme={scheme} ntx={s.ntx}", flush=True)
    st_mix, mix = serve(scheme, s)
    st_alone, alone = serve(scheme, s, max_launch_txs=1)
    assert st_mix["launches"] < st_alone["launches"] == len(s.batch_sizes)
    check_replay(s.ops, mix, ROWS)
    check_replay(s.ops, alone, ROWS)
    if scheme == "gacco":
        assert np.array_equal(mix["table"], alone["table"])
        assert np.array_equal(mix["reads"], alone["reads"])
    # conflict-free: every transaction touches rows no other one touches
    ntx, L = 240, 16
    keys = np.random.default_rng(7).permutation(ROWS)[: ntx * L].astype(np.uint64)
    cols = np.random.default_rng(8).integers(0, 10, ntx * L).astype(np.uint64)
    wr = np.random.default_rng(9).random(ntx * L) < 0.3
    keys = np.sort(keys.reshape(ntx, L), axis=1).reshape(-1)
    ops = g.Ops(g.make_ops(keys, cols, wr), np.arange(0, ntx * L + 1, L, dtype=np.uint64))
    cf = wl.Stream(ops, [40] * 6)
    _, a = serve(scheme, cf)
    _, b = serve(scheme, cf, max_launch_txs=1)
    assert np.array_equal(a["table"], b["table"]) and np.array_equal(a["reads"], b["reads"])


def hot_then_cold(hot_txs=20000, cold_batches=5, cold_txs=100, length=8, seed=g.SEED + 3):
    """Batch 0: hot_txs transactions on the first 64 rows (theta 0.99, read
    ratio 0.5: a long, contended batch); then cold batches of transactions
    on rows that no other transaction touches."""
    hot = g.generate(64, hot_txs, theta=0.99, read_ratio=0.5, len_min=length, len_max=length, seed=seed)
    n = cold_batches * cold_txs
    rng = np.random.default_rng(seed)
    keys = (64 + rng.permutation(ROWS - 64)[: n * length]).astype(np.uint64)
    keys = np.sort(keys.reshape(n, length), axis=1).reshape(-1)
    cols = rng.integers(0, 10, n * length).astype(np.uint64)
    cold = g.Ops(g.make_ops(keys, cols, rng.random(n * length) < 0.3),
                 np.arange(0, n * length + 1, length, dtype=np.uint64))
    return wl.Stream(wl.concat([hot, cold]), [hot_txs] + [cold_txs] * cold_batches)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["silo", "mvcc", "tpl_nw", "tictoc", "gacco"])
def test_serve_continuous(scheme):
    """Continuous batching: batches that arrive while the launch runs join
    it (one launch serves every batch), the late batches commit while the
    first batch's contended transactions are still running, and the whole
    history equals its serial replay."""
    s = hot_then_cold()
    arr = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0])
    print(f"[continuous] scheme={scheme} ntx={s.ntx} batches={len(s.batch_sizes)}", flush=True)
    st, res = serve(scheme, s, arr)
    assert st["committed"] == s.ntx and st["error_flags"] == 0
    assert st["launches"] == 1 and st["admitted_live"] == len(s.batch_sizes) - 1 and st["deferred"] == 0
    batch_of = np.repeat(np.arange(len(s.batch_sizes)), s.batch_sizes)
    commit_ms = st["latency_ms"] + arr[batch_of]
    hot = s.batch_sizes[0]
    assert commit_ms[hot:].min() < commit_ms[:hot].max(), "late batches should commit beside the running batch"
    check_replay(s.ops, res, ROWS)


@pytest.mark.gpu
def test_serve_pool_admission():
    """MVCC: a version-pool budget smaller than two batches' writes defers
    arrived batches to later launches; a batch larger than the pool is
    refused."""
    s = small_stream(0.6, nb=10)
    writes = [int(g.op_is_write(s.ops.ops[s.ops.tx_off[a]:s.ops.tx_off[a + n]]).sum())
              for a, n in zip(np.cumsum([0] + s.batch_sizes[:-1]), s.batch_sizes)]
    budget = 24 * max(writes)
    print(f"[pool] writes={writes} budget={budget}", flush=True)
    st, res = serve("mvcc", s, pool_cap_bytes=budget)
    assert st["committed"] == s.ntx and st["deferred"] > 0
    check_replay(s.ops, res, ROWS)
    with g.Context("mvcc", rows=ROWS, warp_density=3, block_size=8, batch_size=1 << 20) as ctx:
        ctx.seed_table()
        ctx.load_ops(s.ops)
        with pytest.raises(g.GcctbError):
            ctx.serve(s.batch_sizes, pool_cap_bytes=24 * (max(writes) - 1))


@pytest.mark.gpu
def test_serve_poisson_latency():
    """Poisson arrivals: latency counts from arrival, so no transaction's
    latency is below its launch's own run time and the run lasts at least
    until the last arrival."""
    s = small_stream(0.9, nb=40)
    arr = wl.poisson_arrivals(s.batch_sizes, txps=2e6)
    st, res = serve("silo", s, arr)
    assert st["committed"] == s.ntx
    assert st["tot_ms"] >= arr[-1]
    assert np.all(st["latency_ms"] > 0)
    check_replay(s.ops, res, ROWS)


@pytest.mark.gpu
@pytest.mark.skip(reason="the adaptive window is not built: this test faulted the GPU with Silo (out-of-bounds "
                         "access) and once returned reads that differ from the serial replay; the cause was not "
                         "found and the window's code was taken out of the kernels")
@pytest.mark.parametrize("scheme", ["mvcc", "silo"])
def test_serve_adaptive(scheme):
    """Serving with the adaptive admission window in every launch."""
    s = small_stream(0.99)
    with g.Context(scheme, rows=ROWS, warp_density=3, block_size=8, batch_size=1 << 20,
                   persist_adaptive=True) as ctx:
        ctx.seed_table()
        ctx.load_ops(s.ops)
        st = ctx.serve(s.batch_sizes, max_launch_txs=3000)
        res = ctx.fetch()
    assert st["committed"] == s.ntx and st["error_flags"] == 0
    check_replay(s.ops, res, ROWS)
