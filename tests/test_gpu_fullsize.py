"""GPU parity at the benchmark configurations' full sizes.

Every run is replayed by the C oracle (serial replay in the order the GPU
reports; GaccO in transaction-id order; GPUTx against the oracle's one-pass
ranks, src/gputx.cu:5-38) and must reproduce the final table and every read
bit for bit, as the small cases do.  The configurations:

  C1  1,048,576 rows, 65,536 tx x 16 ops, theta 0.0, read 0.9, tpl_nw,
      batch 65,536, warp_density 5, block_size 32
  C2  10,485,760 rows, 1,048,576 tx x 16 ops, theta 0.1, read 0.9, all eight
      schemes, batch 65,536, plain and persistent forms
  C3  the C2 table and size at theta 0.99 (read 0.5 and 0.9, batch 4,096 and
      262,144) and at theta 0.6 / 0.8, tpl_nw silo tictoc mvcc gacco gputx,
      plain form
  C4  the 256-batch stream (sizes log-uniform in [1,024, 262,144], 4-32 ops,
      per-batch theta in [0, 0.99], read 0.5 / 0.9 / 1.0) on 10,485,760 rows,
      persistent form, all eight schemes
  C5  67,108,864 rows, about 1,000,000 requests of 16 ops (theta 0.9, read
      0.8) in batches of log-uniform size in [256, 65,536] arriving as a
      Poisson process, serving form, mvcc and silo

The table, the workloads and the replays are shared through module fixtures;
the replay is split over host threads by rows (orc_replay_mt).
"""
import os

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc
from gpu_cc_tb_amd import autotune, workloads

ROWS = 10485760
NTX = 1048576
THREADS = max(1, min(16, os.cpu_count() or 1))
SEED = g.SEED


@pytest.fixture(scope="module")
def table10m():
    print(f"[fullsize] seeding the {ROWS}-row oracle table", flush=True)
    return orc.seed_table(SEED, ROWS, 10)


_workloads = {}


def workload(theta, rr):
    key = (theta, rr)
    if key not in _workloads:
        _workloads[key] = g.generate(ROWS, NTX, theta=theta, read_ratio=rr, len_min=16)
    return _workloads[key]


def check_replay(w, res, rows, init, what):
    exp_t, exp_r, ties = orc.replay(rows, 10, SEED, init, w.ops, w.tx_off, res["order"], threads=THREADS)
    assert ties == 0, f"{what}: serialization keys must be distinct"
    isr = ~g.op_is_write(w.ops)
    bad = int(np.count_nonzero((res["reads"] != exp_r) & isr))
    assert bad == 0, f"{what}: {bad} reads differ from the serial replay"
    assert np.array_equal(res["table"], exp_t), f"{what}: final table differs from the serial replay"


def run_and_check(scheme, w, init, *, rows, wd, bs, batch, form, window=0, what=""):
    print(f"[fullsize] {what} scheme={scheme} form={form} wd={wd} bs={bs} batch={batch} window={window} "
          f"ntx={w.ntx} nops={w.nops}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=wd, block_size=bs, batch_size=batch,
                   persist_window=window) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run(form)
        res = ctx.fetch()
        ranks = ctx.fetch_ranks() if scheme == "gputx" else None
    assert st["committed"] == w.ntx and st["error_flags"] == 0 and st["index_miss"] == 0
    check_replay(w, res, rows, init, what)
    if scheme == "gacco":
        # closed form: serial execution in transaction-id order
        assert np.array_equal(res["order"][1::3], np.arange(w.ntx, dtype=np.uint64))
    if scheme == "gputx":
        rank_o, counts_o = orc.gputx_rank(rows, w.ops, w.tx_off)
        assert np.array_equal(ranks[0], rank_o), "GPUTx ranks differ from the oracle's one-pass rank"
        assert np.array_equal(ranks[1], counts_o)
        assert np.array_equal(res["order"][1::3], rank_o.astype(np.uint64))
        assert np.array_equal(res["order"][2::3], np.arange(w.ntx, dtype=np.uint64))
    return st


@pytest.mark.gpu
def test_c1_tpl_nw_shape():
    """C1's shape (the CPU baseline's workload) on the GPU path."""
    rows, ntx = 1048576, 65536
    w = g.generate(rows, ntx, theta=0.0, read_ratio=0.9, len_min=16)
    init = orc.seed_table(SEED, rows, 10)
    run_and_check("tpl_nw", w, init, rows=rows, wd=5, bs=32, batch=65536, form="plain", what="C1")


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_c2_low_contention(scheme, form, table10m):
    w = workload(0.1, 0.9)
    if form == "plain":
        wd, bs = autotune.DEFAULT_SHAPE[scheme]
        window = 0
    else:
        shape = autotune.lookup(scheme, 0.1, 0.9)
        wd, bs, window = shape["warp_density"], shape["block_size"], shape["window"]
    run_and_check(scheme, w, table10m, rows=ROWS, wd=wd, bs=bs, batch=65536, form=form, window=window, what="C2")


C3_CASES = [(0.99, rr, b) for rr in (0.5, 0.9) for b in (4096, 262144)] + [(0.6, 0.9, 16384), (0.8, 0.5, 65536)]


@pytest.mark.gpu
@pytest.mark.parametrize("theta,rr,batch", C3_CASES)
@pytest.mark.parametrize("scheme", ["tpl_nw", "silo", "tictoc", "mvcc", "gacco", "gputx"])
def test_c3_high_contention(scheme, theta, rr, batch, table10m):
    if scheme == "gputx" and batch == 262144:
        pytest.skip("GPUTx batches by rank level, not by batch_size (include/gputx.cuh:289-294): "
                    "its batch-4,096 case at the same skew and read ratio runs the same launches")
    w = workload(theta, rr)
    shape = autotune.lookup(scheme, theta, rr)
    run_and_check(scheme, w, table10m, rows=ROWS, wd=shape["warp_density"], bs=shape["block_size"], batch=batch,
                  form="plain", what=f"C3 theta={theta} rr={rr}")


@pytest.mark.gpu
@pytest.mark.parametrize("rr", [0.9, 0.5])
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_c3_persistent_theta99(scheme, rr, table10m):
    """The persistent form at theta 0.99 and its tuned shape, as the bench's
    theta 0.99 leg runs it: with the hot-writer deferral (2PL, T/O, MVCC) and
    GaccO's row combiners engaged, replayed like every other run."""
    w = workload(0.99, rr)
    shape = autotune.lookup(scheme, 0.99, rr)
    st = run_and_check(scheme, w, table10m, rows=ROWS, wd=shape["warp_density"], bs=shape["block_size"],
                       batch=65536, form="persistent", window=shape["window"], what=f"C3 persistent rr={rr}")
    if scheme in ("tpl_nw", "tpl_wd", "to", "mvcc"):
        assert st["deferred"] > 0
    if scheme == "gacco":
        assert st["combined"] > 0


@pytest.fixture(scope="module")
def stream_c4():
    print("[fullsize] generating the 256-batch stream", flush=True)
    return workloads.config4(rows=ROWS)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_c4_batched_stream(scheme, stream_c4, table10m):
    s = stream_c4
    shape = autotune.lookup(scheme, 0.5, 0.9, key="stream")
    print(f"[fullsize] C4 scheme={scheme} ntx={s.ntx} nops={s.ops.nops} batches={len(s.batch_sizes)}", flush=True)
    with g.Context(scheme, rows=ROWS, warp_density=shape["warp_density"], block_size=shape["block_size"],
                   batch_size=1 << 30, persist_window=shape["window"]) as ctx:
        ctx.seed_table()
        ctx.load_ops(s.ops)
        ctx.set_batches(s.batch_sizes)
        st = ctx.run("persistent")
        res = ctx.fetch()
        ranks = ctx.fetch_ranks() if scheme == "gputx" else None
    assert st["committed"] == s.ntx and st["error_flags"] == 0
    check_replay(s.ops, res, ROWS, table10m, "C4")
    if scheme == "gacco":
        assert np.array_equal(res["order"][1::3], np.arange(s.ntx, dtype=np.uint64))
    if scheme == "gputx":
        rank_o, counts_o = orc.gputx_rank(ROWS, s.ops.ops, s.ops.tx_off)
        assert np.array_equal(ranks[0], rank_o)
        assert np.array_equal(ranks[1], counts_o)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["mvcc", "silo"])
def test_c5_serving(scheme):
    rows = 67108864
    s = workloads.serving_stream(rows, 4096, 256, 65536, theta=0.9, read_ratio=0.8, max_txs=1000000)
    arr = workloads.poisson_arrivals(s.batch_sizes, 2.0e6)
    shape = autotune.lookup(scheme, 0.9, 0.8)
    print(f"[fullsize] C5 scheme={scheme} requests={s.ntx} batches={len(s.batch_sizes)}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=shape["warp_density"], block_size=shape["block_size"],
                   batch_size=1 << 30, persist_window=shape["window"]) as ctx:
        ctx.seed_table()
        ctx.load_ops(s.ops)
        st = ctx.serve(s.batch_sizes, arr, max_launch_txs=262144, pool_cap_bytes=32 << 30)
        res = ctx.fetch()
    assert st["committed"] == s.ntx and st["error_flags"] == 0 and st["launches"] > 1
    init = orc.seed_table(SEED, rows, 10)
    check_replay(s.ops, res, rows, init, "C5")
