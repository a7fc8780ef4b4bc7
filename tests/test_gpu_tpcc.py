"""GPU parity for TPC-C payment and new_order (benchmark/tpcc.cu) under every
scheme: the GPU's final table and every op's result (the value its update
read) must equal the oracle's serial replay in the order the GPU reports.
At larger sizes the integer fields give order-free checks: each district
hands out the order ids 3001.. exactly once, and the stock, customer and
district counters grow by the totals of the transactions."""
import os
import re
import subprocess

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

SCHEMES = list(g.SCHEMES)
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu_cc_tb_amd", "bin", "test_tpcc")
CASES = {
    # name: (warehouses, ntx, new_order, batch, wd, bs)
    "pay1": (1, 600, False, 200, 2, 8),     # every payment writes warehouse 0
    "pay4": (4, 2000, False, 500, 0, 32),   # one lane per transaction
    "no1": (1, 600, True, 200, 4, 8),       # ten hot districts
    "no4": (4, 2000, True, 500, 3, 16),     # two slots per lane for 9..16 ops
}


def run(scheme, W, ntx, new_order, batch, wd, bs, form, remote_ratio=0.1):
    cells, price = g.tpcc_seed_tables(W)
    w = g.tpcc_generate(W, ntx, new_order=new_order, remote_ratio=remote_ratio)
    with g.Context(scheme, rows=g.tpcc_rows(W), cols=g.TPCC_COLS, warp_density=wd, block_size=bs,
                   batch_size=batch, tpcc=True) as ctx:
        ctx.load_table(cells)
        ctx.load_ops(w)
        ctx.load_params(w.params)
        st = ctx.run(form)
        res = ctx.fetch()
        ranks = ctx.fetch_ranks() if scheme == "gputx" else None
    return cells, price, w, st, res, ranks


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "graph"])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_tpcc_replay(scheme, case, form):
    W, ntx, no, batch, wd, bs = CASES[case]
    print(f"[tpcc] scheme={scheme} case={case} form={form} W={W} ntx={ntx}", flush=True)
    cells, price, w, st, res, ranks = run(scheme, W, ntx, no, batch, wd, bs, form)
    assert st["committed"] == ntx and st["error_flags"] == 0
    rows = g.tpcc_rows(W)
    exp_t, exp_r, ties = orc.tpcc_replay(rows, cells, w.ops, w.params, w.tx_off, res["order"])
    assert ties == 0, "serialization keys must be distinct"
    bad = np.nonzero(res["reads"] != exp_r)[0]
    assert bad.size == 0, f"{bad.size} op results differ from the serial replay (first op {bad[:1]})"
    assert np.array_equal(res["table"], exp_t), "final table differs from the serial replay"
    if scheme == "gacco":
        assert np.array_equal(res["order"][1::3], np.arange(ntx, dtype=np.uint64))
    if scheme == "gputx":
        rank_o, _ = orc.gputx_rank(rows, w.ops, w.tx_off)
        assert np.array_equal(ranks[0], rank_o)
    if no:
        total, oid = g.tpcc_new_order_totals(W, cells, price, w, res["reads"])
        assert np.all(np.isfinite(total)) and np.all(total > 0)
        assert oid.min() >= 3001


def _lo(x):
    return (x & np.uint64(0xFFFFFFFF)).astype(np.int64)


def _hi(x):
    return (x >> np.uint64(32)).astype(np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", SCHEMES)
def test_tpcc_counters_large(scheme):
    """16 warehouses, 40000 new_orders then 40000 payments in one context
    each: order ids per district are a permutation of 3001..; stock and
    customer counters add up; the replay matches."""
    W = 16
    rows = g.tpcc_rows(W)
    print(f"[tpcc-large] scheme={scheme} W={W}", flush=True)
    cells, price, w, st, res, _ = run(scheme, W, 40000, True, 16384, 4, 16, "plain", remote_ratio=0.2)
    assert st["committed"] == w.ntx and st["error_flags"] == 0
    t = res["table"].reshape(rows, 5)
    c0 = cells.reshape(rows, 5)
    key = g.op_key(w.ops).astype(np.int64)
    first = w.tx_off[:-1].astype(np.int64)
    drow = key[first]
    oid = res["reads"][first].astype(np.int64)
    for d in np.unique(drow):
        got = np.sort(oid[drow == d])
        assert np.array_equal(got, 3001 + np.arange(got.size)), f"district row {d}"
    dist = slice(W, 11 * W)
    assert np.array_equal(_lo(t[dist, 2]) - 3001, np.bincount(drow - W, minlength=10 * W))
    stock = np.ones(key.size, dtype=bool)
    stock[first] = False
    s_rows, qty, rem = key[stock], (w.params[stock] & 0xFFFF).astype(np.int64), (w.params[stock] >> 16).astype(np.int64)
    s0 = 30011 * W
    assert np.array_equal(_hi(t[s0:, 1]) - _hi(c0[s0:, 1]), np.bincount(s_rows - s0, weights=qty, minlength=rows - s0))
    assert np.array_equal(_lo(t[s0:, 2]), np.bincount(s_rows - s0, minlength=rows - s0))
    assert np.array_equal(_hi(t[s0:, 2]), np.bincount(s_rows - s0, weights=rem, minlength=rows - s0))
    exp_t, exp_r, ties = orc.tpcc_replay(rows, cells, w.ops, w.params, w.tx_off, res["order"])
    assert ties == 0 and np.array_equal(res["reads"], exp_r) and np.array_equal(res["table"], exp_t)

    cells, price, w, st, res, _ = run(scheme, W, 40000, False, 16384, 2, 32, "plain")
    assert st["committed"] == w.ntx and st["error_flags"] == 0
    t = res["table"].reshape(rows, 5)
    cust = g.op_key(w.ops).astype(np.int64)[2::3]
    c_beg = 11 * W
    assert np.array_equal(_lo(t[c_beg:s0, 4]) - 1, np.bincount(cust - c_beg, minlength=30000 * W))
    exp_t, exp_r, ties = orc.tpcc_replay(rows, cells, w.ops, w.params, w.tx_off, res["order"])
    assert ties == 0 and np.array_equal(res["reads"], exp_r) and np.array_equal(res["table"], exp_t)


@pytest.mark.gpu
def test_tpcc_refusals():
    W = 1
    w = g.tpcc_generate(W, 10, new_order=False)
    with g.Context("silo", rows=g.tpcc_rows(W), cols=5, tpcc=True) as ctx:
        ctx.load_table(g.tpcc_seed_tables(W)[0])
        ctx.load_ops(w)
        with pytest.raises(g.GcctbError) as e:
            ctx.run("plain")  # no parameters loaded
        assert e.value.code == -4
        with pytest.raises(g.GcctbError) as e:
            ctx.load_params(w.params[:-1])
        assert e.value.code == -1
        ctx.load_params(w.params)
        with pytest.raises(g.GcctbError) as e:
            ctx.run("persistent")
        assert e.value.code == -1
        with pytest.raises(g.GcctbError) as e:
            ctx.serve([10])
        assert e.value.code == -1
        assert ctx.run("plain")["committed"] == 10
        ctx.load_ops(w)  # new ops need their parameters again
        with pytest.raises(g.GcctbError) as e:
            ctx.run("plain")
        assert e.value.code == -4
    with g.Context("silo", rows=64, cols=5) as ctx:
        with pytest.raises(g.GcctbError) as e:
            ctx.load_params(np.zeros(1, np.uint32))
        assert e.value.code == -1


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,isno", [("silo", 1), ("tpl_nw", 0), ("gputx", 1), ("mvcc", 0)])
def test_cli_debug_run(scheme, isno, tmp_path):
    """The drop-in TPC-C driver: the reference's argv (isno, warehouse_cnt,
    txcnt, cc, warp_density, block_size, batch_size, debug), its TRY line,
    the serial replay OK, new_order's totals, and the 15-field CSV line
    appended to tpcc_no_wh<W>.txt / tpcc_pm_wh<W>.txt."""
    assert os.path.exists(CLI), "build the tools first (gpu_cc_tb_amd/build.py)"
    args = [CLI, str(isno), "2", "6000", scheme, "3", "16", "2048", "1"]
    print(f"[cli-tpcc] {' '.join(args[1:])}", flush=True)
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    assert p.stdout.startswith(f"TRY {scheme},3,16,2048\n")
    assert "REPLAY OK" in p.stdout
    if isno:
        assert re.search(r"ORDER 0 o_id 3\d{3} total \d+\.\d\d", p.stdout), p.stdout[-500:]
    name = f"tpcc_{'no' if isno else 'pm'}_wh2.txt"
    fields = (tmp_path / name).read_text().strip().splitlines()[-1].split(",")
    assert len(fields) == 15 and fields[0] == scheme and float(fields[4]) > 0
