"""GaccO's hot-row combiners (persistent form): a row whose access queue
holds at least GCCTB_COMBINE_MIN accesses is run in queue order by one
workgroup before the launch.  The outcome must be GaccO's: the serial
execution in transaction-id order, bit for bit, and the same tables and
reads as the turn-by-turn run with the combiners off."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

CASES = {
    # name: (rows, ntx, theta, read_ratio, len_min, len_max, cols)
    "reads": (4096, 8192, 0.99, 0.9, 16, 16, 10),
    "writes": (4096, 6000, 0.99, 0.3, 8, 8, 10),
    "varlen_wide": (2048, 5000, 0.9, 0.5, 4, 32, 64),  # 64 columns: one per combiner lane
    "one_hot_row": (64, 3000, 0.99, 0.5, 4, 4, 3),
}


def run(rows, cols, w, wd=3, bs=8):
    with g.Context("gacco", rows=rows, cols=cols, warp_density=wd, block_size=bs, batch_size=65536,
                   persist_window=2048) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run("persistent")
        res = ctx.fetch()
    return st, res


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_combiners_match_serial_and_turns(case, monkeypatch):
    rows, ntx, theta, rr, lmin, lmax, cols = CASES[case]
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, cols=cols)
    keys = g.op_key(w.ops).astype(np.int64)
    per_row = np.bincount(keys, minlength=rows)
    expect = int(per_row[per_row >= 256].sum())
    print(f"[combine] case={case} rows with >= 256 accesses: {(per_row >= 256).sum()}", flush=True)
    st, res = run(rows, cols, w)
    assert st["committed"] == ntx and st["error_flags"] == 0
    assert st["combined"] == expect and expect > 0
    init = orc.seed_table(g.SEED, rows, cols)
    t2, r2 = orc.replay_txid(rows, cols, g.SEED, init, w.ops, w.tx_off)
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["table"], t2), "table differs from the serial run in transaction-id order"
    assert np.array_equal(res["reads"][isr], r2[isr]), "reads differ from the serial run in transaction-id order"
    assert np.array_equal(res["order"][1::3], np.arange(ntx, dtype=np.uint64))
    monkeypatch.setenv("GCCTB_COMBINE_MIN", "0")
    print(f"[combine] case={case} combiners off", flush=True)
    st0, res0 = run(rows, cols, w)
    assert st0["combined"] == 0
    assert np.array_equal(res0["table"], res["table"]) and np.array_equal(res0["reads"], res["reads"])


@pytest.mark.gpu
def test_combiners_threshold_and_plain_form(monkeypatch):
    """A lower threshold moves more rows to combiners with the same outcome;
    the plain form never combines."""
    rows, cols = 4096, 10
    w = g.generate(rows, 4000, theta=0.9, read_ratio=0.5, len_min=16, cols=cols)
    per_row = np.bincount(g.op_key(w.ops).astype(np.int64), minlength=rows)
    monkeypatch.setenv("GCCTB_COMBINE_MIN", "16")
    print("[combine] threshold 16", flush=True)
    st, res = run(rows, cols, w)
    assert st["combined"] == int(per_row[per_row >= 16].sum())
    init = orc.seed_table(g.SEED, rows, cols)
    t2, r2 = orc.replay_txid(rows, cols, g.SEED, init, w.ops, w.tx_off)
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["table"], t2) and np.array_equal(res["reads"][isr], r2[isr])
    with g.Context("gacco", rows=rows, cols=cols, warp_density=3, block_size=8, batch_size=1000) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        stp = ctx.run("plain")
        resp = ctx.fetch()
    assert stp["combined"] == 0
    assert np.array_equal(resp["table"], t2) and np.array_equal(resp["reads"][isr], r2[isr])
