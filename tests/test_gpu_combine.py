"""GaccO's hot-row combiners (persistent form): a row whose access queue
holds at least GCCTB_COMBINE_MIN accesses is run in queue order by one
workgroup before the launch.  The outcome must be GaccO's: the serial
execution in transaction-id order, bit for bit, and the same tables and
reads as the turn-by-turn run with the combiners off."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

COMBINE_MIN = 128  # the default threshold (GCCTB_COMBINE_MIN)
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
    expect = int(per_row[per_row >= COMBINE_MIN].sum())
    print(f"[combine] case={case} rows with >= {COMBINE_MIN} accesses: {(per_row >= COMBINE_MIN).sum()}", flush=True)
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


@pytest.mark.gpu
def test_every_queue_combined_and_index_misses(monkeypatch):
    """Threshold 1 sends every row queue to a combiner (the launch only
    finishes transactions); ops whose keys the index lacks miss in both runs
    and are never combined."""
    rows, cols, ntx = 1024, 10, 3000
    w = g.generate(rows, ntx, theta=0.9, read_ratio=0.5, len_min=4, len_max=12, cols=cols)
    keys = np.arange(rows, dtype=np.uint32)
    present = keys[keys % 7 != 3]  # every seventh key is missing from the index
    outs = []
    for cmin in ("1", "0"):
        monkeypatch.setenv("GCCTB_COMBINE_MIN", cmin)
        print(f"[combine-edge] threshold {cmin}", flush=True)
        with g.Context("gacco", rows=rows, cols=cols, warp_density=3, block_size=8, batch_size=65536,
                       persist_window=512) as ctx:
            ctx.seed_table()
            ctx.build_index(present, present)
            ctx.load_ops(w)
            st = ctx.run("persistent", check=False)
            outs.append((st, ctx.fetch()))
    (st1, r1), (st0, r0) = outs
    misses = int(np.count_nonzero(g.op_key(w.ops) % 7 == 3))
    assert misses > 0 and st1["index_miss"] == misses == st0["index_miss"]
    assert st1["combined"] == w.nops - misses and st0["combined"] == 0
    assert st1["committed"] == ntx == st0["committed"]
    assert np.array_equal(r1["table"], r0["table"]) and np.array_equal(r1["reads"], r0["reads"])


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["tpl_nw", "to", "mvcc"])
def test_deferral_tiny_and_empty_transactions(scheme):
    """The deferral pass on one transaction, and on transactions without ops
    beside ones that all write the same key."""
    rows = 64
    cases = [
        (g.make_ops([5, 9], [0, 1], [True, False]), [0, 2]),
        (g.make_ops([3] * 6, [0] * 6, [True] * 6), [0, 0, 1, 1, 2, 3, 3, 4, 5, 6, 6]),
    ]
    for ops, off in cases:
        w = g.Ops(ops, np.array(off, dtype=np.uint64))
        print(f"[defer-edge] scheme={scheme} ntx={w.ntx} nops={w.nops}", flush=True)
        with g.Context(scheme, rows=rows, warp_density=0, block_size=1, batch_size=16, persist_window=4) as ctx:
            ctx.seed_table()
            ctx.load_ops(w)
            st = ctx.run("persistent")
            res = ctx.fetch()
        assert st["committed"] == w.ntx
        init = orc.seed_table(g.SEED, rows, 10)
        exp_t, exp_r, ties = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, res["order"])
        assert ties == 0 and np.array_equal(res["table"], exp_t)
        isr = ~g.op_is_write(w.ops)
        assert np.array_equal(res["reads"][isr], exp_r[isr])
