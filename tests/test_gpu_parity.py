"""GPU parity: every scheme against the CPU serial-replay oracle.

The GPU reports, per transaction, its place in the serialization order
(commit ticket, timestamp, (commit_ts, ticket), transaction id or (rank, id)).
The oracle replays the committed transactions serially in that order from the
same seeded table and must reproduce the final table and every read value
bit for bit.  GaccO is additionally checked against its closed form (serial
execution in transaction-id order), GPUTx against the CPU restatement of its
one-pass rank (src/gputx.cu:5-38) and (rank, id) replay.
"""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

SCHEMES = list(g.SCHEMES)
CASES = {
    # name: (rows, ntx, len_min, len_max, theta, read_ratio, batch, wd, bs)
    "uniform": (4096, 2048, 16, 16, 0.0, 0.9, 512, 5, 32),
    "hot": (64, 1024, 8, 8, 0.99, 0.5, 256, 4, 8),
    "varlen": (1024, 1000, 4, 32, 0.8, 0.5, 333, 5, 16),
    "tiny": (8, 64, 4, 4, 0.9, 0.5, 16, 3, 2),
    "onelane": (512, 300, 10, 10, 0.6, 0.7, 128, 0, 1),
}


def run_case(scheme, case, seed=g.SEED):
    rows, ntx, lmin, lmax, theta, rr, batch, wd, bs = CASES[case]
    print(f"[parity] scheme={scheme} case={case} rows={rows} ntx={ntx} wd={wd} bs={bs}", flush=True)
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, seed=seed)
    with g.Context(scheme, rows=rows, warp_density=wd, block_size=bs, batch_size=batch, seed=seed) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run_plain()
        res = ctx.fetch(aborts=True)
        ranks = ctx.fetch_ranks() if scheme == "gputx" else None
        cc = ctx.fetch_cc() if scheme not in ("gacco", "gputx") else None
    return w, st, res, ranks, cc, (rows, seed)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_replay_parity(scheme, case):
    w, st, res, ranks, cc, (rows, seed) = run_case(scheme, case)
    assert st["committed"] == w.ntx
    assert st["error_flags"] == 0
    assert int(res["aborts"].sum()) == st["aborts"]
    init = orc.seed_table(seed, rows, 10)
    exp_t, exp_r, ties = orc.replay(rows, 10, seed, init, w.ops, w.tx_off, res["order"])
    assert ties == 0, "serialization keys must be distinct"
    is_read = ~g.op_is_write(w.ops)
    bad_reads = np.nonzero((res["reads"] != exp_r) & is_read)[0]
    assert bad_reads.size == 0, f"{bad_reads.size} reads differ from the serial replay"
    assert np.array_equal(res["table"], exp_t), "final table differs from the serial replay"
    if scheme == "gacco":
        t2, r2 = orc.replay_txid(rows, 10, seed, init, w.ops, w.tx_off)
        assert np.array_equal(res["table"], t2)
        assert np.array_equal(res["reads"][is_read], r2[is_read])
    if scheme == "gputx":
        rank_o, counts_o = orc.gputx_rank(rows, w.ops, w.tx_off)
        rank_g, counts_g = ranks
        assert np.array_equal(rank_g, rank_o)
        assert np.array_equal(counts_g, counts_o)
        order = np.zeros(3 * w.ntx, dtype=np.uint64)
        order[1::3] = rank_o
        order[2::3] = np.arange(w.ntx)
        t3, r3, _ = orc.replay(rows, 10, seed, init, w.ops, w.tx_off, order)
        assert np.array_equal(res["table"], t3)
        # same-rank conflicts: adjacent accesses of a row in (row, rank, op) order
        # with equal rank where either writes (the identity index maps key to row)
        txof = np.repeat(np.arange(w.ntx), np.diff(w.tx_off.astype(np.int64)))
        row, rk, wr = g.op_key(w.ops).astype(np.int64), rank_o[txof], g.op_is_write(w.ops)
        srt = np.lexsort((np.arange(row.size), rk, row))
        row, rk, wr = row[srt], rk[srt], wr[srt]
        same = (row[1:] == row[:-1]) & (rk[1:] == rk[:-1]) & (wr[1:] | wr[:-1])
        assert st["same_rank_conflicts"] == int(same.sum())
    if cc is not None:
        if scheme in ("tpl_nw", "tpl_wd"):
            assert not np.any((cc >> np.uint64(1)) & np.uint64(0x7FFFFFFF)), "a lock is still held"
        else:
            assert not np.any(cc), "CC words are reset after every batch"


# Whole-row mode (the reference's 104-byte Item copies, benchmark/ycsb.cu:57,86-91):
# rows of 13 u64 cells, every op moves its whole row, reads return row digests.
ROW_CASES = {
    # name: (rows, ntx, len_min, len_max, theta, read_ratio, batch, wd, bs)
    "item": (2048, 1500, 4, 16, 0.8, 0.5, 500, 4, 8),
    "hot": (64, 600, 8, 8, 0.99, 0.5, 200, 3, 8),
}


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "graph"])
@pytest.mark.parametrize("case", list(ROW_CASES))
@pytest.mark.parametrize("scheme", SCHEMES)
def test_row_mode_replay(scheme, case, form):
    rows, ntx, lmin, lmax, theta, rr, batch, wd, bs = ROW_CASES[case]
    cols = 13
    print(f"[row-mode] scheme={scheme} case={case} form={form}", flush=True)
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, cols=cols)
    with g.Context(scheme, rows=rows, cols=cols, warp_density=wd, block_size=bs, batch_size=batch,
                   row_ops=True) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run(form)
        res = ctx.fetch()
        ranks = ctx.fetch_ranks() if scheme == "gputx" else None
    assert st["committed"] == ntx and st["error_flags"] == 0
    init = orc.seed_table(g.SEED, rows, cols)
    exp_t, exp_r, ties = orc.replay_rows(rows, cols, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert ties == 0
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr]), "row digests differ from the whole-row replay"
    assert np.array_equal(res["table"], exp_t), "table differs from the whole-row replay"
    if scheme == "gacco":
        assert np.array_equal(res["order"][1::3], np.arange(ntx, dtype=np.uint64))
    if scheme == "gputx":
        rank_o, _ = orc.gputx_rank(rows, w.ops, w.tx_off)
        assert np.array_equal(ranks[0], rank_o)


@pytest.mark.gpu
def test_row_mode_refuses_persistent_form():
    with g.Context("silo", rows=64, cols=13, row_ops=True) as ctx:
        ctx.seed_table()
        ctx.load_ops(g.generate(64, 10, theta=0.5, read_ratio=0.5, len_min=4, cols=13))
        with pytest.raises(g.GcctbError) as e:
            ctx.run("persistent")
        assert e.value.code == -1
