"""Tables of any width.  Every scheme, in the plain and persistent forms,
against the serial replay for 1, 13, 14 and 20 columns per row; and a host
table loaded, fetched back unchanged and run over."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("cols", [1, 13, 14, 20])
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_row_blocks_any_width(scheme, cols, form):
    rows, ntx = 512, 600
    w = g.generate(rows, ntx, theta=0.8, read_ratio=0.6, len_min=4, len_max=12, cols=cols)
    print(f"[layout] scheme={scheme} cols={cols} form={form}", flush=True)
    with g.Context(scheme, rows=rows, cols=cols, warp_density=3, block_size=8, batch_size=200) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run(form)
        res = ctx.fetch()
        cc = ctx.fetch_cc() if scheme not in ("gacco", "gputx") else None
    assert st["committed"] == ntx and st["error_flags"] == 0
    init = orc.seed_table(g.SEED, rows, cols)
    exp_t, exp_r, ties = orc.replay(rows, cols, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert ties == 0
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr])
    assert np.array_equal(res["table"], exp_t)
    if cc is not None:
        if scheme in ("tpl_nw", "tpl_wd"):
            assert not np.any((cc >> np.uint64(1)) & np.uint64(0x7FFFFFFF)), "a lock is still held"
        else:
            assert not np.any(cc), "CC words are reset after every batch"


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["silo", "mvcc", "gacco"])
def test_load_table_round_trip(scheme):
    """A 14-column host table: fetched back unchanged, then transactions run
    over it match the replay from it."""
    rows, cols, ntx = 700, 14, 500
    table = np.random.default_rng(3).integers(0, 2**63, rows * cols, dtype=np.uint64)
    w = g.generate(rows, ntx, theta=0.7, read_ratio=0.5, len_min=6, len_max=10, cols=cols)
    print(f"[load-table] scheme={scheme}", flush=True)
    with g.Context(scheme, rows=rows, cols=cols, warp_density=3, block_size=8, batch_size=250) as ctx:
        ctx.load_table(table)
        assert np.array_equal(ctx.fetch(reads=False, order=False)["table"], table)
        ctx.load_ops(w)
        ctx.run("plain")
        res = ctx.fetch()
    exp_t, exp_r, _ = orc.replay(rows, cols, g.SEED, table, w.ops, w.tx_off, res["order"])
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr])
    assert np.array_equal(res["table"], exp_t)
