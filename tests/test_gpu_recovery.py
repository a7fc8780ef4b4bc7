"""Context state across calls: a captured graph after the index is rebuilt,
CC words after a launch that the watchdog stopped, and the index's
distinct-row contract."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["silo", "tpl_nw", "gacco"])
def test_graph_after_index_rebuild(scheme):
    """run_graph, then gcctb_build_index with other rows for the same keys,
    then run_graph again without reloading anything: the second replay of
    the captured batches must use the new index (the graph is captured
    again), and both runs equal the serial replay on the rows their index
    names."""
    rows, ntx = 2048, 1500
    w = g.generate(rows, ntx, theta=0.7, read_ratio=0.7, len_min=8, len_max=16)
    keys = np.arange(rows, dtype=np.uint32)
    perm = np.random.default_rng(5).permutation(rows).astype(np.uint32)
    ent = g.op_key(w.ops).astype(np.int64)
    col, isw = g.op_col(w.ops), g.op_is_write(w.ops)
    print(f"[graph-index] scheme={scheme}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=500) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st1 = ctx.run("graph")
        r1 = ctx.fetch()
        ctx.build_index(keys, perm)
        st2 = ctx.run("graph")
        r2 = ctx.fetch()
    assert st1["committed"] == ntx and st2["committed"] == ntx and st2["index_miss"] == 0
    init = orc.seed_table(g.SEED, rows, 10)
    t1, e1, _ = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, r1["order"])
    assert np.array_equal(r1["table"], t1)
    assert np.array_equal(r1["reads"][~isw], e1[~isw])
    on_rows = g.make_ops(perm[ent], col, isw)
    t2, e2, ties = orc.replay(rows, 10, g.SEED, r1["table"], on_rows, w.tx_off, r2["order"])
    assert ties == 0
    assert np.array_equal(r2["reads"][~isw], e2[~isw])
    assert np.array_equal(r2["table"], t2)


@pytest.mark.gpu
def test_index_values_must_be_distinct():
    with g.Context("tpl_nw", rows=64) as ctx:
        ctx.seed_table()
        with pytest.raises(g.GcctbError) as e:
            ctx.build_index(np.array([1, 2, 3], np.uint32), np.array([4, 5, 4], np.uint32))
        assert e.value.code == -1 and "one row" in str(e.value)
        w = g.Ops(g.make_ops([3, 9, 3], [0, 1, 2], [True, False, False]), np.array([0, 3], np.uint64))
        with pytest.raises(g.GcctbError) as e:
            ctx.load_ops(w)
        assert e.value.code == -1


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("scheme", ["tpl_nw", "tpl_wd", "to", "mvcc", "silo"])
def test_watchdog_stop_leaves_clean_words(scheme, form, monkeypatch):
    """A watchdog of 20 us (the kernels' own bounded wait, set from the host)
    stops waves of a hot workload early: the call reports GCCTB_EKERNEL with
    the watchdog flag, every CC word is free afterwards (2PL locks included,
    which only a commit releases), and the context runs again."""
    monkeypatch.setenv("GCCTB_WATCHDOG_S", "0.00002")
    rows, ntx = 64, 2048
    w = g.generate(rows, ntx, theta=0.99, read_ratio=0.5, len_min=8, len_max=8)
    print(f"[watchdog] scheme={scheme} form={form}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=1024) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        for attempt in range(2):
            st = ctx.run(form, check=False)
            if st["rc"] == 0:
                assert st["committed"] == ntx
            else:
                assert st["rc"] == -6 and st["error_flags"] & 16, st
            assert not np.any(ctx.fetch_cc()), f"CC words held after run {attempt}"
