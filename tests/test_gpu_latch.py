"""The latch schemes (slow_tpl_nw ... slow_tictoc, include/slow_*.cuh): the
same protocols with every access to a row's CC word made under the row's
latch.  Checked like the lock-free schemes: the serial replay of the order
the kernels report reproduces every read and the final table bit for bit,
every transaction commits, and every CC word is free after the run."""
import os
import subprocess

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu_cc_tb_amd", "bin", "test_ycsb")
CASES = {"hot": (256, 2000, 0.95, 0.5, 8, 8), "uniform": (4096, 2048, 0.0, 0.9, 16, 16),
         "varlen": (1024, 1500, 0.8, 0.7, 4, 32)}


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "graph"])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("scheme", list(g.LATCH_SCHEMES))
def test_latch_replay(scheme, case, form):
    rows, ntx, theta, rr, lmin, lmax = CASES[case]
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, seed=g.SEED + 5)
    print(f"[latch] scheme={scheme} case={case} form={form}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=500) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run(form)
        res = ctx.fetch()
        cc = ctx.fetch_cc()
    assert st["committed"] == ntx and st["error_flags"] == 0 and st["index_miss"] == 0
    init = orc.seed_table(g.SEED, rows, 10)
    exp_t, exp_r, ties = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert ties == 0
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr])
    assert np.array_equal(res["table"], exp_t)
    if scheme in ("slow_tpl_nw", "slow_tpl_wd"):
        assert not np.any((cc >> np.uint64(1)) & np.uint64(0x7FFFFFFF)), "a lock is still held"
    else:
        assert not np.any(cc), "CC words are reset after every batch"


@pytest.mark.gpu
def test_latch_schemes_plain_and_graph_only():
    w = g.generate(512, 100, theta=0.5, read_ratio=0.5, len_min=4, len_max=4)
    with g.Context("slow_silo", rows=512, warp_density=3, block_size=8) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        with pytest.raises(g.GcctbError) as e:
            ctx.run("persistent")
        assert e.value.code == -1


@pytest.mark.gpu
def test_latch_cost_against_lock_free():
    """The latch schemes against their lock-free forms on one contended
    workload (the reference's comparison, out_data/ycsb_hc_latch.txt): both
    commit everything; the rates are printed."""
    rows, ntx = 1 << 20, 65536
    w = g.generate(rows, ntx, theta=0.9, read_ratio=0.9, len_min=16, len_max=16)
    for scheme in ("silo", "tictoc", "tpl_nw", "mvcc"):
        rate = {}
        for name in (scheme, "slow_" + scheme):
            with g.Context(name, rows=rows, warp_density=3, block_size=8, batch_size=65536) as ctx:
                ctx.seed_table()
                ctx.load_ops(w)
                ctx.run("plain")  # warm-up
                ctx.seed_table()
                st = ctx.run("plain")
            assert st["committed"] == ntx and st["error_flags"] == 0
            rate[name] = ntx / (st["tot_ms"] / 1e3)
        print(f"[latch-cost] {scheme}: {rate[scheme] / 1e6:.2f} M Tx/s, slow_{scheme}: "
              f"{rate['slow_' + scheme] / 1e6:.2f} M Tx/s", flush=True)


@pytest.mark.gpu
def test_cli_latch_scheme(tmp_path):
    """test_ycsb with a latch scheme name, as the reference's driver accepts it."""
    assert os.path.exists(CLI), "build the tools first (gpu_cc_tb_amd/build.py)"
    dataset = "gen:theta=0.8,rr=0.6,len=12"
    args = [CLI, dataset, "2048", "4096", "slow_tictoc", "3", "8", "1024", "1"]
    print(f"[cli-latch] {' '.join(args[1:])}", flush=True)
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "TRY slow_tictoc,3,8,1024" in p.stdout and "REPLAY OK" in p.stdout
    fields = (tmp_path / (dataset + ".txt")).read_text().strip().splitlines()[-1].split(",")
    assert len(fields) == 15 and fields[0] == "slow_tictoc"
