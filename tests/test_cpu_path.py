"""The CPU path behind the C ABI (gcctb_cpu_run): every scheme on the host's
cores, checked like the GPU path, by the oracle's serial replay of the order
it reports (bit for bit: every read and the final table), GaccO against its
transaction-id closed form and GPUTx against the oracle's ranks.  These run
without a GPU."""
import os
import subprocess

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu_cc_tb_amd", "bin", "test_ycsb")
COLS = 10


def _replay_ok(rows, cols, w, init, res, row_ops=False):
    rep = orc.replay_rows if row_ops else orc.replay
    exp_t, exp_r, ties = rep(rows, cols, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert ties == 0, "serialization keys must be distinct"
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr]), "reads differ from the serial replay"
    assert np.array_equal(res["table"], exp_t), "table differs from the serial replay"


@pytest.mark.parametrize("scheme", list(g.SCHEMES))
@pytest.mark.parametrize("case", ["hot", "uniform", "varlen"])
def test_cpu_path_replay(scheme, case):
    rows, ntx = {"hot": (256, 3000), "uniform": (4096, 3000), "varlen": (1024, 2000)}[case]
    theta, rr, lmin, lmax = {"hot": (0.95, 0.5, 8, 8), "uniform": (0.0, 0.9, 16, 16),
                             "varlen": (0.8, 0.7, 4, 32)}[case]
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, seed=g.SEED + 3)
    init = orc.seed_table(g.SEED, rows, COLS)
    st, res = g.cpu_run(scheme, rows, w, batch_size=700, threads=4)
    assert st["committed"] == ntx and st["index_miss"] == 0
    assert st["batches"] == (st["ranks"] if scheme == "gputx" else (ntx + 699) // 700)
    assert st["tick_khz"] == 1e6 and st["tot_ms"] > 0
    _replay_ok(rows, COLS, w, init, res)
    assert g.check_host(rows, COLS, w, init, res) == 0
    if scheme == "gacco":
        exp_t, exp_r = orc.replay_txid(rows, COLS, g.SEED, init, w.ops, w.tx_off)
        assert np.array_equal(res["table"], exp_t)
    if scheme == "gputx":
        rank, counts = orc.gputx_rank(rows, w.ops, w.tx_off)
        assert np.array_equal(res["order"][1::3], rank.astype(np.uint64))
        assert st["ranks"] == len(counts)


@pytest.mark.parametrize("scheme", ["tpl_wd", "mvcc", "silo", "tictoc", "gputx"])
def test_cpu_path_whole_rows(scheme):
    """Whole-row mode on 13-cell (104-byte) rows: reads return row digests."""
    rows, ntx, cols = 512, 2000, 13
    w = g.generate(rows, ntx, theta=0.8, read_ratio=0.6, len_min=4, len_max=15, cols=cols)
    init = orc.seed_table(g.SEED, rows, cols)
    st, res = g.cpu_run(scheme, rows, w, cols=cols, batch_size=500, threads=4, row_ops=True)
    assert st["committed"] == ntx
    _replay_ok(rows, cols, w, init, res, row_ops=True)
    assert g.check_host(rows, cols, w, init, res, row_ops=True) == 0


def test_cpu_path_index_miss_and_errors():
    rows, ntx = 128, 200
    w = g.generate(rows, ntx, theta=0.5, read_ratio=0.5, len_min=4, len_max=4)
    ops = w.ops.copy()
    ops[::5] = (ops[::5] & ~np.uint64(0xFFFFFFFF)) | np.uint64(rows + 7)  # keys outside the table
    miss = g.Ops(ops, w.tx_off)
    st, res = g.cpu_run("silo", rows, miss, batch_size=64, threads=2)
    assert st["index_miss"] == len(ops[::5])
    assert np.all(res["reads"][::5] == 0)
    with pytest.raises(g.GcctbError):
        g.cpu_run("slow_silo", rows, w)
    with pytest.raises(g.GcctbError):
        g.cpu_run("silo", rows, w, batch_size=0)
    with pytest.raises(g.GcctbError):
        g.cpu_run("silo", rows, w, cols=20, row_ops=True)
    dup = w.ops.copy()
    dup[1] = dup[0]  # a repeated key inside transaction 0
    with pytest.raises(g.GcctbError):
        g.cpu_run("silo", rows, g.Ops(dup, w.tx_off))


def test_cli_cpu_device(tmp_path):
    """The drop-in driver on the CPU path (GCCTB_DEVICE=cpu), with the
    reference's argv: TRY line, the stdout block, REPLAY OK, and the CSV line."""
    assert os.path.exists(CLI), "build the tools first (gpu_cc_tb_amd/build.py)"
    dataset = "gen:theta=0.0,rr=0.9,len=16"
    args = [CLI, dataset, "65536", "8192", "tpl_nw", "5", "32", "4096", "1"]
    env = dict(os.environ, GCCTB_DEVICE="cpu", OMP_NUM_THREADS="4")
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stderr
    assert "TRY tpl_nw,5,32,4096" in p.stdout and "REPLAY OK" in p.stdout
    assert "device cpu" in p.stderr
    fields = (tmp_path / (dataset + ".txt")).read_text().strip().splitlines()[-1].split(",")
    assert len(fields) == 15 and fields[0] == "tpl_nw" and float(fields[4]) > 0
