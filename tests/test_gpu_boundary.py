"""GPU tests of the drop-in boundary: the sorted-array index with arbitrary
keys against the oracle's Find, transactions resolved through such an index,
the debug-mode serializability check, and the test_ycsb command line."""
import os
import subprocess

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu_cc_tb_amd", "bin", "test_ycsb")


def idx_bucket(keys, nb):
    """Home bucket of each key in an index of nb buckets (Fibonacci hashing,
    then the high half of a 32 x 32 product)."""
    h = (np.asarray(keys, dtype=np.uint64) * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
    return ((h * np.uint64(nb)) >> np.uint64(32)).astype(np.int64)


def clustered_keys(n, ncluster, seed):
    """n distinct keys, ncluster of them in the index's last home bucket: its
    probe chain runs past the end and wraps to bucket 0.  Also returns more
    keys of that bucket that are not in the index."""
    nb = n // 4 + 1  # buckets of an index over n keys
    cand = np.arange(1, 1 << 22, dtype=np.uint64)
    home = idx_bucket(cand, nb)
    last, rest = cand[home == nb - 1], cand[home != nb - 1]
    rng = np.random.default_rng(seed)
    pick = rng.choice(last.size, ncluster + 50, replace=False)
    keys = np.concatenate([last[pick[:ncluster]], rng.choice(rest, n - ncluster, replace=False)])
    rng.shuffle(keys)
    return keys.astype(np.uint32), last[pick[ncluster:]].astype(np.uint32)


def random_index(rows, n, seed):
    r = np.random.default_rng(seed)
    keys = r.choice(np.uint64(0xFFFFFFFE), size=n, replace=False).astype(np.uint32)
    vals = r.permutation(rows)[:n].astype(np.uint32)
    return keys, vals


@pytest.mark.gpu
def test_probe_custom_index():
    """Probe of an index over arbitrary keys: hits return the value, misses
    the index length (Find, include/index.cuh:74-91)."""
    rows, n = 5000, 3000
    keys, vals = random_index(rows, n, 11)
    order = np.argsort(keys, kind="stable")
    sk, sv = keys[order], vals[order].astype(np.uint64)
    probes = np.concatenate([keys[:1000], np.random.default_rng(12).integers(0, 2**32 - 1, 1000).astype(np.uint32),
                             np.array([0, 0xFFFFFFFE], dtype=np.uint32)])
    print(f"[probe] n={n} probes={probes.size}", flush=True)
    with g.Context("tpl_nw", rows=rows) as ctx:
        ctx.seed_table()
        ctx.build_index(keys, vals)
        got = ctx.probe(probes)
    exp = np.array([orc.find(sk, sv, int(k)) for k in probes], dtype=np.uint64)
    assert np.array_equal(got.astype(np.uint64), exp)


@pytest.mark.gpu
def test_probe_long_chains():
    """Keys crowded into one home bucket: hits along a chain that wraps past
    the last bucket, and misses that walk the chain to a free slot."""
    n = 400
    keys, absent = clustered_keys(n, 120, 5)
    vals = np.random.default_rng(6).permutation(n).astype(np.uint32)
    probes = np.concatenate([keys, absent, np.array([0, 0xFFFFFFFE], dtype=np.uint32)])
    print(f"[chains] n={n} probes={probes.size}", flush=True)
    with g.Context("tpl_nw", rows=n) as ctx:
        ctx.seed_table()
        ctx.build_index(keys, vals)
        got = ctx.probe(probes).astype(np.int64)
    exp = np.concatenate([vals.astype(np.int64), np.full(absent.size, n),
                          [n if 0 not in keys else int(vals[np.nonzero(keys == 0)[0][0]]), n]])
    assert np.array_equal(got, exp)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("keyset", ["random", "chains", "short"])
@pytest.mark.parametrize("scheme", ["silo", "mvcc", "tpl_wd", "gacco", "gputx"])
def test_run_through_custom_index(scheme, keyset, form):
    """Transactions whose keys are arbitrary 32-bit values, resolved to rows
    through the index (inside the execution kernel in the plain form, by the
    probe pass in the persistent form): the result equals the serial replay on
    the rows the index maps them to.  "short": an index of fewer keys than
    the table has rows, whose values reach the last rows."""
    rows, ntx = 2048, 1500
    nkeys = 700 if keyset == "short" else rows
    w = g.generate(nkeys, ntx, theta=0.8, read_ratio=0.7, len_min=8, len_max=16)
    rng = np.random.default_rng(21)
    if keyset == "chains":
        keys, _ = clustered_keys(nkeys, 300, 22)
    else:
        keys = rng.choice(np.uint64(0xFFFFFFFE), size=nkeys, replace=False).astype(np.uint32)
    vals = np.sort(rng.choice(rows, nkeys, replace=False)) if keyset == "short" else np.arange(rows)
    vals = vals.astype(np.uint32)
    if keyset == "short":
        assert vals.max() >= 2048 - 64 and nkeys < 1024  # rows past the index length's bit width
    ent = g.op_key(w.ops).astype(np.int64)  # index entry of each op
    col, isw = g.op_col(w.ops), g.op_is_write(w.ops)
    mapped = g.Ops(g.make_ops(keys[ent], col, isw), w.tx_off)
    on_rows = g.make_ops(vals[ent], col, isw)  # the same ops addressed by table row
    print(f"[index-run] scheme={scheme} keyset={keyset} form={form}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=500) as ctx:
        ctx.seed_table()
        ctx.build_index(keys, vals)
        ctx.load_ops(mapped)
        st = ctx.run(form)
        res = ctx.fetch()
    assert st["committed"] == ntx and st["index_miss"] == 0
    init = orc.seed_table(g.SEED, rows, 10)
    exp_t, exp_r, ties = orc.replay(rows, 10, g.SEED, init, on_rows, w.tx_off, res["order"])
    assert ties == 0
    isr = ~isw
    assert np.array_equal(res["reads"][isr], exp_r[isr])
    assert np.array_equal(res["table"], exp_t)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["tpl_nw", "tictoc", "gputx"])
def test_check_detects_bad_history(scheme):
    """gcctb_check (the driver's debug check): clean on the true initial
    table, and it counts disagreements when given a different one."""
    rows, ntx = 1024, 2000
    w = g.generate(rows, ntx, theta=0.9, read_ratio=0.5, len_min=8, len_max=8)
    print(f"[check] scheme={scheme}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=1000) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        ctx.run("plain")
        init = orc.seed_table(g.SEED, rows, 10)
        assert ctx.check(init) == (0, 0)
        wrong = init.copy()
        wrong[::7] ^= np.uint64(1)
        bad, unc = ctx.check(wrong)
    assert bad > 0 and unc == 0


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["silo", "gacco"])
def test_cli_debug_run(scheme, tmp_path):
    """The drop-in driver: same argv as the reference's test_ycsb, its TRY
    line, the debug check's GOOD, and the 15-field CSV line appended to
    <dataset>.txt."""
    assert os.path.exists(CLI), "build the tools first (gpu_cc_tb_amd/build.py)"
    dataset = "gen:theta=0.6,rr=0.5,len=16"
    args = [CLI, dataset, "4096", "8192", scheme, "3", "8", "2048", "1"]
    print(f"[cli] {' '.join(args[1:])}", flush=True)
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert f"TRY {scheme},3,8,2048" in p.stdout
    assert "SCAN OK" in p.stdout and "GOOD" in p.stdout and "BAD" not in p.stdout
    assert "REPLAY OK" in p.stdout
    line = (tmp_path / (dataset + ".txt")).read_text().strip().splitlines()[-1]
    fields = line.split(",")
    assert len(fields) == 15 and fields[0] == scheme
    assert float(fields[4]) > 0  # throughput


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["silo", "tictoc", "mvcc"])
def test_cli_reference_txs_file(scheme, tmp_path):
    """A reference-format .txs file (136-byte YCSBTx records with request_cnt
    below 16, as the reference's generator leaves after dropping repeated
    keys) through the drop-in driver: whole-row ops on 104-byte rows, the
    serial replay OK, the conflict check GOOD where the stamps are exact, and
    the CSV line."""
    rows, ntx = 4096, 3000
    w = g.generate(rows, ntx, theta=0.7, read_ratio=0.8, len_min=3, len_max=15, cols=13)
    assert int(np.diff(w.tx_off.astype(np.int64)).max()) < 16
    path = str(tmp_path / "ref.txs")
    g.write_txs(path, w)
    assert os.path.getsize(path) == 136 * ntx
    args = [CLI, path, str(rows), str(ntx), scheme, "3", "8", "1024", "1"]
    print(f"[cli-txs] {' '.join(args[1:])}", flush=True)
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "SCAN OK" in p.stdout and "REPLAY OK" in p.stdout
    if scheme in EXACT_STAMPS:
        assert "GOOD" in p.stdout and "BAD" not in p.stdout, p.stdout[-3000:]
    fields = open(path + ".txt").read().strip().splitlines()[-1].split(",")
    assert len(fields) == 15 and fields[0] == scheme and float(fields[4]) > 0


# Schemes whose event stamps order every pair of conflicting accesses as the
# accesses themselves: 2PL (stamped while every lock is held), T/O (a read
# between the word it saw and the rts raise, a write after its mark), Silo (a
# read at its validation, a write after its lock), GaccO and GPUTx (between
# taking the turn and passing it on).  TicToc's reads whose rts already covers
# the commit are never re-checked, so they are stamped at the snapshot, next
# to the access rather than with it.  MVCC is left out for a reason of its
# own: the conflict graph (src/runtime.cu:23-71) is single-version, so a
# reader that reads the older version of a row after a younger transaction
# installed an uncommitted newer one (legal under MVCC: the reader serializes
# first) counts as reading that writer's data, and two such transactions form
# a cycle in the graph; the reference's checker reports the same history BAD.
# For MVCC and TicToc the serial replay is the verdict.
EXACT_STAMPS = ("tpl_nw", "tpl_wd", "to", "silo", "gacco", "gputx")


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "graph"])
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_event_log_conflict_check(scheme, form):
    """The debug event log (AddEvent, include/runtime.cuh:64-81): every op of
    every committed attempt and every commit is stamped; the reference's
    conflict-graph check over it (src/runtime.cu:23-158) finds no
    uncommitted transaction and, where the stamps order conflicting accesses
    exactly, no cycle; the run still equals its serial replay."""
    rows, ntx = 512, 1500
    w = g.generate(rows, ntx, theta=0.8, read_ratio=0.6, len_min=4, len_max=12)
    print(f"[events] scheme={scheme} form={form}", flush=True)
    with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=500, debug_events=True) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run(form)
        res = ctx.fetch()
        ev = ctx.fetch_events()
    assert st["committed"] == ntx
    assert ev.size == w.nops + ntx
    assert np.unique(ev["id"]).size == ev.size and ev["id"].min() >= 1
    # a commit is stamped after every op of its transaction
    commit = np.zeros(ntx, dtype=np.uint64)
    c = ev[ev["type"] == 2]
    commit[c["tid"]] = c["id"]
    ops = ev[ev["type"] != 2]
    assert np.all(ops["id"] < commit[ops["tid"]])
    r = g.conflict_check(ev, ntx)
    assert r["uncommitted"] == 0
    print(f"[events] {scheme} {form}: cycle of {len(r['cycle'])}", flush=True)
    if scheme in EXACT_STAMPS:
        assert r["cycle"] == [], r["report"][-2000:]
        assert r["report"].endswith("GOOD\n")
    init = orc.seed_table(g.SEED, rows, 10)
    exp_t, exp_r, _ = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert np.array_equal(res["table"], exp_t)
