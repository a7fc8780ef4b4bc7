"""CPU tests that pin the oracle before it is trusted.

Known-answer values, closed forms, hand-derived fixtures and brute force pin
each oracle function; the product's host-side generator is then checked bit
for bit against the oracle's independent restatement.
"""
import itertools
import json
import os

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- hashing
def test_splitmix64_known_answers():
    # splitmix64 (Steele, Lea, Flood 2014; Vigna's reference code): the first
    # outputs of a generator whose state starts at 0 and 1234567.
    L = orc.lib()
    state, outs = 0, []
    for _ in range(3):
        outs.append(L.orc_splitmix64(state))
        state = (state + 0x9E3779B97F4A7C15) & (2**64 - 1)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert L.orc_splitmix64(1234567) == 6457827717110365317


def test_seeded_table_matches_definition():
    t = orc.seed_table(g.SEED, 5, 10)
    L = orc.lib()
    for r in range(5):
        for c in range(10):
            assert t[r * 10 + c] == L.orc_splitmix64(g.SEED ^ (r * 10 + c))


# ---------------------------------------------------------------- zeta / zipf
@pytest.mark.parametrize("n", [1, 2, 3, 100, 65536, 65537, 200001])
def test_zeta_closed_forms(n):
    L = orc.lib()
    assert L.orc_zeta(n, 0.0) == float(n)  # theta = 0: sum of ones
    assert L.orc_zeta(1, 0.7) == 1.0
    assert L.orc_zeta(2, 0.7) == 1.0 + 2 ** -0.7
    i = np.arange(1, n + 1, dtype=np.float64)
    for theta in (0.1, 0.6, 0.99):
        ref = float(np.sum((1.0 / i) ** theta))
        assert L.orc_zeta(n, theta) == pytest.approx(ref, rel=1e-12)
        assert g.zeta(n, theta) == L.orc_zeta(n, theta)  # product uses the same blocking


def test_zipf_exact_head_probabilities():
    # gen_ycsb.cu:97-100: rank 1 iff u*zetan < 1, rank 2 iff u*zetan < 1 + 0.5^theta
    L = orc.lib()
    n, theta = 1000, 0.99
    z = L.orc_zeta(n, theta)
    assert L.orc_zipf_rank(n, theta, z, 0.0) == 1
    assert L.orc_zipf_rank(n, theta, z, (1 - 1e-12) / z) == 1
    assert L.orc_zipf_rank(n, theta, z, (1 + 1e-9) / z) == 2
    assert L.orc_zipf_rank(n, theta, z, (1 + 0.5**theta + 1e-9) / z) >= 3
    us = np.linspace(0, 1, 20001, endpoint=False)
    ranks = np.array([L.orc_zipf_rank(n, theta, z, u) for u in us])
    assert ranks.min() == 1 and ranks.max() <= n
    assert np.all(np.diff(ranks) >= 0), "rank is non-decreasing in u"
    p1 = np.mean(ranks == 1)
    assert p1 == pytest.approx(1.0 / z, abs=2e-4)


def test_zipf_theta0_is_uniform():
    # theta = 0: alpha = 1, eta = 1, rank = 1 + floor(n u) above the head
    L = orc.lib()
    n = 97
    z = L.orc_zeta(n, 0.0)
    for u in np.linspace(2.0 / n, 0.999, 50):
        assert L.orc_zipf_rank(n, 0.0, z, u) == 1 + int(n * u)


# ---------------------------------------------------------------- generator
GEN_CASES = [(4096, 0.0, 0.9, 16, 16), (64, 0.99, 0.5, 8, 8), (1024, 0.8, 0.5, 4, 32), (8, 0.9, 0.5, 4, 4),
             (10, 0.5, 0.5, 10, 10), (10485760, 0.99, 0.9, 16, 16)]


@pytest.mark.parametrize("rows,theta,rr,lmin,lmax", GEN_CASES)
def test_generator_matches_oracle(rows, theta, rr, lmin, lmax):
    ntx = 100
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, tx_base=12345)
    z = orc.lib().orc_zeta(rows, theta)
    for i in range(ntx):
        exp = orc.gen_tx(rows, 10, lmin, lmax, theta, rr, g.SEED, 12345 + i, z)
        got = w.ops[w.tx_off[i]:w.tx_off[i + 1]]
        assert np.array_equal(exp, got)
        keys = g.op_key(got)
        assert len(got) >= lmin and len(got) <= lmax
        assert np.all(np.diff(keys.astype(np.int64)) > 0), "distinct keys in ascending order"
        assert np.all(keys < rows)


def test_generator_statistics():
    rows, ntx = 100000, 20000
    w = g.generate(rows, ntx, theta=0.0, read_ratio=0.9)
    wr = g.op_is_write(w.ops)
    assert np.mean(~wr) == pytest.approx(0.9, abs=0.01)
    cols = np.bincount(g.op_col(w.ops), minlength=10)
    assert cols.min() > 0.9 * len(w.ops) / 10
    # one op per transaction, so no draw is dropped as a duplicate: row 0 is
    # rank 1 and appears with probability 1/zeta(n, theta), row 1 with 2^-theta/zeta
    w2 = g.generate(rows, 200000, theta=0.99, read_ratio=0.5, len_min=1)
    keys = g.op_key(w2.ops)
    z = g.zeta(rows, 0.99)
    assert np.mean(keys == 0) == pytest.approx(1 / z, rel=0.03)
    assert np.mean(keys == 1) == pytest.approx(2**-0.99 / z, rel=0.05)
    w3 = g.generate(rows, 5000, theta=0.5, read_ratio=0.5, len_min=4, len_max=32)
    lens = np.diff(w3.tx_off)
    assert lens.min() == 4 and lens.max() == 32


def test_generator_deterministic_and_thread_independent():
    a = g.generate(5000, 3000, theta=0.9, read_ratio=0.5, threads=1)
    b = g.generate(5000, 3000, theta=0.9, read_ratio=0.5, threads=7)
    assert np.array_equal(a.ops, b.ops) and np.array_equal(a.tx_off, b.tx_off)
    # a stream position shifts the transaction ids
    c = g.generate(5000, 1000, theta=0.9, read_ratio=0.5, tx_base=2000)
    assert np.array_equal(c.ops, a.ops[a.tx_off[2000]:])


# ---------------------------------------------------------------- index Find
def test_find_matches_binary_search():
    rng = np.random.default_rng(1)
    keys = np.unique(rng.integers(0, 1 << 20, 5000).astype(np.uint32))
    vals = rng.integers(0, 1 << 40, len(keys)).astype(np.uint64)
    for k in list(keys[:50]) + list(keys[-50:]) + [0, 1, (1 << 20) + 5]:
        pos = np.searchsorted(keys, k)
        exp = vals[pos] if pos < len(keys) and keys[pos] == k else len(keys)
        assert orc.find(keys, vals, int(k)) == exp
    assert orc.find(keys[:1], vals[:1], int(keys[0])) == vals[0]


# ---------------------------------------------------------------- serial replay
def _mk(txs, cols=10):
    """txs: list of [(key, col, write)] -> Ops"""
    ops, off = [], [0]
    for t in txs:
        ops += [int(g.make_ops([k], [c], [w])[0]) for k, c, w in t]
        off.append(len(ops))
    return np.array(ops, dtype=np.uint64), np.array(off, dtype=np.uint64)


def test_replay_hand_example():
    seed, rows = 7, 4
    ops, off = _mk([[(1, 2, 1)], [(1, 2, 0), (3, 0, 1)], [(3, 0, 0), (1, 2, 0)]])
    init = orc.seed_table(seed, rows, 10)
    L = orc.lib()
    # order T2, T0, T1 (keys 2, 0, 1 within batch 0)
    order = np.array([0, 5, 0, 0, 9, 0, 0, 1, 0], dtype=np.uint64)
    t, r, ties = orc.replay(rows, 10, seed, init, ops, off, order)
    assert ties == 0
    # T2 first: reads initial row3 col0 and row1 col2
    assert r[3] == init[30] and r[4] == init[12]
    # then T0 writes row1 col2, then T1 reads it
    w0 = L.orc_write_value(seed, 0, 0)
    assert r[1] == w0
    assert t[12] == w0 and t[30] == L.orc_write_value(seed, 1, 1)
    assert r[0] == 0 and r[2] == 0  # writes report 0


def test_replay_detects_ties_and_txid_form():
    ops, off = _mk([[(0, 0, 1)], [(0, 0, 1)]])
    init = orc.seed_table(1, 2, 10)
    order = np.zeros(6, dtype=np.uint64)
    _, _, ties = orc.replay(2, 10, 1, init, ops, off, order)
    assert ties == 1
    order = np.array([0, 0, 0, 0, 1, 1], dtype=np.uint64)
    t1, r1, _ = orc.replay(2, 10, 1, init, ops, off, order)
    t2, r2 = orc.replay_txid(2, 10, 1, init, ops, off)
    assert np.array_equal(t1, t2) and np.array_equal(r1, r2)


def test_replay_brute_force_orders():
    # 3 transactions on 2 rows: every permutation gives a different history;
    # replaying a permutation must equal executing it by hand.
    seed, rows = 3, 2
    txs = [[(0, 1, 1), (1, 1, 0)], [(0, 1, 0), (1, 1, 1)], [(0, 1, 1), (1, 1, 1)]]
    ops, off = _mk(txs)
    init = orc.seed_table(seed, rows, 10)
    L = orc.lib()
    seen = set()
    for perm in itertools.permutations(range(3)):
        order = np.zeros(9, dtype=np.uint64)
        for pos, t in enumerate(perm):
            order[3 * t + 1] = pos
        t, r, _ = orc.replay(rows, 10, seed, init, ops, off, order)
        table = init.copy()
        exp_r = np.zeros(len(ops), dtype=np.uint64)
        for tx in perm:
            for j, (k, c, w) in enumerate(txs[tx]):
                i = int(off[tx]) + j
                if w:
                    table[k * 10 + c] = L.orc_write_value(seed, tx, j)
                else:
                    exp_r[i] = table[k * 10 + c]
        assert np.array_equal(t, table) and np.array_equal(r, exp_r)
        seen.add(tuple(r))
    assert len(seen) > 1


# ---------------------------------------------------------------- GPUTx rank
def _golden_rank_cases():
    with open(os.path.join(GOLDEN, "gputx_rank_cases.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden_rank_cases(), ids=lambda c: c["name"])
def test_gputx_rank_golden(case):
    ntx = max(t for t, _, _ in case["ops"]) + 1
    txs = [[] for _ in range(ntx)]
    for t, k, w in case["ops"]:
        txs[t].append((k, 0, w))
    ops, off = _mk(txs)
    rank, counts = orc.gputx_rank(case["rows"], ops, off)
    assert list(rank) == case["rank"]
    assert list(counts) == case["counts"]


def _rank_by_definition(rows, ops, off):
    """src/gputx.cu:17-27 restated with Python loops (small inputs only):
    accesses of a row in transaction-id order, the local rank counts the
    neighbour pairs before an access of which one writes."""
    ntx = len(off) - 1
    per_row = {}
    for t in range(ntx):
        for i in range(int(off[t]), int(off[t + 1])):
            r = int(ops[i]) & 0xFFFFFFFF
            if r < rows:
                per_row.setdefault(r, []).append((t, bool((int(ops[i]) >> 40) & 1)))
    rank = [0] * ntx
    for r in sorted(per_row):
        acc = per_row[r]
        cur = 0
        for i, (t, wr) in enumerate(acc):
            rank[t] = max(rank[t], cur)
            if i + 1 < len(acc) and (wr or acc[i + 1][1]):
                cur += 1
    return np.array(rank, dtype=np.uint32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gputx_rank_matches_definition(seed):
    """The counting-sort rank equals the definition walked access by access,
    including keys past the table (skipped) and empty transactions."""
    rows = 40
    w = g.generate(rows + 8, 300, theta=0.7, read_ratio=0.6, len_min=1, len_max=12, seed=seed)
    off = w.tx_off.copy()
    off[100:] = off[100:] - (off[100] - off[99])  # transaction 99 becomes empty
    ops = np.delete(w.ops, np.arange(int(w.tx_off[99]), int(w.tx_off[100])))
    rank, counts = orc.gputx_rank(rows, ops, off)
    exp = _rank_by_definition(rows, ops, off)
    assert np.array_equal(rank, exp)
    assert np.array_equal(counts, np.bincount(exp, minlength=int(exp.max()) + 1))


@pytest.mark.parametrize("threads", [2, 3, 8, 64])
def test_replay_mt_matches_serial(threads):
    """The row-partitioned replay gives the serial replay's table, reads and
    tie count for any thread count (more threads than rows included), with
    keys past the table and tied order keys in the input."""
    rows = 50
    w = g.generate(rows + 5, 2000, theta=0.9, read_ratio=0.5, len_min=1, len_max=20, seed=threads)
    init = orc.seed_table(g.SEED, rows, 10)
    order = np.random.default_rng(threads).integers(0, 40, 3 * w.ntx).astype(np.uint64)
    order[0::3] &= np.uint64(1)  # batch 0 or 1: many tied keys
    t1, r1, ties1 = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, order)
    t2, r2, ties2 = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, order, threads=threads)
    assert ties1 == ties2 > 0
    assert np.array_equal(t1, t2)
    assert np.array_equal(r1, r2)


def test_gputx_rank_properties_random():
    # every pair of conflicting accesses in a row's tx order has non-decreasing
    # ranks... is NOT guaranteed (no propagation); what is: a transaction's rank
    # is at least its local rank in every row it touches, and the rank count
    # covers every rank.
    w = g.generate(50, 400, theta=0.9, read_ratio=0.6, len_min=3, len_max=6)
    rank, counts = orc.gputx_rank(50, w.ops, w.tx_off)
    assert counts.sum() == 400 and len(counts) == rank.max() + 1
    keys = g.op_key(w.ops)
    wr = g.op_is_write(w.ops)
    txof = np.repeat(np.arange(400), np.diff(w.tx_off).astype(np.int64))
    for row in range(50):
        idx = np.nonzero(keys == row)[0]
        idx = idx[np.argsort(txof[idx], kind="stable")]
        cur = 0
        for a, i in enumerate(idx):
            assert rank[txof[i]] >= cur
            if a + 1 < len(idx) and (wr[i] or wr[idx[a + 1]]):
                cur += 1


# ---------------------------------------------------------------- CPU path
CPU_CASES = [(4096, 2048, 16, 16, 0.0, 0.9, 512), (64, 1024, 8, 8, 0.99, 0.5, 256),
             (1024, 1000, 4, 32, 0.8, 0.5, 333)]


@pytest.mark.parametrize("case", CPU_CASES, ids=["uniform", "hot", "varlen"])
@pytest.mark.parametrize("scheme", list(g.SCHEMES))
def test_cpu_path_matches_serial_replay(scheme, case):
    rows, ntx, lmin, lmax, theta, rr, batch = case
    w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax)
    init = orc.seed_table(g.SEED, rows, 10)
    rc, table, reads, order, st = orc.cpu_run(g.SCHEMES.index(scheme), rows, 10, g.SEED, init, w.ops,
                                              w.tx_off, batch, 4)
    assert rc == 0 and st.committed == ntx
    exp_t, exp_r, ties = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, order)
    assert ties == 0
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(table, exp_t)
    assert np.array_equal(reads[isr], exp_r[isr])
    if scheme == "gacco":
        t2, r2 = orc.replay_txid(rows, 10, g.SEED, init, w.ops, w.tx_off)
        assert np.array_equal(table, t2)


GOLDEN = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def _mix64(z):
    # the splitmix64 finalizer, through the pinned splitmix64: mix64(z) = splitmix64(z - golden)
    return orc.lib().orc_splitmix64((z - GOLDEN) & M64)


def test_row_digest_and_write_cell_forms():
    """Whole-row mode's digest folds the cells with the splitmix64 finalizer
    from "ROW", and a written cell is the finalizer of (write value + column)."""
    init = 0x524F570000000000
    a, b = 0x0123456789ABCDEF, 0xFEDCBA9876543210
    assert orc.row_digest([a]) == _mix64(init ^ a)
    assert orc.row_digest([a, b]) == _mix64(_mix64(init ^ a) ^ b)
    assert orc.row_digest([b, a]) != orc.row_digest([a, b])
    for t, j, c in [(0, 0, 0), (5, 3, 12), (123456, 31, 7)]:
        w = orc.lib().orc_write_value(g.SEED, t, j)
        assert orc.lib().orc_write_cell(g.SEED, t, j, c) == _mix64((w + c) & M64)


def test_replay_rows_one_column_matches_column_replay():
    """With one column, whole-row replay is column replay seen through the
    finalizer: a written cell is mix64 of the column-mode value, and a read's
    digest is mix64(ROW ^ the row's cell at the time)."""
    rows, ntx = 64, 400
    w = g.generate(rows, ntx, theta=0.9, read_ratio=0.6, len_min=1, len_max=8, cols=1, seed=11)
    order = np.zeros(3 * ntx, dtype=np.uint64)
    order[1::3] = np.random.default_rng(3).permutation(ntx)
    init = orc.seed_table(g.SEED, rows, 1)
    tc, rc, _ = orc.replay(rows, 1, g.SEED, init, w.ops, w.tx_off, order)
    tr, rr, ties = orc.replay_rows(rows, 1, g.SEED, init, w.ops, w.tx_off, order)
    assert ties == 0
    isw = g.op_is_write(w.ops)
    written = np.zeros(rows, dtype=bool)
    written[g.op_key(w.ops)[isw].astype(np.int64)] = True
    for r in range(rows):
        assert int(tr[r]) == (_mix64(int(tc[r])) if written[r] else int(init[r]))
    keys = g.op_key(w.ops).astype(np.int64)
    for i in np.nonzero(~isw)[0]:
        v = int(rc[i])
        cell = v if v == int(init[keys[i]]) else _mix64(v)
        assert int(rr[i]) == _mix64(0x524F570000000000 ^ cell)


@pytest.mark.parametrize("threads", [2, 5])
def test_replay_rows_mt_matches_serial(threads):
    rows, ntx, cols = 300, 500, 13
    w = g.generate(rows, ntx, theta=0.8, read_ratio=0.5, len_min=4, len_max=16, cols=cols, seed=12)
    order = np.zeros(3 * ntx, dtype=np.uint64)
    order[1::3] = np.random.default_rng(4).permutation(ntx)
    init = orc.seed_table(g.SEED, rows, cols)
    t1, r1, _ = orc.replay_rows(rows, cols, g.SEED, init, w.ops, w.tx_off, order, threads=1)
    t2, r2, _ = orc.replay_rows(rows, cols, g.SEED, init, w.ops, w.tx_off, order, threads=threads)
    assert np.array_equal(t1, t2) and np.array_equal(r1, r2)
    # every written row holds one transaction's write_cell values in all of its cells
    isw = g.op_is_write(w.ops)
    t1 = t1.reshape(rows, cols)
    for r in set(g.op_key(w.ops)[isw].astype(np.int64).tolist()):
        assert int(t1[r, 1]) != int(t1[r, 0])
