"""TPC-C on the host (no GPU): the oracle's updates against hand-worked rows,
the generator's distributions and layout, the initial tables, and
new_order's totals against an independent float32 computation."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

F = np.float32


def bits(x) -> int:
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def fl(u) -> float:
    return float(np.array([u], dtype=np.uint32).view(np.float32)[0])


def pack(lo, hi) -> int:
    return (int(lo) & 0xFFFFFFFF) | ((int(hi) & 0xFFFFFFFF) << 32)


def test_oracle_updates_hand_worked():
    """Each update of benchmark/tpcc.cu on a row worked by hand."""
    # payment, warehouse: ytd 300000 + 12.5 (tpcc.cu:102)
    res, r = orc.tpcc_apply(0, bits(12.5), [pack(3, bits(0.125)), pack(bits(300000.0), 0), 0, 0, 0])
    assert res == bits(300000.0) and r[1] == pack(bits(300012.5), 0) and r[0] == pack(3, bits(0.125))
    # payment, district: ytd (high half of cell 1) 30000 + 7.25, tax untouched (tpcc.cu:125)
    res, r = orc.tpcc_apply(1, bits(7.25), [pack(2, 5), pack(bits(0.0625), bits(30000.0)), pack(3001, 0), 0, 0])
    assert res == bits(30000.0) and r[1] == pack(bits(0.0625), bits(30007.25)) and r[2] == pack(3001, 0)
    # payment, customer: balance -10 - 100, ytd_payment 10 + 100, payment_cnt 1 + 1 (tpcc.cu:152-154)
    res, r = orc.tpcc_apply(2, bits(100.0), [0, 0, 0, pack(bits(-10.0), bits(10.0)), pack(1, 0)])
    assert res == bits(-10.0) and r[3] == pack(bits(-110.0), bits(110.0)) and r[4] == pack(2, 0)
    # new_order, district: the order takes next_o_id 3001, the district moves to 3002 (tpcc.cu:270-271)
    res, r = orc.tpcc_apply(3, 0, [0, 0, pack(3001, 7), 0, 0])
    assert res == 3001 and r[2] == pack(3002, 7)
    # new_order, stock (tpcc.cu:319-327): 50 > 5 + 10 -> 45; ytd + 5, order_cnt + 1, remote_cnt + 1
    res, r = orc.tpcc_apply(4, 5 | (1 << 16), [0, pack(50, 20), pack(3, 4), 0, 0])
    assert res == 50 and r[1] == pack(45, 25) and r[2] == pack(4, 5)
    # 15 is not > 5 + 10: 15 - (5 - 91) = 101 in unsigned arithmetic; local line
    res, r = orc.tpcc_apply(4, 5, [0, pack(15, 0), pack(0, 0), 0, 0])
    assert res == 15 and r[1] == pack(101, 5) and r[2] == pack(1, 0)
    res, r = orc.tpcc_apply(4, 10, [0, pack(12, 0), 0, 0, 0])
    assert r[1] == pack(93, 10)


def test_oracle_replay_order():
    """The replay applies transactions in (batch, key0, key1) order: two
    payments to one warehouse in either order give the float sums in that
    order, and the results show which came first."""
    rows = g.tpcc_rows(1)
    cells, _ = g.tpcc_seed_tables(1)
    a, b = bits(1.5), bits(2.25)
    ops = g.make_ops([0, 0], [0, 0], [True, True])
    off = np.array([0, 1, 2], dtype=np.uint64)
    params = np.array([a, b], dtype=np.uint32)
    w0 = int(cells[1]) & 0xFFFFFFFF
    for order, first in [([0, 5, 0, 0, 9, 0], 0), ([0, 9, 0, 0, 5, 0], 1)]:
        t, res, ties = orc.tpcc_replay(rows, cells, ops, params, off, np.array(order, dtype=np.uint64))
        assert ties == 0
        assert int(res[first]) == w0
        exp = F(F(fl(w0)) + F(fl(params[first]))) + F(fl(params[1 - first]))
        assert int(t[1]) & 0xFFFFFFFF == bits(exp)
        assert int(res[1 - first]) == bits(F(fl(w0)) + F(fl(params[first])))


def test_seed_tables_fields():
    W = 2
    cells, price = g.tpcc_seed_tables(W)
    rows = g.tpcc_rows(W)
    assert rows == W * (1 + 10 + 30000 + 100000) and cells.size == rows * 5
    c = cells.reshape(rows, 5)
    lo = (c & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    hi = (c >> np.uint64(32)).astype(np.uint32)
    # warehouses: id, tax in [0, 0.2), ytd 300000
    assert list(lo[:W, 0]) == list(range(W))
    tax = hi[:W, 0].view(np.float32)
    assert np.all((tax >= 0) & (tax < 0.2)) and np.all(lo[:W, 1].view(np.float32) == 300000.0)
    # districts: (id, w_id), ytd 30000, next_o_id 3001
    d = slice(W, W + 10 * W)
    assert np.all(lo[d, 2] == 3001) and np.all(hi[d, 1].view(np.float32) == 30000.0)
    assert list(lo[W:W + 10, 0]) == list(range(10)) and np.all(hi[W:W + 10, 0] == 0)
    # customers: balance -10, ytd_payment 10, payment_cnt 1, discount in [0, 0.5)
    cs = slice(11 * W, 11 * W + 30000 * W)
    assert np.all(lo[cs, 3].view(np.float32) == -10.0) and np.all(hi[cs, 3].view(np.float32) == 10.0)
    assert np.all(lo[cs, 4] == 1)
    disc = hi[cs, 2].view(np.float32)
    assert np.all((disc >= 0) & (disc < 0.5)) and disc.std() > 0.1
    # stock: (i_id, w_id), quantity in [10, 100)
    st = slice(30011 * W, rows)
    q = lo[st, 1]
    assert q.min() >= 10 and q.max() < 100 and q.min() < 15 and q.max() > 94
    assert np.all(lo[st, 0][:100000] == np.arange(100000)) and np.all(hi[st, 0][100000:] == 1)
    assert price.min() >= 1 and price.max() < 100


def test_payment_generator():
    W, n = 4, 20000
    w = g.tpcc_generate(W, n, new_order=False)
    assert w.nops == 3 * n and np.all(np.diff(w.tx_off.astype(np.int64)) == 3)
    key, kind = g.op_key(w.ops).astype(np.int64), g.op_col(w.ops)
    assert np.all(g.op_is_write(w.ops))
    assert list(kind[:6]) == [0, 1, 2, 0, 1, 2]
    wh, dist, cust = key[0::3], key[1::3], key[2::3]
    assert wh.max() < W and np.all((dist >= W) & (dist < 11 * W)) and np.all((cust >= 11 * W) & (cust < 30011 * W))
    assert np.all((dist - W) // 10 == wh)  # the district belongs to the paying warehouse
    c_w = (cust - 11 * W) // 30000
    remote = c_w != wh
    # uniform_int(0, 100) > 85: 14% pay through another warehouse (gen_tpcc.cu:122-126)
    assert 0.13 < remote.mean() < 0.15
    d_local = (dist - W) % 10
    c_d = ((cust - 11 * W) % 30000) // 3000
    assert np.all(d_local[remote] != c_d[remote]) and np.all(d_local[~remote] == c_d[~remote])
    c_id = (cust - 11 * W) % 3000
    assert c_id.min() == 0 and c_id.max() == 2999
    amt = w.params[0::3].view(np.float32)
    assert np.all(w.params[0::3] == w.params[1::3]) and np.all(w.params[0::3] == w.params[2::3])
    assert amt.min() >= 1.0 and amt.max() <= 5000.0
    assert np.array_equal(w.tx_customer, cust.astype(np.uint64))
    w2 = g.tpcc_generate(W, n, new_order=False)
    assert np.array_equal(w.ops, w2.ops) and np.array_equal(w.params, w2.params)
    w3 = g.tpcc_generate(W, n, new_order=False, seed=g.SEED + 1)
    assert not np.array_equal(w.ops, w3.ops)


def test_new_order_generator():
    W, n, rr = 4, 5000, 0.1
    w = g.tpcc_generate(W, n, new_order=True, remote_ratio=rr)
    lens = np.diff(w.tx_off.astype(np.int64))
    assert lens.min() == 6 and lens.max() == 16  # the district op and 5..15 lines
    s0 = 30011 * W
    remote_lines = lines = 0
    for t in range(n):
        o0, o1 = int(w.tx_off[t]), int(w.tx_off[t + 1])
        key, kind = g.op_key(w.ops[o0:o1]).astype(np.int64), g.op_col(w.ops[o0:o1])
        assert kind[0] == 3 and np.all(kind[1:] == 4)
        home = (key[0] - W) // 10
        sw, item = (key[1:] - s0) // 100000, (key[1:] - s0) % 100000
        assert np.all(np.diff(sw * 100000 + item) > 0)  # sorted by (supply warehouse, item), distinct
        qty, rem = w.params[o0 + 1:o1] & 0xFFFF, w.params[o0 + 1:o1] >> 16
        assert qty.min() >= 1 and qty.max() <= 10 and w.params[o0] == 0
        assert np.array_equal(rem.astype(bool), sw != home)
        remote_lines += int(rem.sum())
        lines += len(item)
        assert (int(w.tx_customer[t]) - 11 * W) // 3000 == key[0] - W  # customer of the ordering district
    assert abs(remote_lines / lines - rr) < 0.02


def test_new_order_totals_independent():
    """new_order's totals (tpcc.cu:340,371) against float32 arithmetic done
    here from the tables, the ops and a set of results."""
    W, n = 2, 50
    cells, price = g.tpcc_seed_tables(W)
    w = g.tpcc_generate(W, n, new_order=True)
    results = np.arange(w.nops, dtype=np.uint64) + np.uint64(3001)
    total, oid = g.tpcc_new_order_totals(W, cells, price, w, results)
    c = cells.reshape(-1, 5)
    for t in range(n):
        o0, o1 = int(w.tx_off[t]), int(w.tx_off[t + 1])
        drow = int(g.op_key(w.ops[o0:o0 + 1])[0])
        wh = (drow - W) // 10
        w_tax = fl(int(c[wh, 0]) >> 32)
        d_tax = fl(int(c[drow, 1]) & 0xFFFFFFFF)
        disc = fl(int(c[int(w.tx_customer[t]), 2]) >> 32)
        s = F(0)
        for i in range(o0 + 1, o1):
            item = (int(g.op_key(w.ops[i:i + 1])[0]) - 30011 * W) % 100000
            s = F(s + F(F(int(w.params[i]) & 0xFFFF) * price[item]))
        exp = F(F(s * F(F(1) - F(disc))) * F(F(F(1) + F(w_tax)) + F(d_tax)))
        assert bits(total[t]) == bits(exp), t
        assert oid[t] == results[o0]


def test_tpcc_argument_checks():
    with pytest.raises(g.GcctbError):
        g.tpcc_generate(0, 10, new_order=True)
    with pytest.raises(g.GcctbError):
        g.tpcc_generate(2, 10, new_order=True, remote_ratio=1.5)
    with pytest.raises(g.GcctbError):
        g.tpcc_seed_tables(0)
    for kw in [dict(cols=10), dict(row_ops=True), dict(debug_events=True)]:
        with pytest.raises(g.GcctbError) as e:
            g.Context("silo", rows=g.tpcc_rows(1), tpcc=True, **dict(dict(cols=5), **kw))
        assert e.value.code == -1
    with pytest.raises(g.GcctbError) as e:
        g.Context("slow_silo", rows=g.tpcc_rows(1), cols=5, tpcc=True)
    assert e.value.code == -1


@pytest.mark.parametrize("new_order", [False, True])
def test_product_check_against_oracle_replay(new_order):
    """gcctb_tpcc_check (the CLI's debug check) passes the oracle's serial
    execution and counts what a wrong order, a wrong result or a wrong cell
    changes."""
    W, n = 2, 400
    cells, _ = g.tpcc_seed_tables(W)
    w = g.tpcc_generate(W, n, new_order=new_order)
    rng = np.random.default_rng(5)
    order = np.zeros(3 * n, dtype=np.uint64)
    order[1::3] = rng.permutation(n)
    fin, res, ties = orc.tpcc_replay(g.tpcc_rows(W), cells, w.ops, w.params, w.tx_off, order)
    assert ties == 0
    assert g.tpcc_check(W, cells, fin, w, order, res) == 0
    bad_res = res.copy()
    bad_res[7] ^= np.uint64(1)
    assert g.tpcc_check(W, cells, fin, w, order, bad_res) == 1
    bad_fin = fin.copy()
    bad_fin[5 * 3 + 1] += np.uint64(1)
    assert g.tpcc_check(W, cells, bad_fin, w, order, res) == 1
    tied = order.copy()
    tied[1::3][1] = tied[1::3][0]
    assert g.tpcc_check(W, cells, fin, w, tied, res) >= 1
    # another order of conflicting transactions: the results no longer match
    swapped = order.copy()
    swapped[1::3] = order[1::3][::-1]
    assert g.tpcc_check(W, cells, fin, w, swapped, res) > 0
