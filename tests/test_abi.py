"""The C-ABI library loads and exports every entry point of include/gcctb.h;
host-only entry points work without a GPU; GPU entry points fail loudly."""
import os
import re
import tempfile

import numpy as np
import pytest

import gpu_cc_tb_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EINVAL = -1  # GCCTB_EINVAL


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gcctb.h")).read()
    return sorted(set(re.findall(r"\b(gcctb_[a-z_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = g.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s


def test_bad_config_is_rejected_before_device_use():
    for kw in [dict(warp_density=6), dict(block_size=33), dict(block_size=0), dict(batch_size=0),
               dict(persist_adaptive=True)]:
        with pytest.raises(g.GcctbError) as e:
            g.Context("tpl_nw", rows=16, **kw)
        assert e.value.code == -1
    with pytest.raises(g.GcctbError) as e:
        g.Context("no_such_scheme", rows=16)
    assert e.value.code == -1


def test_no_cpu_fallback_without_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(g.GcctbError) as e:
        g.Context("tpl_nw", rows=16)
    assert e.value.code == -3


def test_txs_file_round_trip():
    w = g.generate(1000, 300, theta=0.8, read_ratio=0.5, len_min=1, len_max=16)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.txs")
        g.write_txs(p, w)
        assert os.path.getsize(p) == 136 * 300
        r = g.read_txs(p, 300)
    assert np.array_equal(r.tx_off, w.tx_off)
    assert np.array_equal(g.op_key(r.ops), g.op_key(w.ops))
    assert np.array_equal(g.op_is_write(r.ops), g.op_is_write(w.ops))
    assert np.all(g.op_col(r.ops) < 10)


def test_generator_rejects_bad_input():
    for kw in [dict(theta=1.0), dict(theta=-0.1), dict(read_ratio=1.5), dict(len_min=0), dict(len_min=33)]:
        args = dict(theta=0.5, read_ratio=0.5)
        args.update(kw)
        with pytest.raises(g.GcctbError):
            g.generate(100, 10, **args)
    with pytest.raises(g.GcctbError):
        g.generate(4, 10, theta=0.5, read_ratio=0.5, len_min=8)


def test_sort_pairs_arguments():
    """Argument checks run before any device call; an empty input is a no-op."""
    import ctypes as C
    lib = g.load_library()
    k = np.zeros(4, dtype=np.uint32)
    v = np.zeros(4, dtype=np.uint32)
    kp, vp = k.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p)
    assert lib.gcctb_sort_pairs(kp, vp, 4, 2, 8, None) == EINVAL
    assert lib.gcctb_sort_pairs(kp, vp, 4, 4, 33, None) == EINVAL
    assert lib.gcctb_sort_pairs(None, vp, 4, 4, 8, None) == EINVAL
    assert lib.gcctb_sort_pairs(kp, vp, 0, 4, 8, None) == 0


def test_check_ops_rejects_repeated_keys():
    """load_ops' host checks (gcctb_check_ops): a transaction that names one
    key twice, a column past the table, or more than 32 ops is refused with
    GCCTB_EINVAL before anything reaches the device."""
    w = g.generate(1000, 50, theta=0.5, read_ratio=0.5, len_min=8, len_max=8)
    assert g.check_ops(w) == int(np.count_nonzero(~g.op_is_write(w.ops)))
    dup = w.ops.copy()
    dup[8 * 7 + 3] = g.make_ops([int(g.op_key(dup[8 * 7 + 5:8 * 7 + 6])[0])], [0], [False])[0]
    with pytest.raises(g.GcctbError) as e:
        g.check_ops(g.Ops(dup, w.tx_off))
    assert e.value.code == EINVAL and "twice" in str(e.value) and "transaction 7" in str(e.value)
    with pytest.raises(g.GcctbError) as e:
        g.check_ops(w, cols=int(g.op_col(w.ops).max()))
    assert e.value.code == EINVAL
    ops = g.make_ops(np.arange(33), np.zeros(33), np.zeros(33, dtype=bool))
    with pytest.raises(g.GcctbError) as e:
        g.check_ops(g.Ops(ops, np.array([0, 33], dtype=np.uint64)))
    assert e.value.code == EINVAL
    # the same key in two different transactions is fine
    two = g.make_ops([5, 5], [1, 2], [True, False])
    assert g.check_ops(g.Ops(two, np.array([0, 1, 2], dtype=np.uint64))) == 1


def test_static_entry_point_checks_before_device_use():
    """gcctb_run_plain_static rejects sizes beyond its static array and the
    schemes it does not run before it touches a device."""
    kw = dict(cells=1, cols=10, ops=1, tx_off=1, ntx=1, nops=1, reads=1, order=1)
    for scheme, rows, extra in [("silo", (1 << 20) + 1, {}), ("gacco", 16, {}), ("gputx", 16, {}),
                                ("tpl_nw", 16, {"ntx": (1 << 17) + 1}), ("to", 16, {"nops": (1 << 20) + 1}),
                                ("mvcc", 16, {"warp_density": 6}), ("nope", 16, {})]:
        args = dict(kw, **extra)
        st = g.run_plain_static(scheme, rows=rows, check=False, **args)
        assert st["rc"] == EINVAL, (scheme, rows, extra)
    if not os.path.exists("/dev/kfd"):
        st = g.run_plain_static("silo", rows=16, check=False, **kw)
        assert st["rc"] == -3  # GCCTB_EDEVICE: no GPU, no fallback


def _events(rows):
    return np.array([(i + 1, oid, tid, typ) for i, (tid, typ, oid) in enumerate(rows)], dtype=g.EVENT_DTYPE)


def test_conflict_check_serial_history_is_good():
    # T0 writes x then commits; T1 reads x and writes y; T2 reads y; commits in order
    ev = _events([(0, 1, 5), (0, 2, 0), (1, 0, 5), (1, 1, 7), (1, 2, 0), (2, 0, 7), (2, 2, 0)])
    r = g.conflict_check(ev, 3)
    assert r["cycle"] == [] and r["uncommitted"] == 0
    assert r["report"].startswith("SCAN OK\n") and r["report"].endswith("GOOD\n")


def test_conflict_check_reports_injected_cycle():
    """A non-serializable interleaving (src/runtime.cu:23-127): T0 and T1 both
    write x and y while the other is still active, in opposite orders, so
    T1 -> T0 on x and T0 -> T1 on y; the check prints BAD with the cycle's
    transactions and their events."""
    ev = _events([(0, 1, 10),   # T0 writes x
                  (1, 1, 20),   # T1 writes y
                  (1, 1, 10),   # T1 writes x while T0 is active: T1 -> T0
                  (0, 1, 20),   # T0 writes y while T1 is active: T0 -> T1
                  (0, 2, 0), (1, 2, 0),
                  (2, 0, 30), (2, 2, 0)])
    r = g.conflict_check(ev, 3)
    assert sorted(r["cycle"]) == [0, 1]
    rep = r["report"]
    assert "!!!!!!!!!" in rep and "BAD TX 0" in rep and "BAD TX 1" in rep and rep.endswith("BAD\n")
    assert "BAD ST" in rep and "tid 0 oid10 w" in rep


def test_conflict_check_lists_uncommitted_and_read_edges():
    # T1 reads x while T0 (active) wrote it: T1 -> T0 only (no cycle); T2 never commits
    ev = _events([(0, 1, 4), (1, 0, 4), (0, 2, 0), (1, 2, 0), (2, 0, 9)])
    r = g.conflict_check(ev, 3)
    assert r["uncommitted"] == 1 and "UNCOMMITTED 2\n" in r["report"] and r["cycle"] == []
    # read-read never conflicts; a read after an active writer does, in a cycle with a write-after-read
    ev = _events([(0, 0, 1), (1, 0, 1), (0, 2, 0), (1, 2, 0)])
    assert g.conflict_check(ev, 2)["cycle"] == []
    ev = _events([(0, 1, 1), (1, 0, 1), (1, 0, 2), (0, 1, 2), (0, 2, 0), (1, 2, 0)])
    assert sorted(g.conflict_check(ev, 2)["cycle"]) == [0, 1]


def test_conflict_check_long_chain_no_recursion():
    # 200,000 transactions in a chain of read-after-write edges: the search is iterative
    n = 200000
    rows = []
    for t in range(n):
        rows.append((t, 1, t))          # T_t writes object t
        if t:
            rows.append((t, 0, t - 1))  # reads object t-1 while T_{t-1} is active
    for t in range(n):
        rows.append((t, 2, 0))
    r = g.conflict_check(_events(rows), n, report_len=4096)
    assert r["cycle"] == [] and r["uncommitted"] == 0


def test_output_line_golden():
    """The driver's result block (ExecInfo::Analyse, include/runtime.cuh:169-214)
    from fixed statistics, against the text in tests/golden/output_lines.json
    (one case worked by hand, the rest from an independent float32
    restatement of the reference's arithmetic and iostream formatting)."""
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "output_lines.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 5
    for c in cases:
        st = dict(c["stats"])
        csv, text = g.format_result(st, c["cc_type"], c["warp_density"], c["block_size"], c["batch_size"],
                                    c["txcnt"])
        assert csv == c["csv"], (c["cc_type"], csv)
        assert text == c["text"], (c["cc_type"], text)
        assert len(csv.strip().split(",")) == 15


def test_txs_reader_hand_packed_records():
    """Records packed here byte by byte in the reference's YCSBTx layout
    (include/ycsb.cuh:21-33: u64 request_cnt, then 16 x {i32 read, u32 key}),
    independent of the project's writer: counts, keys and read flags come
    back in order; request_cnt 0 and 16 are both legal."""
    import struct
    recs = [(3, [(1, 7), (0, 4000000000), (1, 0)]), (0, []), (16, [(k % 2, 100 + k) for k in range(16)]),
            (1, [(0, 42)])]
    blob = b""
    for cnt, reqs in recs:
        body = b"".join(struct.pack("<iI", rd, key) for rd, key in reqs)
        blob += struct.pack("<Q", cnt) + body + b"\0" * (8 * (16 - len(reqs)))
    assert len(blob) == 136 * len(recs)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "hand.txs")
        with open(p, "wb") as f:
            f.write(blob)
        r = g.read_txs(p, len(recs), cols=13)
        with pytest.raises(g.GcctbError):
            g.read_txs(p, len(recs) + 1, cols=13)  # short file
        bad = bytearray(blob)
        bad[136:144] = struct.pack("<Q", 17)
        with open(p, "wb") as f:
            f.write(bytes(bad))
        with pytest.raises(g.GcctbError):
            g.read_txs(p, len(recs), cols=13)  # request_cnt above 16
    assert r.tx_off.tolist() == [0, 3, 3, 19, 20]
    keys = [key for _, reqs in recs for _, key in reqs]
    writes = [rd == 0 for _, reqs in recs for rd, _ in reqs]
    assert g.op_key(r.ops).tolist() == keys
    assert g.op_is_write(r.ops).tolist() == writes
    assert np.all(g.op_col(r.ops) < 13)


def test_latch_scheme_names_checked_before_device_use():
    """slow_* names are the reference's latch schemes (benchmark_cpu/ycsb.cu:
    170-190); there is no slow_gacco / slow_gputx, and the latch schemes run
    in column mode without the event log."""
    for name in ["slow_gacco", "slow_gputx", "slow_", "slow_slow_silo"]:
        with pytest.raises(g.GcctbError) as e:
            g.Context(name, rows=16)
        assert e.value.code == -1, name
    for kw in [dict(row_ops=True), dict(debug_events=True)]:
        with pytest.raises(g.GcctbError) as e:
            g.Context("slow_silo", rows=16, **kw)
        assert e.value.code == -1
