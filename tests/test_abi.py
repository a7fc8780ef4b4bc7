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
        g.Context("slow_silo", rows=16)
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
