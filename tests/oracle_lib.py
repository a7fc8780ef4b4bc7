"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class CpuStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("aborts", C.c_uint64), ("committed", C.c_uint64),
                ("threads", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = C.CDLL(ORACLE_SO)
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_hash3.restype = C.c_uint64
        L.orc_hash3.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_cell_init.restype = C.c_uint64
        L.orc_cell_init.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_write_value.restype = C.c_uint64
        L.orc_write_value.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_seed_table.restype = None
        L.orc_seed_table.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _u64p]
        L.orc_zeta.restype = C.c_double
        L.orc_zeta.argtypes = [C.c_uint64, C.c_double]
        L.orc_zipf_rank.restype = C.c_uint64
        L.orc_zipf_rank.argtypes = [C.c_uint64, C.c_double, C.c_double, C.c_double]
        L.orc_gen_tx.restype = C.c_uint32
        L.orc_gen_tx.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                 C.c_uint64, C.c_double, C.c_uint64, _u64p]
        L.orc_find.restype = C.c_uint64
        L.orc_find.argtypes = [_u32p, _u64p, C.c_uint32, C.c_uint32]
        L.orc_replay.restype = C.c_int64
        L.orc_replay.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p, _u64p, C.c_uint64, _u64p,
                                 _u64p]
        L.orc_replay_mt.restype = C.c_int64
        L.orc_replay_mt.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p, _u64p, C.c_uint64, _u64p,
                                    _u64p, C.c_int]
        L.orc_replay_rows.restype = C.c_int64
        L.orc_replay_rows.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p, _u64p, C.c_uint64, _u64p,
                                      _u64p, C.c_int]
        L.orc_write_cell.restype = C.c_uint64
        L.orc_write_cell.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_row_digest.restype = C.c_uint64
        L.orc_row_digest.argtypes = [_u64p, C.c_uint32]
        L.orc_replay_txid.restype = None
        L.orc_replay_txid.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p, _u64p, C.c_uint64,
                                      _u64p]
        L.orc_gputx_rank.restype = C.c_int
        L.orc_gputx_rank.argtypes = [C.c_uint64, _u64p, _u64p, C.c_uint64, _u32p, _u32p,
                                     C.POINTER(C.c_uint64)]
        L.orc_cpu_run.restype = C.c_int
        L.orc_cpu_run.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p, _u64p, C.c_uint64,
                                  C.c_uint64, C.c_int, _u64p, _u64p, C.POINTER(CpuStats)]
        L.orc_tpcc_apply.restype = C.c_uint64
        L.orc_tpcc_apply.argtypes = [C.c_uint32, C.c_uint32, _u64p]
        L.orc_tpcc_replay.restype = C.c_int64
        L.orc_tpcc_replay.argtypes = [C.c_uint64, _u64p, _u64p, _u32p, _u64p, C.c_uint64, _u64p, _u64p]
        _lib = L
    return _lib


def _a64(x):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    return x if x.size else np.zeros(1, dtype=np.uint64)


def seed_table(seed: int, rows: int, cols: int) -> np.ndarray:
    t = np.zeros(rows * cols, dtype=np.uint64)
    lib().orc_seed_table(seed, rows, cols, t)
    return t


def replay(rows, cols, seed, table, ops, tx_off, order, threads=1, in_place=False):
    """Serial replay in (batch, key0, key1) order -> (final table, reads, ties).
    threads > 1 splits the replay over threads by blocks of rows
    (orc_replay_mt, same result); in_place updates `table` itself."""
    t = table if in_place else np.array(table, dtype=np.uint64, copy=True)
    assert t.dtype == np.uint64 and t.flags.c_contiguous
    reads = np.zeros(max(1, len(ops)), dtype=np.uint64)
    ntx = len(tx_off) - 1
    if threads > 1:
        ties = lib().orc_replay_mt(rows, cols, seed, t, _a64(ops), _a64(tx_off), ntx, _a64(order), reads, threads)
    else:
        ties = lib().orc_replay(rows, cols, seed, t, _a64(ops), _a64(tx_off), ntx, _a64(order), reads)
    assert ties >= 0, "oracle allocation failed"
    return t, reads[: len(ops)], ties


def replay_rows(rows, cols, seed, table, ops, tx_off, order, threads=1):
    """Serial replay in whole-row mode (every op moves its whole row; reads
    return row digests) -> (final table, reads, ties)."""
    t = np.array(table, dtype=np.uint64, copy=True)
    reads = np.zeros(max(1, len(ops)), dtype=np.uint64)
    ties = lib().orc_replay_rows(rows, cols, seed, t, _a64(ops), _a64(tx_off), len(tx_off) - 1, _a64(order), reads,
                                 threads)
    assert ties >= 0, "oracle allocation failed"
    return t, reads[: len(ops)], ties


def row_digest(row) -> int:
    r = _a64(row)
    return lib().orc_row_digest(r, len(row))


def replay_txid(rows, cols, seed, table, ops, tx_off):
    t = np.array(table, dtype=np.uint64, copy=True)
    reads = np.zeros(max(1, len(ops)), dtype=np.uint64)
    lib().orc_replay_txid(rows, cols, seed, t, _a64(ops), _a64(tx_off), len(tx_off) - 1, reads)
    return t, reads[: len(ops)]


def gputx_rank(rows, ops, tx_off):
    ntx = len(tx_off) - 1
    rank = np.zeros(max(1, ntx), dtype=np.uint32)
    counts = np.zeros(ntx + 2, dtype=np.uint32)
    n = C.c_uint64(0)
    rc = lib().orc_gputx_rank(rows, _a64(ops), _a64(tx_off), ntx, rank, counts, C.byref(n))
    assert rc == 0
    return rank[:ntx], counts[: n.value]


def gen_tx(rows, cols, len_min, len_max, theta, read_ratio, seed, tx, zetan=None):
    if zetan is None:
        zetan = lib().orc_zeta(rows, theta)
    out = np.zeros(64, dtype=np.uint64)
    n = lib().orc_gen_tx(rows, cols, len_min, len_max, theta, read_ratio, seed, zetan, tx, out)
    return out[:n]


def find(keys, vals, key):
    return lib().orc_find(np.ascontiguousarray(keys, dtype=np.uint32), _a64(vals), len(keys), key)


def cpu_run(scheme: int, rows, cols, seed, table, ops, tx_off, batch_size, threads=0):
    t = np.array(table, dtype=np.uint64, copy=True)
    ntx = len(tx_off) - 1
    order = np.zeros(max(1, 3 * ntx), dtype=np.uint64)
    reads = np.zeros(max(1, len(ops)), dtype=np.uint64)
    st = CpuStats()
    rc = lib().orc_cpu_run(scheme, rows, cols, seed, t, _a64(ops), _a64(tx_off), ntx, batch_size, threads,
                           order, reads, C.byref(st))
    return rc, t, reads[: len(ops)], order[: 3 * ntx], st


def tpcc_apply(kind: int, param: int, row) -> tuple:
    """One TPC-C op on a 5-cell row: (result, updated row)."""
    r = np.array(row, dtype=np.uint64).copy()
    res = lib().orc_tpcc_apply(kind, param, r)
    return int(res), r


def tpcc_replay(rows, table, ops, params, tx_off, order):
    """Serial TPC-C replay in (batch, key0, key1) order -> (final table, results, ties)."""
    t = np.array(table, dtype=np.uint64, copy=True)
    res = np.zeros(max(1, len(ops)), dtype=np.uint64)
    p = np.ascontiguousarray(params, dtype=np.uint32)
    ties = lib().orc_tpcc_replay(rows, t, _a64(ops), p if p.size else np.zeros(1, dtype=np.uint32), _a64(tx_off),
                                 len(tx_off) - 1, _a64(order), res)
    assert ties >= 0
    return t, res[: len(ops)], ties
