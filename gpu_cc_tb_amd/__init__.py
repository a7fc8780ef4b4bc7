"""Python binding of the MI355X YCSB transaction-execution library (libgcctb.so).

A thin ctypes layer over the C ABI declared in include/gcctb.h, with numpy
buffers.  It mirrors the reference driver's knobs (cc_scheme, warp_density,
block_size, batch_size; benchmark_cpu/ycsb.cu:90-98).  There is no CPU
fallback: without the built library or without a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgcctb.so")

SCHEMES = ("tpl_nw", "tpl_wd", "to", "mvcc", "silo", "tictoc", "gacco", "gputx")
# the latch schemes (benchmark_cpu/ycsb.cu:170-190): plain and graph forms, column mode
LATCH_SCHEMES = ("slow_tpl_nw", "slow_tpl_wd", "slow_to", "slow_mvcc", "slow_silo", "slow_tictoc")
SEED = 141919810  # the reference's SEED (benchmark_cpu/ycsb.cu:27, generation/gen_ycsb.cu:12)

GCCTB_OK = 0


class GcctbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gcctb error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("cc_scheme", C.c_char_p),
        ("warp_density", C.c_int32),
        ("block_size", C.c_int32),
        ("batch_size", C.c_uint64),
        ("rows", C.c_uint64),
        ("cols", C.c_uint32),
        ("seed", C.c_uint64),
        ("backoff_ticks", C.c_uint32),
        ("persist_window", C.c_uint32),
        ("persist_adaptive", C.c_int32),
        ("row_ops", C.c_int32),
        ("debug_events", C.c_int32),
        ("tpcc", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("txs", C.c_uint64), ("committed", C.c_uint64), ("aborts", C.c_uint64),
        ("ts_ticks", C.c_uint64), ("wait_ticks", C.c_uint64), ("abort_ticks", C.c_uint64),
        ("manager_ticks", C.c_uint64), ("tot_ticks", C.c_uint64), ("index_ticks", C.c_uint64),
        ("tick_khz", C.c_double), ("pre_ms", C.c_double), ("proc_ms", C.c_double),
        ("tot_ms", C.c_double), ("exec_kernel_ms", C.c_double),
        ("batches", C.c_uint64), ("ranks", C.c_uint64), ("same_rank_conflicts", C.c_uint64),
        ("index_miss", C.c_uint64), ("error_flags", C.c_uint64), ("total_ops", C.c_uint64),
        ("read_ops", C.c_uint64),
        ("deferred", C.c_uint64),
        ("combined", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeConfig(C.Structure):
    _fields_ = [("max_launch_txs", C.c_uint64), ("pool_cap_bytes", C.c_uint64)]


class ServeStats(C.Structure):
    _fields_ = [
        ("txs", C.c_uint64), ("committed", C.c_uint64), ("aborts", C.c_uint64), ("launches", C.c_uint64),
        ("deferred", C.c_uint64), ("error_flags", C.c_uint64), ("tot_ms", C.c_double), ("busy_ms", C.c_double),
        ("txps", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
        ("mean_ms", C.c_double), ("mean_launch_txs", C.c_double), ("admitted_live", C.c_uint64),
        ("pool_nodes", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Workload(C.Structure):
    _fields_ = [
        ("rows", C.c_uint64), ("cols", C.c_uint32), ("ntx", C.c_uint64), ("tx_base", C.c_uint64),
        ("len_min", C.c_uint32), ("len_max", C.c_uint32), ("theta", C.c_double),
        ("read_ratio", C.c_double), ("seed", C.c_uint64), ("threads", C.c_int32),
    ]


class TpccWorkload(C.Structure):
    _fields_ = [("warehouses", C.c_uint32), ("new_order", C.c_int32), ("remote_ratio", C.c_double),
                ("ntx", C.c_uint64), ("seed", C.c_uint64)]


class StaticRun(C.Structure):
    """gcctb_static_run: the plain form over caller-owned device buffers."""
    _fields_ = [
        ("cc_scheme", C.c_char_p), ("warp_density", C.c_int32), ("block_size", C.c_int32),
        ("batch_size", C.c_uint64), ("cells", C.c_void_p), ("rows", C.c_uint64), ("cols", C.c_uint32),
        ("seed", C.c_uint64), ("backoff_ticks", C.c_uint32), ("ops", C.c_void_p), ("tx_off", C.c_void_p),
        ("ntx", C.c_uint64), ("nops", C.c_uint64), ("reads", C.c_void_p), ("order", C.c_void_p),
        ("stream", C.c_void_p),
    ]


STATIC_SCRATCH_BYTES = 24156672  # GCCTB_STATIC_SCRATCH_BYTES

# gcctb_event: a read (0) / write (1) of object oid by transaction tid, or its commit (2)
EVENT_DTYPE = np.dtype([("id", "<u8"), ("oid", "<u8"), ("tid", "<i4"), ("type", "<i4")])


_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_lib = None


def load_library() -> C.CDLL:
    """Load libgcctb.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GcctbError(-3, f"{LIB_PATH} is missing; build it with gpu_cc_tb_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "gcctb_create": (C.c_int, [C.POINTER(Config), C.POINTER(P)]),
        "gcctb_destroy": (None, [P]),
        "gcctb_last_error": (C.c_char_p, []),
        "gcctb_load_table": (C.c_int, [P, _u64p, C.c_uint64, C.c_uint32]),
        "gcctb_seed_table": (C.c_int, [P]),
        "gcctb_build_index_identity": (C.c_int, [P]),
        "gcctb_build_index": (C.c_int, [P, _u32p, _u32p, C.c_uint64]),
        "gcctb_probe": (C.c_int, [P, _u32p, C.c_uint64, _u32p]),
        "gcctb_load_ops": (C.c_int, [P, _u64p, _u64p, C.c_uint64]),
        "gcctb_check_ops": (C.c_int, [_u64p, _u64p, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]),
        "gcctb_run_plain": (C.c_int, [P, C.POINTER(Stats)]),
        "gcctb_run_graph": (C.c_int, [P, C.POINTER(Stats)]),
        "gcctb_run_persistent": (C.c_int, [P, C.POINTER(Stats)]),
        "gcctb_set_batches": (C.c_int, [P, _u64p, C.c_uint64]),
        "gcctb_serve": (C.c_int, [P, C.POINTER(ServeConfig), _u64p, P, C.c_uint64, P, C.POINTER(ServeStats)]),
        "gcctb_fetch": (C.c_int, [P, P, P, P, P]),
        "gcctb_fetch_ranks": (C.c_int, [P, P, P, C.POINTER(C.c_uint64)]),
        "gcctb_fetch_cc": (C.c_int, [P, _u64p]),
        "gcctb_generate": (C.c_int, [C.POINTER(Workload), _u64p, _u64p, C.POINTER(C.c_uint64)]),
        "gcctb_zeta": (C.c_double, [C.c_uint64, C.c_double]),
        "gcctb_read_txs_file": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.c_uint64, _u64p, _u64p,
                                          C.POINTER(C.c_uint64)]),
        "gcctb_write_txs_file": (C.c_int, [C.c_char_p, _u64p, _u64p, C.c_uint64]),
        "gcctb_cpu_seed_table": (C.c_int, [_u64p, C.c_uint64, C.c_uint32, C.c_uint64]),
        "gcctb_cpu_run": (C.c_int, [C.POINTER(Config), C.c_int32, _u64p, _u64p, _u64p, C.c_uint64, _u64p, _u64p,
                                    C.POINTER(Stats)]),
        "gcctb_check_host": (C.c_int, [C.POINTER(Config), _u64p, _u64p, _u64p, _u64p, C.c_uint64, _u64p, _u64p,
                                       C.POINTER(C.c_uint64)]),
        "gcctb_load_params": (C.c_int, [P, _u32p, C.c_uint64]),
        "gcctb_tpcc_rows": (C.c_uint64, [C.c_uint32]),
        "gcctb_tpcc_seed_tables": (C.c_int, [C.c_uint32, C.c_uint64, _u64p, P]),
        "gcctb_tpcc_generate": (C.c_int, [C.POINTER(TpccWorkload), _u64p, _u32p, _u64p, P, C.POINTER(C.c_uint64)]),
        "gcctb_tpcc_new_order_totals": (C.c_int, [C.c_uint32, _u64p, P, _u64p, _u32p, _u64p, C.c_uint64, _u64p,
                                                  _u64p, P, P]),
        "gcctb_tpcc_check": (C.c_int, [C.c_uint32, _u64p, _u64p, _u64p, _u32p, _u64p, C.c_uint64, _u64p, _u64p,
                                       C.POINTER(C.c_uint64)]),
        "gcctb_format_result": (C.c_int, [C.POINTER(Stats), C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                          C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint64]),
        "gcctb_device_info": (C.c_int, [C.c_char_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
        "gcctb_check": (C.c_int, [P, _u64p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "gcctb_microbench": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
        "gcctb_sort_pairs": (C.c_int, [P, P, C.c_uint64, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
        "gcctb_run_plain_static": (C.c_int, [C.POINTER(StaticRun), C.POINTER(Stats)]),
        "gcctb_conflict_check": (C.c_int, [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), P, C.c_uint64,
                                           C.POINTER(C.c_uint64), C.c_char_p, C.c_uint64]),
        "gcctb_fetch_events": (C.c_int, [P, P, C.c_uint64, C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != GCCTB_OK:
        raise GcctbError(rc, load_library().gcctb_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# host-only helpers (no GPU needed)
# ---------------------------------------------------------------------------
@dataclass
class Ops:
    """Transactions in CSR form: ops[tx_off[t]:tx_off[t+1]] belong to transaction t."""
    ops: np.ndarray
    tx_off: np.ndarray

    @property
    def ntx(self) -> int:
        return len(self.tx_off) - 1

    @property
    def nops(self) -> int:
        return int(self.tx_off[-1])


def generate(rows: int, ntx: int, *, theta: float, read_ratio: float, len_min: int = 16,
             len_max: int | None = None, cols: int = 10, seed: int = SEED, tx_base: int = 0,
             threads: int = 0) -> Ops:
    """Deterministic YCSB transactions (Zipf keys, distinct and sorted per transaction)."""
    lib = load_library()
    len_max = len_min if len_max is None else len_max
    w = Workload(rows, cols, ntx, tx_base, len_min, len_max, theta, read_ratio, seed, threads)
    ops = np.zeros(max(1, ntx * len_max), dtype=np.uint64)
    off = np.zeros(ntx + 1, dtype=np.uint64)
    n = C.c_uint64(0)
    _check(lib.gcctb_generate(C.byref(w), ops, off, C.byref(n)))
    return Ops(ops[: n.value].copy() if n.value else np.zeros(0, dtype=np.uint64), off)


class TpccOps(Ops):
    """TPC-C transactions: CSR ops of the TPC-C row space, one update
    parameter per op, and each transaction's customer row."""

    def __init__(self, ops, tx_off, params, tx_customer):
        super().__init__(ops, tx_off)
        self.params = params
        self.tx_customer = tx_customer


TPCC_COLS = 5


def tpcc_rows(warehouses: int) -> int:
    return int(load_library().gcctb_tpcc_rows(warehouses))


def tpcc_seed_tables(warehouses: int, seed: int = SEED) -> tuple:
    """(initial cells of the TPC-C row space, item prices)."""
    cells = np.zeros(tpcc_rows(warehouses) * TPCC_COLS, dtype=np.uint64)
    price = np.zeros(100000, dtype=np.float32)
    _check(load_library().gcctb_tpcc_seed_tables(warehouses, seed, cells, price.ctypes.data_as(C.c_void_p)))
    return cells, price


def tpcc_generate(warehouses: int, ntx: int, *, new_order: bool, remote_ratio: float = 0.1,
                  seed: int = SEED) -> TpccOps:
    """Deterministic payment or new_order transactions (generation/gen_tpcc.cu)."""
    per = 16 if new_order else 3
    ops = np.zeros(max(1, ntx * per), dtype=np.uint64)
    params = np.zeros(max(1, ntx * per), dtype=np.uint32)
    off = np.zeros(ntx + 1, dtype=np.uint64)
    cust = np.zeros(max(1, ntx), dtype=np.uint64)
    n = C.c_uint64(0)
    wl = TpccWorkload(warehouses, 1 if new_order else 0, remote_ratio, ntx, seed)
    _check(load_library().gcctb_tpcc_generate(C.byref(wl), ops, params, off, cust.ctypes.data_as(C.c_void_p),
                                               C.byref(n)))
    return TpccOps(ops[: n.value].copy(), off, params[: n.value].copy(), cust[:ntx].copy())


def tpcc_new_order_totals(warehouses: int, init_cells: np.ndarray, price: np.ndarray, w: TpccOps,
                          results: np.ndarray) -> tuple:
    """new_order's output: (total amount per order, o_id per order)."""
    total = np.zeros(max(1, w.ntx), dtype=np.float32)
    oid = np.zeros(max(1, w.ntx), dtype=np.uint32)
    _check(load_library().gcctb_tpcc_new_order_totals(
        warehouses, np.ascontiguousarray(init_cells), np.ascontiguousarray(price, dtype=np.float32).ctypes.data_as(
            C.c_void_p), np.ascontiguousarray(w.ops), np.ascontiguousarray(w.params), np.ascontiguousarray(w.tx_off),
        w.ntx, np.ascontiguousarray(w.tx_customer), np.ascontiguousarray(results, dtype=np.uint64),
        total.ctypes.data_as(C.c_void_p), oid.ctypes.data_as(C.c_void_p)))
    return total[: w.ntx], oid[: w.ntx]


def tpcc_check(warehouses: int, init_cells: np.ndarray, final_cells: np.ndarray, w: TpccOps, order: np.ndarray,
               results: np.ndarray) -> int:
    """Serial re-execution of a TPC-C run in its reported order: the count of
    op results and cells that differ, plus tied order keys (0: the run is
    the serial execution in that order)."""
    bad = C.c_uint64(0)
    a64 = lambda x: np.ascontiguousarray(x, dtype=np.uint64) if np.size(x) else np.zeros(1, np.uint64)
    _check(load_library().gcctb_tpcc_check(
        warehouses, a64(init_cells), a64(final_cells), a64(w.ops),
        np.ascontiguousarray(w.params, dtype=np.uint32) if w.params.size else np.zeros(1, np.uint32), a64(w.tx_off),
        w.ntx, a64(order), a64(results), C.byref(bad)))
    return bad.value


def check_ops(w: Ops, cols: int = 10) -> int:
    """The host-side input checks of load_ops (distinct keys per transaction,
    at most 32 ops, columns in range); returns the read count or raises."""
    ops = np.ascontiguousarray(w.ops, dtype=np.uint64)
    if ops.size == 0:
        ops = np.zeros(1, dtype=np.uint64)
    n = C.c_uint64(0)
    _check(load_library().gcctb_check_ops(ops, np.ascontiguousarray(w.tx_off, dtype=np.uint64), w.ntx, cols,
                                          C.byref(n)))
    return n.value


def zeta(n: int, theta: float) -> float:
    return load_library().gcctb_zeta(n, theta)


def read_txs(path: str, ntx: int, cols: int = 10, seed: int = SEED) -> Ops:
    lib = load_library()
    ops = np.zeros(max(1, 16 * ntx), dtype=np.uint64)
    off = np.zeros(ntx + 1, dtype=np.uint64)
    n = C.c_uint64(0)
    _check(lib.gcctb_read_txs_file(path.encode(), ntx, cols, seed, ops, off, C.byref(n)))
    return Ops(ops[: n.value].copy(), off)


def write_txs(path: str, w: Ops) -> None:
    _check(load_library().gcctb_write_txs_file(path.encode(), np.ascontiguousarray(w.ops),
                                               np.ascontiguousarray(w.tx_off), w.ntx))


def cpu_run(scheme: str, rows: int, w: "Ops", *, cols: int = 10, batch_size: int = 65536, threads: int = 0,
            row_ops: bool = False, seed: int = SEED, table: np.ndarray = None) -> tuple:
    """The CPU path (gcctb_cpu_run) on a seeded table, or on `table` (updated
    in place).  Returns (stats dict, {"table", "reads", "order"})."""
    lib = load_library()
    if table is None:
        table = np.zeros(rows * cols, dtype=np.uint64)
        _check(lib.gcctb_cpu_seed_table(table, rows, cols, seed))
    cfg = Config(cc_scheme=scheme.encode(), batch_size=batch_size, rows=rows, cols=cols, seed=seed,
                 row_ops=1 if row_ops else 0)
    reads = np.zeros(max(1, w.nops), dtype=np.uint64)
    order = np.zeros(max(1, 3 * w.ntx), dtype=np.uint64)
    st = Stats()
    _check(lib.gcctb_cpu_run(C.byref(cfg), threads, table, np.ascontiguousarray(w.ops), np.ascontiguousarray(w.tx_off),
                             w.ntx, reads, order, C.byref(st)))
    return st.as_dict(), {"table": table, "reads": reads[: w.nops], "order": order[: 3 * w.ntx]}


def check_host(rows: int, cols: int, w: "Ops", init: np.ndarray, res: dict, *, row_ops: bool = False,
               seed: int = SEED) -> int:
    """gcctb_check_host: reads and cells that differ from the serial replay."""
    cfg = Config(cc_scheme=b"tpl_nw", batch_size=1, rows=rows, cols=cols, seed=seed, row_ops=1 if row_ops else 0)
    bad = C.c_uint64(0)
    _check(load_library().gcctb_check_host(C.byref(cfg), np.ascontiguousarray(init), res["table"],
                                           np.ascontiguousarray(w.ops), np.ascontiguousarray(w.tx_off), w.ntx,
                                           res["order"], np.ascontiguousarray(res["reads"]), C.byref(bad)))
    return bad.value


def format_result(stats: dict, cc_type: str, warp_density: int, block_size: int, batch_size: int,
                  txcnt: int) -> tuple:
    """The driver's result block (ExecInfo::Analyse): (CSV line, stdout text)."""
    st = Stats(**{k: stats[k] for k, _ in Stats._fields_ if k in stats})
    csv, text = C.create_string_buffer(4096), C.create_string_buffer(8192)
    _check(load_library().gcctb_format_result(C.byref(st), cc_type.encode(), warp_density, block_size, batch_size,
                                              txcnt, csv, len(csv), text, len(text)))
    return csv.value.decode(), text.value.decode()


def op_key(ops: np.ndarray) -> np.ndarray:
    return (ops & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def op_col(ops: np.ndarray) -> np.ndarray:
    return ((ops >> np.uint64(32)) & np.uint64(0xFF)).astype(np.uint32)


def op_is_write(ops: np.ndarray) -> np.ndarray:
    return ((ops >> np.uint64(40)) & np.uint64(1)).astype(bool)


def make_ops(keys, cols, writes) -> np.ndarray:
    keys = np.asarray(keys, dtype=np.uint64)
    cols = np.asarray(cols, dtype=np.uint64)
    writes = np.asarray(writes, dtype=np.uint64)
    return keys | (cols << np.uint64(32)) | (writes << np.uint64(40))


def device_info() -> dict:
    lib = load_library()
    name = C.create_string_buffer(256)
    khz = C.c_double(0)
    cus = C.c_int32(0)
    _check(lib.gcctb_device_info(name, 256, C.byref(khz), C.byref(cus)))
    return {"name": name.value.decode(), "wall_clock_khz": khz.value, "cus": cus.value}


MICROBENCH = {"cas_random": 0, "add_random": 1, "add_same": 2, "load_coherent": 3, "load_cached": 4, "copy": 5,
              "load_pair_one_line": 6, "load_pair_two_lines": 7, "ts_read_one_line": 8, "ts_read_two_lines": 9,
              "add_hot16_one_line": 10, "add_hot16_16_lines": 11, "add_polled_same_line": 12,
              "add_polled_other_lines": 13}


def microbench(kind: str, n_words: int, threads: int = 1 << 20, iters: int = 16) -> dict:
    """Memory-system ceiling: rate in ops/s (bytes/s for 'copy')."""
    lib = load_library()
    rate, ms = C.c_double(0), C.c_double(0)
    _check(lib.gcctb_microbench(MICROBENCH[kind], n_words, threads, iters, C.byref(rate), C.byref(ms)))
    return {"kind": kind, "n_words": n_words, "threads": threads, "iters": iters, "rate": rate.value,
            "ms": ms.value}


def sort_pairs(keys: np.ndarray, vals: np.ndarray, key_bits: int | None = None) -> tuple[np.ndarray, np.ndarray, float]:
    """Stable device sort of (key, value) pairs by the low `key_bits` bits of
    the key (all of them by default): the sort behind the GaccO and GPUTx
    preparation.  Returns sorted copies and the device time in ms."""
    keys = np.ascontiguousarray(keys)
    if keys.dtype not in (np.uint32, np.uint64):
        raise TypeError("keys must be uint32 or uint64")
    k = keys.copy()
    v = np.ascontiguousarray(vals, dtype=np.uint32).copy()
    if v.shape != k.shape or k.ndim != 1:
        raise ValueError("keys and vals must be 1-D and of one length")
    nbits = k.itemsize * 8 if key_bits is None else key_bits
    lib = load_library()
    ms = C.c_double(0)
    _check(lib.gcctb_sort_pairs(k.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), k.size, k.itemsize,
                                nbits, C.byref(ms)))
    return k, v, ms.value


# ---------------------------------------------------------------------------
# GPU context
# ---------------------------------------------------------------------------
class Context:
    """One scheme instance on the GPU (table, index, CC state, transactions)."""

    def __init__(self, cc_scheme: str, *, rows: int, cols: int = 10, warp_density: int = 5,
                 block_size: int = 32, batch_size: int = 65536, seed: int = SEED,
                 backoff_ticks: int = 20, persist_window: int = 0, persist_adaptive: bool = False,
                 row_ops: bool = False, debug_events: bool = False, tpcc: bool = False):
        self._lib = load_library()
        self.scheme = cc_scheme
        self.rows, self.cols, self.seed, self.row_ops = rows, cols, seed, row_ops
        cfg = Config(cc_scheme.encode(), warp_density, block_size, batch_size, rows, cols, seed,
                     backoff_ticks, persist_window, 1 if persist_adaptive else 0, 1 if row_ops else 0,
                     1 if debug_events else 0, 1 if tpcc else 0)
        h = C.c_void_p()
        _check(self._lib.gcctb_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.ntx = 0
        self.nops = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.gcctb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed_table(self) -> None:
        _check(self._lib.gcctb_seed_table(self._h))

    def load_table(self, cells: np.ndarray) -> None:
        cells = np.ascontiguousarray(cells, dtype=np.uint64).reshape(-1)
        _check(self._lib.gcctb_load_table(self._h, cells, self.rows, self.cols))

    def build_index(self, keys: np.ndarray, vals: np.ndarray) -> None:
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        vals = np.ascontiguousarray(vals, dtype=np.uint32)
        _check(self._lib.gcctb_build_index(self._h, keys, vals, len(keys)))

    def probe(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.zeros(max(1, len(keys)), dtype=np.uint32)
        _check(self._lib.gcctb_probe(self._h, keys, len(keys), out))
        return out[: len(keys)]

    def load_ops(self, w: Ops) -> None:
        ops = np.ascontiguousarray(w.ops, dtype=np.uint64)
        if ops.size == 0:
            ops = np.zeros(1, dtype=np.uint64)
        _check(self._lib.gcctb_load_ops(self._h, ops, np.ascontiguousarray(w.tx_off, dtype=np.uint64), w.ntx))
        self.ntx, self.nops = w.ntx, w.nops

    def load_params(self, params: np.ndarray) -> None:
        """TPC-C: one update parameter per loaded op."""
        p = np.ascontiguousarray(params, dtype=np.uint32)
        _check(self._lib.gcctb_load_params(self._h, p if p.size else np.zeros(1, dtype=np.uint32), p.size))

    def run(self, form: str = "plain", check: bool = True) -> dict:
        """form: 'plain' (per-batch launches), 'graph' (captured batch sequence)
        or 'persistent' (one launch over the whole stream)."""
        st = Stats()
        fn = {"plain": self._lib.gcctb_run_plain, "graph": self._lib.gcctb_run_graph,
              "persistent": self._lib.gcctb_run_persistent}[form]
        rc = fn(self._h, C.byref(st))
        if check:
            _check(rc)
        d = st.as_dict()
        d["rc"] = rc
        return d

    def run_plain(self, check: bool = True) -> dict:
        return self.run("plain", check)

    def serve(self, batch_sizes, arrival_ms=None, *, max_launch_txs: int = 0, pool_cap_bytes: int = 0,
              latencies: bool = False, check: bool = True) -> dict:
        """Serving form: the loaded transactions arrive as batches (sizes in
        order) at arrival_ms after the call starts (None: all at once).
        Returns the serving statistics; with latencies=True also the
        per-transaction commit latency (ms) under "latency_ms"."""
        sizes = np.ascontiguousarray(batch_sizes, dtype=np.uint64)
        arr = None if arrival_ms is None else np.ascontiguousarray(arrival_ms, dtype=np.float64)
        if arr is not None and arr.size != sizes.size:
            raise ValueError("one arrival time per batch")
        lat = np.zeros(max(1, self.ntx), dtype=np.float64) if latencies else None
        cfg = ServeConfig(max_launch_txs, pool_cap_bytes)
        st = ServeStats()
        rc = self._lib.gcctb_serve(self._h, C.byref(cfg), sizes if sizes.size else np.zeros(1, np.uint64),
                                   _ptr(arr), sizes.size, _ptr(lat), C.byref(st))
        if check:
            _check(rc)
        d = st.as_dict()
        d["rc"] = rc
        if latencies:
            d["latency_ms"] = lat[: self.ntx]
        return d

    def set_batches(self, sizes) -> None:
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        _check(self._lib.gcctb_set_batches(self._h, sizes if sizes.size else np.zeros(1, np.uint64), sizes.size))

    def fetch(self, table: bool = True, reads: bool = True, order: bool = True, aborts: bool = False):
        t = np.zeros(self.rows * self.cols, dtype=np.uint64) if table else None
        r = np.zeros(max(1, self.nops), dtype=np.uint64) if reads else None
        o = np.zeros(max(1, 3 * self.ntx), dtype=np.uint64) if order else None
        a = np.zeros(max(1, self.ntx), dtype=np.uint32) if aborts else None
        _check(self._lib.gcctb_fetch(self._h, _ptr(t), _ptr(r), _ptr(o), _ptr(a)))
        return {
            "table": t,
            "reads": None if r is None else r[: self.nops],
            "order": None if o is None else o[: 3 * self.ntx],
            "aborts": None if a is None else a[: self.ntx],
        }

    def fetch_events(self) -> np.ndarray:
        """Event log of the last plain / graph run (debug_events contexts)."""
        out = np.zeros(max(1, self.nops + self.ntx), dtype=EVENT_DTYPE)
        n = C.c_uint64(0)
        _check(self._lib.gcctb_fetch_events(self._h, out.ctypes.data_as(C.c_void_p), out.size, C.byref(n)))
        return out[: n.value]

    def fetch_ranks(self):
        rank = np.zeros(max(1, self.ntx), dtype=np.uint32)
        n = C.c_uint64(0)
        _check(self._lib.gcctb_fetch_ranks(self._h, _ptr(rank), None, C.byref(n)))
        counts = np.zeros(max(1, n.value), dtype=np.uint32)
        _check(self._lib.gcctb_fetch_ranks(self._h, None, _ptr(counts), C.byref(n)))
        return rank[: self.ntx], counts[: n.value]

    def check(self, init_table: np.ndarray) -> tuple[int, int]:
        """Debug-mode serializability check of the last run: (bad, uncommitted)."""
        bad, unc = C.c_uint64(0), C.c_uint64(0)
        _check(self._lib.gcctb_check(self._h, np.ascontiguousarray(init_table, dtype=np.uint64), C.byref(bad),
                                     C.byref(unc)))
        return bad.value, unc.value

    def fetch_cc(self) -> np.ndarray:
        w = np.zeros(self.rows, dtype=np.uint64)
        _check(self._lib.gcctb_fetch_cc(self._h, w))
        return w


def run_plain_static(cc_scheme: str, *, cells: int, rows: int, cols: int, ops: int, tx_off: int, ntx: int,
                     nops: int, reads: int, order: int, warp_density: int = 3, block_size: int = 8,
                     batch_size: int = 1 << 30, seed: int = SEED, backoff_ticks: int = 20, stream: int = 0,
                     check: bool = True) -> dict:
    """gcctb_run_plain_static over device pointers (integers, e.g. a torch
    tensor's data_ptr()): the plain form that allocates nothing.  Returns the
    stats as a dict with the return code under "rc"."""
    lib = load_library()
    r = StaticRun(cc_scheme.encode(), warp_density, block_size, batch_size, cells, rows, cols, seed, backoff_ticks,
                  ops, tx_off, ntx, nops, reads, order, stream or None)
    st = Stats()
    rc = lib.gcctb_run_plain_static(C.byref(r), C.byref(st))
    if check:
        _check(rc)
    d = st.as_dict()
    d["rc"] = rc
    return d


def conflict_check(events: np.ndarray, ntx: int, report_len: int = 1 << 20) -> dict:
    """gcctb_conflict_check on the host: {"uncommitted", "cycle" (transaction
    ids of the first cycle found, empty when acyclic), "report" (the text the
    reference's debug check prints)}."""
    lib = load_library()
    ev = np.ascontiguousarray(events, dtype=EVENT_DTYPE)
    unc, clen = C.c_uint64(0), C.c_uint64(0)
    cyc = np.zeros(max(1, ntx), dtype=np.uint64)
    rep = C.create_string_buffer(report_len)
    _check(lib.gcctb_conflict_check(ev.ctypes.data_as(C.c_void_p) if ev.size else None, ev.size, ntx, C.byref(unc),
                                    cyc.ctypes.data_as(C.c_void_p), cyc.size, C.byref(clen), rep, report_len))
    return {"uncommitted": unc.value, "cycle": cyc[: min(clen.value, cyc.size)].tolist(),
            "report": rep.value.decode()}
