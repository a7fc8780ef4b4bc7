"""gcctb_run_plain_static, the plain form that allocates nothing, against
gcctb_run_plain (the entry point it is a variant of), through the C ABI.

Device buffers come from the HIP runtime the library itself uses (hipMalloc
and hipMemcpy through ctypes: test plumbing).  Two calls in a row on one GPU,
the second with a smaller input: each must pass the serial-replay oracle bit
for bit, as gcctb_run_plain does on the same input, and on a conflict-free
input (every row touched once, so no interleaving can change a result) both
entry points must return the same table and reads bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc

SCHEMES6 = ["tpl_nw", "tpl_wd", "to", "mvcc", "silo", "tictoc"]
COLS = 10
# (rows, ntx, len_min, len_max, theta, read ratio, batch, warp_density, block_size)
CALLS = [(4096, 3000, 8, 16, 0.8, 0.7, 1024, 3, 8), (1024, 700, 4, 8, 0.5, 0.5, 256, 2, 4)]
H2D, D2H = 1, 2  # hipMemcpyKind


def _hip():
    """The HIP runtime libgcctb.so is linked against (already loaded)."""
    g.load_library()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert paths, "libamdhip64 is not loaded"
    return C.CDLL(sorted(paths)[0])


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.n, self.p = hip, max(8, nbytes), C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(self.n)) == 0

    @property
    def ptr(self):
        return self.p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, count, dtype=np.uint64):
        a = np.zeros(count, dtype=dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def run_static(scheme, rows, w, batch, wd, bs):
    hip = _hip()
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    init = orc.seed_table(g.SEED, rows, COLS)
    bufs = [DevBuf(hip, init.nbytes), DevBuf(hip, w.ops.nbytes), DevBuf(hip, w.tx_off.nbytes),
            DevBuf(hip, 8 * w.nops), DevBuf(hip, 16 * w.ntx)]
    cells, ops, off, reads, order = bufs
    try:
        cells.put(init)
        ops.put(w.ops)
        off.put(w.tx_off)
        reads.put(np.zeros(w.nops, np.uint64))
        st = g.run_plain_static(scheme, cells=cells.ptr, rows=rows, cols=COLS, ops=ops.ptr, tx_off=off.ptr,
                                ntx=w.ntx, nops=w.nops, reads=reads.ptr, order=order.ptr, warp_density=wd,
                                block_size=bs, batch_size=batch)
        table, rd, o2 = cells.get(rows * COLS), reads.get(w.nops), order.get(2 * w.ntx)
    finally:
        for b in bufs:
            b.free()
    order3 = np.zeros(3 * w.ntx, dtype=np.uint64)
    order3[0::3] = np.arange(w.ntx, dtype=np.uint64) // np.uint64(batch)
    order3[1::3] = o2[: w.ntx]
    order3[2::3] = o2[w.ntx:]
    return st, {"table": table, "reads": rd, "order": order3}


def run_existing(scheme, rows, w, batch, wd, bs):
    with g.Context(scheme, rows=rows, warp_density=wd, block_size=bs, batch_size=batch) as ctx:
        ctx.seed_table()
        ctx.load_ops(w)
        st = ctx.run_plain()
        res = ctx.fetch()
    return st, res


def check_replay(rows, w, res, what):
    init = orc.seed_table(g.SEED, rows, COLS)
    exp_t, exp_r, ties = orc.replay(rows, COLS, g.SEED, init, w.ops, w.tx_off, res["order"])
    assert ties == 0, f"{what}: serialization keys must be distinct"
    isr = ~g.op_is_write(w.ops)
    assert np.array_equal(res["reads"][isr], exp_r[isr]), f"{what}: reads differ from the serial replay"
    assert np.array_equal(res["table"], exp_t), f"{what}: table differs from the serial replay"


def conflict_free(rows, ntx, length, seed):
    rng = np.random.default_rng(seed)
    keys = rng.permutation(rows)[: ntx * length].astype(np.uint32)
    cols = rng.integers(0, COLS, keys.size).astype(np.uint32)
    writes = rng.random(keys.size) < 0.4
    return g.Ops(g.make_ops(keys, cols, writes), np.arange(ntx + 1, dtype=np.uint64) * np.uint64(length))


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", SCHEMES6)
def test_static_matches_run_plain(scheme):
    for i, (rows, ntx, lmin, lmax, theta, rr, batch, wd, bs) in enumerate(CALLS):
        print(f"[no-alloc] scheme={scheme} call={i} rows={rows} ntx={ntx} batch={batch} wd={wd} bs={bs}", flush=True)
        w = g.generate(rows, ntx, theta=theta, read_ratio=rr, len_min=lmin, len_max=lmax, seed=g.SEED + i)
        st_s, res_s = run_static(scheme, rows, w, batch, wd, bs)
        assert st_s["committed"] == ntx and st_s["error_flags"] == 0 and st_s["index_miss"] == 0
        assert st_s["batches"] == (ntx + batch - 1) // batch
        check_replay(rows, w, res_s, f"static call {i}")
        st_e, res_e = run_existing(scheme, rows, w, batch, wd, bs)
        assert st_e["committed"] == ntx
        check_replay(rows, w, res_e, f"run_plain call {i}")
        # conflict-free input: results cannot depend on the interleaving
        cf = conflict_free(rows, rows // (2 * lmax), lmax, 7 + i)
        _, a = run_static(scheme, rows, cf, batch, wd, bs)
        _, b = run_existing(scheme, rows, cf, batch, wd, bs)
        isr = ~g.op_is_write(cf.ops)
        assert np.array_equal(a["table"], b["table"]), f"call {i}: tables differ between the two entry points"
        assert np.array_equal(a["reads"][isr], b["reads"][isr]), f"call {i}: reads differ between the two entry points"
        check_replay(rows, cf, a, f"static conflict-free call {i}")


@pytest.mark.gpu
def test_static_rejects_beyond_limits():
    with pytest.raises(g.GcctbError) as e:
        g.run_plain_static("silo", cells=1, rows=(1 << 20) + 1, cols=10, ops=1, tx_off=1, ntx=1, nops=1, reads=1,
                           order=1)
    assert e.value.code == -1
    with pytest.raises(g.GcctbError) as e:
        g.run_plain_static("gacco", cells=1, rows=16, cols=10, ops=1, tx_off=1, ntx=1, nops=1, reads=1, order=1)
    assert e.value.code == -1
