"""The dense index: the identity index (keys 0 .. n-1, key k on row k, the
sorted array of dense keys that the workload generator's tables have) is
direct addressing, with no table read per lookup.  Find's contract must
hold as with the hash table: hits return the key's row, misses the index
length; GCCTB_INDEX_HASH=1 builds the same identity as a hash table, and
every run must come out the same either way."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g
import oracle_lib as orc


def probe(rows, probes, build=None, hash_env=False, monkeypatch=None):
    if monkeypatch is not None:
        if hash_env:
            monkeypatch.setenv("GCCTB_INDEX_HASH", "1")
        else:
            monkeypatch.delenv("GCCTB_INDEX_HASH", raising=False)
    with g.Context("tpl_nw", rows=rows) as ctx:
        ctx.seed_table()
        if build is not None:
            ctx.build_index(*build)
        return ctx.probe(probes).astype(np.int64)


@pytest.mark.gpu
def test_dense_probe_hits_and_misses(monkeypatch):
    rows = 5000
    rng = np.random.default_rng(3)
    probes = np.concatenate([np.arange(rows, dtype=np.uint32), rng.integers(rows, 2**32 - 1, 500).astype(np.uint32),
                             np.array([rows, 0xFFFFFFFE], dtype=np.uint32)])
    print(f"[dense] identity over {rows} rows, {probes.size} probes", flush=True)
    exp = np.where(probes < rows, probes.astype(np.int64), rows)
    got = probe(rows, probes, monkeypatch=monkeypatch)
    assert np.array_equal(got, exp)
    got_hash = probe(rows, probes, hash_env=True, monkeypatch=monkeypatch)
    assert np.array_equal(got_hash, exp)
    # identity keys handed over explicitly, shorter than the table: the same
    # dense index of length n (keys from n on miss and return n)
    n = 1200
    keys = rng.permutation(n).astype(np.uint32)
    short = probe(rows, probes, build=(keys, keys), monkeypatch=monkeypatch)
    assert np.array_equal(short, np.where(probes < n, probes.astype(np.int64), n))
    sk = np.arange(n, dtype=np.uint32)
    exp_find = np.array([orc.find(sk, sk.astype(np.uint64), int(k)) for k in probes[::7]], dtype=np.int64)
    assert np.array_equal(short[::7], exp_find)


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["plain", "persistent"])
@pytest.mark.parametrize("scheme", ["tpl_nw", "to", "mvcc", "silo", "tictoc", "gacco", "gputx"])
def test_dense_and_hash_identity_runs(scheme, form, monkeypatch):
    """Each run replays serially in its own order; GaccO and GPUTx fix their
    order, so their results are also identical between the two indexes."""
    rows = 4096
    w = g.generate(rows, 3000, theta=0.9, read_ratio=0.7, len_min=4, len_max=20)
    out = {}
    for kind in ("dense", "hash"):
        if kind == "hash":
            monkeypatch.setenv("GCCTB_INDEX_HASH", "1")
        else:
            monkeypatch.delenv("GCCTB_INDEX_HASH", raising=False)
        print(f"[dense-run] scheme={scheme} form={form} index={kind}", flush=True)
        with g.Context(scheme, rows=rows, warp_density=3, block_size=8, batch_size=700, persist_window=512) as ctx:
            ctx.seed_table()
            ctx.load_ops(w)
            st = ctx.run(form)
            res = ctx.fetch()
        assert st["committed"] == w.ntx and st["index_miss"] == 0
        init = orc.seed_table(g.SEED, rows, 10)
        exp_t, exp_r, ties = orc.replay(rows, 10, g.SEED, init, w.ops, w.tx_off, res["order"])
        isr = ~g.op_is_write(w.ops)
        assert ties == 0
        assert np.array_equal(res["table"], exp_t) and np.array_equal(res["reads"][isr], exp_r[isr])
        out[kind] = res
    if scheme in ("gacco", "gputx"):
        assert np.array_equal(out["dense"]["table"], out["hash"]["table"])
        assert np.array_equal(out["dense"]["reads"], out["hash"]["reads"])
