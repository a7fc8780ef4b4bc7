"""The device's stable key-value sort (the GaccO / GPUTx preparation sort)
against numpy's stable argsort, over tile-edge sizes, duplicate-heavy keys,
partial key widths and 64-bit keys."""
import numpy as np
import pytest

import gpu_cc_tb_amd as g

TILE = 4096  # elements per sort tile


def expect(keys, vals, bits):
    mask = (1 << bits) - 1
    order = np.argsort(keys & keys.dtype.type(mask), kind="stable")
    return keys[order], vals[order]


def keyset(kind, n, dtype, rng):
    top = np.iinfo(dtype).max
    if kind == "uniform":
        return rng.integers(0, top, n, dtype=dtype, endpoint=True)
    if kind == "few":
        return rng.integers(0, 5, n).astype(dtype)
    if kind == "equal":
        return np.full(n, 7, dtype=dtype)
    if kind == "zipf":
        return np.minimum(rng.zipf(1.3, n), 1 << 24).astype(dtype)
    raise ValueError(kind)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 100, TILE - 1, TILE, TILE + 1, 3 * TILE + 17, (1 << 20) + 123])
@pytest.mark.parametrize("kind", ["uniform", "few", "zipf"])
def test_sort_u32(n, kind):
    rng = np.random.default_rng(n * 7 + len(kind))
    keys = keyset(kind, n, np.uint32, rng)
    vals = np.arange(n, dtype=np.uint32)
    print(f"[sort-u32] n={n} kind={kind}", flush=True)
    k, v, _ = g.sort_pairs(keys, vals)
    ek, ev = expect(keys, vals, 32)
    assert np.array_equal(k, ek) and np.array_equal(v, ev)


@pytest.mark.gpu
@pytest.mark.parametrize("bits", [1, 8, 13, 24])
def test_sort_low_bits_only(bits):
    """Only the low key_bits bits order the pairs; ties keep input order."""
    rng = np.random.default_rng(bits)
    n = 5 * TILE + 3
    keys = rng.integers(0, 2**32 - 1, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    print(f"[sort-bits] bits={bits}", flush=True)
    k, v, _ = g.sort_pairs(keys, vals, key_bits=bits)
    ek, ev = expect(keys, vals, bits)
    assert np.array_equal(k, ek) and np.array_equal(v, ev)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,bits", [("uniform", 64), ("uniform", 40), ("few", 64), ("equal", 64)])
def test_sort_u64(kind, bits):
    rng = np.random.default_rng(bits)
    n = 2 * TILE + 999
    keys = keyset(kind, n, np.uint64, rng)
    vals = rng.integers(0, 2**32 - 1, n, dtype=np.uint32)
    print(f"[sort-u64] kind={kind} bits={bits}", flush=True)
    k, v, _ = g.sort_pairs(keys, vals, key_bits=bits)
    ek, ev = expect(keys, vals, bits)
    assert np.array_equal(k, ek) and np.array_equal(v, ev)


@pytest.mark.gpu
def test_sort_large_rows():
    """The preparation's shape: 16,777,216 row ids below 2^24, 24-bit sort."""
    rng = np.random.default_rng(5)
    n = 1 << 24
    keys = rng.integers(0, 10485760, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    print(f"[sort-large] n={n}", flush=True)
    k, v, ms = g.sort_pairs(keys, vals, key_bits=24)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v, order.astype(np.uint32)) and np.array_equal(k, keys[order])
    print(f"[sort-large] {ms:.3f} ms", flush=True)
