"""Host logic of the autotuner's table (no GPU)."""
from gpu_cc_tb_amd import autotune as at


def test_buckets():
    assert at.bucket(0.0, 0.9) == "t0-rr"
    assert at.bucket(0.99, 0.5) == "t4-rw"
    assert at.bucket(0.95, 1.0) == "t4-ro"
    assert at.bucket(0.7, 0.8) in ("t1-rr", "t2-rr")


def test_lookup_default_and_nearest():
    assert at.lookup("silo", 0.9, 0.9, table={}) == {"warp_density": 3, "block_size": 8, "window": 0,
                                                      "adaptive": False}
    table = {"silo": {"t0-rr": {"warp_density": 4, "block_size": 16, "window": 0, "txps": 1},
                      "t4-rr": {"warp_density": 3, "block_size": 8, "window": 1024, "txps": 1},
                      "t2-ro": {"warp_density": 3, "block_size": 8, "window": 0, "adaptive": True, "txps": 1},
                      "t4-rw": {"warp_density": 2, "block_size": 4, "window": 4096, "txps": 1}}}
    assert at.lookup("silo", 0.99, 0.9, table)["window"] == 1024
    assert at.lookup("silo", 0.99, 0.5, table)["window"] == 4096
    # no t3 entry: nearest skew with the same read bucket
    assert at.lookup("silo", 0.9, 0.9, table)["window"] == 1024
    assert at.lookup("silo", 0.05, 0.9, table)["block_size"] == 16
    assert at.lookup("silo", 0.8, 1.0, table)["adaptive"] is True
    # another scheme falls back to its default shape
    assert at.lookup("gacco", 0.99, 0.9, table)["warp_density"] == 2


def test_lookup_plain_entries():
    table = {"silo": {"plain-t4-rr": {"warp_density": 4, "block_size": 4, "batch_size": 262144, "txps": 1},
                      "plain-t1-rw": {"warp_density": 2, "block_size": 16, "batch_size": 4096, "txps": 1},
                      "t4-rr": {"warp_density": 3, "block_size": 8, "window": 1024, "txps": 1}}}
    assert at.lookup_plain("silo", 0.99, 0.9, table)["batch_size"] == 262144
    # nearest skew with the same read bucket
    assert at.lookup_plain("silo", 0.8, 0.9, table)["warp_density"] == 4
    assert at.lookup_plain("silo", 0.6, 0.5, table)["block_size"] == 16
    # no plain entry: the default plain shape at batch 65,536
    assert at.lookup_plain("tictoc", 0.99, 0.9, table) == {"warp_density": 3, "block_size": 8, "batch_size": 65536}
    # persistent lookups ignore plain entries
    assert at.lookup("silo", 0.99, 0.9, table)["window"] == 1024
