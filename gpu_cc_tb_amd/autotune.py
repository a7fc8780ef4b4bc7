"""Launch-shape autotuner.

Chooses, per scheme and contention bucket, the lanes per transaction
(warp_density), the workgroup size (block_size, in 32-lane units as the
reference counts them) and the persistent form's admission window
(transactions in flight), by measuring the persistent form on a sample
workload of that bucket.  The search is coordinate descent from the scheme's
default shape: window first (the parameter that moves most with contention),
then warp_density, then block_size, then the window once more.

The result is a small JSON table keyed by scheme, then by bucket name
"t<theta bucket>-r<read-ratio bucket>" (or "stream" for the batched stream
of mixed skew), each entry {"warp_density", "block_size", "window",
"adaptive", "txps"}.  `lookup` maps a workload's skew and read ratio to the
nearest measured bucket.

The plain per-batch form is tuned too, with the batch size as a fourth
parameter (the reference's batch_size argument, benchmark_cpu/ycsb.cu:96):
entries "plain-t<theta>-r<read>" hold {"warp_density", "block_size",
"batch_size", "txps"}, and `lookup_plain` reads them.

    python -m gpu_cc_tb_amd.autotune --out gpu_cc_tb_amd/tuned.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import SCHEMES, SEED, Context, generate

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuned.json")

# skew and read-ratio buckets: representative value of each
THETAS = {"0": 0.1, "1": 0.6, "2": 0.8, "3": 0.9, "4": 0.99}
READS = {"w": 0.5, "r": 0.9, "o": 1.0}

WD = [1, 2, 3, 4, 5]
BS = [1, 2, 4, 8]
# plain form: lanes per transaction, block size (in 32-lane units) and batch size
PLAIN_WD = [1, 2, 3, 4, 5]
PLAIN_BS = [4, 8, 16]
PLAIN_BATCHES = [4096, 16384, 65536, 262144]
# admission windows (transactions in flight; 0 = device-filling).  The
# adaptive window (-1) is left out: the library refuses it.
WINDOWS = [0, 32768, 16384, 8192, 4096, 1024, 512, 256]

# starting shapes (from sweeps of the plain form at low contention)
DEFAULT_SHAPE = {"tpl_nw": (3, 8), "tpl_wd": (3, 8), "to": (4, 8), "mvcc": (3, 8), "silo": (3, 8),
                 "tictoc": (3, 8), "gacco": (2, 8), "gputx": (3, 8)}


def theta_bucket(theta: float) -> str:
    return min(THETAS, key=lambda k: abs(THETAS[k] - theta))


def read_bucket(rr: float) -> str:
    return min(READS, key=lambda k: abs(READS[k] - rr))


def bucket(theta: float, read_ratio: float) -> str:
    return f"t{theta_bucket(theta)}-r{read_bucket(read_ratio)}"


def load_table(path: str = TABLE_PATH) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(scheme: str, theta: float, read_ratio: float, table: dict | None = None, key: str | None = None) -> dict:
    """Tuned {warp_density, block_size, window, adaptive} for a workload (or
    for the named bucket `key`, e.g. "stream", when the table has it); the
    default shape with a device-filling window when the table has no entry."""
    table = load_table() if table is None else table
    ent = table.get(scheme, {})
    if key is None or key not in ent:
        key = bucket(theta, read_ratio)
    if key not in ent and ent:
        # nearest measured skew with the same read bucket, else any
        buckets = [k for k in ent if k.startswith("t")]
        same_r = [k for k in buckets if k.endswith("-r" + read_bucket(read_ratio))] or buckets
        if same_r:
            key = min(same_r, key=lambda k: abs(THETAS[k[1]] - theta))
    if key in ent:
        e = ent[key]
        return {"warp_density": e["warp_density"], "block_size": e["block_size"], "window": e["window"],
                "adaptive": bool(e.get("adaptive", False))}
    wd, bs = DEFAULT_SHAPE[scheme]
    return {"warp_density": wd, "block_size": bs, "window": 0, "adaptive": False}


def lookup_plain(scheme: str, theta: float, read_ratio: float, table: dict | None = None) -> dict:
    """Tuned {warp_density, block_size, batch_size} of the plain form for a
    workload: the nearest measured skew with the same read bucket (else any
    plain entry), or the default shape at batch 65,536."""
    table = load_table() if table is None else table
    ent = {k[len("plain-"):]: v for k, v in table.get(scheme, {}).items() if k.startswith("plain-t")}
    key = bucket(theta, read_ratio)
    if key not in ent and ent:
        same_r = [k for k in ent if k.endswith("-r" + read_bucket(read_ratio))] or list(ent)
        key = min(same_r, key=lambda k: abs(THETAS[k[1]] - theta))
    if key in ent:
        e = ent[key]
        return {"warp_density": e["warp_density"], "block_size": e["block_size"], "batch_size": e["batch_size"]}
    wd, bs = DEFAULT_SHAPE[scheme]
    return {"warp_density": wd, "block_size": bs, "batch_size": 65536}


def measure_plain(scheme: str, rows: int, ops, wd: int, bs: int, batch: int) -> float:
    """Committed transactions per second of one plain run (0 on a failed run)."""
    with Context(scheme, rows=rows, warp_density=wd, block_size=bs, batch_size=batch) as c:
        c.seed_table()
        c.load_ops(ops)
        st = c.run("plain", check=False)
    if st["rc"] != 0 or st["committed"] != ops.ntx:
        return 0.0
    return ops.ntx / (st["tot_ms"] / 1e3)


def tune_plain(scheme: str, rows: int, ops, log=print) -> dict:
    """Coordinate descent over batch size, warp_density, block_size, then the
    batch size again, for the plain form."""
    wd, bs = DEFAULT_SHAPE[scheme]
    batch = 65536
    cache: dict[tuple, float] = {}

    def score(w, b, n):
        k = (w, b, n)
        if k not in cache:
            t0 = time.time()
            cache[k] = measure_plain(scheme, rows, ops, w, b, n)
            log(json.dumps({"scheme": scheme, "form": "plain", "wd": w, "bs": b, "batch": n,
                            "txps": round(cache[k]), "s": round(time.time() - t0, 2)}))
        return cache[k]

    batch = max(PLAIN_BATCHES, key=lambda n: score(wd, bs, n))
    wd = max(PLAIN_WD, key=lambda w: score(w, bs, batch))
    bs = max(PLAIN_BS, key=lambda b: score(wd, b, batch))
    batch = max(PLAIN_BATCHES, key=lambda n: score(wd, bs, n))
    return {"warp_density": wd, "block_size": bs, "batch_size": batch, "txps": round(score(wd, bs, batch))}


def measure(scheme: str, rows: int, ops, batch_sizes, wd: int, bs: int, window: int, form: str) -> float:
    """Committed transactions per second of one run (0 on a failed run);
    window -1 selects the adaptive window."""
    with Context(scheme, rows=rows, warp_density=wd, block_size=bs, batch_size=1 << 30,
                 persist_window=max(0, window), persist_adaptive=window < 0) as c:
        c.seed_table()
        c.load_ops(ops)
        c.set_batches(batch_sizes)
        st = c.run(form, check=False)
    if st["rc"] != 0 or st["committed"] != ops.ntx:
        return 0.0
    return ops.ntx / (st["tot_ms"] / 1e3)


def tune(scheme: str, rows: int, ops, batch_sizes, form: str = "persistent", log=print) -> dict:
    """Coordinate descent over window, warp_density, block_size."""
    wd, bs = DEFAULT_SHAPE[scheme]
    win = 0
    cache: dict[tuple, float] = {}

    def score(w, b, n):
        k = (w, b, n)
        if k not in cache:
            t0 = time.time()
            cache[k] = measure(scheme, rows, ops, batch_sizes, w, b, n, form)
            log(json.dumps({"scheme": scheme, "wd": w, "bs": b, "window": n, "txps": round(cache[k]),
                            "s": round(time.time() - t0, 2)}))
        return cache[k]

    windows = WINDOWS if form == "persistent" else [0]
    win = max(windows, key=lambda n: score(wd, bs, n))
    wd = max(WD, key=lambda w: score(w, bs, win))
    bs = max(BS, key=lambda b: score(wd, b, win))
    win = max(windows, key=lambda n: score(wd, bs, n))
    return {"warp_density": wd, "block_size": bs, "window": max(0, win), "adaptive": win < 0,
            "txps": round(score(wd, bs, win))}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--schemes", default=",".join(SCHEMES))
    ap.add_argument("--thetas", default=",".join(THETAS))
    ap.add_argument("--reads", default="w,r")
    ap.add_argument("--rows", type=int, default=10485760)
    ap.add_argument("--ntx", type=int, default=131072, help="sample transactions per bucket")
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--stream-batches", type=int, default=64,
                    help="also tune bucket \"stream\" on this many batches of the batched stream (0 = skip)")
    ap.add_argument("--forms", default="persistent", help="persistent, plain or both (comma-separated)")
    args = ap.parse_args(argv)
    forms = [f for f in args.forms.split(",") if f]
    table = load_table(args.out)
    if args.stream_batches and "persistent" in forms:
        from . import workloads
        full = workloads.config4(rows=args.rows)
        nb = min(args.stream_batches, len(full.batch_sizes))
        ntx = int(sum(full.batch_sizes[:nb]))
        ops = type(full.ops)(full.ops.ops[: int(full.ops.tx_off[ntx])], full.ops.tx_off[: ntx + 1])
        for scheme in args.schemes.split(","):
            best = tune(scheme, args.rows, ops, full.batch_sizes[:nb], log=lambda s: print(s, flush=True))
            table.setdefault(scheme, {})["stream"] = best
            print(json.dumps({"scheme": scheme, "bucket": "stream", "best": best}), flush=True)
            with open(args.out, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
    for tb in [x for x in args.thetas.split(",") if x]:
        for rb in args.reads.split(","):
            w = generate(args.rows, args.ntx, theta=THETAS[tb], read_ratio=READS[rb], seed=SEED)
            sizes = [min(args.batch, args.ntx - s) for s in range(0, args.ntx, args.batch)]
            for scheme in args.schemes.split(","):
                if "persistent" in forms:
                    best = tune(scheme, args.rows, w, sizes, log=lambda s: print(s, flush=True))
                    best["source"] = (f"autotune coordinate descent on {args.ntx} tx x 16 ops, theta "
                                      f"{THETAS[tb]}, read ratio {READS[rb]}, batch {args.batch}")
                    table.setdefault(scheme, {})[f"t{tb}-r{rb}"] = best
                    print(json.dumps({"scheme": scheme, "bucket": f"t{tb}-r{rb}", "best": best}), flush=True)
                if "plain" in forms:
                    best = tune_plain(scheme, args.rows, w, log=lambda s: print(s, flush=True))
                    best["source"] = (f"autotune coordinate descent (plain form) on {args.ntx} tx x 16 ops, theta "
                                      f"{THETAS[tb]}, read ratio {READS[rb]}")
                    table.setdefault(scheme, {})[f"plain-t{tb}-r{rb}"] = best
                    print(json.dumps({"scheme": scheme, "bucket": f"plain-t{tb}-r{rb}", "best": best}), flush=True)
                with open(args.out, "w") as f:
                    json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
