#!/usr/bin/env python3
"""Benchmark: committed transactions per second of the YCSB transaction path on
one MI355X, under the eight CC schemes (tpl_nw tpl_wd to mvcc silo tictoc gacco
gputx).

Workload (default): 10,485,760 rows x 10 u64 columns seeded from (seed, row,
col); 1,048,576 transactions of 16 ops with Zipf(theta 0.1) keys, distinct per
transaction; read ratio 0.9; batch_size 65,536 -- the "plain method on GPU, low
contention" configuration.  Synthetic data, generated deterministically.

A step is one pass of every scheme over all 1,048,576 transactions with the
reference's launch structure (per batch: reset/prepare, one kernel launch,
synchronize).  The timed region holds exactly --steps steps; inputs are already
resident in HBM.  `value` is whole-job committed Tx/s (8 x 1,048,576 x steps
transactions / time).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpu_cc_tb_amd as g  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--rows", type=int, default=10485760)
    p.add_argument("--txs", type=int, default=1048576)
    p.add_argument("--ops", type=int, default=16)
    p.add_argument("--theta", type=float, default=0.1)
    p.add_argument("--read-ratio", type=float, default=0.9)
    p.add_argument("--batch", type=int, default=65536)
    p.add_argument("--schemes", default=",".join(g.SCHEMES))
    p.add_argument("--wd", type=int, default=-1, help="warp_density; -1 = tuned default per scheme")
    p.add_argument("--bs", type=int, default=-1, help="block_size; -1 = tuned default per scheme")
    p.add_argument("--cpu-sample", type=int, default=262144, help="transactions per scheme for the CPU baseline")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


# launch shape per scheme (warp_density, block_size): best of the wd x bs sweep
# on this workload (wd 1..4 x bs 4, 8) on MI355X
DEFAULT_SHAPE = {"tpl_nw": (3, 8), "tpl_wd": (3, 8), "to": (4, 8), "mvcc": (4, 4), "silo": (3, 8),
                 "tictoc": (3, 8), "gacco": (2, 8), "gputx": (3, 8)}


def algorithmic_bytes(ntx: int, nops: int, nreads: int) -> int:
    # per transaction 8 B header; per op 8 B op word + 8 B cell + 16 B CC word
    # load/store + 16 B index leaf entry; per read 8 B result
    return 8 * ntx + 48 * nops + 8 * nreads


def cpu_baseline(w_full: g.Ops, args, schemes):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    n = min(args.cpu_sample, w_full.ntx)
    ops = w_full.ops[: int(w_full.tx_off[n])]
    off = w_full.tx_off[: n + 1]
    init = orc.seed_table(g.SEED, args.rows, 10)
    tot_s, tot_tx, threads = 0.0, 0, 0
    per = {}
    for s in schemes:
        rc, _, _, _, st = orc.cpu_run(g.SCHEMES.index(s), args.rows, 10, g.SEED, init, ops, off, args.batch,
                                      args.cpu_threads)
        if rc != 0:
            raise RuntimeError(f"cpu path failed for {s}: {rc}")
        tot_s += st.seconds
        tot_tx += n
        threads = st.threads
        per[s] = round(n / st.seconds)
    return {
        "value": round(tot_tx / tot_s),
        "unit": "Tx/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} transactions of the same workload per scheme ({len(schemes)} schemes), "
                  f"oracle/cpu_cc.c with {threads} threads, per-batch CC reset included",
        "per_scheme": per,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path has no multi-GPU form in this build.",
              file=sys.stderr)
        sys.exit(2)
    schemes = [s for s in args.schemes.split(",") if s]
    t0 = time.time()
    w = g.generate(args.rows, args.txs, theta=args.theta, read_ratio=args.read_ratio, len_min=args.ops)
    nreads = int(np.count_nonzero(~g.op_is_write(w.ops)))
    gen_s = time.time() - t0
    info = g.device_info()
    per = {}
    tot_ms = 0.0
    kern_ms = 0.0
    launches = 0
    for s in schemes:
        wd, bs = DEFAULT_SHAPE[s]
        if args.wd >= 0:
            wd = args.wd
        if args.bs >= 0:
            bs = args.bs
        with g.Context(s, rows=args.rows, warp_density=wd, block_size=bs, batch_size=args.batch) as ctx:
            ctx.seed_table()
            ctx.load_ops(w)
            for _ in range(args.warmup):
                ctx.run_plain()
            s_ms = s_kern = 0.0
            s_launch = s_abort = 0
            wall0 = time.perf_counter()
            for _ in range(args.steps):
                st = ctx.run_plain()
                s_ms += st["tot_ms"]
                s_kern += st["exec_kernel_ms"]
                s_launch += st["batches"]
                s_abort += st["aborts"]
            wall = time.perf_counter() - wall0
        per[s] = {
            "txps": round(args.txs * args.steps / (s_ms / 1e3)),
            "wall_txps": round(args.txs * args.steps / wall),
            "ms_per_pass": round(s_ms / args.steps, 3),
            "kernel_ms_per_pass": round(s_kern / args.steps, 3),
            "aborts_per_pass": s_abort // args.steps,
            "launches_per_pass": s_launch // args.steps,
            "wd": wd, "bs": bs,
        }
        tot_ms += s_ms
        kern_ms += s_kern
        launches += s_launch
    committed = args.txs * args.steps * len(schemes)
    value = committed / (tot_ms / 1e3)
    bytes_total = algorithmic_bytes(args.txs, w.nops, nreads) * args.steps * len(schemes)
    achieved = bytes_total / (kern_ms / 1e3) / 1e9
    line = {
        "metric": "committed Tx/s on 1 MI355X (plain per-batch form, 8 CC schemes)",
        "value": round(value),
        "unit": "Tx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tot_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (deterministic Zipf YCSB generator, seeded table)",
        "config": {
            "workload": f"ycsb plain: {args.rows} rows x 10 u64, {args.txs} tx x {args.ops} ops, "
                        f"theta {args.theta}, read ratio {args.read_ratio}, batch {args.batch}, "
                        f"schemes {'+'.join(schemes)}",
            "rows": args.rows, "txs": args.txs, "ops_per_tx": args.ops, "theta": args.theta,
            "read_ratio": args.read_ratio, "batch_size": args.batch, "schemes": schemes,
        },
        "per_scheme": per,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "kernel": "txn_exec (all schemes)",
            "bytes_per_tx": round(algorithmic_bytes(args.txs, w.nops, nreads) / args.txs, 1),
            "avg_launch_ms": round(kern_ms / max(1, launches), 4),
        },
        "device": info["name"],
        "gen_s": round(gen_s, 2),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w, args, schemes)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
