"""Experiment drivers for the workloads beyond the headline benchmark.  Each
prints one JSON object per run, so that a long study reports as it goes.

    python -m gpu_cc_tb_amd.study sweep   # high contention: theta x read ratio x batch size
    python -m gpu_cc_tb_amd.study stream  # batched stream: plain / graph / persistent forms
    python -m gpu_cc_tb_amd.study serve   # serving: Poisson arrivals at fractions of peak

sweep:  10,485,760 rows; transactions of 16 ops; theta in {0.6, 0.8, 0.99},
        read ratio in {0.5, 0.9}, batch size in {4,096 .. 262,144}; plain
        form; per run Tx/s, aborts and the per-phase time per transaction
        (timestamp, wait, abort, index; ms per transaction like the
        reference's output line).
stream: 256 batches of log-uniform size in [1,024, 262,144], length
        U[4, 32], per-batch theta U[0, 0.99] and read ratio in {0.5, 0.9, 1.0}
        over 10,485,760 rows; every form on the same stream.
serve:  67,108,864 rows; batches of log-uniform size in [256, 65,536] of
        16-op transactions, theta 0.9, read ratio 0.8; peak = all requests
        present at once; then Poisson arrivals at the given fractions of peak.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import SCHEMES, Context, generate
from . import autotune, workloads


def _phases(st: dict) -> dict:
    n = max(1, st["txs"])
    k = st["tick_khz"] or 1.0
    ms = lambda t: round(t / (n * k), 6)  # noqa: E731  ticks -> ms per transaction
    return {"ts_d": ms(st["ts_ticks"]), "wait_d": ms(st["wait_ticks"]), "abort_d": ms(st["abort_ticks"]),
            "mng_d": ms(st["manager_ticks"]), "tot_d": ms(st["tot_ticks"]), "idx_d": ms(st["index_ticks"])}


def sweep(args) -> None:
    thetas = [float(x) for x in args.thetas.split(",")]
    reads = [float(x) for x in args.reads.split(",")]
    batches = [int(x) for x in args.batches.split(",")]
    for th in thetas:
        for rr in reads:
            w = generate(args.rows, args.txs, theta=th, read_ratio=rr)
            for scheme in args.schemes.split(","):
                shape = autotune.lookup(scheme, th, rr)
                for b in batches:
                    with Context(scheme, rows=args.rows, warp_density=shape["warp_density"],
                                 block_size=shape["block_size"], batch_size=b) as c:
                        c.seed_table()
                        c.load_ops(w)
                        st = c.run("plain", check=False)
                    if st["tot_ms"] <= 0:
                        print(json.dumps({"study": "sweep", "scheme": scheme, "rc": st["rc"]}), flush=True)
                        continue
                    print(json.dumps({"study": "sweep", "scheme": scheme, "theta": th, "read_ratio": rr, "batch": b,
                                      "wd": shape["warp_density"], "bs": shape["block_size"],
                                      "txps": round(w.ntx / (st["tot_ms"] / 1e3)), "aborts": st["aborts"],
                                      "pre_ms": round(st["pre_ms"], 2), "proc_ms": round(st["proc_ms"], 2),
                                      "rc": st["rc"], **_phases(st)}), flush=True)


def stream(args) -> None:
    t0 = time.time()
    s = workloads.config4(nbatches=args.batches, rows=args.rows)
    print(json.dumps({"study": "stream", "ntx": s.ntx, "nops": s.ops.nops, "batches": len(s.batch_sizes),
                      "gen_s": round(time.time() - t0, 1)}), flush=True)
    forms = args.forms.split(",")
    for scheme in args.schemes.split(","):
        shape = autotune.lookup(scheme, 0.5, 0.9, key="stream")
        res = {}
        for f in forms:
            parts = f.split(":")
            if parts[1:] == ["tuned"]:
                form, win, adapt = parts[0], shape["window"], shape["adaptive"]
            else:
                form, win, adapt = parts[0], int(parts[1]) if len(parts) > 1 else 0, parts[2:] == ["a"]
            with Context(scheme, rows=args.rows, warp_density=shape["warp_density"], block_size=shape["block_size"],
                         batch_size=1 << 30, persist_window=win, persist_adaptive=adapt) as c:
                c.seed_table()
                c.load_ops(s.ops)
                c.set_batches(s.batch_sizes)
                st = c.run(form, check=False)
            if st["rc"] != 0 and st["tot_ms"] <= 0:
                print(json.dumps({"study": "stream", "scheme": scheme, "form": f, "rc": st["rc"]}), flush=True)
                continue
            res[f] = s.ntx / (st["tot_ms"] / 1e3)
            print(json.dumps({"study": "stream", "scheme": scheme, "form": f, "txps": round(res[f]),
                              "tot_ms": round(st["tot_ms"], 2), "pre_ms": round(st["pre_ms"], 2),
                              "aborts": st["aborts"], "launches": st["batches"], "rc": st["rc"]}), flush=True)
        if "plain" in res:
            best = max(res, key=res.get)
            print(json.dumps({"study": "stream", "scheme": scheme, "best": best,
                              "best_over_plain": round(res[best] / res["plain"], 3)}), flush=True)


def serve(args) -> None:
    s = workloads.serving_stream(args.rows, 4096, 256, 65536, theta=0.9, read_ratio=0.8, max_txs=args.txs)
    print(json.dumps({"study": "serve", "ntx": s.ntx, "batches": len(s.batch_sizes)}), flush=True)
    for scheme in args.schemes.split(","):
        shape = autotune.lookup(scheme, 0.9, 0.8)
        with Context(scheme, rows=args.rows, warp_density=shape["warp_density"], block_size=shape["block_size"],
                     batch_size=1 << 30, persist_window=shape["window"], persist_adaptive=shape["adaptive"]) as c:
            c.seed_table()
            c.load_ops(s.ops)
            c.serve(s.batch_sizes, max_launch_txs=args.max_launch)
            peak = c.serve(s.batch_sizes, max_launch_txs=args.max_launch)
            print(json.dumps({"study": "serve", "scheme": scheme, "load": "peak", "txps": round(peak["txps"]),
                              "launches": peak["launches"], "p50_ms": round(peak["p50_ms"], 3),
                              "p99_ms": round(peak["p99_ms"], 3)}), flush=True)
            for load in [float(x) for x in args.loads.split(",")]:
                arr = workloads.poisson_arrivals(s.batch_sizes, load * peak["txps"])
                st = c.serve(s.batch_sizes, arr, max_launch_txs=args.max_launch)
                print(json.dumps({"study": "serve", "scheme": scheme, "load": load, "offered_txps":
                                  round(load * peak["txps"]), "txps": round(st["txps"]),
                                  "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                                  "max_ms": round(st["max_ms"], 3), "launches": st["launches"],
                                  "mean_launch_txs": round(st["mean_launch_txs"]), "deferred": st["deferred"],
                                  "busy_ms": round(st["busy_ms"], 1), "tot_ms": round(st["tot_ms"], 1)}),
                      flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("sweep")
    a.add_argument("--rows", type=int, default=10485760)
    a.add_argument("--txs", type=int, default=1048576)
    a.add_argument("--thetas", default="0.6,0.8,0.99")
    a.add_argument("--reads", default="0.5,0.9")
    a.add_argument("--batches", default="4096,16384,65536,262144")
    a.add_argument("--schemes", default="tpl_nw,silo,tictoc,mvcc,gacco,gputx")
    a.set_defaults(fn=sweep)
    b = sub.add_parser("stream")
    b.add_argument("--rows", type=int, default=10485760)
    b.add_argument("--batches", type=int, default=256)
    b.add_argument("--schemes", default=",".join(SCHEMES))
    b.add_argument("--forms", default="plain,graph,persistent:tuned,persistent:0,persistent:4096,persistent:1024",
                   help="form[:window[:a]]; persistent:tuned takes the autotuned stream entry")
    b.set_defaults(fn=stream)
    c = sub.add_parser("serve")
    c.add_argument("--rows", type=int, default=67108864)
    c.add_argument("--txs", type=int, default=2000000)
    c.add_argument("--loads", default="0.25,0.5,0.9")
    c.add_argument("--schemes", default="mvcc,silo")
    c.add_argument("--max-launch", type=int, default=262144)
    c.set_defaults(fn=serve)
    args = ap.parse_args(argv)
    args.fn(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
