#!/usr/bin/env python3
"""Benchmark: committed transactions per second of the YCSB transaction path on
one MI355X, batched (persistent) form against the plain per-batch form, under
the eight CC schemes (tpl_nw tpl_wd to mvcc silo tictoc gacco gputx), plus the
p50/p99 commit latency of the serving form.

Workload (throughput): 10,485,760 rows x 10 u64 columns seeded from (seed,
row, col); 1,048,576 transactions of 16 ops with Zipf(theta 0.1) keys,
distinct per transaction; read ratio 0.9; batch_size 65,536.  Synthetic data,
generated deterministically.

A step is one pass of every scheme over all 1,048,576 transactions in the
batched form: the scheme's preparation (GaccO's and GPUTx's ordering passes;
the sampled hot-key count of the writer deferral for 2PL, T/O and MVCC), one
persistent launch over the 16 batches whose groups refill as transactions
commit, one restore of the touched CC words.
The timed region holds exactly --steps steps; inputs are already resident in
HBM.  `value` is whole-job committed Tx/s of the batched form (8 x 1,048,576 x
steps transactions / time).  The plain form (the reference's per batch
reset + launch + synchronize) is timed the same way and reported beside it.

Batched-vs-plain legs (one pass of each form per scheme, preparation inside
the timed region, at full size):
  stream:  the 256-batch stream (sizes log-uniform in [1,024, 262,144], 4-32
           ops, per-batch theta U[0, 0.99], read ratio 0.5 / 0.9 / 1.0) on
           10,485,760 rows: persistent form against per-batch plain launches.
  theta99: 1,048,576 transactions of 16 ops at Zipf theta 0.99, read ratio 0.9
           and 0.5, on 10,485,760 rows: persistent over plain at batch 65,536
           on the faster of two launch shapes (the persistent form's and the
           plain form's own tuned one), and over plain at its own tuned shape
           and batch size.

Latency: the serving form on a 67,108,864-row table with requests of 16 ops,
theta 0.9, read ratio 0.8, in batches of log-uniform size in [256, 65,536]
arriving as a seeded Poisson process at a fraction of the measured peak.

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
from gpu_cc_tb_amd import autotune, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--rows", type=int, default=10485760)
    p.add_argument("--txs", type=int, default=1048576)
    p.add_argument("--ops", type=int, default=16)
    p.add_argument("--theta", type=float, default=0.1)
    p.add_argument("--read-ratio", type=float, default=0.9)
    p.add_argument("--batch", type=int, default=65536)
    p.add_argument("--schemes", default=",".join(g.SCHEMES))
    p.add_argument("--wd", type=int, default=-1, help="warp_density; -1 = autotuned per scheme")
    p.add_argument("--bs", type=int, default=-1, help="block_size; -1 = autotuned per scheme")
    p.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    p.add_argument("--serve-rows", type=int, default=67108864)
    p.add_argument("--serve-txs", type=int, default=1000000,
                   help="requests per serving run (enough launches for stable percentiles)")
    p.add_argument("--serve-loads", default="0.25,0.5,0.9", help="offered loads, fractions of measured peak")
    p.add_argument("--serve-schemes", default="mvcc,silo")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-stream", action="store_true", help="skip the 256-batch stream leg")
    p.add_argument("--no-theta99", action="store_true", help="skip the theta 0.99 leg")
    p.add_argument("--leg-schemes", default=",".join(g.SCHEMES), help="schemes of the stream and theta 0.99 legs")
    p.add_argument("--cpu-sample", type=int, default=1048576, help="transactions per scheme for the CPU baseline")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def algorithmic_bytes(ntx: int, nops: int, nreads: int) -> int:
    # per transaction 8 B header; per op 8 B op word + 8 B cell + 16 B CC word
    # load/store + 16 B index leaf entry; per read 8 B result
    return 8 * ntx + 48 * nops + 8 * nreads


def traffic_per_launch():
    """HBM bytes per launch of the execution kernel on this workload, from the
    committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), or None when no summary is present."""
    path = os.path.join(PROFILES, "txn_exec_config2_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip().split()[0])
    return None


def atomic_roofline(args, avg_launch_ms):
    """Atomic throughput of the execution kernel against the random 64-bit
    CAS ceiling, measured here on this GPU over a rows x 8 B array."""
    ceiling = g.microbench("cas_random", args.rows, threads=1 << 20, iters=16)["rate"]
    per_launch = atomics_per_launch()
    achieved = per_launch / (avg_launch_ms / 1e3) if per_launch else None
    return {
        "bound": "atomic",
        "achieved": round(achieved) if achieved else None,
        "ceiling": round(ceiling),
        "unit": "atomics/s",
        "frac": round(achieved / ceiling, 5) if achieved else None,
        "per_launch": per_launch,
        "ceiling_kernel": f"random 64-bit CAS over {args.rows} x 8 B (gcctb_microbench cas_random)",
    }


def atomics_per_launch():
    """Atomic requests per launch of the execution kernel on this workload,
    from the committed rocprofv3 PMC summary (TCC_ATOMIC_sum, scaled by what
    the counter reads for the random-CAS kernel's known count), or None."""
    path = os.path.join(PROFILES, "txn_exec_config2_atomics.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("atomics_per_launch:"):
            return float(line.split(":", 1)[1].strip().split()[0])
    return None


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


def run_form(ctx, form: str, steps: int):
    tot = kern = 0.0
    launches = aborts = 0
    for _ in range(steps):
        st = ctx.run(form)
        tot += st["tot_ms"]
        kern += st["exec_kernel_ms"]
        launches += st["batches"]
        aborts += st["aborts"]
    return tot, kern, launches, aborts


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _one_pass(scheme, rows, ops, shape, form, *, batch=65536, batch_sizes=None):
    """One timed pass of `form` (preparation included) at the given shape."""
    with g.Context(scheme, rows=rows, warp_density=shape["warp_density"], block_size=shape["block_size"],
                   batch_size=batch, persist_window=shape["window"] if form == "persistent" else 0) as ctx:
        ctx.seed_table()
        ctx.load_ops(ops)
        if batch_sizes is not None:
            ctx.set_batches(batch_sizes)
        st = ctx.run(form)
    return st


def stream_leg(args, table, schemes):
    """The 256-batch stream: persistent (tuned stream shape) against the plain
    per-batch launches at the faster of that shape and the plain form's own,
    per scheme."""
    t0 = time.time()
    s = workloads.config4(rows=args.rows)
    _progress(f"stream: {s.ntx} tx, {s.ops.nops} ops, {len(s.batch_sizes)} batches generated in {time.time() - t0:.1f} s")
    out = {}
    for sc in schemes:
        shape = autotune.lookup(sc, 0.5, 0.9, table, key="stream")
        # the plain launches on the faster of two shapes: the persistent
        # form's and the plain form's own at moderate skew
        own = autotune.lookup_plain(sc, 0.6, 0.9, table)
        p = _one_pass(sc, args.rows, s.ops, shape, "persistent", batch=1 << 30, batch_sizes=s.batch_sizes)
        q = _one_pass(sc, args.rows, s.ops, shape, "plain", batch=1 << 30, batch_sizes=s.batch_sizes)
        if (own["warp_density"], own["block_size"]) != (shape["warp_density"], shape["block_size"]):
            q2 = _one_pass(sc, args.rows, s.ops, dict(own, window=0), "plain", batch=1 << 30,
                           batch_sizes=s.batch_sizes)
            q = q2 if q2["tot_ms"] < q["tot_ms"] else q
        pv, qv = s.ntx / (p["tot_ms"] / 1e3), s.ntx / (q["tot_ms"] / 1e3)
        out[sc] = {"persistent_txps": round(pv), "plain_txps": round(qv), "ratio": round(pv / qv, 3),
                   "deferred": p["deferred"], "combined_ops": p["combined"], "plain_launches": q["batches"], "wd": shape["warp_density"], "bs": shape["block_size"],
                   "window": shape["window"]}
        _progress(f"stream {sc}: persistent {pv / 1e6:.2f} M, plain {qv / 1e6:.2f} M Tx/s ({pv / qv:.2f}x)")
    ratios = [v["ratio"] for v in out.values()]
    return {"workload": f"256 batches log-uniform in [1024, 262144] tx, 4-32 ops, theta U[0, 0.99], "
                        f"read ratio 0.5/0.9/1.0, {args.rows} rows", "ntx": s.ntx, "nops": s.ops.nops,
            "per_scheme": out, "min_ratio": min(ratios) if ratios else None}


def theta99_leg(args, table, schemes):
    """Theta 0.99, 1,048,576 tx x 16 ops: persistent at the tuned theta-0.99
    shape against the plain form, per scheme, at read ratio 0.9 and 0.5: at
    batch 65,536 (the reference's per-batch structure) on whichever launch
    shape runs it faster, the persistent form's or its own tuned one (both
    reported), and at the plain form's own tuned shape and batch size (best
    against best).  The persistent pass includes its
    hot-key handling, counted per scheme: transactions deferred behind the
    others (2PL, T/O, MVCC) and ops run by GaccO's row combiners."""
    out = {}
    for rr in (0.9, 0.5):
        w = g.generate(args.rows, args.txs, theta=0.99, read_ratio=rr, len_min=16)
        res = {}
        for sc in schemes:
            shape = autotune.lookup(sc, 0.99, rr, table)
            pshape = autotune.lookup_plain(sc, 0.99, rr, table)
            p = _one_pass(sc, args.rows, w, shape, "persistent")
            q = _one_pass(sc, args.rows, w, shape, "plain")
            same = (pshape["warp_density"], pshape["block_size"]) == (shape["warp_density"], shape["block_size"])
            q2 = q if same else _one_pass(sc, args.rows, w, dict(pshape, window=0), "plain")
            t = _one_pass(sc, args.rows, w, dict(pshape, window=0), "plain", batch=pshape["batch_size"])
            pv, qv, q2v, tv = (w.ntx / (x["tot_ms"] / 1e3) for x in (p, q, q2, t))
            # the plain form at batch 65,536 at whichever launch shape suits it
            # better: the persistent form's, or its own tuned one
            bv = max(qv, q2v)
            res[sc] = {"persistent_txps": round(pv), "plain_txps": round(bv), "ratio": round(pv / bv, 3),
                       "plain_same_shape_txps": round(qv), "plain_own_shape_txps": round(q2v),
                       "plain_tuned_txps": round(tv), "ratio_vs_tuned_plain": round(pv / tv, 3),
                       "persistent_aborts": p["aborts"], "plain_aborts": q["aborts"],
                       "deferred": p["deferred"], "combined_ops": p["combined"],
                       "wd": shape["warp_density"], "bs": shape["block_size"], "window": shape["window"],
                       "plain_tuned": [pshape["warp_density"], pshape["block_size"], pshape["batch_size"]]}
            _progress(f"theta99 r{rr} {sc}: persistent {pv / 1e6:.2f} M, plain {bv / 1e6:.2f} M Tx/s ({pv / bv:.2f}x; "
                      f"same shape {qv / 1e6:.2f} M, own shape {q2v / 1e6:.2f} M), plain tuned {tv / 1e6:.2f} M ({pv / tv:.2f}x)")
        out[f"read_{rr}"] = {"per_scheme": res, "min_ratio": min(v["ratio"] for v in res.values()) if res else None,
                             "min_ratio_vs_tuned_plain": (min(v["ratio_vs_tuned_plain"] for v in res.values())
                                                          if res else None)}
    return {"workload": f"{args.rows} rows, {args.txs} tx x 16 ops, theta 0.99, batch 65536 (plain tuned: its own "
                        f"batch size)", **out}


SERVE_POOL_BYTES = 32 << 30  # the serving configuration's MVCC version-pool cap


def serving(args, table):
    """p50/p99 commit latency of the serving form (continuous batching) at
    each --serve-loads fraction of the measured peak (peak = every request
    present at once)."""
    # bounded sample: the first batches holding about --serve-txs requests
    s = workloads.serving_stream(args.serve_rows, 4096, 256, 65536, theta=0.9, read_ratio=0.8,
                                 max_txs=args.serve_txs)
    sizes, ops, ntx, nb = s.batch_sizes, s.ops, s.ntx, len(s.batch_sizes)
    loads = [float(x) for x in args.serve_loads.split(",") if x]
    out = {"workload": f"{args.serve_rows} rows, {ntx} requests of 16 ops (theta 0.9, read ratio 0.8) in {nb} "
                       f"batches of log-uniform size in [256, 65536], Poisson arrivals, MVCC pool cap 32 GB"}
    for scheme in [x for x in args.serve_schemes.split(",") if x]:
        shape = autotune.lookup(scheme, 0.9, 0.8, table)
        with g.Context(scheme, rows=args.serve_rows, warp_density=shape["warp_density"],
                       block_size=shape["block_size"], batch_size=1 << 30,
                       persist_window=shape["window"]) as ctx:
            ctx.seed_table()
            ctx.load_ops(ops)
            ctx.serve(sizes, pool_cap_bytes=SERVE_POOL_BYTES)  # warm-up
            peak = ctx.serve(sizes, pool_cap_bytes=SERVE_POOL_BYTES)
            res = {"peak_txps": round(peak["txps"]), "loads": {}}
            for load in loads:
                arr = workloads.poisson_arrivals(sizes, load * peak["txps"])
                st = ctx.serve(sizes, arr, pool_cap_bytes=SERVE_POOL_BYTES)
                res["loads"][str(load)] = {
                    "offered_txps": round(load * peak["txps"]), "txps": round(st["txps"]),
                    "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                    "max_ms": round(st["max_ms"], 3), "launches": st["launches"],
                    "admitted_live": st["admitted_live"], "deferred": st["deferred"]}
                _progress(f"serve {scheme} load {load}: p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
                          f"{st['launches']} launch(es), {st['admitted_live']} batches joined a running launch")
        out[scheme] = res
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the library has no multi-GPU form.",
              file=sys.stderr)
        sys.exit(2)
    schemes = [s for s in args.schemes.split(",") if s]
    table = autotune.load_table()
    t0 = time.time()
    w = g.generate(args.rows, args.txs, theta=args.theta, read_ratio=args.read_ratio, len_min=args.ops)
    nreads = int(np.count_nonzero(~g.op_is_write(w.ops)))
    gen_s = time.time() - t0
    info = g.device_info()
    per = {}
    b_ms = b_kern = p_ms = p_kern = 0.0
    b_launch = 0
    for s in schemes:
        shape = autotune.lookup(s, args.theta, args.read_ratio, table)
        wd = args.wd if args.wd >= 0 else shape["warp_density"]
        bs = args.bs if args.bs >= 0 else shape["block_size"]
        # the plain form runs at its own best low-contention shape
        pwd, pbs = (wd, bs) if args.wd >= 0 or args.bs >= 0 else autotune.DEFAULT_SHAPE[s]
        plain_here = not args.no_plain and (pwd, pbs) == (wd, bs)
        sp_ms = sp_kern = 0.0
        with g.Context(s, rows=args.rows, warp_density=wd, block_size=bs, batch_size=args.batch,
                       persist_window=shape["window"]) as ctx:
            ctx.seed_table()
            ctx.load_ops(w)
            for _ in range(args.warmup):
                ctx.run("persistent")
                if plain_here:
                    ctx.run("plain")
            sb_ms, sb_kern, sb_launch, sb_abort = run_form(ctx, "persistent", args.steps)
            if plain_here:
                sp_ms, sp_kern, _, _ = run_form(ctx, "plain", args.steps)
        if not args.no_plain and not plain_here:
            with g.Context(s, rows=args.rows, warp_density=pwd, block_size=pbs, batch_size=args.batch) as ctx:
                ctx.seed_table()
                ctx.load_ops(w)
                for _ in range(args.warmup):
                    ctx.run("plain")
                sp_ms, sp_kern, _, _ = run_form(ctx, "plain", args.steps)
        _progress(f"bench workload {s}: batched {args.txs * args.steps / (sb_ms / 1e3) / 1e6:.1f} M Tx/s")
        per[s] = {
            "batched_txps": round(args.txs * args.steps / (sb_ms / 1e3)),
            "plain_txps": round(args.txs * args.steps / (sp_ms / 1e3)) if sp_ms else None,
            "batched_ms_per_pass": round(sb_ms / args.steps, 3),
            "kernel_ms_per_pass": round(sb_kern / args.steps, 3),
            "plain_kernel_ms_per_pass": round(sp_kern / args.steps, 3) if sp_ms else None,
            "aborts_per_pass": sb_abort // args.steps,
            "wd": wd, "bs": bs, "window": shape["window"], "plain_wd": pwd, "plain_bs": pbs,
        }
        b_ms += sb_ms
        b_kern += sb_kern
        b_launch += sb_launch
        p_ms += sp_ms
        p_kern += sp_kern
    committed = args.txs * args.steps * len(schemes)
    value = committed / (b_ms / 1e3)
    bytes_launch = algorithmic_bytes(args.txs, w.nops, nreads)
    avg_launch_ms = b_kern / max(1, b_launch)
    achieved = bytes_launch / (avg_launch_ms / 1e3) / 1e9
    p_achieved = bytes_launch * args.steps * len(schemes) / (p_kern / 1e3) / 1e9 if p_kern else 0.0
    line = {
        "metric": "committed Tx/s on 1 MI355X (batched vs plain) + p99 commit latency",
        "value": round(value),
        "unit": "Tx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (deterministic Zipf YCSB generator, seeded table)",
        "config": {
            "workload": f"ycsb low contention, batched form: {args.rows} rows x 10 u64, {args.txs} tx x {args.ops} ops, "
                        f"theta {args.theta}, read ratio {args.read_ratio}, batch {args.batch}, "
                        f"schemes {'+'.join(schemes)}",
            "rows": args.rows, "txs": args.txs, "ops_per_tx": args.ops, "theta": args.theta,
            "read_ratio": args.read_ratio, "batch_size": args.batch, "schemes": schemes,
        },
        "plain": {"value": round(committed / (p_ms / 1e3)) if p_ms else None,
                  "batched_over_plain": round(p_ms / b_ms, 3) if p_ms else None,
                  # the same algorithmic bytes per pass, over the plain form's kernel time
                  # (every scheme's launches of one pass together)
                  "roofline": ({"bound": "hbm", "achieved": round(p_achieved, 2), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(p_achieved / HBM_PEAK_GBS, 5),
                                "kernel_ms_per_pass": round(p_kern / args.steps / len(schemes), 4)}
                               if p_kern else None)},
        "per_scheme": per,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic_per_launch(),
            "kernel": "txn_exec, persistent form (one launch per scheme per pass)",
            "bytes_per_launch": bytes_launch,
            "avg_launch_ms": round(avg_launch_ms, 4),
        },
        "atomic_roofline": atomic_roofline(args, avg_launch_ms),
        "device": info["name"],
        "gen_s": round(gen_s, 2),
    }
    leg_schemes = [x for x in args.leg_schemes.split(",") if x]
    if not args.no_stream:
        line["stream"] = stream_leg(args, table, leg_schemes)
    if not args.no_theta99:
        line["theta99"] = theta99_leg(args, table, leg_schemes)
    if not args.no_serve:
        line["latency"] = serving(args, table)
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(w, args, schemes)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
