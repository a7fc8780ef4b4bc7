"""Writes output_lines.json: the driver's result block (the CSV line and the
stdout text of ExecInfo::Analyse, include/runtime.cuh:169-214) for fixed
run statistics.

The expected text is computed here independently of the library: the
reference's arithmetic restated with numpy float32 (inv = 1.0f / (txcnt *
clock_rate) with an integer product, tick sums times inv, float milliseconds
divided by 1000.f, throughput 1000.0f / tot_time * txcnt, the abort average
in double) and C++ iostream's default float output, which is printf's %g at
six significant digits.  The first case is worked by hand (its values are
exact in binary, so the text can be read off directly); the others exercise
rounding, exponents and large counters.

Run from the repository root:  python tests/golden/make_output_golden.py
"""
import json
import os

import numpy as np

F = np.float32


def g6(x) -> str:
    return "%g" % float(x)


def expected(st: dict, cc_type: str, wd: int, bs: int, batch: int, txcnt: int):
    inv = F(1.0) / F(txcnt * int(st["tick_khz"]))
    d = {k: F(st[k + "_ticks"]) * inv for k in ("ts", "wait", "abort", "manager", "tot", "index")}
    pre, proc, tot = F(st["pre_ms"]), F(st["proc_ms"]), F(st["tot_ms"])
    through = F(1000.0) / tot * F(txcnt)
    k1000 = F(1000.0)
    fields = [cc_type, str(wd), str(bs), str(batch), g6(through), str(st["aborts"]), g6(d["ts"]), g6(d["wait"]),
              g6(d["abort"]), g6(d["manager"]), g6(d["tot"]), g6(d["index"]), g6(pre / k1000), g6(proc / k1000),
              g6(tot / k1000)]
    csv = ",".join(fields) + "\n"
    diff = (2 ** 31 - 1) - (st["aborts"] + batch)
    text = (f"{cc_type},valid_txn_bitoffset {wd},\nwarp_cnt {bs},\nbatch_size {batch},\n"
            f"through {g6(through)},\nabort {st['aborts']} diff {diff} avr {g6(st['aborts'] * 1.0 / batch)},\n"
            f"ts_d {g6(d['ts'])},\nwait_d {g6(d['wait'])},\nabort_d {g6(d['abort'])},\n"
            f"mng_d {g6(d['manager'])},\ntot_d {g6(d['tot'])},\nidx_d {g6(d['index'])},\n"
            f"pre {g6(pre / k1000)},\nproc {g6(proc / k1000)},\ntot {g6(tot / k1000)}\n")
    return csv, text


def stats(aborts, ts, wait, abort, mng, tot, idx, khz, pre, proc, tot_ms):
    return {"aborts": aborts, "ts_ticks": ts, "wait_ticks": wait, "abort_ticks": abort, "manager_ticks": mng,
            "tot_ticks": tot, "index_ticks": idx, "tick_khz": khz, "pre_ms": pre, "proc_ms": proc, "tot_ms": tot_ms}


CASES = [
    # hand-worked: 1000 tx at 1000 kHz; 500,000 ticks -> 0.5 ms per tx, 4 s -> 250 Tx/s
    ("tpl_nw", 5, 32, 100, 1000, stats(10, 500000, 250000, 0, 1000000, 2000000, 125000, 1000.0, 1000.0, 2000.0,
                                       4000.0)),
    # a full-size plain run at 100 MHz
    ("silo", 3, 8, 65536, 1048576, stats(123456, 9876543210, 1234567890, 987654321, 55555555555, 77777777777,
                                          3333333333, 100000.0, 0.8125, 3.1415926, 4.2001953)),
    # a high-contention run with many aborts and long waits
    ("mvcc", 2, 4, 4096, 262144, stats(98765432, 7 * 10 ** 12, 3 * 10 ** 12, 11 * 10 ** 11, 9 * 10 ** 12,
                                        2 * 10 ** 13, 10 ** 9, 100000.0, 12.5, 345.678, 358.2)),
    # aborts above 2^31 (negative diff), sub-microsecond figures
    ("gputx", 4, 16, 1024, 3000, stats(3000000000, 17, 3, 0, 101, 250, 9, 1800000.0, 0.0123, 0.456, 0.5)),
    # the reference's own tick rate (clock_rate in kHz of a 2.52 GHz part)
    ("tictoc", 5, 32, 65536, 65536, stats(77, 1234567, 7654321, 111111, 98765432, 123456789, 4242, 2520000.0,
                                          1.25, 19.75, 21.0)),
]


def main():
    out = []
    for cc, wd, bs, batch, txcnt, st in CASES:
        csv, text = expected(st, cc, wd, bs, batch, txcnt)
        out.append({"cc_type": cc, "warp_density": wd, "block_size": bs, "batch_size": batch, "txcnt": txcnt,
                    "stats": st, "csv": csv, "text": text})
    # the hand-worked case, stated literally
    assert out[0]["csv"] == "tpl_nw,5,32,100,250,10,0.5,0.25,0,1,2,0.125,1,2,4\n"
    assert out[0]["text"].startswith("tpl_nw,valid_txn_bitoffset 5,\nwarp_cnt 32,\nbatch_size 100,\nthrough 250,\n"
                                     "abort 10 diff 2147483537 avr 0.1,\nts_d 0.5,\n")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "output_lines.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
