"""Summaries of rocprofv3 output for the committed profiles.

    python3 -m ntustcfdlab_amd.rocprof summarize DIR > SUMMARY.json
        kernel trace (per kernel: calls, mean/min/max/total duration, and the
        span of the trace) and counter collection (per kernel: dispatches and
        the mean of each counter per dispatch; the calibration copies --
        k_stream_copy, made by bench.py --calibrate -- one by one) from the
        CSV files under DIR.

    python3 -m ntustcfdlab_amd.rocprof traffic N READS.json WRITES.json OUT.json [CAP]
        HBM bytes per dispatch on the N^3 cavity from two counter passes
        (READS: TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum;
        WRITES: TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum), summed over the 16
        L2 channels of each of the 8 XCDs.  On gfx950 an RDREQ without a size
        flag is a 128-byte request, so
            read  = 128 (RDREQ - RDREQ_64B - RDREQ_32B) + 64 RDREQ_64B
                    + 32 RDREQ_32B
            write = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B)
        with no scale factor; the decode is accepted ("validated") only when
        it reproduces the calibration copies' sizes within 5 %.  CAP: the
        sweep cap (itmax) of the profiled runs, recorded with the figures --
        a two-pass half-sweep moves the same bytes at any cap, the
        lexicographic step launches do not (the first and last launches of
        a batch hold fewer cells).

    python3 -m ntustcfdlab_amd.rocprof markdown TRACE.json [TRAFFIC.json] [ROWS]
        Markdown tables of a trace summary (the ROWS kernels with the most
        GPU time, default 16) and of a traffic file (bytes per dispatch and
        per cell of the grid).

The raw CSV files stay where rocprofv3 wrote them; only the summaries are
kept.
"""
from __future__ import annotations

import collections
import csv
import glob
import json
import os
import sys

# kernels of a time step whose bytes the summary lists (rocprofv3 names
# carry the namespace and the argument list)
STEP_KERNELS = (
    "k_rb_sweep32<0, 2>", "k_rb_sweep32<1, 2>", "k_rb_sweep32<0, 1>", "k_rb_sweep32<1, 1>",
    "k_rb_sweep32<0, 0>", "k_rb_sweep32<1, 0>", "k_rb_sweep<0, 2>", "k_rb_sweep<1, 2>",
    "k_rb_sweep<0, 1>", "k_rb_sweep<1, 1>", "k_rb_sweep<0, 0>", "k_rb_sweep<1, 0>",
    "k_lex_sweep<0, false>", "k_lex_sweep<1, false>", "k_lex_sweep<0, true>",
    "k_lex_sweep<1, true>", "k_predict_quick", "k_predict_quick_march<float>",
    "k_smagorinsky", "k_smagorinsky_fast<float>", "k_project", "k_rb_rhs",
    "k_rb_correct<0, true>", "k_rb_correct<1, true>", "k_copy_shell", "k_force_z")

# the headline kernels' per-launch entries: both colours averaged
HEADLINE = {
    "k_rb_sweep32<C>": ("k_rb_sweep32<0, 2>", "k_rb_sweep32<1, 2>"),
    "k_rb_sweep<C>": ("k_rb_sweep<0, 2>", "k_rb_sweep<1, 2>"),
    "k_lex_sweep<C>": ("k_lex_sweep<0, false>", "k_lex_sweep<1, false>"),
}


def summarize(d: str) -> dict:
    out: dict = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        agg = collections.defaultdict(list)
        span = []
        with open(f) as fh:
            for r in csv.DictReader(fh):
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                agg[r["Kernel_Name"]].append((t1 - t0) / 1e3)
                span.append((t0, t1))
        out["trace"] = {k: {"calls": len(v), "avg_us": sum(v) / len(v), "min_us": min(v),
                            "max_us": max(v), "total_ms": sum(v) / 1e3}
                        for k, v in agg.items()}
        out["span_ms"] = ((max(b for _, b in span) - min(a for a, _ in span)) / 1e6
                          if span else None)
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        bydisp = collections.defaultdict(dict)
        names = {}
        for r in rows:
            did = r.get("Dispatch_Id") or r.get("Correlation_Id")
            c = r["Counter_Name"]
            bydisp[did][c] = bydisp[did].get(c, 0.0) + float(r["Counter_Value"])
            names[did] = r["Kernel_Name"]
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        copies = []
        for did, cs in bydisp.items():
            for c, v in cs.items():
                agg[names[did]][c].append(v)
            if "k_stream_copy" in names[did]:
                copies.append({"dispatch": did, **cs})
        out["counters"] = {k: {"dispatches": len(next(iter(v.values()))),
                               **{c: sum(x) / len(x) for c, x in v.items()}}
                           for k, v in agg.items()}
        out["copies"] = copies
    return out


def read_bytes(c: dict) -> float:
    return (128.0 * (c["TCC_EA0_RDREQ_sum"] - c["TCC_EA0_RDREQ_64B_sum"]
                     - c["TCC_EA0_RDREQ_32B_sum"])
            + 64.0 * c["TCC_EA0_RDREQ_64B_sum"] + 32.0 * c["TCC_EA0_RDREQ_32B_sum"])


def write_bytes(c: dict) -> float:
    return 64.0 * c["TCC_EA0_WRREQ_64B_sum"] + 32.0 * (c["TCC_EA0_WRREQ_sum"]
                                                       - c["TCC_EA0_WRREQ_64B_sum"])


def traffic(n: int, reads: dict, writes: dict, cap: float | None = None) -> dict:
    cells = n ** 3
    cal = []
    for a, b in zip(reads["copies"], writes["copies"]):
        # copy sizes are whole MiB: the nearest one to the decoded read
        nbytes = round(read_bytes(a) / (1 << 20)) << 20
        cal.append({"bytes_each_way": nbytes, "read_bytes": read_bytes(a),
                    "write_bytes": write_bytes(b), "read_err": read_bytes(a) / nbytes - 1,
                    "write_err": write_bytes(b) / nbytes - 1})
    validated = bool(cal) and all(abs(c["read_err"]) < 0.05 and abs(c["write_err"]) < 0.05
                                  for c in cal)
    per = {}
    for k, c in reads["counters"].items():
        for key in STEP_KERNELS:
            if ("ntust::" + key + "(") in k and k in writes["counters"]:
                cw = writes["counters"][k]
                per[key] = {"dispatches": int(c["dispatches"]), "read_bytes": read_bytes(c),
                            "write_bytes": write_bytes(cw),
                            "read_B_per_cell": read_bytes(c) / cells,
                            "write_B_per_cell": write_bytes(cw) / cells}
    head = {}
    for name, (a, b) in HEADLINE.items():
        if a in per and b in per:
            head[name] = {"read_bytes": (per[a]["read_bytes"] + per[b]["read_bytes"]) / 2,
                          "write_bytes": (per[a]["write_bytes"] + per[b]["write_bytes"]) / 2}
    return {"workload": f"LES cavity Re=1000 {n}^3", "cells": cells, "validated": validated,
            "sweep_cap": cap,
            "decode": ("read = 128 B x (TCC_EA0_RDREQ - _64B - _32B) + 64 B x _64B + 32 B x "
                       "_32B; write = 64 B x TCC_EA0_WRREQ_64B + 32 B x the rest; all _sum "
                       "over 16 channels x 8 XCDs; no scale factor"),
            "calibration_copy": cal, "per_launch": head, "kernels": per}


def _short(name: str) -> str:
    name = name.split("(")[0]
    for pre in ("void ", "ntust::"):
        name = name.replace(pre, "")
    return name


def markdown(trace: dict, traffic_: dict | None = None, rows: int = 16) -> str:
    out = []
    tr = trace.get("trace", {})
    tot = sum(v["total_ms"] for v in tr.values()) or 1.0
    out.append("| kernel | calls | avg µs | min µs | max µs | total ms | % of GPU time |")
    out.append("|---|---:|---:|---:|---:|---:|---:|")
    for k, v in sorted(tr.items(), key=lambda kv: -kv[1]["total_ms"])[:rows]:
        out.append(f"| `{_short(k)}` | {v['calls']} | {v['avg_us']:.2f} | {v['min_us']:.2f} | "
                   f"{v['max_us']:.2f} | {v['total_ms']:.1f} | {100 * v['total_ms'] / tot:.1f} |")
    if trace.get("span_ms"):
        out.append("")
        out.append(f"Trace span {trace['span_ms']:.0f} ms; kernels busy {tot:.0f} ms.")
    if traffic_:
        out.append("")
        out.append("| calibration copy MiB each way | read MB | write MB | read err | write err |")
        out.append("|---:|---:|---:|---:|---:|")
        for c in traffic_["calibration_copy"]:
            out.append(f"| {c['bytes_each_way'] >> 20} | {c['read_bytes'] / 1e6:.2f} | "
                       f"{c['write_bytes'] / 1e6:.2f} | {c['read_err']:+.1e} | "
                       f"{c['write_err']:+.1e} |")
        out.append("")
        out.append("| kernel | dispatches | read MB | write MB | read B/cell | write B/cell |")
        out.append("|---|---:|---:|---:|---:|---:|")
        for k, v in traffic_["kernels"].items():
            out.append(f"| `{k}` | {v['dispatches']} | {v['read_bytes'] / 1e6:.2f} | "
                       f"{v['write_bytes'] / 1e6:.2f} | {v['read_B_per_cell']:.2f} | "
                       f"{v['write_B_per_cell']:.2f} |")
    return "\n".join(out)


def main(argv):
    if len(argv) >= 2 and argv[0] == "summarize":
        json.dump(summarize(argv[1]), sys.stdout, indent=1)
        print()
        return 0
    if len(argv) >= 5 and argv[0] == "traffic":
        cap = float(argv[5]) if len(argv) > 5 else None
        out = traffic(int(argv[1]), json.load(open(argv[2])), json.load(open(argv[3])), cap)
        with open(argv[4], "w") as fh:
            json.dump(out, fh, indent=1)
            fh.write("\n")
        for c in out["calibration_copy"]:
            print(json.dumps(c))
        for k, v in out["kernels"].items():
            print(f"{k:30s} {v['dispatches']:6d} read {v['read_B_per_cell']:7.2f} "
                  f"write {v['write_B_per_cell']:7.2f} B/cell")
        return 0 if out["validated"] else 1
    if len(argv) >= 2 and argv[0] == "markdown":
        tr = json.load(open(argv[1]))
        tf = json.load(open(argv[2])) if len(argv) > 2 and argv[2] != "-" else None
        print(markdown(tr, tf, int(argv[3]) if len(argv) > 3 else 16))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
