"""Summaries of rocprofv3 runs: counters for the compositor roofline, kernel-trace statistics.

    python -m svraster_amd.profiling compositors --out FILE.json \
        --calib-fetch DIR --calib-write DIR --run DIR [--run DIR ...] [--command TEXT ...]
    python -m svraster_amd.profiling stats --run DIR --out FILE.txt [--title TEXT] [--command TEXT ...]

Each DIR is the output directory of one `rocprofv3 --pmc ... --output-format csv -d DIR` run.
The counter means per dispatch are merged over the runs, per kernel.  HBM bytes per launch are
FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same counters read for the plain
1 GiB copy of `bench.py --copy-calibration` (the calibration runs): bytes = FETCH_SIZE * 1024 *
(copy bytes read / copy FETCH bytes) + WRITE_SIZE * 1024 * (copy bytes written / copy WRITE bytes).
VALU instructions per launch are SQ_INSTS_VALU.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import sys

COPY_BYTES = 1 << 30
COPY_KERNEL = "MulFunctor"   # the copy launches of `bench.py --copy-calibration` (not its fill)


def counter_means(run_dir: str, name_filter: str = "") -> dict:
    """{kernel name (without arguments): {counter: mean value per dispatch, "dispatches": n}}."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                k = row.get("Kernel_Name", "")
                if name_filter not in k:
                    continue
                acc[k.split("(")[0].replace("void ", "").replace("svr::", "")][row["Counter_Name"]].append(
                    float(row["Counter_Value"]))
    out = {}
    for k, cs in acc.items():
        out[k] = {c: sum(v) / len(v) for c, v in cs.items()}
        out[k]["dispatches"] = max(len(v) for v in cs.values())
    return out


def copy_scale(run_dir: str, counter: str) -> float:
    """Bytes per counter unit for the 1 GiB copy kernel (counter in KiB per dispatch)."""
    for k, cs in counter_means(run_dir).items():
        if COPY_KERNEL in k and counter in cs and cs[counter] > 0:
            return COPY_BYTES / (cs[counter] * 1024.0)
    raise RuntimeError(f"no {counter} of the copy kernel under {run_dir}")


def compositors(args) -> dict:
    fscale = copy_scale(args.calib_fetch, "FETCH_SIZE")
    wscale = copy_scale(args.calib_write, "WRITE_SIZE")
    merged = collections.defaultdict(dict)
    for d in args.run:
        for k, cs in counter_means(d, "svr::").items():
            merged[k].update(cs)
    kernels = {}
    for k, cs in sorted(merged.items()):
        e = {c: (int(v) if float(v).is_integer() else v) for c, v in cs.items()}
        if "FETCH_SIZE" in cs and "WRITE_SIZE" in cs:
            e["hbm_bytes_per_launch"] = int(cs["FETCH_SIZE"] * 1024.0 * fscale + cs["WRITE_SIZE"] * 1024.0 * wscale)
        if "SQ_INSTS_VALU" in cs:
            e["valu_insts_per_launch"] = int(cs["SQ_INSTS_VALU"])
        kernels[k] = e
    return {"commands": args.command, "fetch_scale_bytes_per_kib": fscale * 1024.0,
            "write_scale_bytes_per_kib": wscale * 1024.0, "kernels": kernels}


def kernel_stats(args) -> str:
    """Text table of a `rocprofv3 --kernel-trace --stats --output-format csv` run: per kernel the
    number of dispatches, total and mean duration (ms) and share of the kernel time."""
    paths = glob.glob(os.path.join(args.run, "**", "*kernel_stats.csv"), recursive=True)
    if not paths:
        raise RuntimeError(f"no kernel_stats.csv under {args.run}")
    rows = []
    with open(paths[0]) as f:
        for r in csv.DictReader(f):
            rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]) * 1e-6,
                         float(r["AverageNs"]) * 1e-6, float(r["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    lines = [f"# rocprofv3 kernel statistics: {args.title}"]
    lines += [f"# command: {c}" for c in args.command]
    lines.append("")
    lines.append(f"{'kernel':<96} {'calls':>7} {'total_ms':>10} {'avg_ms':>9} {'pct':>6}")
    for name, calls, tot, avg, pct in rows[:args.top]:
        lines.append(f"{name[:96]:<96} {calls:>7} {tot:>10.3f} {avg:>9.4f} {pct:>6.2f}")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    c = sub.add_parser("compositors")
    c.add_argument("--out", required=True)
    c.add_argument("--calib-fetch", required=True)
    c.add_argument("--calib-write", required=True)
    c.add_argument("--run", action="append", required=True)
    c.add_argument("--command", action="append", default=[])
    t = sub.add_parser("stats")
    t.add_argument("--run", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--title", default="")
    t.add_argument("--top", type=int, default=40)
    t.add_argument("--command", action="append", default=[])
    args = ap.parse_args(argv)
    if args.what == "stats":
        text = kernel_stats(args)
        with open(args.out, "w") as f:
            f.write(text)
        print(f"wrote {args.out}", file=sys.stderr)
        return
    res = compositors(args)
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {args.out}: {len(res['kernels'])} kernels", file=sys.stderr)


if __name__ == "__main__":
    main()
