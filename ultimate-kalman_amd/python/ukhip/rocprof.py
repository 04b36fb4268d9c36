"""Summaries of rocprofv3 CSV output for libukhip's kernels.

    python -m ukhip.rocprof stats DIR [--per N]
        per-kernel launch counts and durations from DIR/run_kernel_stats.csv
        or DIR/run_results.db (--per N divides totals by N smoothings)
    python -m ukhip.rocprof traffic --fetch DIR1 --write DIR2 --copy-bytes B
            [--per N] [--workload NAME --json PATH]
        HBM bytes per kernel from a FETCH_SIZE run (DIR1) and a WRITE_SIZE run
        (DIR2), each calibrated by what the counter reads for the plain copy
        kernel of B bytes that the profiled program ran (bench.py
        --copy-calibration); --json merges the result into PATH under NAME.

Counter values are in KiB per dispatch. The calibration factor of a counter is
B / (1024 * the counter's value on the copy kernel).
"""
import argparse
import collections
import csv
import json
import os
import re
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"^void ", "", name)
    return re.sub(r"\(.*", "", name).strip()


def base_name(name):
    """The kernel without its launch-variant template arguments
    ("k_phase1<48, double, 2, 2, true>" -> "k_phase1<48, double>"): the
    variants move the same bytes, so traffic is keyed by the base name."""
    m = re.match(r"^(k_phase\d<\d+, (?:double|float))(?:, [^>]*)?>$", name)
    return m.group(1) + ">" if m else name


def kernel_stats(d):
    """(name, calls, total ns, average ns, percent) per kernel, from
    run_kernel_stats.csv or, for rocprofv3's SQLite output, run_results.db."""
    rows = []
    csv_path = os.path.join(d, "run_kernel_stats.csv")
    if os.path.exists(csv_path):
        with open(csv_path) as f:
            for r in csv.DictReader(f):
                rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                             float(r["AverageNs"]), float(r["Percentage"])))
        return rows
    import sqlite3
    db = os.path.join(d, "run_results.db")
    con = sqlite3.connect(db)
    acc = collections.defaultdict(lambda: [0, 0.0])
    for name, dur in con.execute("select name, duration from kernels"):
        e = acc[short_name(name)]
        e[0] += 1
        e[1] += float(dur)
    total = sum(v[1] for v in acc.values()) or 1.0
    for name, (calls, tot) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        rows.append((name, calls, tot, tot / calls, 100.0 * tot / total))
    return rows


def counter_sums(d):
    s = collections.defaultdict(lambda: [0, 0.0])
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            e = s[base_name(short_name(r["Kernel_Name"]))]
            e[0] += 1
            e[1] += float(r["Counter_Value"])
    return s


def traffic(fetch_dir, write_dir, copy_bytes, per=1):
    fs, ws = counter_sums(fetch_dir), counter_sums(write_dir)
    cf = copy_bytes / (1024.0 * fs[COPY_KERNEL][1])
    cw = copy_bytes / (1024.0 * ws[COPY_KERNEL][1])
    out = {}
    for k in fs:
        if k == COPY_KERNEL or k.startswith("at::"):
            continue
        fb = fs[k][1] * 1024.0 * cf / per
        wb = ws.get(k, [0, 0.0])[1] * 1024.0 * cw / per
        out[k] = {"fetch_bytes": fb, "write_bytes": wb, "bytes_per_step": fb + wb, "launches_per_step": fs[k][0] / per}
    return out, cf, cw


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--per", type=int, default=1)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--per", type=int, default=1)
    t.add_argument("--workload")
    t.add_argument("--json")
    t.add_argument("--algorithmic", action="append", default=[], help="KERNEL=BYTES per step")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(f"{'kernel':40s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'ms/step':>9s} {'%':>6s}")
        rows = kernel_stats(args.dir)
        for name, calls, tot, avg, pct in rows:
            print(f"{name[:40]:40s} {calls:6d} {tot / 1e6:10.3f} {avg / 1e6:9.4f} {tot / 1e6 / args.per:9.3f} {pct:6.2f}")
        # a phase kernel launched with different per-level flags appears once per
        # flag; its launches together are what a per-kernel HIP-event average covers
        groups = collections.defaultdict(lambda: [0, 0.0, 0.0, set()])
        for name, calls, tot, avg, pct in rows:
            m = re.match(r"^(k_phase\d<\d+, (?:double|float), \d+, \d+)(, (?:true|false|\d))?>$", name)
            if m and m.group(2):
                g = groups[m.group(1) + ", *>"]
                g[0] += calls
                g[1] += tot
                g[2] += pct
                g[3].add(name)
        multi = {k: v for k, v in groups.items() if len(v[3]) > 1}
        if multi:
            print("\nall launches of a kernel over its per-level flags:")
            for k, (calls, tot, pct, _) in sorted(multi.items(), key=lambda kv: -kv[1][1]):
                print(f"{k[:40]:40s} {calls:6d} {tot / 1e6:10.3f} {tot / calls / 1e6:9.4f} {tot / 1e6 / args.per:9.3f} {pct:6.2f}")
        return 0
    res, cf, cw = traffic(args.fetch, args.write, args.copy_bytes, args.per)
    alg = dict(x.split("=") for x in args.algorithmic)
    print(f"calibration (copy kernel of {args.copy_bytes:.0f} B): FETCH_SIZE x{cf:.4f}, WRITE_SIZE x{cw:.4f}")
    print(f"{'kernel':28s} {'launches':>8s} {'fetch GB':>10s} {'write GB':>10s} {'total GB':>10s} {'algorithmic GB':>15s}")
    for k, e in sorted(res.items(), key=lambda kv: -kv[1]["bytes_per_step"]):
        if k in alg:
            e["algorithmic_bytes_per_step"] = float(alg[k])
        a_ = e.get("algorithmic_bytes_per_step")
        print(f"{k[:28]:28s} {e['launches_per_step']:8.0f} {e['fetch_bytes'] / 1e9:10.3f} {e['write_bytes'] / 1e9:10.3f} "
              f"{e['bytes_per_step'] / 1e9:10.3f} {'' if a_ is None else f'{a_ / 1e9:15.3f}'}")
    if args.json and args.workload:
        data = {}
        if os.path.exists(args.json):
            with open(args.json) as f:
                data = json.load(f)
        data[args.workload] = {k: dict(v, fetch_factor=cf, write_factor=cw) for k, v in res.items()}
        with open(args.json, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
