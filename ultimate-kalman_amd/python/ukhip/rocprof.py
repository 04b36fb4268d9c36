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
    python -m ukhip.rocprof counters DIR [DIR ...] [--cus 256] [--xcds 8]
        per-kernel utilisation from SQ/GRBM counter runs (each DIR one
        rocprofv3 --pmc run; the kernels' dispatches summed over the runs'
        counters).

Counter values are in KiB per dispatch. The calibration factor of a counter is
B / (1024 * the counter's value on the copy kernel). The HIP runtime's own fill
kernels (the calibration's buffer initialisation) are left out of the tables.
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
        if k == COPY_KERNEL or k.startswith("at::") or k.startswith("__amd_rocclr_"):
            continue
        fb = fs[k][1] * 1024.0 * cf / per
        wb = ws.get(k, [0, 0.0])[1] * 1024.0 * cw / per
        out[k] = {"fetch_bytes": fb, "write_bytes": wb, "bytes_per_step": fb + wb, "launches_per_step": fs[k][0] / per}
    return out, cf, cw


def counter_table(dirs, cus=256, xcds=8):
    """Per kernel (short name): dispatches, device ms (dispatch timestamps) and
    utilisation figures from the SQ/GRBM counters found in `dirs`:
      valu_busy   100 * SQ_ACTIVE_INST_VALU / CUs / GRBM cycles (rocprof's
                  VALUBusy: VALU issue against one wave64 instruction per
                  4 cycles per SIMD);
      waves_cu    resident waves per CU, 4 * SQ_WAVE_CYCLES / (GRBM cycles * CUs);
      lds_util    100 * SQ_LDS_IDX_ACTIVE / (GRBM cycles * CUs);
      bank_conf   SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - conflicts);
      mfma_busy   100 * SQ_VALU_MFMA_BUSY_CYCLES / (GRBM cycles * CUs);
      valu_per_wave, lds_per_wave, fma_share (FMA F32 + F64 of all VALU).
    GRBM_GUI_ACTIVE is reported summed over the XCDs, so GRBM cycles are its
    value / xcds."""
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    for d in dirs:
        seen = set()
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = short_name(r["Kernel_Name"])
                e = acc[k]
                e[r["Counter_Name"] + "@" + d] += float(r["Counter_Value"])
                key = (d, r["Dispatch_Id"])
                if key not in seen:
                    seen.add(key)
                    e["dispatches@" + d] += 1
                    e["ns@" + d] += float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
    out = {}
    for k, e in acc.items():
        def tot(c):
            v = [x for kk, x in e.items() if kk.split("@")[0] == c]
            return sum(v) if v else None

        def per_dir(c, d):
            return e.get(c + "@" + d)

        row = {"dispatches": int(max(e.get("dispatches@" + d, 0) for d in dirs)),
               "ms": max(e.get("ns@" + d, 0.0) for d in dirs) / 1e6}
        for d in dirs:  # ratios within one run (its own GRBM cycles)
            g = per_dir("GRBM_GUI_ACTIVE", d)
            if not g:
                continue
            cyc = g / xcds
            for name, c, scale in (("valu_busy", "SQ_ACTIVE_INST_VALU", 100.0 / cus),
                                   ("waves_cu", "SQ_WAVE_CYCLES", 4.0 / cus),
                                   ("lds_util", "SQ_LDS_IDX_ACTIVE", 100.0 / cus),
                                   ("mfma_busy", "SQ_VALU_MFMA_BUSY_CYCLES", 100.0 / cus)):
                v = per_dir(c, d)
                if v is not None:
                    row[name] = v * scale / cyc
            w, vi, li = per_dir("SQ_WAVES", d), per_dir("SQ_INSTS_VALU", d), per_dir("SQ_INSTS_LDS", d)
            if w:
                if vi is not None:
                    row["valu_per_wave"] = vi / w
                if li is not None:
                    row["lds_per_wave"] = li / w
            bc, ia = per_dir("SQ_LDS_BANK_CONFLICT", d), per_dir("SQ_LDS_IDX_ACTIVE", d)
            if bc is not None and ia:
                row["bank_conf"] = bc / max(ia - bc, 1.0)
        f32, f64, vall = tot("SQ_INSTS_VALU_FMA_F32"), tot("SQ_INSTS_VALU_FMA_F64"), tot("SQ_INSTS_VALU")
        if f32 is not None and f64 is not None and vall:
            row["fma_share"] = (f32 + f64) / vall
        out[k] = row
    return out


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
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--xcds", type=int, default=8)
    c.add_argument("--top", type=int, default=24)
    args = ap.parse_args(argv)
    if args.cmd == "counters":
        tab = counter_table(args.dirs, args.cus, args.xcds)
        cols = ("valu_busy", "waves_cu", "lds_util", "bank_conf", "mfma_busy", "valu_per_wave", "lds_per_wave",
                "fma_share")
        print(f"{'kernel':44s} {'disp':>5s} {'ms':>8s} " + " ".join(f"{c:>13s}" for c in cols))
        rows = [(k, v) for k, v in tab.items() if not k.startswith("__amd_rocclr_")]
        for k, v in sorted(rows, key=lambda kv: -kv[1]["ms"])[:args.top]:
            cells = " ".join(f"{'' if v.get(c) is None else format(v[c], '.3f' if c in ('bank_conf', 'fma_share') else '.1f'):>13s}"
                             for c in cols)
            print(f"{k[:44]:44s} {v['dispatches']:5d} {v['ms']:8.2f} {cells}")
        return 0
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
