"""Turn rocprofv3 CSV output into the short text summaries kept in profiles/.

    python -m us_lab4a_amd.profsum trace KERNEL_STATS_CSV OUT.md --title T --frames F --command CMD
    python -m us_lab4a_amd.profsum sq OUT.md DIR [DIR ...] --kernel SUBSTR --units N --title T --command CMD
    python -m us_lab4a_amd.profsum traffic FETCH_DIR WRITE_DIR OUT.md --kernel SUBSTR
            --calib SUBSTR --calib-bytes N --title T --command CMD

trace:   per-kernel calls / average / total / per-frame time from
         rocprofv3 --kernel-trace --stats (<prefix>_kernel_stats.csv).
sq:      instruction mix of the kernel matching --kernel from rocprofv3 --pmc
         runs of SQ counters (any number of directories), per 64 of
         --units (pixel-pairs per dispatch); writes a line "per_wave_pair:
         valu=.. salu=.. lds=.. lds_conflict=.." that bench.py reads.
traffic: HBM-side bytes of one dispatch of the kernel matching --kernel from
         two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch),
         each scaled by what the counter reads for a calibration kernel
         (--calib) that streams a known number of bytes (--calib-bytes) in and
         out.  Writes a line "traffic_bytes_per_launch: N" that bench.py reads.
"""
import argparse
import csv
import glob
import os
import sys


def _clean(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("mista::", "")
    return name.split("(")[0]


def trace(args):
    rows = list(csv.DictReader(open(args.csv)))
    out = [f"# {args.title}", "", "Command (repo root, MI355X):", "", f"    {args.command}", ""]
    if args.note:
        out += [args.note, ""]
    out += ["| kernel | calls | avg (ms) | total (ms) | per frame (ms) |", "|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        tot = float(r["TotalDurationNs"]) / 1e6
        n = int(r["Calls"])
        per = f"{tot / args.frames:.3f}" if n >= args.frames else "-"
        out.append(f"| `{_clean(r['Name'])[:90]}` | {n} | {float(r['AverageNs']) / 1e6:.3f} | {tot:.2f} | {per} |")
    open(args.out, "w").write("\n".join(out) + "\n")


def _pmc(d, counter):
    files = glob.glob(os.path.join(d, "*counter_collection.csv"))
    if not files:
        sys.exit(f"no counter_collection.csv under {d}")
    vals = {}
    rows = [r for r in csv.DictReader(open(files[0])) if r["Counter_Name"] == counter]
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return vals


def _pick(vals, sub):
    hits = [(k, v) for k, v in vals.items() if sub in k]
    if not hits:
        sys.exit(f"no kernel matching {sub!r}")
    k, v = hits[0]
    return k, sum(v) / len(v)


def traffic(args):
    fetch = _pmc(args.fetch_dir, "FETCH_SIZE")
    write = _pmc(args.write_dir, "WRITE_SIZE")
    name, f_kib = _pick(fetch, args.kernel)
    _, w_kib = _pick(write, args.kernel)
    _, cf_kib = _pick(fetch, args.calib)
    _, cw_kib = _pick(write, args.calib)
    f_scale = args.calib_bytes / (cf_kib * 1024.0)
    w_scale = args.calib_bytes / (cw_kib * 1024.0)
    fb = f_kib * 1024.0 * f_scale
    wb = w_kib * 1024.0 * w_scale
    out = [f"# {args.title}", "", "Commands (repo root, MI355X; one counter per run):", ""]
    out += [f"    {c}" for c in args.command]
    out += ["", f"Kernel: `{_clean(name)}`", "",
            "| counter | KiB per dispatch (raw) | scale (calibration) | bytes per dispatch |", "|---|---|---|---|",
            f"| FETCH_SIZE | {f_kib:.1f} | {f_scale:.4f} | {fb:.4g} |",
            f"| WRITE_SIZE | {w_kib:.1f} | {w_scale:.4f} | {wb:.4g} |", "",
            f"Calibration: `{_clean(_pick(fetch, args.calib)[0])}` streams {args.calib_bytes} bytes in and the same "
            f"out; it reads {cf_kib:.1f} KiB FETCH_SIZE and {cw_kib:.1f} KiB WRITE_SIZE per dispatch.", ""]
    if args.algorithmic:
        out += [f"Algorithmic bytes per dispatch: {args.algorithmic} (traffic / algorithmic = "
                f"{(fb + wb) / args.algorithmic:.2f}).", ""]
    out += [f"traffic_bytes_per_launch: {int(round(fb + wb))}"]
    open(args.out, "w").write("\n".join(out) + "\n")


def sq(args):
    vals = {}
    name = None
    for d in args.dirs:
        files = glob.glob(os.path.join(d, "*counter_collection.csv"))
        if not files:
            sys.exit(f"no counter_collection.csv under {d}")
        per = {}
        for r in csv.DictReader(open(files[0])):
            if args.kernel not in r["Kernel_Name"]:
                continue
            name = r["Kernel_Name"]
            per.setdefault(r["Counter_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
            per[r["Counter_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
        for c, disp in per.items():
            vals[c] = sum(disp.values()) / len(disp)
    if not vals:
        sys.exit(f"no kernel matching {args.kernel!r}")
    wp = args.units / 64.0
    out = [f"# {args.title}", "", "Commands (repo root, MI355X; one counter group per run; per-dispatch means):", ""]
    out += [f"    {c}" for c in args.command]
    out += ["", f"Kernel: `{_clean(name)}`", "", f"Per wave-pair = per 64 of {args.units:.4g} units per dispatch.", "",
            "| counter | per dispatch | per wave-pair |", "|---|---|---|"]
    for c in sorted(vals):
        out.append(f"| {c} | {vals[c]:.4g} | {vals[c] / wp:.4g} |")
    key = {"valu": "SQ_INSTS_VALU", "salu": "SQ_INSTS_SALU", "lds": "SQ_INSTS_LDS", "lds_conflict": "SQ_LDS_BANK_CONFLICT",
           "smem": "SQ_INSTS_SMEM", "vmem": "SQ_INSTS_VMEM", "wait_any": "SQ_WAIT_ANY", "wave_cycles": "SQ_WAVE_CYCLES"}
    kv = " ".join(f"{k}={vals[c] / wp:.4g}" for k, c in key.items() if c in vals)
    if args.note:
        out += ["", args.note]
    out += ["", f"per_wave_pair: {kv}"]
    open(args.out, "w").write("\n".join(out) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("out")
    t.add_argument("--title", required=True)
    t.add_argument("--frames", type=int, required=True)
    t.add_argument("--command", required=True)
    t.add_argument("--note", default="")
    r = sub.add_parser("traffic")
    r.add_argument("fetch_dir")
    r.add_argument("write_dir")
    r.add_argument("out")
    r.add_argument("--kernel", required=True)
    r.add_argument("--calib", required=True)
    r.add_argument("--calib-bytes", type=int, required=True)
    r.add_argument("--algorithmic", type=int, default=0)
    r.add_argument("--title", required=True)
    r.add_argument("--command", action="append", required=True)
    q = sub.add_parser("sq")
    q.add_argument("out")
    q.add_argument("dirs", nargs="+")
    q.add_argument("--kernel", required=True)
    q.add_argument("--units", type=float, required=True)
    q.add_argument("--title", required=True)
    q.add_argument("--command", action="append", required=True)
    q.add_argument("--note", default="")
    args = ap.parse_args()
    {"trace": trace, "traffic": traffic, "sq": sq}[args.cmd](args)


if __name__ == "__main__":
    main()
