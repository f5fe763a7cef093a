"""Text summaries of rocprofv3 CSV output for the profiles/ folder.

    python -m flutas_amd.profsum stats DIR [--command CMD]
    python -m flutas_amd.profsum traffic FETCH_DIR WRITE_DIR --copy-kernel S
                                 --copy-bytes B [--command CMD]

`stats` reads the `*kernel_stats.csv` that `rocprofv3 --kernel-trace --stats
--output-format csv` writes and prints a table of calls and average duration
per kernel.

`traffic` reads the `*counter_collection.csv` of two `--pmc` runs (one with
FETCH_SIZE, one with WRITE_SIZE; the counters are in KiB per dispatch),
calibrates each counter on a plain device copy of known size (the dispatches
whose kernel name contains S; B bytes read and B bytes written per dispatch)
and prints, per kernel, the scaled HBM bytes per launch.  The last line is
`TRAFFIC_JSON {...}` (kernel name -> read/write/total bytes per launch) for
programs that read the summary back.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict


def _one(pattern):
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise SystemExit(f"profsum: nothing matches {pattern}")
    return hits[0]


def short_name(name):
    """Kernel name without the argument list and namespaces."""
    base = name.split("(")[0].strip()
    if base.startswith("void "):
        base = base[5:]
    base = re.sub(r"\(anonymous namespace\)::", "", base)
    return base.replace("flutas::", "")


def stats(args):
    path = _one(os.path.join(args.dir, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(path)))
    out = []
    if args.command:
        out.append(f"command: {args.command}")
    out.append(f"source: rocprofv3 kernel stats ({os.path.basename(path)})")
    out.append("")
    out.append(f"{'calls':>6} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'pct':>6}  kernel")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        g = lambda key: float(r.get(key) or 0.0)
        out.append(f"{int(g('Calls')):6d} {g('AverageNs') / 1e3:10.1f} "
                   f"{g('MinNs') / 1e3:10.1f} {g('MaxNs') / 1e3:10.1f} "
                   f"{g('Percentage'):6.2f}  {short_name(r['Name'])}")
    print("\n".join(out))


def _counter_means(d, counter):
    path = _one(os.path.join(d, "**", "*counter_collection.csv"))
    acc = defaultdict(list)
    grid = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        name = r["Kernel_Name"]
        acc[name].append(float(r["Counter_Value"]) * 1024.0)
        grid[name] = max(grid.get(name, 0), int(r["Grid_Size"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def traffic(args):
    fetch, nf = _counter_means(args.fetch_dir, "FETCH_SIZE")
    write, nw = _counter_means(args.write_dir, "WRITE_SIZE")
    copy_f = [v for k, v in fetch.items() if args.copy_kernel in k]
    copy_w = [v for k, v in write.items() if args.copy_kernel in k]
    if not copy_f or not copy_w:
        raise SystemExit(f"profsum: no dispatch of a kernel named *{args.copy_kernel}*")
    sf = args.copy_bytes / max(copy_f)
    sw = args.copy_bytes / max(copy_w)
    out = []
    if args.command:
        out.append(f"command: {args.command}")
    out.append(f"calibration: copy of {args.copy_bytes} B read + {args.copy_bytes} B written "
               f"(kernel *{args.copy_kernel}*): FETCH_SIZE scale {sf:.4f}, WRITE_SIZE scale {sw:.4f}")
    out.append("")
    out.append(f"{'n':>4} {'fetch_KiB':>14} {'write_KiB':>14} {'read_B':>14} {'write_B':>14} {'total_B':>14}  kernel")
    table = {}
    for k in sorted(set(fetch) & set(write), key=lambda k: -(fetch[k] * sf + write[k] * sw)):
        rb, wb = fetch[k] * sf, write[k] * sw
        s = short_name(k)
        out.append(f"{nf[k]:4d} {fetch[k] / 1024:14.1f} {write[k] / 1024:14.1f} "
                   f"{rb:14.0f} {wb:14.0f} {rb + wb:14.0f}  {s}")
        table[s] = {"read": rb, "write": wb, "total": rb + wb}
    out.append("")
    out.append("TRAFFIC_JSON " + json.dumps(table, sort_keys=True))
    print("\n".join(out))


def read_traffic(path):
    """kernel name -> {read, write, total} bytes per launch from a summary."""
    with open(path) as fh:
        for line in fh:
            if line.startswith("TRAFFIC_JSON "):
                return json.loads(line[len("TRAFFIC_JSON "):])
    return {}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m flutas_amd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--command", default="")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--copy-kernel", required=True)
    b.add_argument("--copy-bytes", type=float, required=True)
    b.add_argument("--command", default="")
    args = ap.parse_args(argv)
    {"stats": stats, "traffic": traffic}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
