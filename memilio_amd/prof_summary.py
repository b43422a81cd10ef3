"""Summarise a rocprofv3 kernel trace (CSV output) per kernel and launch shape.

    rocprofv3 --kernel-trace --stats --output-format csv -d DIR -o NAME -- python3 bench.py
    python3 -m memilio_amd.prof_summary DIR/NAME_kernel_trace.csv [--top 20] > summary.md

The same kernel is launched with different grids by different parts of a run (a 1024-sample ensemble
and a single 400-node graph both launch node_phase<6, 0>); the table groups dispatches by (kernel, grid
size) so that the average duration of one launch shape can be compared with the timing a benchmark
measures for that shape. A second table gives the per-kernel totals over all shapes, as --stats does.

PMC mode sums a counter per (kernel, grid) from a `--pmc` run's counter_collection.csv:

    python3 -m memilio_amd.prof_summary DIR/NAME_counter_collection.csv --pmc FETCH_SIZE
"""
import argparse
import collections
import csv
import re
import sys


def short_name(name):
    n = re.sub(r"\(.*\)$", "", name.strip())
    n = n.replace("void ", "").replace("mgo::", "")
    return n


def _col(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    raise KeyError(f"none of {names} in trace columns {sorted(row)}")


def trace_summary(path, top):
    groups = collections.defaultdict(list)
    totals = collections.defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = short_name(_col(row, "Kernel_Name", "KernelName"))
            t = (int(_col(row, "End_Timestamp", "EndNs")) - int(_col(row, "Start_Timestamp", "BeginNs"))) * 1e-3
            grid = int(_col(row, "Grid_Size_X", "Grid_Size", "grd"))
            wg = int(_col(row, "Workgroup_Size_X", "Workgroup_Size", "wgr"))
            groups[(name, grid // max(wg, 1), wg)].append(t)
            totals[name].append(t)
    all_us = sum(sum(v) for v in totals.values())
    out = ["| kernel | blocks x threads | calls | total ms | avg us | min us | max us | % time |",
           "|---|---|---|---|---|---|---|---|"]
    rows = sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:top]
    for (name, blocks, wg), ts in rows:
        s = sum(ts)
        out.append(f"| `{name}` | {blocks} x {wg} | {len(ts)} | {s / 1e3:.2f} | {s / len(ts):.2f} | {min(ts):.2f} | "
                   f"{max(ts):.2f} | {100 * s / all_us:.1f} |")
    out += ["", "| kernel (all shapes) | calls | total ms | avg us | % time |", "|---|---|---|---|---|"]
    for name, ts in sorted(totals.items(), key=lambda kv: -sum(kv[1]))[:top]:
        s = sum(ts)
        out.append(f"| `{name}` | {len(ts)} | {s / 1e3:.2f} | {s / len(ts):.2f} | {100 * s / all_us:.1f} |")
    out.append("")
    out.append(f"total kernel time {all_us / 1e3:.1f} ms over {sum(len(v) for v in totals.values())} dispatches")
    return "\n".join(out)


def pmc_summary(path, counter, top):
    groups = collections.defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if _col(row, "Counter_Name", "counter") != counter:
                continue
            name = short_name(_col(row, "Kernel_Name", "KernelName"))
            grid = int(_col(row, "Grid_Size", "Grid_Size_X"))
            wg = int(_col(row, "Workgroup_Size", "Workgroup_Size_X"))
            disp = _col(row, "Dispatch_Id", "Correlation_Id")
            groups[(name, grid // max(wg, 1), wg)].append((disp, float(_col(row, "Counter_Value", "value"))))
    out = [f"| kernel | blocks x threads | dispatches | {counter} per dispatch (avg) | min | max |",
           "|---|---|---|---|---|---|"]
    rows = sorted(groups.items(), key=lambda kv: -sum(v for _, v in kv[1]))[:top]
    for (name, blocks, wg), vals in rows:
        per = collections.defaultdict(float)   # a dispatch may report one value per XCD / instance
        for d, v in vals:
            per[d] += v
        vs = list(per.values())
        out.append(f"| `{name}` | {blocks} x {wg} | {len(vs)} | {sum(vs) / len(vs):.1f} | {min(vs):.1f} | {max(vs):.1f} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--pmc", default=None, help="counter name: summarise a counter_collection.csv")
    a = ap.parse_args(argv)
    print(pmc_summary(a.csv, a.pmc, a.top) if a.pmc else trace_summary(a.csv, a.top))


if __name__ == "__main__":
    sys.exit(main())
