"""Summaries of rocprofv3 CSV output for the profiles of this package.

Usage:
  python -m cm1_amd.profsum trace <run_kernel_stats.csv>   # per-kernel time table
  python -m cm1_amd.profsum pmc <fetch counter csv> <write counter csv>
      # per-dispatch HBM bytes, calibrated on the plain copies that
      # ``bench.py --copy-calibration`` runs first (1 GiB each)

Kernel names are shortened to the device function each launcher runs
(``k_one<NT, F>`` and ``k_one_w<NT, W, F>``: ``F``'s host function and
lambda number)."""
import csv
import re
import sys
from collections import defaultdict


def short(name):
    """k_one<256, for3<cm1g_solve2::{lambda(int, int, int)#1}>(...)> -> for3<cm1g_solve2 λ1>."""
    m = re.match(r"void k_one(?:_w)?<\d+, (?:\d+, )?(.*)>\(", name)
    if not m:
        return name.split("(")[0]
    body = m.group(1).replace("(anonymous namespace)::", "")
    # for3<host::{lambda ...#n}>(...)::{lambda#1}  or  host(...)::{lambda(...)#n}
    f = re.match(r"(for[23])<([\w:]+?(?:<\d+>)?)(?:\(.*?\))?::\{lambda\(int, int(?:, int)?\)#(\d+)\}", body)
    if f:
        return f"{f.group(1)}<{f.group(2).split('::')[-1]} λ{f.group(3)}>"
    t = re.match(r"(\w+)<([^>]*)>\(", body)
    if t and t.group(1) in ("advs_launch", "advm_launch"):
        return {"advs_launch": "advs_col", "advm_launch": "advm_col"}[t.group(1)] + f"<{t.group(2)}>"
    h = re.match(r"([\w:]+)(?:\(.*?\))?::\{lambda\(int, int(?:, int)?\)#(\d+)\}", body)
    if h:
        host = h.group(1).split("::")[-1]
        names = {("sound_run", "1"): "sound_coef", ("sound_run", "2"): "sound_sub",
                 ("advs_run", "1"): "pdef_xy", ("advs_run", "2"): "pdef_z",
                 ("cm1g_solve2", "2"): "massflux_div", ("advm_run", "1"): "ten_edges"}
        return names.get((host, h.group(2)), f"{host} λ{h.group(2)}")
    return body.split("(")[0]


def trace(path):
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | dispatches | total ms | avg us | share |")
    print("|---|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        print(f"| `{short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {100.0 * float(r['TotalDurationNs']) / tot:.1f}% |")
    print(f"\nWhole process kernel time {tot / 1e6:.1f} ms.")


def counters(path):
    per = defaultdict(list)
    for r in csv.DictReader(open(path)):
        per[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return per


def pmc(fpath, wpath):
    f, w = counters(fpath), counters(wpath)
    gib = 1024.0 ** 3
    # calibration: the copies of 1 GiB the bench ran first
    cf = [v for n, vs in f.items() if "copyBuffer" in n for v in vs if v * 1024 > 0.25 * gib]
    cw = [v for n, vs in w.items() if "copyBuffer" in n for v in vs if v * 1024 > 0.5 * gib]
    rf = (gib / (min(cf) * 1024)) if cf else 1.0
    rw = (gib / (min(cw) * 1024)) if cw else 1.0
    print(f"Calibration: a plain device copy of 1 GiB reads FETCH_SIZE {min(cf) * 1024 / gib:.3f} GiB and "
          f"WRITE_SIZE {min(cw) * 1024 / gib:.3f} GiB, so bytes = FETCH_SIZE x {rf:.3f} + WRITE_SIZE x {rw:.3f} "
          f"(KiB counters x 1024).\n")
    rows = []
    for n in f:
        if n not in w or "k_one" not in n:
            continue
        rd = sum(f[n]) / len(f[n]) * 1024 * rf / 1e9
        wr = sum(w[n]) / len(w[n]) * 1024 * rw / 1e9
        rows.append((short(n), len(f[n]), rd, wr))
    print("| kernel | dispatches | read GB | write GB | traffic GB / dispatch |")
    print("|---|---:|---:|---:|---:|")
    for n, c, rd, wr in sorted(rows, key=lambda r: -(r[2] + r[3])):
        if rd + wr < 0.05:
            continue
        print(f"| `{n}` | {c} | {rd:.3f} | {wr:.3f} | {rd + wr:.3f} |")
    for n, c, rd, wr in rows:
        if n == "sound_sub":
            print(f"\ntraffic sound_sub bytes_per_launch={(rd + wr) * 1e9:.1f}")


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(sys.argv[2])
    else:
        pmc(sys.argv[2], sys.argv[3])
