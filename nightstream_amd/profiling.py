"""Summaries of rocprofv3 output for the prover's kernels (text tables kept beside the code).

    python -m nightstream_amd.profiling stats  <kernel_stats.csv> --cmd "<command>"
    python -m nightstream_amd.profiling traffic <fetch_counters.csv> <write_counters.csv> \\
        <copy_fetch_counters.csv> <copy_write_counters.csv> --copy-bytes N --cmd "<command>"

`traffic` turns the per-dispatch FETCH_SIZE / WRITE_SIZE counters (KiB) into HBM bytes per launch.
Each counter is scaled by the ratio (bytes a plain element-wise copy kernel moves) / (what the same
counter reads for it), measured in the same way on the same GPU.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys


def short_name(name: str) -> str:
    n = name.strip().strip('"')
    if "rocprim" in n:  # library kernels: the algorithm's name is enough
        m = re.search(r"(radix_sort_block_sort|radix_sort|merge_sort_block_merge|merge_sort|scan|transform|"
                      r"init_lookback)", n)
        return "rocprim " + (m.group(1) if m else "kernel")
    n = re.sub(r"^void\s+", "", n)
    n = n.replace("ns::", "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def stats_table(path: str, cmd: str, top: int = 25) -> str:
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["# Kernel time summary (rocprofv3 --kernel-trace --stats)", "", f"Command: `{cmd}`", "",
           "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        out.append(f"| {short_name(r['Name'])} | {int(r['Calls'])} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['Percentage']):.2f} |")
    out += ["", f"Total kernel time: {total / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} dispatches."]
    return "\n".join(out) + "\n"


def _per_kernel(path: str, counter: str):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            agg[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]) * 1024.0)
    return agg


def traffic_table(fetch: str, write: str, copy_fetch: str, copy_write: str, copy_bytes: float, cmd: str,
                  copy_cmd: str, top: int = 20) -> str:
    cf = max(_per_kernel(copy_fetch, "FETCH_SIZE").items(), key=lambda kv: sum(kv[1]))
    cw = max(_per_kernel(copy_write, "WRITE_SIZE").items(), key=lambda kv: sum(kv[1]))
    f_scale = copy_bytes / (sum(cf[1]) / len(cf[1]))
    w_scale = copy_bytes / (sum(cw[1]) / len(cw[1]))
    F = _per_kernel(fetch, "FETCH_SIZE")
    W = _per_kernel(write, "WRITE_SIZE")
    out = ["# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE)", "",
           f"Command (one run per counter): `{cmd}`", "",
           f"Calibration: `{copy_cmd}` -- kernel `{cf[0]}` reads and writes {copy_bytes:.0f} B per launch; "
           f"FETCH_SIZE reads {sum(cf[1]) / len(cf[1]):.0f} B (scale {f_scale:.4f}), "
           f"WRITE_SIZE reads {sum(cw[1]) / len(cw[1]):.0f} B (scale {w_scale:.4f}).", "",
           "| kernel | launches | fetch B/launch | write B/launch | traffic B/launch |", "|---|---|---|---|---|"]
    names = sorted(set(F) | set(W), key=lambda k: -(sum(F.get(k, [0])) * f_scale + sum(W.get(k, [0])) * w_scale))
    for k in names[:top]:
        fv, wv = F.get(k, [0.0]), W.get(k, [0.0])
        fa, wa = sum(fv) / len(fv) * f_scale, sum(wv) / len(wv) * w_scale
        out.append(f"| {k} | {len(fv)} | {fa:.0f} | {wa:.0f} | {fa + wa:.0f} |")
    return "\n".join(out) + "\n"


def load_traffic(md_path: str, kernels):
    """Traffic bytes per launch of a kernel class from a table written by traffic_table, or None.
    `kernels` is one kernel name or several (a class with more than one kernel): the launch-
    weighted mean over the rows found."""
    if isinstance(kernels, str):
        kernels = (kernels,)
    try:
        text = open(md_path).read()
    except OSError:
        return None
    launches = total = 0.0
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] in kernels:
            try:
                n, b = float(cells[1]), float(cells[4])
            except ValueError:
                return None
            launches += n
            total += n * b
    return total / launches if launches else None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("copy_fetch")
    t.add_argument("copy_write")
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--cmd", required=True)
    t.add_argument("--copy-cmd", required=True)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        sys.stdout.write(stats_table(a.csv, a.cmd))
    else:
        sys.stdout.write(traffic_table(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, a.cmd, a.copy_cmd))


if __name__ == "__main__":
    main()
