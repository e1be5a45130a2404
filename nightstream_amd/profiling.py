"""Summaries of rocprofv3 output for the prover's kernels (text tables kept beside the code).

    python -m nightstream_amd.profiling stats  <kernel_stats.csv> --cmd "<command>" [--trace <kernel_trace.csv>]
    python -m nightstream_amd.profiling traffic <fetch_counters.csv> <write_counters.csv> \\
        <copy_fetch_counters.csv> <copy_write_counters.csv> --copy-bytes N --cmd "<command>"
    python -m nightstream_amd.profiling sq <results.db> --cmd "<command>"

`stats --trace` adds the median duration of each fused Pi_CCS round (k_round_fused launches in
order, --rounds-per-step = ell_n - 3 of them per fold step).  `sq` tabulates the SQ counters (waves, wave cycles, issue and
wait cycles, VALU instructions) of each fused round from a rocpd database.

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


def fused_rounds_line(trace_csv: str, period: int, name: str = "k_round_fused", first: int = 4) -> str:
    """Median duration of each fused Pi_CCS round over the fold steps of a kernel trace (the
    fused launches in start order, `period` of them per fold step: ell_n - 4 from 2^9 rows, where
    rounds 2 and 3 are not fused, else ell_n - 3)."""
    import statistics
    rows = [r for r in csv.DictReader(open(trace_csv)) if short_name(r["Kernel_Name"]) == name]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows]
    per = collections.defaultdict(list)
    for i, v in enumerate(d):
        per[i % period].append(v)
    return (f"{name} per round (median over {len(d) // period} steps, us): " +
            ", ".join(f"round {first + i}: {statistics.median(v):.1f}" for i, v in sorted(per.items())))


def tail_line(trace_csv: str, prefix: str, count: int) -> str:
    """Average duration of the last `count` dispatches of the kernels whose short name starts with
    `prefix`: the timed region of a run whose first commitments of each shape are the tuner's
    tile-shape and K-split trials."""
    rows = [r for r in csv.DictReader(open(trace_csv)) if short_name(r["Kernel_Name"]).startswith(prefix)]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    last = rows[-count:]
    d = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in last]
    names = collections.Counter(short_name(r["Kernel_Name"]) for r in last)
    return (f"{prefix}* over the last {len(d)} dispatches (the timed region): avg {sum(d) / max(1, len(d)):.1f} us "
            f"(min {min(d):.1f}, max {max(d):.1f}); " + ", ".join(f"{k} x{v}" for k, v in sorted(names.items())))


def sq_table(db: str, cmd: str, name: str = "k_round_fused", first: int = 4) -> str:
    """SQ counters of each fused round (the last fold step of the run), per wave."""
    import sqlite3
    cur = sqlite3.connect(db).cursor()
    d = collections.defaultdict(dict)
    dur = {}
    for disp, kname, cname, val, du in cur.execute(
            "select dispatch_id, kernel_name, counter_name, value, duration from counters_collection"):
        if short_name(kname) == name:
            d[disp][cname] = d[disp].get(cname, 0) + val
            dur[disp] = du
    ids = sorted(d)
    # the launches of one step: from the last launch whose wave count is the maximum onward
    w = [d[k].get("SQ_WAVES", 0) for k in ids]
    last = max(i for i in range(len(ids)) if w[i] == max(w))
    step = ids[last:]
    out = [f"# SQ counters per fused Pi_CCS round ({name}, one fold step with k = 13 claims)", "",
           f"Command: `{cmd}`", "",
           "Cycles are shader clocks (SQ quad-cycles x 4), per wave.", "",
           "| round | waves | duration us | wave cycles | issuing any / VALU | waiting (any) | waiting (inst) | "
           "VALU insts |", "|---|---|---|---|---|---|---|---|"]
    for i, k in enumerate(step):
        c = d[k]
        nw = max(c.get("SQ_WAVES", 1), 1)
        wc = 4 * c.get("SQ_WAVE_CYCLES", 0) / nw
        act, valu = 4 * c.get("SQ_ACTIVE_INST_ANY", 0) / nw, 4 * c.get("SQ_ACTIVE_INST_VALU", 0) / nw
        out.append(f"| {first + i} | {nw:.0f} | {dur[k] / 1e3:.1f} | {wc:.0f} | {100 * act / max(wc, 1):.0f}% / "
                   f"{100 * valu / max(wc, 1):.0f}% | {4 * c.get('SQ_WAIT_ANY', 0) / nw:.0f} | "
                   f"{4 * c.get('SQ_WAIT_INST_ANY', 0) / nw:.0f} | {c.get('SQ_INSTS_VALU', 0) / nw:.0f} |")
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
    s.add_argument("--trace", default=None)
    s.add_argument("--rounds-per-step", type=int, default=12)
    s.add_argument("--first-round", type=int, default=4)
    s.add_argument("--tail-prefix", default=None)
    s.add_argument("--tail-count", type=int, default=17)
    q = sub.add_parser("sq")
    q.add_argument("db")
    q.add_argument("--cmd", required=True)
    q.add_argument("--first-round", type=int, default=4)
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
        if a.trace:
            sys.stdout.write("\n" + fused_rounds_line(a.trace, a.rounds_per_step, first=a.first_round) + "\n")
            if a.tail_prefix:
                sys.stdout.write("\n" + tail_line(a.trace, a.tail_prefix, a.tail_count) + "\n")
    elif a.mode == "sq":
        sys.stdout.write(sq_table(a.db, a.cmd, first=a.first_round))
    else:
        sys.stdout.write(traffic_table(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, a.cmd, a.copy_cmd))


if __name__ == "__main__":
    main()
