"""Summarise rocprofv3 output into the short text tables kept under ``profiles/``.

    python -m hqt.profsummary stats <dir>            # kernel-trace --stats run (CSV or rocpd .db)
    python -m hqt.profsummary pmc <dir> <COUNTER>    # one --pmc run: per-kernel average counter value

Run from ``ads4-holographic-quantum-turbulence_amd`` (or with it on ``sys.path``).
"""

from __future__ import annotations

import csv
import glob
import os
import sqlite3
import sys
from collections import defaultdict


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name.replace("hqt::", "").replace("void ", "").split("(")[0]


def kernel_rows(d: str):
    """(name, calls, total_ns, avg_ns, min_ns, max_ns) per kernel from a kernel-trace run."""
    rows = []
    stats = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    if stats:
        with open(stats[0]) as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                             float(r["MinNs"]), float(r["MaxNs"])))
        return sorted(rows, key=lambda r: -r[2])
    dbs = glob.glob(os.path.join(d, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no kernel stats under {d}")
    cur = sqlite3.connect(dbs[0]).cursor()
    for name, n, tot, avg, mn, mx in cur.execute(
            "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) from kernels "
            "group by name order by sum(duration) desc"):
        rows.append((name, n, float(tot), float(avg), float(mn), float(mx)))
    return rows


def stats_table(d: str) -> str:
    rows = kernel_rows(d)
    tot = sum(r[2] for r in rows) or 1.0
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for name, n, t, a, mn, mx in rows:
        out.append(f"| `{_short(name)}` | {n} | {t / 1e6:.3f} | {a / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                   f"{100 * t / tot:.1f} |")
    return "\n".join(out)


def pmc_rows(d: str, counter: str):
    """{kernel: (dispatches, mean counter value)} from a --pmc run (CSV counter_collection)."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    acc = defaultdict(lambda: [0, 0.0])
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                k = r.get("Kernel_Name", "")
                acc[k][0] += 1
                acc[k][1] += float(r["Counter_Value"])
    return {k: (n, v / n) for k, (n, v) in acc.items() if n}


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(stats_table(argv[1]))
    elif len(argv) >= 3 and argv[0] == "pmc":
        for k, (n, v) in sorted(pmc_rows(argv[1], argv[2]).items(), key=lambda kv: -kv[1][1]):
            print(f"| `{_short(k)}` | {n} | {v:.1f} |")
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
