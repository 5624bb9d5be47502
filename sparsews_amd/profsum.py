"""Summarise a rocprofv3 ``*_kernel_stats.csv`` as a short markdown table.

    python -m sparsews_amd.profsum STATS_CSV --title T --command "CMD" [--top N]
"""
from __future__ import annotations

import argparse
import csv
import re


def short_name(name: str) -> str:
    name = re.sub(r"\(.*\)$", "", name)
    name = name.replace("void ", "").replace("spws::", "")
    return name


def summarise(path: str, title: str, command: str, top: int = 15) -> str:
    rows = list(csv.DictReader(open(path)))
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             "| kernel | calls | avg (us) | total (ms) | % |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        lines.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("stats_csv")
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--top", type=int, default=15)
    ap.add_argument("--note", default="")
    a = ap.parse_args(argv)
    text = summarise(a.stats_csv, a.title, a.command, a.top)
    if a.note:
        text += "\n" + a.note + "\n"
    print(text, end="")


if __name__ == "__main__":
    main()
