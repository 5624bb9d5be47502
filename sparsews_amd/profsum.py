"""Summarise rocprofv3 output as short markdown tables.

    python -m sparsews_amd.profsum STATS_CSV --title T --command "CMD" [--top N]
    python -m sparsews_amd.profsum --pmc RESULTS_DB [RESULTS_DB ...] --title T --command "CMD"

The second form reads the counter values of ``rocprofv3 --pmc`` runs from
their ``*_results.db`` files and prints, per kernel, the dispatch count and
the mean and total of each counter (FETCH_SIZE / WRITE_SIZE are in KiB per
dispatch).
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


def pmc_per_kernel(db_path: str) -> dict:
    """{(kernel, counter): [dispatches, total]} from a rocprofv3 results db."""
    import sqlite3
    con = sqlite3.connect(db_path)
    q = ("select s.display_name, i.name, e.value from rocpd_pmc_event e "
         "join rocpd_info_pmc i on e.pmc_id = i.id "
         "join rocpd_kernel_dispatch d on d.event_id = e.event_id "
         "join rocpd_info_kernel_symbol s on d.kernel_id = s.id")
    out = {}
    for kname, counter, value in con.execute(q):
        slot = out.setdefault((short_name(kname), counter), [0, 0.0])
        slot[0] += 1
        slot[1] += float(value)
    return out


def summarise_pmc(paths, title: str, command: str) -> str:
    agg = {}
    for p in paths:
        agg.update(pmc_per_kernel(p))
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             "| kernel | counter | dispatches | mean per dispatch | total |", "|---|---|---:|---:|---:|"]
    for (k, c), (n, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{k}` | {c} | {n} | {tot / n:.1f} | {tot:.1f} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("stats_csv", nargs="?")
    ap.add_argument("--pmc", nargs="+", default=None)
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--top", type=int, default=15)
    ap.add_argument("--note", default="")
    a = ap.parse_args(argv)
    if a.pmc:
        text = summarise_pmc(a.pmc, a.title, a.command)
    else:
        text = summarise(a.stats_csv, a.title, a.command, a.top)
    if a.note:
        text += "\n" + a.note + "\n"
    print(text, end="")


if __name__ == "__main__":
    main()
