"""Summarise rocprofv3 output as short markdown tables.

    python -m sparsews_amd.profsum STATS_CSV --title T --command "CMD" [--top N]
    python -m sparsews_amd.profsum --trace RESULTS_DB --title T --command "CMD" [--top N]
    python -m sparsews_amd.profsum --pmc RESULTS_DB [RESULTS_DB ...] --title T --command "CMD"
    python -m sparsews_amd.profsum --traffic FETCH_DB WRITE_DB --kernel K [--copy-kernel C]
                                   [--copy-bytes N] [--launch-name L] --title T --command "CMD"

The second form reads the counter values of ``rocprofv3 --pmc`` runs from
their ``*_results.db`` files and prints, per kernel, the dispatch count and
the mean and total of each counter (FETCH_SIZE / WRITE_SIZE are in KiB per
dispatch).  The third turns a FETCH_SIZE run and a WRITE_SIZE run into HBM
bytes per launch of one kernel, calibrated by a plain device copy of known
size in the same runs; its last line (``bench-traffic: ...``) is what
bench.py reports as ``roofline.traffic``.
"""
from __future__ import annotations

import argparse
import csv
import re


def short_name(name: str) -> str:
    """Kernel name without return type, namespaces and parameter list."""
    name = name.replace("(anonymous namespace)::", "")
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


def trace_summary(db_path: str, title: str, command: str, top: int = 15) -> str:
    """Per-kernel calls / average / total from a rocprofv3 --kernel-trace
    results db (the same figures as its kernel_stats.csv)."""
    import sqlite3
    con = sqlite3.connect(db_path)
    q = ("select s.display_name, count(*), avg(d.end - d.start), sum(d.end - d.start) "
         "from rocpd_kernel_dispatch d join rocpd_info_kernel_symbol s on d.kernel_id = s.id "
         "group by s.display_name order by 4 desc")
    rows = list(con.execute(q))
    total = sum(r[3] for r in rows) or 1
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             "| kernel | calls | avg (us) | total (ms) | % |", "|---|---:|---:|---:|---:|"]
    for name, calls, avg, tot in rows[:top]:
        lines.append(f"| `{short_name(name)}` | {calls} | {avg / 1e3:.2f} | {tot / 1e6:.3f} | {100.0 * tot / total:.1f} |")
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


def traffic_summary(fetch_db: str, write_db: str, kernel: str, copy_kernel: str, copy_bytes: int,
                    title: str, command: str, workload: str = "", launch_name: str = "") -> str:
    """HBM bytes per launch of `kernel` from a FETCH_SIZE run and a WRITE_SIZE
    run, each scaled by what the counter reads for a plain copy of
    `copy_bytes` bytes (which reads and writes exactly that many bytes).
    `launch_name` (default: the kernel name) is the name the library's own
    per-kernel timer gives the launch, written in the bench-traffic line."""
    f = pmc_per_kernel(fetch_db)
    w = pmc_per_kernel(write_db)
    fk = f[(kernel, "FETCH_SIZE")]
    wk = w[(kernel, "WRITE_SIZE")]
    fc = f[(copy_kernel, "FETCH_SIZE")]
    wc = w[(copy_kernel, "WRITE_SIZE")]
    copy_kib = copy_bytes / 1024.0
    f_scale = copy_kib / (fc[1] / fc[0])
    w_scale = copy_kib / (wc[1] / wc[0])
    fetch_b = fk[1] / fk[0] * f_scale * 1024.0
    write_b = wk[1] / wk[0] * w_scale * 1024.0
    lines = [f"# {title}", "", "Commands:", "", "```", command, "```", "",
             "| | dispatches | counter mean per dispatch (KiB) | scale (copy calibration) | bytes per launch |",
             "|---|---:|---:|---:|---:|",
             f"| `{kernel}` FETCH_SIZE | {fk[0]} | {fk[1] / fk[0]:.1f} | {f_scale:.4f} | {fetch_b:.4g} |",
             f"| `{kernel}` WRITE_SIZE | {wk[0]} | {wk[1] / wk[0]:.1f} | {w_scale:.4f} | {write_b:.4g} |",
             f"| `{copy_kernel}` FETCH_SIZE ({copy_bytes} B copy) | {fc[0]} | {fc[1] / fc[0]:.1f} | | |",
             f"| `{copy_kernel}` WRITE_SIZE ({copy_bytes} B copy) | {wc[0]} | {wc[1] / wc[0]:.1f} | | |",
             "",
             f"bench-traffic: workload={workload} kernel={launch_name or kernel} bytes_per_launch={fetch_b + write_b:.0f} "
             f"fetch_bytes={fetch_b:.0f} write_bytes={write_b:.0f}"]
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("stats_csv", nargs="?")
    ap.add_argument("--pmc", nargs="+", default=None)
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DB", "WRITE_DB"), default=None)
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--copy-kernel", default="k_copy16")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--workload", default="")
    ap.add_argument("--launch-name", default="", help="the library timer's name for the kernel's launches")
    ap.add_argument("--trace", default=None, help="rocprofv3 --kernel-trace results db")
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--top", type=int, default=15)
    ap.add_argument("--note", default="")
    a = ap.parse_args(argv)
    if a.trace:
        text = trace_summary(a.trace, a.title, a.command, a.top)
    elif a.traffic:
        text = traffic_summary(a.traffic[0], a.traffic[1], a.kernel, a.copy_kernel, a.copy_bytes, a.title, a.command,
                               a.workload, a.launch_name)
    elif a.pmc:
        text = summarise_pmc(a.pmc, a.title, a.command)
    else:
        text = summarise(a.stats_csv, a.title, a.command, a.top)
    if a.note:
        text += "\n" + a.note + "\n"
    print(text, end="")


if __name__ == "__main__":
    main()
