"""Markdown and JSON summaries of rocprofv3 runs of the hydro stage kernels.

    python -m torch_amd.prof_summary --grid 256 --stats S/bench_kernel_stats.csv \\
        --fetch F/bench_counter_collection.csv --write W/bench_counter_collection.csv \\
        --fetch-copy FC/copy_counter_collection.csv --write-copy WC/copy_counter_collection.csv \\
        --sq Q/bench_counter_collection.csv --commands cmds.txt \\
        --md profiles/hydro_stage_turb256.md --json profiles/hydro_stage_turb256_traffic.json

    python -m torch_amd.prof_summary db RESULTS.db [REGEX]   # kernel table of a rocprofv3 database
    python -m torch_amd.prof_summary stats KERNEL_STATS.csv [REGEX]   # the same from a --stats CSV

Inputs are the CSV files rocprofv3 writes for `--kernel-trace --stats` and
for `--pmc` runs of bench.py and of torch_amd.hbm_probe (a device copy of
--copy-mib MiB, read and written once per dispatch).  FETCH_SIZE and
WRITE_SIZE (KiB per dispatch) are scaled by the ratio of the copy's known
bytes to what the counters read for it.  The SQ run gives instructions per
wave and the VALU issue utilisation SQ_INSTS_VALU / (CUs x GRBM_GUI_ACTIVE /
XCDs): GRBM_GUI_ACTIVE is summed over the XCDs, and a CU issues at most one
wave64 VALU instruction per cycle (four SIMDs, four cycles each).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

KEEP = re.compile(r"hydro_stage|q_float|q_from_u|u_from_q|step_begin|step_end|cfl_kernel|scale_rotate|state_max")


def short(name: str) -> str:
    """Kernel name without its argument list."""
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def counters(path: str):
    d = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(path) as fh:
        for r in csv.DictReader(fh):
            d[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in d.items()}


def copy_factor(path: str, counter: str, copy_bytes: float) -> float:
    c = counters(path)
    vals = [cs[counter] for k, cs in c.items() if "copy" in k.lower() and counter in cs]
    if not vals:
        raise SystemExit(f"no copy kernel with {counter} in {path}")
    return copy_bytes / (max(vals) * 1024.0)


def db_kernel_table(path: str, keep: str = ".") -> list[str]:
    """Markdown rows (calls, total, average, min, max in µs) per kernel from a
    rocprofv3 SQLite database (its `kernels` view), for kernel names matching
    the regular expression `keep`, largest total first."""
    import sqlite3

    rows = sqlite3.connect(path).execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) from kernels "
        "group by name order by sum(duration) desc").fetchall()
    out = ["| kernel | calls | total µs | avg µs | min µs | max µs |", "|---|---|---|---|---|---|"]
    pat = re.compile(keep)
    for name, calls, tot, avg, mn, mx in rows:
        k = short(name)
        if pat.search(k):
            out.append(f"| `{k}` | {calls} | {tot / 1e3:.1f} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} |")
    return out


def stats_kernel_table(path: str, keep: str = ".") -> list[str]:
    """The same rows from the `*_kernel_stats.csv` of a `--kernel-trace
    --stats --output-format csv` run (rocprofv3 sorts it by total time)."""
    out = ["| kernel | calls | total µs | avg µs | min µs | max µs |", "|---|---|---|---|---|---|"]
    pat = re.compile(keep)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            k = short(r["Name"])
            if pat.search(k):
                out.append(f"| `{k}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
                           f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                           f"{float(r['MaxNs']) / 1e3:.2f} |")
    return out


def main():
    if len(sys.argv) > 1 and sys.argv[1] in ("db", "stats"):
        # python -m torch_amd.prof_summary db RESULTS.db [REGEX]
        # python -m torch_amd.prof_summary stats KERNEL_STATS.csv [REGEX]
        table = db_kernel_table if sys.argv[1] == "db" else stats_kernel_table
        print("\n".join(table(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else ".")))
        return
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--grid", type=int, required=True)
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--fetch-copy", required=True)
    ap.add_argument("--write-copy", required=True)
    ap.add_argument("--sq", default=None)
    ap.add_argument("--lds", default=None,
                    help="PMC CSV with SQ_LDS_BANK_CONFLICT, SQ_LDS_IDX_ACTIVE and GRBM_GUI_ACTIVE")
    ap.add_argument("--copy-mib", type=float, default=1024.0)
    ap.add_argument("--cus", type=int, default=256)
    ap.add_argument("--xcds", type=int, default=8)
    ap.add_argument("--commands", default=None, help="text file with the commands that made the inputs")
    ap.add_argument("--algorithmic", default="{}",
                    help='JSON {kernel: bytes per cell} for the algorithmic-bytes column')
    ap.add_argument("--workload", default="sigma_ln rho = 1, rms Mach 2",
                    help="field statistics of the turbulent box the runs used")
    ap.add_argument("--driver", default="bench.py", help="the program the runs profiled")
    ap.add_argument("--md", required=True)
    ap.add_argument("--json", required=True)
    a = ap.parse_args()
    ncell = a.grid ** 3
    algo = json.loads(a.algorithmic)
    copy_bytes = a.copy_mib * (1 << 20)

    lines = [f"# hydro stage kernels, {a.grid}^3 turbulent box: rocprofv3 summaries", ""]
    lines += [f"Workload: `{a.driver}` ({a.workload}, RotatedHLLC + van Albada, outflow); "
              "FP64 parity kernel `hydro_stage_kernel<solver, stage>`, mixed-precision kernel "
              "`hydro_stage_mixed_kernel<solver, albada, stage, a32>` (solver 2 = RotatedHLLC; stage 1 predictor, "
              "2 corrector + CFL; a32: 32-bit byte offsets into the state, grids below 2^29 cells).", ""]
    if a.commands:
        with open(a.commands) as fh:
            cmds = fh.read().strip()
        lines += ["Commands (repository root, one MI355X; then this module on the outputs):", "", "```", cmds,
                  "```", ""]

    lines += ["## Kernel trace (--stats)", "", "| kernel | calls | avg µs | min µs | max µs |", "|---|---|---|---|---|"]
    with open(a.stats) as fh:
        for r in csv.DictReader(fh):
            k = short(r["Name"])
            if not KEEP.search(k):
                continue
            lines.append(f"| `{k}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                         f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} |")
    lines.append("")

    ff = copy_factor(a.fetch_copy, "FETCH_SIZE", copy_bytes)
    wf = copy_factor(a.write_copy, "WRITE_SIZE", copy_bytes)
    fetch, write = counters(a.fetch), counters(a.write)
    lines += ["## HBM traffic (PMC), per dispatch", "",
              f"Calibration: a {a.copy_mib:.0f} MiB device copy (read and written once) scales FETCH_SIZE by "
              f"{ff:.3f} and WRITE_SIZE by {wf:.3f}.", "",
              "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB | write GB | traffic GB | algorithmic GB |",
              "|---|---|---|---|---|---|---|"]
    out = {"grid": a.grid, "calibration": {"fetch_factor": ff, "write_factor": wf}, "kernels": {}}
    for k in sorted(fetch):
        if "hydro_stage" not in k or k not in write:
            continue
        fk, wk = fetch[k]["FETCH_SIZE"], write[k]["WRITE_SIZE"]
        rb, wb = fk * 1024.0 * ff, wk * 1024.0 * wf
        out["kernels"][k] = {"read_bytes": rb, "write_bytes": wb}
        al = algo.get(k)
        lines.append(f"| `{k}` | {fk:.0f} | {wk:.0f} | {rb / 1e9:.3f} | {wb / 1e9:.3f} | {(rb + wb) / 1e9:.3f} | "
                     f"{'' if al is None else f'{al * ncell / 1e9:.3f}'} |")
    lines.append("")

    if a.sq:
        sq = counters(a.sq)
        lines += ["## Instruction counters (PMC), per dispatch", "",
                  "| kernel | SQ_WAVES | VALU instr / wave | SALU instr / wave | LDS instr / wave | kernel cycles | "
                  "VALU utilisation |", "|---|---|---|---|---|---|---|"]
        for k in sorted(sq):
            c = sq[k]
            if "hydro_stage" not in k or "SQ_WAVES" not in c:
                continue
            cyc = c["GRBM_GUI_ACTIVE"] / a.xcds
            util = c["SQ_INSTS_VALU"] / (a.cus * cyc)
            w = c["SQ_WAVES"]
            lines.append(f"| `{k}` | {w:.0f} | {c['SQ_INSTS_VALU'] / w:.0f} | {c['SQ_INSTS_SALU'] / w:.0f} | "
                         f"{c['SQ_INSTS_LDS'] / w:.0f} | {cyc:.3g} | {util:.2f} |")
        lines.append("")
    if a.lds:
        ld = counters(a.lds)
        lines += ["## LDS counters (PMC), per dispatch", "",
                  "Bank-conflict cycles per conflict-free LDS cycle (SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - "
                  "SQ_LDS_BANK_CONFLICT)), and LDS-active and bank-conflict cycles as a share of the kernel's CU "
                  "cycles (CUs x GRBM_GUI_ACTIVE / XCDs).", "",
                  "| kernel | conflict / conflict-free | LDS active share | bank-conflict share |",
                  "|---|---|---|---|"]
        for k in sorted(ld):
            c = ld[k]
            if "hydro_stage" not in k or "SQ_LDS_BANK_CONFLICT" not in c:
                continue
            cyc = a.cus * c["GRBM_GUI_ACTIVE"] / a.xcds
            act, bc = c["SQ_LDS_IDX_ACTIVE"], c["SQ_LDS_BANK_CONFLICT"]
            ratio = bc / (act - bc) if act > bc else float("nan")
            lines.append(f"| `{k}` | {ratio:.3f} | {act / cyc:.3f} | {bc / cyc:.4f} |")
        lines.append("")
    with open(a.md, "w") as fh:
        fh.write("\n".join(lines))
    with open(a.json, "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
