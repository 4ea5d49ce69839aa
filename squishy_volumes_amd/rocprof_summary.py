"""Summaries of rocprofv3 CSV output for the engine's kernels.

Two reports:

``stats DIR``
    The per-kernel table of a ``rocprofv3 --kernel-trace --stats
    --output-format csv`` run (calls, average and total duration).

``traffic FETCH_DIR WRITE_DIR --copy-mib M``
    HBM bytes per dispatch from two ``rocprofv3 --pmc`` runs, one with
    ``FETCH_SIZE`` and one with ``WRITE_SIZE`` (both in KiB per dispatch).
    Each counter is scaled by the factor that makes the plain copy kernel
    (``k_copy_bytes``, launched by ``bench.py --copy-mib M``) read and write
    exactly M MiB.  With ``--write-json PREFIX`` it also writes
    ``PREFIX<stage>_<workload>.json`` files holding ``bytes_per_launch`` for
    the substep kernels, which ``bench.py`` reports as ``roofline.traffic``.

Usage::

    python -m squishy_volumes_amd.rocprof_summary stats TRACE_DIR --command "..."
    python -m squishy_volumes_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR \\
        --copy-mib 1024 --workload spheres_and_slab --write-json profiles/traffic_
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys
from pathlib import Path

# Kernel name prefix -> substep stage name used by bench.py; k_transfer's
# first template argument is its mode.
STAGES = {"k_grid_update": "grid_update"}
TRANSFER_MODES = {"0": "p2g", "1": "g2p2g", "2": "g2p"}


def stage_of(name: str) -> str | None:
    base = name.split("<")[0]
    if base == "k_transfer" and "<" in name:
        return TRANSFER_MODES.get(name.split("<")[1].split(",")[0].strip())
    return STAGES.get(base)


def short_name(name: str) -> str:
    """k_p2g<true, false, 39u> style name from a demangled HIP kernel name."""
    m = re.search(r"(k_[A-Za-z0-9_]+)", name)
    if not m:
        return name.strip()[:60]
    base = m.group(1)
    t = re.search(r"k_[A-Za-z0-9_]+<([^>]*)>", name)
    return base + (f"<{t.group(1)}>" if t else "")


def _csvs(d: Path, suffix: str) -> list[Path]:
    found = sorted(Path(d).rglob(f"*{suffix}"))
    if not found:
        raise SystemExit(f"no *{suffix} under {d}")
    return found


def stats_table(d: Path) -> list[dict]:
    rows = []
    for f in _csvs(d, "kernel_stats.csv"):
        for r in csv.DictReader(open(f)):
            rows.append(
                {
                    "name": short_name(r["Name"]),
                    "calls": int(r["Calls"]),
                    "total_ns": float(r["TotalDurationNs"]),
                    "avg_ns": float(r["AverageNs"]),
                    "min_ns": float(r["MinNs"]),
                    "max_ns": float(r["MaxNs"]),
                    "percent": float(r["Percentage"]),
                }
            )
    rows.sort(key=lambda r: -r["total_ns"])
    return rows


def counter_means(d: Path, counter: str) -> dict[str, tuple[int, float]]:
    """Per kernel: (dispatches, mean counter value per dispatch)."""
    acc = collections.defaultdict(list)
    for f in _csvs(d, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != counter:
                continue
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def cmd_stats(args) -> None:
    rows = stats_table(Path(args.dir))
    print(f"# rocprofv3 kernel statistics: {args.title}\n")
    print(f"Command:\n\n    {args.command}\n")
    print("| kernel | calls | avg us | min us | max us | total ms | % |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for r in rows:
        print(
            f"| `{r['name']}` | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | {r['min_ns'] / 1e3:.2f} | "
            f"{r['max_ns'] / 1e3:.2f} | {r['total_ns'] / 1e6:.3f} | {r['percent']:.1f} |"
        )


def cmd_traffic(args) -> None:
    fetch = counter_means(Path(args.fetch_dir), "FETCH_SIZE")
    write = counter_means(Path(args.write_dir), "WRITE_SIZE")
    copy = "k_copy_bytes"
    if copy not in fetch or copy not in write:
        raise SystemExit("the copy kernel k_copy_bytes is missing: run bench.py with --copy-mib")
    copy_kib = args.copy_mib * 1024.0
    f_fetch = copy_kib / fetch[copy][1]
    f_write = copy_kib / write[copy][1]
    print(f"# HBM traffic per dispatch: {args.workload}\n")
    print("Commands:\n")
    for c in args.command:
        print(f"    {c}")
    print(
        f"\nCalibration: `k_copy_bytes` over {args.copy_mib} MiB reads FETCH_SIZE {fetch[copy][1]:.0f} KiB "
        f"and WRITE_SIZE {write[copy][1]:.0f} KiB per dispatch; scale factors {f_fetch:.4f} (fetch) and "
        f"{f_write:.4f} (write) make both equal to the {copy_kib:.0f} KiB copied.\n"
    )
    print("| kernel | dispatches | FETCH KiB (raw) | WRITE KiB (raw) | read MB (scaled) | written MB (scaled) |")
    print("|---|---:|---:|---:|---:|---:|")
    best: dict[str, tuple[int, float, float, str]] = {}
    for k in sorted(set(fetch) | set(write)):
        nf, vf = fetch.get(k, (0, float("nan")))
        nw, vw = write.get(k, (0, float("nan")))
        rd = vf * f_fetch * 1024 / 1e6
        wr = vw * f_write * 1024 / 1e6
        print(f"| `{k}` | {max(nf, nw)} | {vf:.0f} | {vw:.0f} | {rd:.2f} | {wr:.2f} |")
        stage = stage_of(k)
        if stage and k in fetch and k in write:
            n = min(nf, nw)
            if stage not in best or n > best[stage][0]:
                best[stage] = (n, rd * 1e6, wr * 1e6, k)
    if args.write_json:
        for stage, (n, rd, wr, k) in best.items():
            out = Path(f"{args.write_json}{stage}_{args.workload}.json")
            out.write_text(
                json.dumps(
                    {
                        "kernel": k,
                        "workload": args.workload,
                        "dispatches": n,
                        "read_bytes_per_launch": rd,
                        "write_bytes_per_launch": wr,
                        "bytes_per_launch": rd + wr,
                        "calibration": {"copy_mib": args.copy_mib, "fetch_factor": f_fetch, "write_factor": f_write},
                        "commands": args.command,
                    },
                    indent=1,
                )
                + "\n"
            )
            print(f"\nwrote {out}", file=sys.stderr)


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-mib", type=int, required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--command", action="append", default=[])
    t.add_argument("--write-json", default="")
    args = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic}[args.cmd](args)


if __name__ == "__main__":
    main()
