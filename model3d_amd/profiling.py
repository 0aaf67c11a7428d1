"""Summaries of rocprofv3 output: kernel-trace statistics and per-dispatch
HBM traffic from the FETCH_SIZE / WRITE_SIZE counters.

    python -m model3d_amd.profiling trace DIR [--title T] [--command C]
        Markdown table per kernel (calls, average / min / max / total
        microseconds, registers, LDS) from a `rocprofv3 --kernel-trace` run
        (the rocpd `*_results.db`, or `*_kernel_trace.csv` with
        `--output-format csv`).

    python -m model3d_amd.profiling sq DIR [--title T] [--command C]
        Markdown table per kernel of the SQ counters of a `rocprofv3 --pmc
        SQ_...` run (per dispatch, averaged over dispatches), with per-wave
        figures: VALU / SALU / LDS instructions per wave, wave cycles per
        wave, and the share of wave cycles spent waiting (SQ_WAIT_INST_ANY).

    python -m model3d_amd.profiling pmc FETCH_DIR WRITE_DIR --copy-bytes N [...]
        JSON: for every kernel the average FETCH_SIZE and WRITE_SIZE per
        dispatch (KiB, as the counters report them), and the same counters of
        the library's two plain streams of N bytes the profiled program ran
        first (one read-only, one write-only), from which the bytes-per-KiB
        calibration factors follow.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sqlite3
import sys

_SHORT = re.compile(r"(k_[A-Za-z_0-9]+(?:<[^>]*>)?|__amd_[A-Za-z_]+|[A-Za-z_]*elementwise_kernel[A-Za-z_]*)")


def short_name(name: str) -> str:
    m = _SHORT.search(name)
    return m.group(1) if m else name[:60]


def kernel_rows(path: str):
    """[(name, duration_ns, vgpr, sgpr, lds, grid, workgroup)] of every dispatch."""
    csvs = glob.glob(os.path.join(path, "**", "*_kernel_trace.csv"), recursive=True)
    if csvs:
        out = []
        for f in csvs:
            for r in csv.DictReader(open(f)):
                out.append((r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                            int(r.get("VGPR_Count", 0) or 0), int(r.get("SGPR_Count", 0) or 0),
                            int(r.get("LDS_Block_Size", 0) or 0), int(r.get("Grid_Size", 0) or 0),
                            int(r.get("Workgroup_Size", 0) or 0)))
        return out
    dbs = glob.glob(os.path.join(path, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no kernel trace under {path}")
    out = []
    for db in dbs:
        c = sqlite3.connect(db)
        for r in c.execute("select name, duration, vgpr_count, sgpr_count, lds_size, grid_x, workgroup_x from kernels"):
            out.append(tuple(r))
    return out


def trace_table(path: str) -> list:
    agg = collections.OrderedDict()
    for name, dur, vgpr, sgpr, lds, grid, wg in kernel_rows(path):
        k = short_name(name)
        a = agg.setdefault(k, {"calls": 0, "total": 0, "min": None, "max": 0, "vgpr": 0, "sgpr": 0, "lds": 0,
                               "durs": []})
        a["calls"] += 1
        a["durs"].append(dur)
        a["total"] += dur
        a["min"] = dur if a["min"] is None else min(a["min"], dur)
        a["max"] = max(a["max"], dur)
        a["vgpr"], a["sgpr"], a["lds"] = max(a["vgpr"], vgpr), max(a["sgpr"], sgpr), max(a["lds"], lds)
    rows = sorted(agg.items(), key=lambda kv: -kv[1]["total"])
    return [(k, a["calls"], a["total"] / a["calls"] / 1e3, sorted(a["durs"])[(len(a["durs"]) - 1) // 2] / 1e3,
             a["min"] / 1e3, a["max"] / 1e3, a["total"] / 1e3, a["vgpr"], a["sgpr"], a["lds"]) for k, a in rows]


def trace_markdown(path: str, title: str = "", command: str = "") -> str:
    rows = trace_table(path)
    total = sum(r[6] for r in rows) or 1.0
    lines = []
    if title:
        lines += [f"# {title}", ""]
    if command:
        lines += ["Command:", "", "```", command, "```", ""]
    lines += ["| kernel | calls | avg us | median us | min us | max us | total us | % | VGPR | SGPR | LDS B |",
              "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for k, calls, avg, med, mn, mx, tot, vgpr, sgpr, lds in rows:
        lines.append(f"| `{k}` | {calls} | {avg:.2f} | {med:.2f} | {mn:.2f} | {mx:.2f} | {tot:.1f} | "
                     f"{100 * tot / total:.1f} | {vgpr} | {sgpr} | {lds} |")
    return "\n".join(lines) + "\n"


def counter_values(path: str, counter: str) -> dict:
    """`counter` per dispatch, by kernel: {kernel: {dispatch id: value}}."""
    files = glob.glob(os.path.join(path, "**", "*_counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter collection under {path}")
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for f in files:
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != counter:
                continue
            per[short_name(r["Kernel_Name"])][r["Dispatch_Id"]] += float(r["Counter_Value"])
    return per


def pmc_summary(fetch_dir: str, write_dir: str, copy_bytes: int) -> dict:
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE per dispatch, and the
    calibration from the program's two streams of `copy_bytes` (the largest
    dispatch of each: k_stream_read reads them once -- its FETCH_SIZE gives
    the bytes per fetched KiB -- and k_stream_write writes them once -- its
    WRITE_SIZE gives the bytes per written KiB)."""
    fetch = counter_values(fetch_dir, "FETCH_SIZE")
    write = counter_values(write_dir, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        f, w = fetch.get(k, {}), write.get(k, {})
        kernels[k] = {"fetch_kib": sum(f.values()) / len(f) if f else None,
                      "write_kib": sum(w.values()) / len(w) if w else None,
                      "dispatches": max(len(f), len(w))}
    reads = [k for k in fetch if k.startswith("k_stream_read")]
    writes = [k for k in write if k.startswith("k_stream_write")]
    if not reads or not writes:
        raise SystemExit("no k_stream_read / k_stream_write dispatch found for the calibration")
    cf = max(fetch[reads[0]].values())
    cw = max(write[writes[0]].values())
    kib = copy_bytes / 1024.0
    return {
        "counters": "FETCH_SIZE, WRITE_SIZE (KiB per dispatch, averaged over dispatches)",
        "calibration": {"read_kernel": reads[0], "write_kernel": writes[0], "bytes": copy_bytes,
                        "fetch_kib_of_read": cf, "write_kib_of_write": cw,
                        "fetch_factor": kib / cf if cf else None,
                        "write_factor": kib / cw if cw else None},
        "kernels": kernels,
    }


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_SMEM", "SQ_WAVE_CYCLES",
               "SQ_BUSY_CYCLES", "SQ_WAIT_INST_ANY")


def sq_markdown(path: str, title: str = "", command: str = "") -> str:
    per = {c: counter_values(path, c) for c in SQ_COUNTERS}
    kernels = sorted({k for c in per.values() for k in c}, key=lambda k: -sum(per["SQ_WAVE_CYCLES"].get(k, {}).values()))
    lines = []
    if title:
        lines += [f"# {title}", ""]
    if command:
        lines += ["Command:", "", "```", command, "```", ""]
    lines += ["Per dispatch, averaged over the kernel's dispatches.  Per-wave columns divide by SQ_WAVES; "
              "wait % is SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES.", "",
              "| kernel | dispatches | " + " | ".join(c[3:] for c in SQ_COUNTERS) +
              " | VALU/wave | SALU/wave | LDS/wave | cycles/wave | wait % |",
              "|---|---:|" + "---:|" * (len(SQ_COUNTERS) + 5)]
    for k in kernels:
        avg = {}
        n = 0
        for c in SQ_COUNTERS:
            d = per[c].get(k, {})
            n = max(n, len(d))
            avg[c] = sum(d.values()) / len(d) if d else 0.0
        w = avg["SQ_WAVES"] or 1.0
        wait = 100.0 * avg["SQ_WAIT_INST_ANY"] / avg["SQ_WAVE_CYCLES"] if avg["SQ_WAVE_CYCLES"] else 0.0
        lines.append(f"| `{k}` | {n} | " + " | ".join(f"{avg[c]:.0f}" for c in SQ_COUNTERS) +
                     f" | {avg['SQ_INSTS_VALU'] / w:.0f} | {avg['SQ_INSTS_SALU'] / w:.0f} | "
                     f"{avg['SQ_INSTS_LDS'] / w:.0f} | {avg['SQ_WAVE_CYCLES'] / w:.0f} | {wait:.1f} |")
    return "\n".join(lines) + "\n"


def traffic_bytes(summary: dict, kernel: str):
    """Calibrated HBM bytes per dispatch of `kernel` (fetch + write), or None."""
    k = summary.get("kernels", {}).get(kernel)
    c = summary.get("calibration", {})
    if not k or k.get("fetch_kib") is None or k.get("write_kib") is None:
        return None
    if not c.get("fetch_factor") or not c.get("write_factor"):
        return None
    return 1024.0 * (k["fetch_kib"] * c["fetch_factor"] + k["write_kib"] * c["write_factor"])


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m model3d_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--title", default="")
    t.add_argument("--command", default="")
    q = sub.add_parser("sq")
    q.add_argument("dir")
    q.add_argument("--title", default="")
    q.add_argument("--command", default="")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_dir")
    p.add_argument("write_dir")
    p.add_argument("--copy-bytes", type=int, required=True)
    p.add_argument("--workload", default="")
    p.add_argument("--config", default="{}", help="JSON of the profiled configuration")
    p.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        sys.stdout.write(trace_markdown(a.dir, a.title, a.command))
    elif a.cmd == "sq":
        sys.stdout.write(sq_markdown(a.dir, a.title, a.command))
    else:
        s = pmc_summary(a.fetch_dir, a.write_dir, a.copy_bytes)
        s = {"workload": a.workload, "config": json.loads(a.config), "commands": a.command, **s}
        json.dump(s, sys.stdout, indent=1)
        sys.stdout.write("\n")


if __name__ == "__main__":
    main()
