"""Summaries of rocprofv3 CSV output for the render path's profiles.

    python -m kestrel_amd.rocprof_summary trace DIR [--blocks K] [--skip-last S]
        per-kernel mean duration over the K blocks before the last S of a
        --kernel-trace run (blocks start at each k_spawn dispatch, else at each
        k_begin), plus the whole-run --stats table;
    python -m kestrel_amd.rocprof_summary pmc DIR --kernel NAME [--last K]
        mean of each counter over the last K dispatches of one kernel;
    python -m kestrel_amd.rocprof_summary traffic --fetch DIR --write DIR
            --copy-fetch DIR --copy-write DIR [--kernel NAME] [--last K]
        HBM bytes per launch of NAME: FETCH_SIZE and WRITE_SIZE (KiB per
        dispatch) each scaled by bytes-known / counter-read for the plain 1 GiB
        copy of ``bench.py --calibrate-copy``; prints markdown and a JSON line.
    python -m kestrel_amd.rocprof_summary valu DIR --frames F --mean-voices V
            [--kernel NAME] [--last K] [--workload W]
        SQ_INSTS_VALU of NAME per launch and per voice-frame (a wave64
        instruction serves the 64 voices of one batch, so per voice-frame it is
        the count x 64 / (V x F)); prints a JSON line.

DIR is the ``-d`` directory of a ``rocprofv3 ... --output-format csv`` run.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import sys

COPY_BYTES = 1 << 30


def _one(d: str, suffix: str) -> str:
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits[-1]


def short(name: str) -> str:
    """'void ns::(anonymous namespace)::k_x<1, true>(ns::Args)' -> 'k_x'."""
    base = name.split("(kst::")[0].split("<")[0]
    if base.startswith("void "):
        base = base[5:]
    return base.rsplit("::", 1)[-1].split("(")[0].strip()


def _rows(path: str):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


# Kernels that wait on the device for another stream's work: their duration
# is mostly that wait.
WAITERS = {"k_sel_begin"}


def trace(d: str, blocks: int, skip: int = 0) -> str:
    rows = _rows(_one(d, "kernel_trace.csv"))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    # A block's chain opens with k_spawn (staged runs begin each block inside
    # the previous one's k_end); without spawns, with k_begin.
    begins = [i for i, r in enumerate(rows) if short(r["Kernel_Name"]) == "k_spawn"]
    if len(begins) < blocks:
        begins = [i for i, r in enumerate(rows) if short(r["Kernel_Name"]) == "k_begin"]
    where = f"last {blocks}" if not skip else f"{blocks} before the last {skip}"
    out = [f"### The {where} blocks (from {os.path.basename(_one(d, 'kernel_trace.csv'))})", "",
           "| kernel | launches | mean us | share of block |", "|---|---:|---:|---:|"]
    if len(begins) >= blocks + skip:
        seg = rows[begins[-blocks - skip]:(begins[-skip] if skip else len(rows))]
        agg = collections.OrderedDict()
        for r in seg:
            name = short(r["Kernel_Name"])
            dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            a = agg.setdefault(name, [0, 0.0])
            a[0] += 1
            a[1] += dur
        # Block period: from the first block's k_spawn to the last one's (the
        # time after the last block runs into whatever the host does next).
        first = int(rows[begins[-blocks - skip]]["Start_Timestamp"])
        last = int(rows[begins[-skip - 1]]["Start_Timestamp"])
        span = (last - first) / 1e3 / max(blocks - 1, 1)
        for name, (n, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            note = " (spins until the render has ended)" if name in WAITERS else ""
            out.append(f"| {name}{note} | {n} | {tot / n:.1f} | {tot / blocks / span:.1%} |")
        out += ["", f"Block period {span:.1f} us (k_spawn to k_spawn)."]
    stats = _rows(_one(d, "kernel_stats.csv"))
    out += ["", "### Whole run (--stats)", "", "| kernel | calls | avg us | total % |",
            "|---|---:|---:|---:|"]
    for r in stats[:20]:
        out.append(f"| {short(r['Name'])} | {r['Calls']} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def pmc_means(d: str, kernel: str | None, last: int):
    rows = _rows(_one(d, "counter_collection.csv"))
    by_disp = collections.OrderedDict()
    for r in rows:
        name = short(r["Kernel_Name"])
        if kernel and name != kernel:
            continue
        by_disp.setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = float(r["Counter_Value"])
        by_disp[int(r["Dispatch_Id"])]["_name"] = name
    disp = list(by_disp.values())[-last:]
    if not disp:
        raise SystemExit(f"no dispatches of {kernel} in {d}")
    keys = [k for k in disp[0] if k != "_name"]
    return {k: sum(x[k] for x in disp) / len(disp) for k in keys}, len(disp), disp[0]["_name"]


def copy_means(d: str):
    """The copy kernel: the last 5 dispatches that are not the initial fill."""
    rows = _rows(_one(d, "counter_collection.csv"))
    by_disp = collections.OrderedDict()
    for r in rows:
        by_disp.setdefault(int(r["Dispatch_Id"]), {})[r["Counter_Name"]] = float(r["Counter_Value"])
    disp = list(by_disp.values())[-5:]
    keys = list(disp[0])
    return {k: sum(x[k] for x in disp) / len(disp) for k in keys}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--blocks", type=int, default=10)
    t.add_argument("--skip-last", type=int, default=0,
                   help="leave out the last S blocks (e.g. bench.py's per-pass replay)")
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("--kernel", default="k_render_fast")
    p.add_argument("--last", type=int, default=10)
    tr = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        tr.add_argument("--" + k.replace("_", "-"), required=True)
    tr.add_argument("--kernel", default="k_render_fast")
    tr.add_argument("--last", type=int, default=10)
    tr.add_argument("--workload", default=None, help="workload name recorded in the JSON")
    tr.add_argument("--mean-voices", type=float, default=None,
                    help="the profiled run's mean voices per block, recorded in the JSON")
    va = sub.add_parser("valu")
    va.add_argument("dir")
    va.add_argument("--kernel", default="k_render_fast")
    va.add_argument("--last", type=int, default=10)
    va.add_argument("--frames", type=int, required=True, help="frames per block")
    va.add_argument("--mean-voices", type=float, required=True,
                    help="the profiled run's mean voices per block")
    va.add_argument("--workload", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "valu":
        m, n, name = pmc_means(a.dir, a.kernel, a.last)
        insts = m["SQ_INSTS_VALU"]
        rec = {"kernel": name, "sq_insts_valu_per_launch": round(insts),
               "valu_per_voice_frame": round(insts * 64.0 / (a.mean_voices * a.frames), 3),
               "frames": a.frames, "mean_voices": a.mean_voices, "launches": n}
        if a.workload:
            rec["workload"] = a.workload
        print(json.dumps(rec))
    elif a.cmd == "trace":
        print(trace(a.dir, a.blocks, a.skip_last))
    elif a.cmd == "pmc":
        m, n, name = pmc_means(a.dir, a.kernel, a.last)
        print(f"{name}: mean over the last {n} dispatches")
        for k, v in m.items():
            print(f"  {k} = {v:.1f}")
    else:
        f, nf, _ = pmc_means(a.fetch, a.kernel, a.last)
        w, nw, _ = pmc_means(a.write, a.kernel, a.last)
        cf = copy_means(a.copy_fetch)["FETCH_SIZE"]
        cw = copy_means(a.copy_write)["WRITE_SIZE"]
        kf = COPY_BYTES / (cf * 1024.0)
        kw = COPY_BYTES / (cw * 1024.0)
        fetch_b = f["FETCH_SIZE"] * 1024.0 * kf
        write_b = w["WRITE_SIZE"] * 1024.0 * kw
        print(f"| quantity | value |\n|---|---:|\n"
              f"| copy FETCH_SIZE (KiB / 1 GiB copy) | {cf:.0f} |\n"
              f"| copy WRITE_SIZE (KiB / 1 GiB copy) | {cw:.0f} |\n"
              f"| fetch scale | {kf:.4f} |\n| write scale | {kw:.4f} |\n"
              f"| {a.kernel} FETCH_SIZE KiB/dispatch (last {nf}) | {f['FETCH_SIZE']:.0f} |\n"
              f"| {a.kernel} WRITE_SIZE KiB/dispatch (last {nw}) | {w['WRITE_SIZE']:.0f} |\n"
              f"| HBM bytes read per launch | {fetch_b:.4g} |\n"
              f"| HBM bytes written per launch | {write_b:.4g} |\n"
              f"| HBM bytes per launch | {fetch_b + write_b:.4g} |")
        rec = {"kernel": a.kernel, "bytes_per_launch": round(fetch_b + write_b),
               "bytes_read": round(fetch_b), "bytes_written": round(write_b),
               "fetch_scale": round(kf, 5), "write_scale": round(kw, 5)}
        if a.workload:
            rec["workload"] = a.workload
        if a.mean_voices:
            rec["mean_voices"] = a.mean_voices
        print(json.dumps(rec))


if __name__ == "__main__":
    sys.exit(main())
