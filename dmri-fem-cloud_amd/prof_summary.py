#!/usr/bin/env python3
"""Text summaries of rocprofv3 CSV output for libdmri runs.

  python3 prof_summary.py trace    <kernel_trace.csv>      [--sample K]
  python3 prof_summary.py counters <counter_collection.csv> [--sample K] [--copy-bytes B]

trace: per kernel the number of dispatches, mean and median duration, and the mean idle time
between the end of the previous dispatch and its start.  For the SpMV kernels it also prints the
mean over the first K dispatches after every residual kernel (the first K Krylov iterations
of each theta step, where every signal of the launch group is still active): the same launches
that the solver's sampled kernel events time (dmri_solver_opts.profile = K).

counters: per kernel the mean counter value per dispatch (KiB), for all dispatches and for the
same first-K sample.  With --copy-bytes B (the size given to dmri_copy_bandwidth) the copy kernel
calibrates the counter: scale = expected KiB per copy dispatch (B/1024 read, B/1024 written) /
measured mean, and scaled values are printed beside the raw ones.
"""
from __future__ import annotations

import argparse
import collections
import csv
import statistics

SPMV = ("spmv_v_kernel", "spmv_t_kernel", "spmv_v_sr_kernel", "spmv_t_sr_kernel", "spmv_v_wave_kernel",
        "spmv_t_wave_kernel", "spmv_v_tile_kernel", "spmv_t_tile_kernel")


# kernels that form a theta step's residual just before its Krylov iterations
STEP_START = ("rhs_residual_kernel", "rhs_history_kernel", "residual_kernel")


def short(name: str) -> str:
    n = name.replace("void ", "").replace("dmri::(anonymous namespace)::", "").replace("dmri::", "")
    return n.split("(")[0]


def base(name: str) -> str:
    return short(name).split("<")[0]


def sampled(rows, k):
    """Indices of the first k SpMV dispatches (per SpMV kernel) after each step's residual kernel."""
    out = set()
    seen = collections.Counter()
    for i, r in enumerate(rows):
        b = base(r["Kernel_Name"])
        if b in STEP_START:
            seen.clear()
        elif b in SPMV:
            key = short(r["Kernel_Name"])  # each template instance (operation) counts on its own
            if seen[key] < k:
                out.add(i)
            seen[key] += 1
    return out


def trace(path, k):
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    samp = sampled(rows, k)
    dur, gap, sdur = collections.defaultdict(list), collections.defaultdict(list), collections.defaultdict(list)
    prev_end = None
    for i, r in enumerate(rows):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        name = short(r["Kernel_Name"])
        dur[name].append((e - s) / 1e3)
        if i in samp:
            sdur[name].append((e - s) / 1e3)
        if prev_end is not None and 0 <= s - prev_end < 1_000_000:
            gap[name].append((s - prev_end) / 1e3)
        prev_end = e
    print(f"{'kernel':44s} {'calls':>7s} {'mean_us':>9s} {'median_us':>9s} {'gap_us':>7s} "
          f"{'first' + str(k) + '_calls':>12s} {'first' + str(k) + '_mean_us':>14s}")
    for name, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        g = statistics.mean(gap[name]) if gap[name] else float("nan")
        sv = sdur.get(name, [])
        sm = f"{statistics.mean(sv):14.2f}" if sv else f"{'':14s}"
        print(f"{name[:44]:44s} {len(v):7d} {statistics.mean(v):9.2f} {statistics.median(v):9.2f} {g:7.2f} "
              f"{len(sv) if sv else '':>12} {sm}")


def counters(path, k, copy_bytes):
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    by_counter = collections.defaultdict(list)
    for r in rows:
        by_counter[r["Counter_Name"]].append(r)
    for cname, crows in by_counter.items():
        # several rows per dispatch when the counter is reported per dimension: sum them
        per = collections.OrderedDict()
        for r in crows:
            key = int(r["Dispatch_Id"])
            if key not in per:
                per[key] = [r, 0.0]
            per[key][1] += float(r["Counter_Value"])
        disp = [v[0] for v in per.values()]
        vals = [v[1] for v in per.values()]
        samp = sampled(disp, k)
        allv, sv = collections.defaultdict(list), collections.defaultdict(list)
        for i, (r, v) in enumerate(zip(disp, vals)):
            allv[short(r["Kernel_Name"])].append(v)
            if i in samp:
                sv[short(r["Kernel_Name"])].append(v)
        scale = None
        if copy_bytes:
            cp = allv.get("copy_kernel")
            if cp:
                scale = (copy_bytes / 1024.0) / statistics.mean(cp)
        print(f"counter {cname}" + (f"  (copy-kernel scale {scale:.4f})" if scale else ""))
        print(f"{'kernel':44s} {'calls':>7s} {'mean_KiB':>12s} {'first' + str(k) + '_mean_KiB':>16s} "
              f"{'scaled_MB':>10s}")
        for name, v in sorted(allv.items(), key=lambda kv: -sum(kv[1])):
            s = sv.get(name, [])
            m = statistics.mean(s) if s else statistics.mean(v)
            scaled = f"{m * scale * 1024 / 1e6:10.2f}" if scale else f"{'':10s}"
            print(f"{name[:44]:44s} {len(v):7d} {statistics.mean(v):12.1f} "
                  f"{(f'{statistics.mean(s):16.1f}' if s else f'{chr(32) * 16}')} {scaled}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counters"])
    ap.add_argument("csv")
    ap.add_argument("--sample", type=int, default=2)
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    a = ap.parse_args()
    if a.mode == "trace":
        trace(a.csv, a.sample)
    else:
        counters(a.csv, a.sample, a.copy_bytes)


if __name__ == "__main__":
    main()
