#!/usr/bin/env python3
"""Text summaries of rocprofv3 CSV output for libdmri runs.

  python3 prof_summary.py trace    <kernel_trace.csv>      [--sample K]
  python3 prof_summary.py steps    <kernel_trace.csv>      [--first KERNEL]
  python3 prof_summary.py counters <counter_collection.csv> [--sample K] [--copy-bytes B]
  python3 prof_summary.py traffic  <FETCH_SIZE csv> <WRITE_SIZE csv> --copy-bytes B
                                   --nnz NNZ --ndof N --mapping M L --workload TEXT --command CMD ... --out F

trace: per kernel the number of dispatches, mean and median duration, and the mean idle time
between the end of the previous dispatch and its start.  For the Krylov iteration kernels (SpMVs,
bicg_p, s-pass, update) it also prints the mean over the first K dispatches after every residual
kernel (the first K Krylov iterations of each theta step, where every signal of the launch group is
still active): the same launches that the solver's sampled kernel events time
(dmri_solver_opts.profile = K).

steps: per theta step (from one dispatch of its first kernel to the next) the span, the time the
kernels cover and the idle rest, as medians over the steps.

counters: per kernel the mean counter value per dispatch (KiB), for all dispatches and for the
same first-K sample.  With --copy-bytes B (the size given to dmri_copy_bandwidth) the copy kernel
calibrates the counter: scale = expected KiB per copy dispatch (B/1024 read, B/1024 written) /
measured mean, and scaled values are printed beside the raw ones.

traffic: the HBM bytes per launch of every kernel of a Krylov iteration (calibrated FETCH_SIZE +
WRITE_SIZE, mean over the full dispatches: those above half the kernel's largest value, which
leaves out the launches queued past convergence that exit at once) as the JSON file bench.py reads
for the roofline's "traffic", keyed by the names of the iteration kernels in dmri_stats (bicg_p,
spmv_v_pre, spmv_v, spass, spmv_t_pre, spmv_t, bicg_update).  bicg_p's figure is the general
iteration's (p = r + beta (p - omega v)), not the first one's (p = r).
"""
from __future__ import annotations

import argparse
import collections
import csv
import statistics

SPMV = ("spmv_v_kernel", "spmv_t_kernel", "spmv_v_sr_kernel", "spmv_t_sr_kernel", "spmv_v_wave_kernel",
        "spmv_t_wave_kernel", "spmv_v_tile_kernel", "spmv_t_tile_kernel", "spmv_v_pair_kernel",
        "spmv_t_pair_kernel", "bicg_p_kernel", "spass_kernel", "bicg_update_kernel", "cocg_p_kernel",
        "cocg_update_kernel")


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


def steps(path, first):
    """Per theta step, from one dispatch of the step's first kernel to the next: the span, the time
    covered by kernels (the union of their intervals), idle = span - covered, and the dispatches.
    Spans over 5x the median (the gap between two solves) are left out; prints the medians."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if base(r["Kernel_Name"]) == first]
    rec = []
    for a, b in zip(starts, starts[1:]):
        t0, t1 = int(rows[a]["Start_Timestamp"]), int(rows[b]["Start_Timestamp"])
        covered, cs, ce = 0, None, None
        for r in rows[a:b]:
            s, e = int(r["Start_Timestamp"]), min(int(r["End_Timestamp"]), t1)
            if ce is None or s > ce:
                covered += 0 if ce is None else ce - cs
                cs, ce = s, e
            else:
                ce = max(ce, e)
        covered += ce - cs
        rec.append(((t1 - t0) / 1e3, covered / 1e3, (t1 - t0 - covered) / 1e3, b - a))
    if not rec:
        print(f"no two dispatches of {first}")
        return None
    med_span = statistics.median(r[0] for r in rec)
    rec = [r for r in rec if r[0] <= 5 * med_span]
    out = {"steps": len(rec), "span_us": statistics.median(r[0] for r in rec),
           "busy_us": statistics.median(r[1] for r in rec), "idle_us": statistics.median(r[2] for r in rec),
           "idle_mean_us": statistics.mean(r[2] for r in rec), "dispatches": statistics.median(r[3] for r in rec)}
    print(f"theta steps {out['steps']} (from {first}): median span {out['span_us']:.1f} us, kernels "
          f"{out['busy_us']:.1f} us, idle {out['idle_us']:.1f} us (mean {out['idle_mean_us']:.1f}), "
          f"{out['dispatches']:.0f} dispatches")
    return out


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


# SpMV kernels of each mapping (dmri_solver_opts.mapping) with the polynomial preconditioner
MAPPING_KERNELS = {1: ("spmv_v_kernel", "spmv_t_kernel"), 2: ("spmv_v_kernel", "spmv_t_kernel"),
                   4: ("spmv_v_wave_kernel", "spmv_t_wave_kernel"), 6: ("spmv_v_pair_kernel", "spmv_t_pair_kernel")}


def iteration_kernel(name: str, mapping: int):
    """dmri_stats name of an iteration kernel from its template instance (SpMV kernels of the given
    mapping only), or None.  The last template argument of the SpMV kernels is the operation: 0 v =
    A ph (or t = A s), 1 the polynomial preconditioner, 2 t = A sh, 3 the preconditioner with the
    complex diagonal; COCG's 4 q = A p (spmv_q) and 5, 6, 7, 8 z = M r (spmv_pre: FP64, complex
    diagonal, FP32 input, with the update of r folded in); 9 q = A p with the p-step folded in."""
    b = base(name)
    if b in ("bicg_p_kernel", "spass_kernel", "bicg_update_kernel", "cocg_p_kernel", "cocg_update_kernel"):
        return b[:-len("_kernel")]
    if b not in MAPPING_KERNELS.get(mapping, ()):
        return None
    side = "spmv_v" if b.startswith("spmv_v") else "spmv_t"
    args = short(name).split("<")[1].rstrip(">").split(",") if "<" in short(name) else []
    op = int(args[-1]) if args and args[-1].strip().lstrip("-").isdigit() else 0
    if op in (4, 9):  # COCG: q = A p (9: with the p-step folded in)
        return "spmv_q"
    if op in (5, 6, 7, 8):  # COCG: z = M r
        return "spmv_pre"
    return side + ("_pre" if op in (1, 3) else "")


def per_dispatch(path):
    """(rows, values) per dispatch of a counter CSV (rows of one dispatch summed)."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    per = collections.OrderedDict()
    for r in rows:
        key = int(r["Dispatch_Id"])
        if key not in per:
            per[key] = [r, 0.0]
        per[key][1] += float(r["Counter_Value"])
    return [v[0] for v in per.values()], [v[1] for v in per.values()]


def traffic(fetch_csv, write_csv, copy_bytes, out, meta, mapping):
    import json

    res, scales = {}, {}
    for which, path in (("fetch", fetch_csv), ("write", write_csv)):
        disp, vals = per_dispatch(path)
        cp = [v for r, v in zip(disp, vals) if short(r["Kernel_Name"]) == "copy_kernel"]
        scale = (copy_bytes / 1024.0) / statistics.mean(cp)
        scales[which] = scale
        acc, acc1 = collections.defaultdict(list), collections.defaultdict(list)
        first = sampled(disp, 1)  # each kernel's first dispatch after a step's residual
        for i, (r, v) in enumerate(zip(disp, vals)):
            nm = iteration_kernel(r["Kernel_Name"], mapping)
            if nm:
                acc[nm].append(v * scale * 1024.0)
                if i in first:
                    acc1[nm].append(v * scale * 1024.0)
        for nm, v in acc.items():
            full = [x for x in v if x > 0.5 * max(v)]
            res.setdefault(nm, {})[which] = statistics.mean(full)
            res[nm][which + "_launches"] = len(full)
        # first iterations of the steps that iterate (a step solved by its initial state stops at once)
        for nm, v in acc1.items():
            full = [x for x in v if x > 0.5 * max(v)]
            res.setdefault(nm, {})[which + "_first"] = statistics.mean(full)
            res[nm][which + "_first_launches"] = len(full)
    doc = dict(meta)
    doc["fetch_scale"] = scales["fetch"]
    doc["write_scale"] = scales["write"]
    doc["read_bytes_per_launch_by_kernel"] = {nm: v.get("fetch", 0.0) for nm, v in res.items()}
    doc["write_bytes_per_launch_by_kernel"] = {nm: v.get("write", 0.0) for nm, v in res.items()}
    doc["traffic_per_launch_by_kernel"] = {nm: v.get("fetch", 0.0) + v.get("write", 0.0) for nm, v in res.items()}
    doc["launches_by_kernel"] = {nm: v.get("fetch_launches", 0) for nm, v in res.items()}
    doc["traffic_per_launch_first_iteration_by_kernel"] = {
        nm: v.get("fetch_first", 0.0) + v.get("write_first", 0.0) for nm, v in res.items() if "fetch_first" in v}
    doc["first_iteration_launches_by_kernel"] = {nm: v.get("fetch_first_launches", 0) for nm, v in res.items()}
    doc["calibration"] = (f"rocprofv3 FETCH_SIZE x {scales['fetch']:.4f} + WRITE_SIZE x {scales['write']:.4f} "
                          f"(copy kernel of {copy_bytes:.0f} B), mean over each kernel's full dispatches")
    with open(out, "w") as fh:
        json.dump(doc, fh, indent=1)
    for nm, v in sorted(doc["traffic_per_launch_by_kernel"].items()):
        v1 = doc["traffic_per_launch_first_iteration_by_kernel"].get(nm, 0.0)
        print(f"{nm:12s} {v / 1e9:8.3f} GB per launch ({v1 / 1e9:.3f} GB in a step's first iteration)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "steps", "counters", "traffic"])
    ap.add_argument("csv")
    ap.add_argument("csv2", nargs="?")
    ap.add_argument("--sample", type=int, default=2)
    ap.add_argument("--copy-bytes", type=float, default=0.0)
    ap.add_argument("--nnz", type=int)
    ap.add_argument("--ndof", type=int)
    ap.add_argument("--mapping", type=int, nargs=2)
    ap.add_argument("--workload", default="")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--out")
    ap.add_argument("--first", default="rhs_history_kernel", help="steps: the kernel that opens a theta step")
    a = ap.parse_args()
    if a.mode == "trace":
        trace(a.csv, a.sample)
    elif a.mode == "steps":
        steps(a.csv, a.first)
    elif a.mode == "counters":
        counters(a.csv, a.sample, a.copy_bytes)
    else:
        meta = {"workload": a.workload, "nnz": a.nnz, "ndof": a.ndof, "mapping_lanes": a.mapping,
                "commands": a.command}
        traffic(a.csv, a.csv2, a.copy_bytes, a.out, meta, a.mapping[0])


if __name__ == "__main__":
    main()
