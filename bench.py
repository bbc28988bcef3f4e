#!/usr/bin/env python3
"""Benchmark: two-compartment permeable Bloch-Torrey solve on one MI355X.

Workload (the headline configuration): concentric cylinders r_in=4 um, r_out=8 um, L=20 um,
~200k-vertex P1 tetrahedral mesh (seeded, 2-D Delaunay + extrusion), D_in=D_out=3e-3 um^2/us,
kappa=1e-5 um/us, T2=inf (reference default 1e16), PGSE delta=10600 us, Delta=43100 us,
dt=200 us (270 theta steps, GCloud loop t < T+k, lagged right-hand side), theta=1/2,
gdir=(0,1,0), b in {0, 500, 1000, 2000, 4000} s/mm^2, FP64, Krylov rtol 1e-12.

One step = one complete job on mesh data already resident in HBM: device assembly
(pattern, colouring, M/K/R/I/Jx/Jy/Jz) + the five signals.  value = signals/s.

Also reported: DOF-steps/s, the SpMV roofline and the CPU baseline.  The SpMV launches of the
first --event-iters Krylov iterations of every theta step inside the timed region are bracketed
by HIP events on the solver's stream (sampled so that the events, which cost GPU time of their
own, barely perturb the timed work); achieved = algorithmic bytes per launch / mean launch time.
After the timed region the same two kernels are replayed back to back on the solver's buffers
(no per-launch events) as a second figure that compares directly with a kernel trace.  CPU baseline: (the FP64 CPU oracle on the host cores,
bounded sample: the first CPU_STEPS theta steps of the b=1000 signal).  The GPU result of
the same truncated sequence is compared with the CPU one (full-size parity on a prefix).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmri-fem-cloud_amd"))

METRIC = "signals/sec & DOF-steps/sec at <=1e-8 rel err vs FP64 CPU; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


def pgse_sequence(delta, Delta, k):
    """GCloud loop (t < T + k, t += k, right-hand side lagged, first step f(0) on both sides)."""
    T = delta + Delta

    def f(s):
        return 1.0 if s < delta else (0.0 if s < Delta else (-1.0 if s < T else 0.0))

    fA, fB = [], []
    t, tp = 0.0, 0.0
    while t < T + k:
        fA.append(f(t))
        fB.append(f(tp))
        tp = t
        t += k
    return np.asarray(fA), np.asarray(fB)


def measured_traffic(info, mapping_lanes):
    """HBM bytes per SpMV launch from the committed FETCH_SIZE/WRITE_SIZE counter runs of this
    workload (copy-kernel calibrated), if the mesh and SpMV mapping are the ones measured."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic_cylinders200k_5sig.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    if (d.get("nnz"), d.get("ndof")) != (info["nnz"], info["ndof"]) or \
            list(d.get("mapping_lanes", [])) != list(mapping_lanes or []):
        return None
    return float(d["traffic_per_launch"])


def q_from_b(b, delta, Delta):
    return float(np.sqrt(b) / np.sqrt(delta * delta * (Delta - delta / 3.0)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nv", type=int, default=200000, help="target vertex count")
    ap.add_argument("--cpu-steps", type=int, default=60, help="theta steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rtol", type=float, default=1e-12)
    ap.add_argument("--lanes", type=int, default=0, help="SpMV lanes per row; 0 = tuned")
    ap.add_argument("--mapping", type=int, default=0, help="SpMV mapping 1 spread, 2 staged, 3 registers; 0 = tuned")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-signal-at-a-time comparison")
    ap.add_argument("--guess", type=int, default=0, help="Krylov initial guess: 0 extrapolated, 1 previous step")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="counter runs only: truncate to the first K theta steps (result line not comparable)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-kernel reference")
    ap.add_argument("--event-iters", type=int, default=2,
                    help="Krylov iterations per theta step whose SpMV launches are event-timed")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus N>1 is not supported "
                                   "(signals shard with no exchange; see the sharding section of DESIGN.md)"}))
        sys.exit(2)

    import dmri_mesh as dm
    import dmri_native as N

    delta, Delta, k = 10600.0, 43100.0, 200.0
    bvals = [0.0, 500.0, 1000.0, 2000.0, 4000.0]
    gdir = (0.0, 1.0, 0.0)
    fA, fB = pgse_sequence(delta, Delta, k)
    if args.profile_steps > 0:
        fA, fB = fA[:args.profile_steps], fB[:args.profile_steps]

    t0 = time.time()
    mesh = dm.cylinder_mesh(8.0, 20.0, args.nv, seed=1001, interface_radii=(4.0,))
    t_mesh = time.time() - t0
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(D=3e-3, invT2=1e-16, kappa=1e-5)

    qs = np.array([q_from_b(b, delta, Delta) for b in bvals])

    def one_step(profile=False):
        ctx.assemble()
        sig, st = ctx.solve_batched(fA, fB, k, qs, gdir, rtol=args.rtol, profile=profile, lanes=args.lanes,
                                    guess=args.guess, mapping=args.mapping)
        return list(sig), [st]

    copy_gbs = ctx.copy_bandwidth(args.copy_bytes, 10) if args.copy_bytes > 0 else None
    tuned = None
    if not args.no_tune and args.lanes == 0 and args.mapping == 0:
        ctx.assemble()
        tuned = ctx.tune(len(bvals), False)[0]
    for _ in range(args.warmup):
        one_step()
    t1 = time.perf_counter()
    all_stats = []
    sig = None
    for _ in range(args.steps):
        sig, st = one_step(profile=args.event_iters)
        all_stats += st
    t2 = time.perf_counter()
    info = ctx.info()
    elapsed = t2 - t1
    nsig = len(bvals) * args.steps
    signals_per_s = nsig / elapsed
    dof_steps = float(info["ndof"]) * len(fA) * nsig
    spmv_ms = sum(s["spmv_ms"] for s in all_stats)
    spmv_n = sum(s["spmv_launches"] for s in all_stats)
    spmv_bytes = all_stats[0]["spmv_bytes"]
    avg_us = 1e3 * spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes / (avg_us * 1e-6) / 1e9 if spmv_n else 0.0
    iters = sum(s["iters_total"] for s in all_stats)
    us_v, us_t, replay_bytes = ctx.time_spmv(len(bvals), False, 50)
    used = tuned if tuned is not None else ([args.mapping, args.lanes] if args.mapping and args.lanes else None)
    traffic = measured_traffic(info, used)
    replay_achieved = replay_bytes / (0.5 * (us_v + us_t) * 1e-6) / 1e9

    # b = 0 must keep the signal at 1 (mass conservation; size-independent parity property)
    b0_err = abs(sig[0].real - 1.0)

    plain = None
    if not args.no_plain:
        tp0 = time.perf_counter()
        ps = [ctx.solve(fA, fB, k, q, gdir, rtol=args.rtol)[0] for q in qs]
        tp = time.perf_counter() - tp0
        plain = {"signals_per_s": len(qs) / tp,
                 "max_rel_diff_vs_batched": max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(ps, sig))}

    cpu = None
    prefix_parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        nst = min(args.cpu_steps, len(fA))
        q = q_from_b(1000.0, delta, Delta)
        S = O.System(mesh.xyz, mesh.tets, mesh.phase, D=3e-3, invT2=1e-16, kappa=1e-5)
        tc = time.perf_counter()
        s_cpu, _, cst = S.solve(gdir, q, k, fA[:nst], fB[:nst], rtol=args.rtol)
        tcpu = time.perf_counter() - tc
        s_gpu, _ = ctx.solve(fA[:nst], fB[:nst], k, q, gdir, rtol=args.rtol)
        prefix_parity = abs(s_gpu.real - s_cpu) / abs(s_cpu)
        per_signal = tcpu * len(fA) / nst
        cpu = {"value": 1.0 / per_signal, "unit": "signals/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"FP64 CPU oracle (C/OpenMP BiCGStab+Jacobi, same mesh, rtol {args.rtol:g}): first {nst} of "
                         f"{len(fA)} theta steps of the b=1000 signal, {tcpu:.2f} s, scaled per signal",
               "dof_steps_per_s": float(info["ndof"]) * nst / tcpu}

    out = {
        "metric": METRIC,
        "value": signals_per_s,
        "unit": "signals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "two-compartment permeable solve: concentric cylinders r_in=4 r_out=8 L=20 um, "
                        "kappa=1e-5, D=3e-3, PGSE 10600/43100 us, dt=200 us, gdir (0,1,0), "
                        "b in {0,500,1000,2000,4000}; step = device assembly + the 5 signals in one batched launch group",
            "nv": info["nv"], "nc": info["nc"], "ndof": info["ndof"], "nnz": info["nnz"],
            "n_iface": info["n_iface"], "theta_steps": len(fA), "b_values": bvals, "rtol": args.rtol,
            "spmv_mapping_lanes": tuned,
        },
        "dof_steps_per_s": dof_steps / elapsed,
        "krylov_iters_per_theta_step": iters / (len(fA) * nsig),
        "setup_ms": info["setup_ms"],
        "mesh_gen_s": t_mesh,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/spmv_traffic_cylinders200k_5sig.json (rocprofv3 FETCH_SIZE x1.9998 + "
                                       "WRITE_SIZE x1.0, copy-kernel calibrated)" if traffic else None,
                     "kernel": "spmv_v/spmv_t (batched complex value-split CSR SpMV, 5 signals, fused reductions)",
                     "avg_launch_us": avg_us, "bytes_per_launch": spmv_bytes, "launches": spmv_n,
                     "events": f"first {args.event_iters} Krylov iterations of every theta step",
                     "replay_us": {"spmv_v": us_v, "spmv_t": us_t}, "replay_achieved": replay_achieved,
                     "copy_kernel_gbs": copy_gbs,
                     "frac_of_copy": (achieved / copy_gbs) if copy_gbs else None},
        "cpu_baseline": cpu,
        "plain_one_signal_at_a_time": plain,
        "parity": {"b0_signal_abs_err": b0_err, "prefix_rel_err_vs_cpu": prefix_parity,
                   "signals": [s.real for s in sig]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
