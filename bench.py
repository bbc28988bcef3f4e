#!/usr/bin/env python3
"""Benchmark: batched Bloch-Torrey FEM solves on one MI355X.

Default workload (--workload c1, the headline configuration): concentric cylinders r_in=4 um,
r_out=8 um, L=20 um, ~200k-vertex P1 tetrahedral mesh (seeded, 2-D Delaunay + extrusion),
D_in=D_out=3e-3 um^2/us, kappa=1e-5 um/us, T2=inf (reference default 1e16), PGSE delta=10600 us,
Delta=43100 us, dt=200 us (270 theta steps, GCloud loop t < T+k, lagged right-hand side),
theta=1/2, gdir=(0,1,0), b in {0, 500, 1000, 2000, 4000} s/mm^2, FP64, Krylov rtol 1e-12.

One step = one complete job on mesh data already resident in HBM: device assembly
(pattern, colouring, M/K/R/I/Jx/Jy/Jz) + every signal of the workload.  value = signals/s.

Other workloads (each prints the same kind of line; --theta-steps cuts the sequence to its
first K steps for a shorter measurement, stated in the line):
  c2  periodic box 30x30x10 um with 20 seeded cylinders (r uniform 2-4 um), ~500k vertices,
      weak pseudo-periodic faces in x and y, kappa=1e-5, 64 Fibonacci-sphere directions x
      8 b-values (linspace 0..4000) = 512 signals in launch groups of 32 (--group).
  c4  neuron-like surrogate (soma r=8 um + 6 seeded dendrite trees, 3 bifurcation levels),
      ~2M vertices, cos-OGSE 4 periods, T=40 ms, dt=100 us (401 steps), 32 directions x
      4 b-values = 128 signals.
  c3  the c1 geometry at ~300k vertices, b=2000, T2=80 ms: the adjoint (forward + checkpointed
      reverse sweep with every gradient) against the forward solve alone.

Also reported: DOF-steps/s, the SpMV roofline and the CPU baseline.  The SpMV launches of the
first --event-iters Krylov iterations of every theta step inside the timed region are bracketed
by HIP events on the solver's stream (sampled so that the events, which cost GPU time of their
own, barely perturb the timed work); achieved = algorithmic bytes per launch / mean launch time.
After the timed region the same two kernels are replayed back to back on the solver's buffers
(no per-launch events) as a second figure that compares directly with a kernel trace.  CPU baseline:
the FP64 CPU oracle on the host cores, bounded sample: the first --cpu-steps theta steps of one
signal, scaled per signal.  The GPU result of the same truncated sequence is compared with the
CPU one (full-size parity on a prefix).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dmri-fem-cloud_amd"))

METRIC = "signals/sec & DOF-steps/sec at <=1e-8 rel err vs FP64 CPU; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


def gcloud_loop(f, T, k):
    """GCloud loop (t < T + k, t += k): (f(t_n), f(t_{n-1})) with t_{-1} = 0, and the list of t_{n-1}."""
    fA, fB, tB = [], [], []
    t, tp = 0.0, 0.0
    while t < T + k:
        fA.append(f(t))
        fB.append(f(tp))
        tB.append(tp)
        tp = t
        t += k
    return np.asarray(fA), np.asarray(fB), np.asarray(tB)


def pgse_sequence(delta, Delta, k):
    """PGSE of GCloudDmriSolver.py:188-193 (exclusive at T) on the GCloud loop."""
    T = delta + Delta

    def f(s):
        return 1.0 if s < delta else (0.0 if s < Delta else (-1.0 if s < T else 0.0))

    fA, fB, _ = gcloud_loop(f, T, k)
    return fA, fB


def pgse_integral_explicit(delta, Delta, k):
    """F(t_{n-1}) = int_0^t f of the PGSE on the GCloud loop (the weak-periodic mirror phase)."""
    T = delta + Delta

    def F(t):
        return t if t < delta else (delta if t < Delta else (delta - (t - Delta) if t < T else 0.0))

    _, FB, _ = gcloud_loop(F, T, k)
    return FB


def q_from_b(b, delta, Delta):
    return float(np.sqrt(b) / np.sqrt(delta * delta * (Delta - delta / 3.0)))


def cos_ogse(nperiod, delta, Delta, t0=100.0, ext=100.0):
    """cos-OGSE of ArbitraryTimeSequence.ipynb (profile 3): omega = 2 pi nperiod / delta,
    tau = (Delta + delta)/2, T = Delta + delta + t0 + ext; returns (f, T, int_0^T F^2)."""
    om = 2.0 * math.pi * nperiod / delta
    tau = 0.5 * (Delta + delta)
    T = Delta + delta + t0 + ext

    def f(s):
        if s < t0:
            return 0.0
        if s <= delta + t0:
            return math.cos(om * (s - t0))
        if s <= tau + t0:
            return 0.0
        if s <= delta + tau + t0:
            return -math.cos(om * (s - t0 - tau))
        return 0.0

    # F = sin(om (s - t0))/om on the first lobe, its negative shifted by tau on the second;
    # with nperiod whole periods F returns to 0 after each lobe: int F^2 = 2 (delta/2)/om^2
    return f, T, delta / (om * om)


def fibonacci_directions(n):
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = math.pi * (1 + 5 ** 0.5) * k
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)


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


# ------------------------------------------------------------------------------------------
# workloads
# ------------------------------------------------------------------------------------------
def workload(name, args):
    import dmri_mesh as dm

    t0 = time.time()
    if name == "c1":
        delta, Delta, k = 10600.0, 43100.0, 200.0
        bvals = [0.0, 500.0, 1000.0, 2000.0, 4000.0]
        mesh = dm.cylinder_mesh(8.0, 20.0, args.nv or 200000, seed=1001, interface_radii=(4.0,))
        fA, fB = pgse_sequence(delta, Delta, k)
        w = dict(mesh=mesh, coeffs=dict(D=3e-3, invT2=1e-16, kappa=1e-5), periodic=None, fA=fA, fB=fB, FB=None,
                 k=k, qs=np.array([q_from_b(b, delta, Delta) for b in bvals]), gdirs=np.tile([0.0, 1.0, 0.0], (5, 1)),
                 cpu_signal=2, desc="two-compartment permeable solve: concentric cylinders r_in=4 r_out=8 L=20 um, "
                 "kappa=1e-5, D=3e-3, PGSE 10600/43100 us, dt=200 us, gdir (0,1,0), b in {0,500,1000,2000,4000}; "
                 "step = device assembly + the 5 signals in one batched launch group",
                 extra={"b_values": bvals})
    elif name == "c2":
        delta, Delta, k = 10600.0, 43100.0, 200.0
        mesh = dm.cylinders_box_mesh(30.0, 30.0, 10.0, ncyl=20, rmin=2.0, rmax=4.0, target_nv=args.nv or 500000,
                                     seed=1002)
        fA, fB = pgse_sequence(delta, Delta, k)
        FB = pgse_integral_explicit(delta, Delta, k)
        dirs = fibonacci_directions(64)
        bvals = np.linspace(0.0, 4000.0, 8)
        qs = np.array([q_from_b(b, delta, Delta) for _ in dirs for b in bvals])
        gd = np.repeat(dirs, len(bvals), axis=0)
        w = dict(mesh=mesh, coeffs=dict(D=3e-3, invT2=1e-16, kappa=1e-5), periodic=(1, 1, 0), fA=fA, fB=fB, FB=FB,
                 k=k, qs=qs, gdirs=gd, cpu_signal=3,
                 desc="periodic box 30x30x10 um, 20 seeded cylinders r~U[2,4] um (two compartments), weak "
                 "pseudo-periodic faces x and y, kappa=1e-5, D=3e-3, PGSE 10600/43100 us, dt=200 us, 64 "
                 f"Fibonacci-sphere directions x 8 b-values linspace(0,4000) = 512 signals in launch groups of {args.group}; "
                 "step = device assembly + all signals", extra={"b_values": bvals.tolist(), "n_dirs": 64})
    elif name == "c4":
        f, T, int4gb = cos_ogse(4, 19900.0, 19900.0)
        k = 100.0
        mesh = dm.neuron_mesh(target_nv=args.nv or 2000000, seed=1004)
        fA, fB, _ = gcloud_loop(f, T, k)
        dirs = fibonacci_directions(32)
        bvals = [1000.0, 2000.0, 3000.0, 4000.0]
        qs = np.array([math.sqrt(b / int4gb) for _ in dirs for b in bvals])
        gd = np.repeat(dirs, len(bvals), axis=0)
        w = dict(mesh=mesh, coeffs=dict(D=3e-3, invT2=1e-16, kappa=0.0), periodic=None, fA=fA, fB=fB, FB=None, k=k,
                 qs=qs, gdirs=gd, cpu_signal=1,
                 desc="neuron-like surrogate (soma r=8 um + 6 dendrite trees, 3 bifurcation levels, voxelised), "
                 "cos-OGSE 4 periods delta=Delta=19900 us, T=40 ms, dt=100 us, D=3e-3, 32 Fibonacci-sphere "
                 f"directions x b in {{1000,2000,3000,4000}} = 128 signals in launch groups of {args.group}; step = device "
                 "assembly + all signals", extra={"b_values": bvals, "n_dirs": 32})
    else:
        raise SystemExit(f"unknown workload {name}")
    w["mesh_gen_s"] = time.time() - t0
    if args.theta_steps and args.theta_steps < len(w["fA"]):
        n = args.theta_steps
        w["full_steps"] = len(w["fA"])
        w["fA"], w["fB"] = w["fA"][:n], w["fB"][:n]
        if w["FB"] is not None:
            w["FB"] = w["FB"][:n]
    return w


def cpu_baseline(args, w, ctx, info, gdir, q):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    m = w["mesh"]
    fA, fB, FB = w["fA"], w["fB"], w["FB"]
    nst = min(args.cpu_steps, len(fA))
    c = w["coeffs"]
    S = O.System(m.xyz, m.tets, m.phase, D=c["D"], invT2=c["invT2"], kappa=c["kappa"], periodic=w["periodic"])
    tc = time.perf_counter()
    s_cpu, _, _ = S.solve(gdir, q, w["k"], fA[:nst], fB[:nst], rtol=args.rtol,
                          FB=None if FB is None else FB[:nst])
    tcpu = time.perf_counter() - tc
    s_gpu, _ = ctx.solve(fA[:nst], fB[:nst], w["k"], q, gdir, rtol=args.rtol,
                         F_explicit=None if FB is None else FB[:nst])
    parity = abs(s_gpu.real - s_cpu) / abs(s_cpu)
    per_signal = tcpu * len(fA) / nst
    cpu = {"value": 1.0 / per_signal, "unit": "signals/s", "cores": O.num_threads(), "kind": "port",
           "sample": f"FP64 CPU oracle (C/OpenMP BiCGStab+Jacobi, same mesh, rtol {args.rtol:g}): first {nst} of "
                     f"{len(fA)} theta steps of one signal (q={q:.4g}), {tcpu:.2f} s, scaled per signal",
           "dof_steps_per_s": float(info["ndof"]) * nst / tcpu}
    return cpu, parity


def parity_block(w, b0, prefix_parity, sig):
    out = {"b0_signal_abs_err_max": max(b0) if b0 else None, "prefix_rel_err_vs_cpu": prefix_parity,
           "signals_first": [s.real for s in sig[:8]]}
    if w["periodic"] and not all(w["periodic"]) and b0:
        # x/y periodic, z faces insulating: the weak periodic term integrates kappa_e (u - u_bc) over every
        # boundary facet, and u_bc is 0 at vertices off the periodic faces (DmriFemLib.py:267-318), so the
        # z-face facets touching a periodic edge remove magnetisation even at b = 0; the oracle shows the
        # same loss (all faces periodic keeps S(b=0) = 1 to 1e-10)
        out["b0_signal_minus_1_max"] = out.pop("b0_signal_abs_err_max")
        out["b0_note"] = ("partially periodic box: S(b=0) < 1 is a property of the weak periodic form, "
                          "reproduced by the oracle")
    return out


def run_forward(args, name):
    import dmri_native as N

    w = workload(name, args)
    mesh, fA, fB, FB, k = w["mesh"], w["fA"], w["fB"], w["FB"], w["k"]
    qs, gd = w["qs"], w["gdirs"]
    nsig = len(qs)
    per_gdir = bool(np.any(np.abs(gd - gd[0]) > 0))
    S_group = min(args.group, nsig)
    if args.profile_steps > 0:
        fA, fB = fA[:args.profile_steps], fB[:args.profile_steps]
        FB = None if FB is None else FB[:args.profile_steps]
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(**w["coeffs"])
    if w["periodic"]:
        ctx.set_periodic(w["periodic"])

    def one_step(profile=0):
        """Device assembly + every signal.  Several launch groups are solved one call each, with a
        progress line on stderr after each (a full C2 job runs for minutes)."""
        ctx.assemble()
        if nsig <= S_group:
            return ctx.solve_batched(fA, fB, k, qs, gd, rtol=args.rtol, profile=profile, lanes=args.lanes,
                                     guess=args.guess, mapping=args.mapping, F_explicit=FB, precond=args.precond)
        sigs, stats = [], []
        t0 = time.perf_counter()
        for g0 in range(0, nsig, S_group):
            sl = slice(g0, min(nsig, g0 + S_group))
            s_, st_ = ctx.solve_batched(fA, fB, k, qs[sl], gd[sl], rtol=args.rtol, profile=profile, lanes=args.lanes,
                                        guess=args.guess, mapping=args.mapping, F_explicit=FB, precond=args.precond)
            sigs.append(s_)
            stats.append(st_)
            print(f"  signals {sl.stop}/{nsig} done, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        merged = dict(stats[0])
        for key in merged:
            if key in ("iters_max", "relres_max"):
                merged[key] = max(s_[key] for s_ in stats)
            elif key not in ("spmv_bytes", "precond_used", "reserved"):
                merged[key] = sum(s_[key] for s_ in stats)
        return np.concatenate(sigs), merged

    copy_gbs = ctx.copy_bandwidth(args.copy_bytes, 10) if args.copy_bytes > 0 else None
    tuned = None
    if not args.no_tune and args.lanes == 0 and args.mapping == 0:
        ctx.assemble()
        tuned = ctx.tune(S_group, per_gdir)[0]
    for _ in range(args.warmup):
        one_step()
    t1 = time.perf_counter()
    all_stats = []
    sig = None
    for _ in range(args.steps):
        sig, st = one_step(profile=args.event_iters)
        all_stats.append(st)
    t2 = time.perf_counter()
    info = ctx.info()
    elapsed = t2 - t1
    njob = nsig * args.steps
    spmv_ms = sum(s["spmv_ms"] for s in all_stats)
    spmv_n = sum(s["spmv_launches"] for s in all_stats)
    spmv_bytes = all_stats[0]["spmv_bytes"]
    avg_us = 1e3 * spmv_ms / max(spmv_n, 1)
    achieved = spmv_bytes / (avg_us * 1e-6) / 1e9 if spmv_n else 0.0
    iters = sum(s["iters_total"] for s in all_stats)
    us_v, us_t, replay_bytes = ctx.time_spmv(S_group, per_gdir, 50)
    used = tuned if tuned is not None else ([args.mapping, args.lanes] if args.mapping and args.lanes else None)
    traffic = measured_traffic(info, used) if name == "c1" else None
    replay_achieved = replay_bytes / (0.5 * (us_v + us_t) * 1e-6) / 1e9
    b0 = [abs(s.real - 1.0) for s, q in zip(sig, qs) if q == 0.0]

    plain = None
    if not args.no_plain:
        idx = list(range(nsig)) if nsig <= 8 else [w["cpu_signal"], nsig - 1]
        tp0 = time.perf_counter()
        ps = [ctx.solve(fA, fB, k, qs[j], gd[j], rtol=args.rtol, F_explicit=FB)[0] for j in idx]
        tp = time.perf_counter() - tp0
        plain = {"signals_per_s": len(idx) / tp, "sample": f"{len(idx)} of {nsig} signals, one launch group each",
                 "batched_over_plain": (njob / elapsed) / (len(idx) / tp),
                 "max_rel_diff_vs_batched": max(abs(a - sig[j]) / max(abs(sig[j]), 1e-300) for a, j in zip(ps, idx))}

    cpu, prefix_parity = (None, None)
    if not args.no_cpu:
        j = w["cpu_signal"]
        cpu, prefix_parity = cpu_baseline(args, w, ctx, info, gd[j], qs[j])

    cfg = {"workload": w["desc"], "nv": info["nv"], "nc": info["nc"], "ndof": info["ndof"], "nnz": info["nnz"],
           "n_iface": info["n_iface"], "n_periodic": info["n_periodic"], "theta_steps": len(fA),
           "signals": nsig, "launch_group": S_group, "rtol": args.rtol, "spmv_mapping_lanes": tuned}
    cfg.update(w["extra"])
    if "full_steps" in w:
        cfg["theta_steps_of_full_sequence"] = f"first {len(fA)} of {w['full_steps']}"
    out = {
        "metric": METRIC,
        "value": njob / elapsed,
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
        "config": cfg,
        "dof_steps_per_s": float(info["ndof"]) * len(fA) * njob / elapsed,
        "krylov_iters_per_theta_step": iters / (len(fA) * njob),
        "preconditioner_used": {0: "jacobi", 1: "polynomial", 2: "polynomial of the real part",
                                3: "polynomial, complex diagonal"}.get(all_stats[-1].get("precond_used")),
        "setup_ms": info["setup_ms"],
        "mesh_gen_s": w["mesh_gen_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/spmv_traffic_cylinders200k_5sig.json (rocprofv3 FETCH_SIZE x1.9999 + "
                                       "WRITE_SIZE x1.0, copy-kernel calibrated)" if traffic else None,
                     "kernel": f"spmv_v/spmv_t (batched complex value-split CSR SpMV, {S_group} signals per launch, "
                               "fused reductions)",
                     "avg_launch_us": avg_us, "bytes_per_launch": spmv_bytes, "launches": spmv_n,
                     "events": f"first {args.event_iters} Krylov iterations of every theta step",
                     "replay_us": {"spmv_v": us_v, "spmv_t": us_t}, "replay_achieved": replay_achieved,
                     "copy_kernel_gbs": copy_gbs,
                     "frac_of_copy": (achieved / copy_gbs) if copy_gbs else None},
        "cpu_baseline": cpu,
        "plain_one_signal_at_a_time": plain,
        "parity": parity_block(w, b0, prefix_parity, sig),
    }
    print(json.dumps(out))


def run_adjoint(args):
    """c3: forward solve alone vs adjoint (forward + checkpointed reverse sweep, all gradients)."""
    import dmri_mesh as dm
    import dmri_native as N

    delta, Delta, k, b = 10600.0, 43100.0, 200.0, 2000.0
    t0 = time.time()
    mesh = dm.cylinder_mesh(8.0, 20.0, args.nv or 300000, seed=1003, interface_radii=(4.0,))
    t_mesh = time.time() - t0
    fA, fB = pgse_sequence(delta, Delta, k)
    if args.theta_steps:
        fA, fB = fA[:args.theta_steps], fB[:args.theta_steps]
    q = q_from_b(b, delta, Delta)
    g = (0.0, 1.0, 0.0)
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(D=3e-3, invT2=1.0 / 80000.0, kappa=1e-5)
    ctx.assemble()
    sqrt_stride = int(math.ceil(math.sqrt(len(fA))))
    for _ in range(args.warmup):
        ctx.solve(fA, fB, k, q, g, rtol=args.rtol)
        ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol)
        ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, checkpoint_stride=sqrt_stride)
    tf = ta = tq = 0.0
    grads = gq = None
    for _ in range(args.steps):
        a0 = time.perf_counter()
        s_f, _ = ctx.solve(fA, fB, k, q, g, rtol=args.rtol)
        a1 = time.perf_counter()
        s_a, grads, _ = ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol)
        a2 = time.perf_counter()
        _, gq, _ = ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, checkpoint_stride=sqrt_stride)
        a3 = time.perf_counter()
        tf += a1 - a0
        ta += a2 - a1
        tq += a3 - a2
    info = ctx.info()
    ratio = ta / tf
    out = {
        "metric": "adjoint gradient cost / forward solve cost (all of dS/dD_c, dS/dkappa, dS/d(1/T2)_c, dS/dq, "
                  "dS/df(t_n))",
        "value": ratio, "unit": "x forward", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * (tf + ta) / args.steps, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "adjoint: concentric cylinders r_in=4 r_out=8 L=20 um (~300k vertices, seed 1003), "
                               "kappa=1e-5, D=3e-3, T2=80 ms, PGSE 10600/43100 us, dt=200 us, b=2000, gdir (0,1,0)",
                   "nv": info["nv"], "ndof": info["ndof"], "nnz": info["nnz"], "theta_steps": len(fA),
                   "rtol": args.rtol, "checkpoints": grads["checkpoints"],
                   "recomputed_steps": grads["recomputed_steps"]},
        "forward_ms": 1e3 * tf / args.steps, "adjoint_ms": 1e3 * ta / args.steps,
        "policy": "automatic: every state kept in HBM (they fit), no recompute",
        "sqrt_checkpointing": {"stride": sqrt_stride, "checkpoints": gq["checkpoints"],
                               "recomputed_steps": gq["recomputed_steps"], "adjoint_ms": 1e3 * tq / args.steps,
                               "ratio": tq / tf,
                               "max_rel_diff_vs_all_states": max(
                                   abs(gq[key_] - grads[key_]) / max(abs(grads[key_]), 1e-300)
                                   for key_ in ("kappa", "q"))},
        "signal_forward": s_f.real, "signal_adjoint": s_a.real, "mesh_gen_s": t_mesh,
        "gradients": {"D": grads["D"].tolist(), "invT2": grads["invT2"].tolist(), "kappa": grads["kappa"],
                      "q": grads["q"]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--nv", type=int, default=0, help="target vertex count (0 = the workload's own)")
    ap.add_argument("--theta-steps", type=int, default=0, help="cut the sequence to its first K theta steps")
    ap.add_argument("--cpu-steps", type=int, default=60, help="theta steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rtol", type=float, default=1e-12)
    ap.add_argument("--precond", type=int, default=0,
                    help="0 = Jacobi + degree-1 polynomial, chosen by the library (default), 1 = Jacobi only, "
                         "2 = the polynomial scaled by the complex diagonal")
    ap.add_argument("--group", type=int, default=32,
                    help="signals per launch group (1..64); 32 was the fastest for the 512- and 128-signal "
                         "workloads (64: 2-4 %% slower, 16: 5-8 %% slower)")
    ap.add_argument("--lanes", type=int, default=0, help="SpMV lanes per row; 0 = tuned")
    ap.add_argument("--mapping", type=int, default=0, help="SpMV mapping 1 spread, 2 staged, 3 registers, 4 wave per row, "
                    "5 tile-compressed; 0 = tuned")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults)")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-signal-at-a-time comparison")
    ap.add_argument("--guess", type=int, default=0, help="Krylov initial guess: 0 residual-minimising projection on the last states, 1 previous "
                    "step, 2 linear, 3 quadratic extrapolation")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="counter runs only: truncate to the first K theta steps (result line not comparable)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-kernel reference")
    ap.add_argument("--event-iters", type=int, default=1,
                    help="Krylov iterations per theta step whose SpMV launches are event-timed")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus N>1 is not supported "
                                   "(signals shard with no exchange; see the sharding section of DESIGN.md)"}))
        sys.exit(2)
    if args.workload == "c3":
        run_adjoint(args)
    else:
        run_forward(args, args.workload)


if __name__ == "__main__":
    main()
