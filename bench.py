#!/usr/bin/env python3
"""Benchmark: batched Bloch-Torrey FEM solves on one MI355X.

Default workload (--workload c4, the largest configuration): neuron-like surrogate mesh, soma
sphere r=8 um + 6 seeded branching dendrites (r 0.5-2 um, 3 bifurcation levels, lengths 20-60 um),
voxelised into ~2M vertices / 9.4M P1 tetrahedra, one compartment, D=3e-3 um^2/us, T2=inf
(reference default 1e16); cos-OGSE with 4 periods (delta = Delta = 19900 us, t0 = seq_ext = 100 us,
T = 40 ms), dt = 100 us (401 theta steps, GCloud loop t < T+k, lagged right-hand side), theta=1/2,
32 Fibonacci-sphere gradient directions x b in {1000, 2000, 3000, 4000} s/mm^2 = 128 signals in
launch groups of 32, FP64, Krylov rtol 1e-12.

One bench step = one theta step of all 128 signals (a resumable batch keeps every launch group's
states, Krylov workspace and initial-guess history in HBM between steps); the mesh is assembled
on the device before the timed window.  value = complex DOF-steps/s (ndof x signals x steps /
time); signals_per_s_equiv divides it by the DOF-steps of one full 401-step signal.

Other workloads (whole jobs per bench step: device assembly + every signal's full sequence;
--theta-steps cuts the sequence to its first K steps, stated in the line):
  c1  two-compartment concentric cylinders r_in=4, r_out=8, L=20 um, ~200k vertices, kappa=1e-5,
      PGSE delta=10600 us, Delta=43100 us, dt=200 us (270 steps), gdir (0,1,0),
      b in {0, 500, 1000, 2000, 4000}: value = signals/s.
  c2  periodic box 30x30x10 um with 20 seeded cylinders (r uniform 2-4 um), ~500k vertices,
      weak pseudo-periodic faces in x and y, kappa=1e-5, 64 Fibonacci-sphere directions x
      8 b-values (linspace 0..4000) = 512 signals in launch groups of 32 (--group).
  c4 --whole-jobs  the default workload timed as whole jobs.
  c3  the c1 geometry at ~300k vertices, b=2000, T2=80 ms: the adjoint (forward + checkpointed
      reverse sweep with every gradient) against the forward solve alone.

Also reported: the roofline and the CPU baseline.  Every kernel of the first --event-iters
Krylov iterations of every theta step inside the timed region is bracketed by kernel-bound HIP
events on the solver's stream (sampled so that the events, which cost GPU time of their own,
barely perturb the timed work); each kernel's achieved bandwidth = its algorithmic bytes per
launch / its mean launch time, and "roofline" is that of the kernel with the most device time.
After the timed region the two operator SpMVs are replayed back to back on the solver's buffers
(no per-launch events) as a second figure that compares directly with a kernel trace.  CPU
baseline: the FP64 CPU oracle on the host cores, bounded sample (c4: theta steps 1..--cpu-steps-1
of one signal after an untimed step 0, and step 1 again on one core; c1/c2: the first --cpu-steps
theta steps of one signal, scaled per signal).  The GPU's plain solve of the same prefix is
compared with the CPU one (full-size parity on a prefix).
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

import dmri_sequences as DS  # noqa: E402

METRIC = "signals/sec & DOF-steps/sec at <=1e-8 rel err vs FP64 CPU; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


def gcloud_loop(f, T, k):
    """(f(t_n), f(t_{n-1}), t_{n-1}) over the GCloud loop (t < T + k, t_{-1} = 0)."""
    fA, fB, tB = DS.gcloud_loop(f, T, k, F=lambda t: t)
    return fA, fB, tB


def pgse_sequence(delta, Delta, k):
    """PGSE of GCloudDmriSolver.py:188-193 (exclusive at T) on the GCloud loop."""
    fA, fB, _ = DS.gcloud_loop(DS.pgse(delta, Delta), delta + Delta, k)
    return fA, fB


def pgse_integral_explicit(delta, Delta, k):
    """F(t_{n-1}) = int_0^t f of the PGSE on the GCloud loop (the weak-periodic mirror phase)."""
    return DS.gcloud_loop(DS.pgse(delta, Delta), delta + Delta, k, F=DS.pgse_integral(delta, Delta))[2]


def q_from_b(b, delta, Delta):
    return DS.q_from_b(b, DS.pgse_int_F2(delta, Delta))


def cos_ogse(nperiod, delta, Delta, t0=100.0, ext=100.0):
    """cos-OGSE of ArbitraryTimeSequence.ipynb (profile 3): returns (f, T, int_0^T F^2)."""
    f, T = DS.cos_ogse(nperiod, delta, Delta, t0, ext)
    return f, T, DS.cos_ogse_int_F2(nperiod, delta, Delta)


fibonacci_directions = DS.fibonacci_directions

KERNEL_NAMES = ("bicg_p", "spmv_v_pre", "spmv_v", "spass", "spmv_t_pre", "spmv_t", "bicg_update")
# the same dmri_stats slots under COCG (slots 3-5 unused)
COCG_NAMES = ("cocg_p", "spmv_pre", "spmv_q", None, None, None, "cocg_update")
KERNEL_DESC = {"bicg_p": "p = r + beta (p - omega v)", "spmv_v_pre": "ph = c (2 p - c A p)", "spmv_v": "v = A ph",
               "spass": "alpha = rho / <r^, v>, s = r - alpha v", "spmv_t_pre": "sh = c (2 s - c A s)",
               "spmv_t": "t = A sh", "bicg_update": "x += alpha ph + omega sh, r = s - omega t",
               "cocg_p": "x += alpha p (previous iteration), p = z + beta p", "spmv_pre": "z = c (2 r - c A r), r^T z",
               "spmv_q": "q = A p, p^T q", "cocg_update": "alpha = rho / p^T q, r -= alpha q"}
# With one gradient direction the library folds COCG's update of r into the preconditioner's SpMV
KERNEL_DESC_FOLDED = {"spmv_pre": "r -= alpha q (gathered), z = c (2 r - c A r), r^T z, |r|^2"}
# With per-signal directions it folds the p-step into the q = A p SpMV
KERNEL_DESC_QFOLDED = {"spmv_q": "x += alpha p, p = z + beta p, q = A z + beta q, p^T q"}


def merge_stats(stats):
    """Sum the per-call solver statistics (max for the per-step maxima, element-wise for the per-kernel
    arrays)."""
    out = dict(stats[0])
    for key in out:
        vals = [s_[key] for s_ in stats]
        if key in ("iters_max", "relres_max"):
            out[key] = max(vals)
        elif key.startswith("kern_"):
            out[key] = [sum(v[j] for v in vals) for j in range(len(vals[0]))]
        elif key not in ("spmv_bytes", "precond_used"):
            out[key] = sum(vals)
    return out


def kernel_table(st, krylov="bicgstab"):
    """Per-kernel device time and algorithmic bandwidth of the event-timed Krylov iterations."""
    total = sum(st["kern_ms"]) or 1.0
    out = {}
    folded = krylov == "cocg" and st["kern_launches"][6] == 0 and st["kern_launches"][1] > 0
    qfolded = krylov == "cocg" and st["kern_launches"][0] == 0 and st["kern_launches"][2] > 0
    for j, name in enumerate(COCG_NAMES if krylov == "cocg" else KERNEL_NAMES):
        nl = st["kern_launches"][j]
        if nl == 0 or name is None:
            continue
        us = 1e3 * st["kern_ms"][j] / nl
        b = st["kern_bytes"][j] / nl
        what = KERNEL_DESC[name]
        if folded:
            what = KERNEL_DESC_FOLDED.get(name, what)
        if qfolded:
            what = KERNEL_DESC_QFOLDED.get(name, what)
        out[name] = {"what": what, "launches": nl, "avg_us": us, "bytes_per_launch": b,
                     "achieved_gbs": b / (us * 1e-6) / 1e9, "frac_of_peak": b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                     "share_of_iteration_time": st["kern_ms"][j] / total}
    return out


def roofline_block(st, copy_gbs, traffic, S_group, event_iters, krylov="bicgstab"):
    """The roofline of the dominant kernel (most device time over the timed iterations) and the table
    of every kernel of an iteration."""
    table = kernel_table(st, krylov)
    if not table:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None}
    dom = max(table, key=lambda k: table[k]["avg_us"] * table[k]["launches"])
    d = table[dom]
    t = traffic.get(dom) if traffic else None
    # the events time each step's first iteration(s); with one, the counters' mean over the steps'
    # first iterations is the traffic of the same launches
    tb, tsrc = None, None
    if t:
        if event_iters == 1 and t.get("bytes_first"):
            tb, tsrc = t["bytes_first"], t["source"] + ", each step's first iteration"
        else:
            tb, tsrc = t["bytes"], t["source"] + ", every iteration"
    return {"bound": "hbm", "achieved": d["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["achieved_gbs"] / HBM_PEAK_GBS, "traffic": tb,
            "traffic_source": tsrc,
            "kernel": f"{dom}: {d['what']} (batched complex CSR, {S_group} signals per launch)",
            "avg_launch_us": d["avg_us"], "bytes_per_launch": d["bytes_per_launch"], "launches": d["launches"],
            "events": f"every kernel of the first {event_iters} Krylov iteration(s) of every theta step, "
                      "kernel-bound HIP events on the solver's stream",
            "copy_kernel_gbs": copy_gbs, "frac_of_copy": (d["achieved_gbs"] / copy_gbs) if copy_gbs else None,
            "kernels": table}


TRAFFIC_FILES = {"c1": "spmv_traffic_cylinders200k_5sig.json", "c2": "spmv_traffic_box500k_32sig.json",
                 "c4": "spmv_traffic_neuron2m_32sig.json"}


def measured_traffic(info, mapping_lanes, name="c1"):
    """HBM bytes per launch of each iteration kernel from the committed FETCH_SIZE/WRITE_SIZE counter
    runs of this workload (copy-kernel calibrated), if the mesh and SpMV mapping are the ones measured:
    {kernel name: {"bytes": ..., "source": ...}}."""
    fname = TRAFFIC_FILES.get(name)
    if fname is None:
        return None
    path = os.path.join(ROOT, "profiles", fname)
    try:
        with open(path) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    if (d.get("nnz"), d.get("ndof")) != (info["nnz"], info["ndof"]) or \
            list(d.get("mapping_lanes", [])) != list(mapping_lanes or []):
        return None
    src = f"profiles/{fname} ({d.get('calibration', 'rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-kernel calibrated')})"
    first = d.get("traffic_per_launch_first_iteration_by_kernel", {})
    return {k: {"bytes": float(v), "bytes_first": first.get(k), "source": src}
            for k, v in d.get("traffic_per_launch_by_kernel", {}).items()}


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
                         F_explicit=None if FB is None else FB[:nst], krylov=args.krylov, pre_fp32=args.pre_fp32)
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
                                     guess=args.guess, mapping=args.mapping, F_explicit=FB, precond=args.precond,
                                     graphs=args.graphs, krylov=args.krylov, pre_fp32=args.pre_fp32)
        sigs, stats = [], []
        t0 = time.perf_counter()
        for g0 in range(0, nsig, S_group):
            sl = slice(g0, min(nsig, g0 + S_group))
            s_, st_ = ctx.solve_batched(fA, fB, k, qs[sl], gd[sl], rtol=args.rtol, profile=profile, lanes=args.lanes,
                                        guess=args.guess, mapping=args.mapping, F_explicit=FB, precond=args.precond,
                                        graphs=args.graphs, krylov=args.krylov, pre_fp32=args.pre_fp32)
            sigs.append(s_)
            stats.append(st_)
            print(f"  signals {sl.stop}/{nsig} done, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        return np.concatenate(sigs), merge_stats(stats)

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
    merged = merge_stats(all_stats)
    iters = merged["iters_total"]
    us_v, us_t, replay_bytes = ctx.time_spmv(S_group, per_gdir, 50)
    used = tuned if tuned is not None else ([args.mapping, args.lanes] if args.mapping and args.lanes else None)
    traffic = measured_traffic(info, used, name)
    replay_achieved = replay_bytes / (0.5 * (us_v + us_t) * 1e-6) / 1e9
    roof = roofline_block(merged, copy_gbs, traffic, S_group, args.event_iters, args.krylov)
    roof["replay_us"] = {"spmv_v": us_v, "spmv_t": us_t}
    roof["replay_achieved"] = replay_achieved
    b0 = [abs(s.real - 1.0) for s, q in zip(sig, qs) if q == 0.0]

    plain = None
    if not args.no_plain:
        idx = list(range(nsig)) if nsig <= 8 else [w["cpu_signal"], nsig - 1]
        tp0 = time.perf_counter()
        ps = [ctx.solve(fA, fB, k, qs[j], gd[j], rtol=args.rtol, F_explicit=FB, krylov=args.krylov, pre_fp32=args.pre_fp32)[0] for j in idx]
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
           "signals": nsig, "launch_group": S_group, "rtol": args.rtol, "spmv_mapping_lanes": tuned, "graphs": args.graphs, "krylov": args.krylov, "pre_fp32": args.pre_fp32}
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
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_one_signal_at_a_time": plain,
        "parity": parity_block(w, b0, prefix_parity, sig),
    }
    print(json.dumps(out))


def run_stepwise(args, name):
    """Theta step by theta step: every signal of the workload advances one theta step per bench step
    through a resumable batch (dmri_batch_advance) that keeps the states, Krylov workspaces and
    initial-guess histories of all launch groups in HBM.  value = complex DOF-steps/s."""
    import dmri_native as N

    w = workload(name, args)
    mesh, fA, fB, FB, k = w["mesh"], w["fA"], w["fB"], w["FB"], w["k"]
    qs, gd = w["qs"], w["gdirs"]
    nsig = len(qs)
    per_gdir = bool(np.any(np.abs(gd - gd[0]) > 0))
    S_group = min(args.group, nsig)
    nfull = len(fA)
    if args.warmup + args.steps > nfull:
        raise SystemExit(f"warmup + steps ({args.warmup + args.steps}) exceed the {nfull} theta steps of the sequence")
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(**w["coeffs"])
    if w["periodic"]:
        ctx.set_periodic(w["periodic"])
    ctx.assemble()
    info = ctx.info()
    copy_gbs = ctx.copy_bandwidth(args.copy_bytes, 10) if args.copy_bytes > 0 else None
    tuned = None
    if not args.no_tune and args.lanes == 0 and args.mapping == 0:
        tuned, times = ctx.tune(S_group, per_gdir)
        print("  SpMV tuning, us per iteration's four SpMVs (mapping/lanes): " +
              ", ".join(f"{m}/{l} {t:.0f}" for (m, l), t in sorted(times.items(), key=lambda kv: kv[1])[:8]),
              file=sys.stderr, flush=True)
    print(f"  mesh nv={info['nv']} ndof={info['ndof']} nnz={info['nnz']}, {nsig} signals in groups of {S_group}, "
          f"SpMV mapping/lanes {tuned}", file=sys.stderr, flush=True)
    tb = time.perf_counter()
    batch = ctx.batch(fA, fB, k, qs, gd, rtol=args.rtol, profile=args.event_iters, lanes=args.lanes,
                      guess=args.guess, mapping=args.mapping, F_explicit=FB, precond=args.precond, group=S_group,
                      graphs=args.graphs, krylov=args.krylov, pre_fp32=args.pre_fp32)
    t_begin = time.perf_counter() - tb
    for i in range(args.warmup):
        batch.advance(1)
        print(f"  warmup theta step {i} done", file=sys.stderr, flush=True)
    all_stats = []
    t1 = time.perf_counter()
    for i in range(args.steps):
        all_stats.append(batch.advance(1))
        print(f"  theta step {args.warmup + i}: {all_stats[-1]['solve_ms']:.1f} ms, "
              f"{all_stats[-1]['iters_total'] / nsig:.1f} Krylov iterations per signal", file=sys.stderr, flush=True)
    t2 = time.perf_counter()
    elapsed = t2 - t1
    sig, done = batch.signals()
    nd = float(info["ndof"])
    value = nd * nsig * args.steps / elapsed
    merged = merge_stats(all_stats)
    iters = merged["iters_total"]
    precond = all_stats[-1].get("precond_used")
    batch.close()
    us_v, us_t, replay_bytes = ctx.time_spmv(S_group, per_gdir, 20)
    replay_achieved = replay_bytes / (0.5 * (us_v + us_t) * 1e-6) / 1e9
    traffic = measured_traffic(info, tuned, name)
    roof = roofline_block(merged, copy_gbs, traffic, S_group, args.event_iters, args.krylov)
    roof["replay_us"] = {"spmv_v": us_v, "spmv_t": us_t}
    roof["replay_achieved"] = replay_achieved

    # one signal at a time on the GPU over the same theta steps (a launch group of one)
    plain = None
    if not args.no_plain:
        j = w["cpu_signal"]
        pb = ctx.batch(fA, fB, k, qs[j:j + 1], gd[j:j + 1], rtol=args.rtol, lanes=0, guess=args.guess,
                       F_explicit=FB, precond=args.precond, group=1, krylov=args.krylov, pre_fp32=args.pre_fp32)
        pb.advance(args.warmup)
        tp0 = time.perf_counter()
        pst = pb.advance(args.steps)
        tp = time.perf_counter() - tp0
        psig, _ = pb.signals()
        pb.close()
        plain_rate = nd * args.steps / tp
        plain = {"dof_steps_per_s": plain_rate, "signals_per_s_equiv": plain_rate / (nd * nfull),
                 "sample": f"signal {j} alone (a launch group of one), theta steps {args.warmup}..."
                           f"{args.warmup + args.steps - 1}", "batched_over_plain": value / plain_rate,
                 "krylov_iters_per_theta_step": pst["iters_total"] / args.steps,
                 "rel_diff_vs_batched": abs(psig[0] - sig[j]) / max(abs(sig[j]), 1e-300)}

    cpu, parity = (None, None)
    if not args.no_cpu:
        cpu, parity = cpu_baseline_steps(args, w, ctx, info)

    cfg = {"workload": w["desc"].replace("step = device assembly + all signals",
                                         "step = one theta step of every signal"),
           "nv": info["nv"], "nc": info["nc"], "ndof": info["ndof"], "nnz": info["nnz"], "n_iface": info["n_iface"],
           "theta_steps_full_sequence": nfull, "signals": nsig, "launch_group": S_group, "rtol": args.rtol,
           "timed_window": f"theta steps {args.warmup}..{args.warmup + args.steps - 1} of {nfull} (after "
                           f"{args.warmup} warm-up steps from u0 = 1), every signal",
           "spmv_mapping_lanes": tuned, "graphs": args.graphs, "krylov": args.krylov, "pre_fp32": args.pre_fp32}
    cfg.update(w["extra"])
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-steps/s",
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
        "signals_per_s_equiv": value / (nd * nfull),
        "krylov_iters_per_theta_step": iters / (args.steps * nsig),
        "preconditioner_used": {0: "jacobi", 1: "polynomial", 2: "polynomial of the real part",
                                3: "polynomial, complex diagonal"}.get(precond),
        "setup_ms": info["setup_ms"],
        "batch_begin_s": t_begin,
        "mesh_gen_s": w["mesh_gen_s"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_one_signal_at_a_time": plain,
        "parity": parity,
        "signals_first": [complex(x).real for x in sig[:4]],
    }
    print(json.dumps(out))


def cpu_baseline_steps(args, w, ctx, info):
    """The FP64 CPU oracle on one signal: theta step 0 untimed, then steps 1..n-1 timed on every host
    core (and step 1 again on one core); the GPU's plain solve of the same n steps is the parity
    check at full size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    m = w["mesh"]
    fA, fB, FB = w["fA"], w["fB"], w["FB"]
    n = max(2, min(args.cpu_steps, len(fA)))
    j = w["cpu_signal"]
    q, g = w["qs"][j], w["gdirs"][j]
    c = w["coeffs"]
    S = O.System(m.xyz, m.tets, m.phase, D=c["D"], invT2=c["invT2"], kappa=c["kappa"], periodic=w["periodic"])
    FBn = None if FB is None else FB
    _, u1, _ = S.solve(g, q, w["k"], fA[:1], fB[:1], rtol=args.rtol, FB=None if FBn is None else FBn[:1])
    cores = O.num_threads()
    tc = time.perf_counter()
    s_cpu, _, st = S.solve(g, q, w["k"], fA[1:n], fB[1:n], rtol=args.rtol, u0=u1,
                           FB=None if FBn is None else FBn[1:n])
    tcpu = time.perf_counter() - tc
    one = None
    if not args.no_cpu_1core:
        O.set_threads(1)
        t1c = time.perf_counter()
        _, _, st1 = S.solve(g, q, w["k"], fA[1:2], fB[1:2], rtol=args.rtol, u0=u1,
                            FB=None if FBn is None else FBn[1:2])
        t1 = time.perf_counter() - t1c
        O.set_threads(cores)
        one = {"value": float(info["ndof"]) / t1, "unit": "DOF-steps/s", "cores": 1,
               "sample": f"theta step 1 of signal {j}, {int(st1['iters'])} BiCGStab iterations, {t1:.2f} s"}
    s_gpu, _ = ctx.solve(fA[:n], fB[:n], w["k"], q, g, rtol=args.rtol, F_explicit=None if FB is None else FB[:n],
                         krylov=args.krylov, pre_fp32=args.pre_fp32)
    cpu = {"value": float(info["ndof"]) * (n - 1) / tcpu, "unit": "DOF-steps/s", "cores": cores, "kind": "port",
           "sample": f"FP64 CPU oracle (C/OpenMP BiCGStab+Jacobi, same mesh, rtol {args.rtol:g}): theta steps "
                     f"1..{n - 1} of signal {j} (q={q:.4g}), {int(st['iters'])} BiCGStab iterations, {tcpu:.2f} s",
           "krylov_iters_per_theta_step": st["iters"] / (n - 1), "one_core": one}
    parity = {"theta_steps": n, "signal_gpu": s_gpu.real, "signal_cpu": s_cpu,
              "rel_err": abs(s_gpu.real - s_cpu) / abs(s_cpu)}
    return cpu, parity


def run_adjoint(args):
    """c3: forward solve alone vs adjoint (forward + checkpointed reverse sweep, all gradients) on the
    300k-vertex cylinders; c4adj: the same on the 2M-vertex neuron surrogate under cos-OGSE (one
    signal: b = 4000, the first of the 32 directions)."""
    import dmri_mesh as dm
    import dmri_native as N

    t0 = time.time()
    if args.workload == "c4adj":
        f, T, int4gb = cos_ogse(4, 19900.0, 19900.0)
        k = 100.0
        mesh = dm.neuron_mesh(target_nv=args.nv or 2000000, seed=1004)
        fA, fB, _ = gcloud_loop(f, T, k)
        q = math.sqrt(4000.0 / int4gb)
        g = tuple(float(x) for x in fibonacci_directions(32)[0])
        coeffs = dict(D=3e-3, invT2=1e-16, kappa=0.0)
        desc = ("adjoint: neuron-like surrogate (~2M vertices, seed 1004), D=3e-3, cos-OGSE 4 periods "
                "delta=Delta=19900 us, dt=100 us, b=4000, the first of 32 Fibonacci directions")
    else:
        delta, Delta, k, b = 10600.0, 43100.0, 200.0, 2000.0
        mesh = dm.cylinder_mesh(8.0, 20.0, args.nv or 300000, seed=1003, interface_radii=(4.0,))
        fA, fB = pgse_sequence(delta, Delta, k)
        q = q_from_b(b, delta, Delta)
        g = (0.0, 1.0, 0.0)
        coeffs = dict(D=3e-3, invT2=1.0 / 80000.0, kappa=1e-5)
        desc = ("adjoint: concentric cylinders r_in=4 r_out=8 L=20 um (~300k vertices, seed 1003), "
                "kappa=1e-5, D=3e-3, T2=80 ms, PGSE 10600/43100 us, dt=200 us, b=2000, gdir (0,1,0)")
    t_mesh = time.time() - t0
    if args.theta_steps:
        fA, fB = fA[:args.theta_steps], fB[:args.theta_steps]
    ctx = N.Context(0)
    ctx.set_mesh(mesh.xyz, mesh.tets, mesh.phase)
    ctx.set_coeffs(**coeffs)
    ctx.assemble()
    sqrt_stride = int(math.ceil(math.sqrt(len(fA))))
    for _ in range(args.warmup):
        ctx.solve(fA, fB, k, q, g, rtol=args.rtol, krylov=args.krylov, pre_fp32=args.pre_fp32)
        ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, krylov=args.krylov, pre_fp32=args.pre_fp32)
        ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, checkpoint_stride=sqrt_stride, krylov=args.krylov, pre_fp32=args.pre_fp32)
    tf = ta = tq = 0.0
    grads = gq = None
    for _ in range(args.steps):
        a0 = time.perf_counter()
        s_f, _ = ctx.solve(fA, fB, k, q, g, rtol=args.rtol, krylov=args.krylov, pre_fp32=args.pre_fp32)
        a1 = time.perf_counter()
        s_a, grads, _ = ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, krylov=args.krylov, pre_fp32=args.pre_fp32)
        a2 = time.perf_counter()
        _, gq, _ = ctx.adjoint(fA, fB, k, q, g, rtol=args.rtol, checkpoint_stride=sqrt_stride,
                                krylov=args.krylov, pre_fp32=args.pre_fp32)
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
        "config": {"workload": desc,
                   "nv": info["nv"], "ndof": info["ndof"], "nnz": info["nnz"], "theta_steps": len(fA),
                   "rtol": args.rtol, "krylov": args.krylov, "checkpoints": grads["checkpoints"],
                   "recomputed_steps": grads["recomputed_steps"]},
        "forward_ms": 1e3 * tf / args.steps, "adjoint_ms": 1e3 * ta / args.steps,
        "policy": "automatic: every state kept in HBM (they fit), no recompute",
        "krylov_iterations": {k_: grads["iters_" + k_] for k_ in ("forward", "recompute", "backward")},
        "sqrt_checkpointing": {"stride": sqrt_stride, "checkpoints": gq["checkpoints"],
                               "recomputed_steps": gq["recomputed_steps"], "adjoint_ms": 1e3 * tq / args.steps,
                               "ratio": tq / tf,
                               "krylov_iterations": {k_: gq["iters_" + k_] for k_ in ("forward", "recompute",
                                                                                      "backward")},
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
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c4adj"],
                    help="c4 (default): the neuron-like surrogate, one theta step of all 128 signals per bench step; "
                         "c1, c2: whole jobs per bench step; c3, c4adj: the adjoint on the cylinders / the neuron")
    ap.add_argument("--whole-jobs", action="store_true",
                    help="c4 only: time whole jobs (assembly + every signal's full sequence) instead of theta steps")
    ap.add_argument("--nv", type=int, default=0, help="target vertex count (0 = the workload's own)")
    ap.add_argument("--theta-steps", type=int, default=0, help="cut the sequence to its first K theta steps")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="theta steps of the CPU baseline sample (0 = the workload's default: 60 for c1/c2, 3 for c4)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-1core", action="store_true", help="skip the one-core CPU figure (c4)")
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
    ap.add_argument("--krylov", default="cocg", choices=["bicgstab", "cocg", "gmres"],
                    help="Krylov method of every theta step (cocg, the default here: the workloads' operators are "
                         "complex symmetric, one preconditioned product per iteration against BiCGStab's two)")
    ap.add_argument("--pre-fp32", type=int, default=0, choices=[0, 1],
                    help="1 (COCG): the preconditioner's SpMV reads r and the operator values in FP32, summing in "
                         "FP64 (the residual test and every other vector stay FP64)")
    ap.add_argument("--graphs", type=int, default=0, choices=[0, 1],
                    help="1: replay the Krylov iteration runs as captured hipGraphs")
    ap.add_argument("--event-iters", type=int, default=1,
                    help="Krylov iterations per theta step whose SpMV launches are event-timed")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus N>1 is not supported "
                                   "(signals are independent: each GPU would take its own launch groups)"}))
        sys.exit(2)
    if args.cpu_steps <= 0:
        args.cpu_steps = 3 if args.workload == "c4" else 60
    if args.workload in ("c3", "c4adj"):
        run_adjoint(args)
    elif args.workload == "c4" and not args.whole_jobs:
        run_stepwise(args, args.workload)
    else:
        run_forward(args, args.workload)


if __name__ == "__main__":
    main()
