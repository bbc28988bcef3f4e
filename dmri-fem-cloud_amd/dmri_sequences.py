"""Gradient waveforms and theta-loop sequences for the solver's time profile (host side).

The solver takes per-step samples of the waveform f (dmri_seq.f_implicit / f_explicit) and, for
the pseudo-periodic terms, of its integral F(t) = int_0^t f.  This module produces them for the
reference's waveforms without sympy:

* PGSE as the sympy Piecewise of GCloudDmriSolver.py:188-193 (exclusive at s = T);
* cos-OGSE, profile 3 of ArbitraryTimeSequence.ipynb:380-392 (t0 = 100 us before the first lobe,
  seq_ext = 100 us after the second, lobes of length delta that start tau = (Delta + delta)/2 apart);
* the GCloud loop of MRI_simulation.solve (DmriFemLib.py:886-911): while t < T + k, t += k, the
  right-hand side lagged one step (t_{-1} = 0);
* q = sqrt(b / int_0^T F^2) (MRI_parameters.convert_b2q, DmriFemLib.py:826-842).

DmriFemLib.MRI_parameters evaluates a user's sympy profile instead, exactly as the reference does.
"""
from __future__ import annotations

import math

import numpy as np


def pgse(delta: float, Delta: float):
    """f of the PGSE Piecewise (1 on [0, delta), 0 on [delta, Delta), -1 on [Delta, Delta + delta))."""
    T = Delta + delta

    def f(s):
        return 1.0 if s < delta else (0.0 if s < Delta else (-1.0 if s < T else 0.0))

    return f


def pgse_integral(delta: float, Delta: float):
    """F(t) = int_0^t f of :func:`pgse`."""
    T = Delta + delta

    def F(t):
        return float(t) if t < delta else (float(delta) if t < Delta else
                                           (float(delta - (t - Delta)) if t < T else 0.0))

    return F


def pgse_int_F2(delta: float, Delta: float) -> float:
    """int_0^T F^2 of the PGSE: delta^2 (Delta - delta/3)."""
    return delta * delta * (Delta - delta / 3.0)


def cos_ogse(nperiod: int, delta: float, Delta: float, t0: float = 100.0, seq_ext: float = 100.0):
    """f and T of the cos-OGSE (omega = 2 pi nperiod / delta; branch ends inclusive as in the
    notebook's Piecewise: s <= delta + t0, s <= tau + t0, s <= delta + tau + t0)."""
    om = 2.0 * nperiod * math.pi / delta
    tau = 0.5 * (Delta + delta)
    T = Delta + delta + t0 + seq_ext

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

    return f, T


def cos_ogse_integral(nperiod: int, delta: float, Delta: float, t0: float = 100.0):
    """F(t) = int_0^t f of :func:`cos_ogse`."""
    om = 2.0 * nperiod * math.pi / delta
    tau = 0.5 * (Delta + delta)
    F1 = math.sin(om * delta) / om

    def F(t):
        if t < t0:
            return 0.0
        if t <= delta + t0:
            return math.sin(om * (t - t0)) / om
        if t <= tau + t0:
            return F1
        if t <= delta + tau + t0:
            return F1 - math.sin(om * (t - t0 - tau)) / om
        return F1 - math.sin(om * delta) / om

    return F


def cos_ogse_int_F2(nperiod: int, delta: float, Delta: float) -> float:
    """int_0^T F^2 of the cos-OGSE with whole periods and non-overlapping lobes (tau >= delta):
    two lobes of int_0^delta sin^2(omega u)/omega^2 du = delta/(2 omega^2) each."""
    if not (0.5 * (Delta + delta) >= delta):
        raise ValueError("the lobes overlap (Delta < delta)")
    om = 2.0 * nperiod * math.pi / delta
    return delta / (om * om)


def q_from_b(b: float, int_F2: float) -> float:
    return math.sqrt(b) / math.sqrt(int_F2)


def gcloud_loop(f, T: float, k: float, F=None):
    """(f(t_n), f(t_{n-1}), F(t_{n-1}) or None) over the GCloud loop: t = 0, k, 2k, ... while t < T + k
    (t accumulated by +=), t_{-1} = 0."""
    fA, fB, FB = [], [], []
    t, tp = 0.0, 0.0
    while t < T + k:
        fA.append(f(t))
        fB.append(f(tp))
        if F is not None:
            FB.append(F(tp))
        tp = t
        t += k
    return np.asarray(fA), np.asarray(fB), (np.asarray(FB) if F is not None else None)


def fibonacci_directions(n: int) -> np.ndarray:
    """n unit vectors on a Fibonacci sphere lattice (the batched workloads' gradient directions)."""
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = math.pi * (1 + 5 ** 0.5) * k
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)
