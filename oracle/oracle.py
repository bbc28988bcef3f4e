"""Python face of the CPU oracle (liboracle.so built from dmri_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of bench.py, always as the checker or the timed CPU baseline, never
as part of the product path.

Besides the ctypes bindings this module restates the reference's time-sequence logic:

* PGSE profile as the sympy Piecewise of GCloudDmriSolver.py:188-193 (exclusive at s=T)
  and as the comri ``FT`` (comri/one-comp/hpc-fenics-cpp/main.cpp:87-92, inclusive at
  t=Delta+delta);
* the GCloud loop ``while t < T + k`` with t accumulated by ``+=`` and the right-hand side
  lagged one step (DmriFemLib.py:886-911; first step uses f(0) on both sides);
* the comri loop ``while (t < T + dt)`` with the same f on both sides (main.cpp:296-318);
* q = sqrt(b / int_0^T F(s)^2 ds) with int F^2 = delta^2 (Delta - delta/3) for PGSE
  (DmriFemLib.py:826-842; main.cpp:187-191).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GAMMA_RATIO = 2.675e8  # DmriFemLib.py:680-686


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64, i32p, dp, i64p = ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)
        L.orc_dof_map.restype = i64
        L.orc_dof_map.argtypes = [i64, i64, i32p, i32p, ctypes.c_int, i32p]
        L.orc_interface_facets.restype = i64
        L.orc_interface_facets.argtypes = [i64, i32p, i32p, i32p, i32p, i64]
        L.orc_boundary_facets.restype = i64
        L.orc_boundary_facets.argtypes = [i64, i32p, i32p, i32p, i64]
        L.orc_pattern.restype = i64
        L.orc_pattern.argtypes = [i64, i64, i64, i32p, i32p, i32p, i64, i32p, i64p, i32p]
        L.orc_assemble.restype = ctypes.c_int
        L.orc_assemble.argtypes = [i64, dp, i64, i32p, i32p, i32p, dp, dp, dp, dp, ctypes.c_double, dp, i64, i32p,
                                   i64p, i32p, i64, dp, dp, dp, dp, dp, dp, dp]
        L.orc_theta_solve.restype = ctypes.c_int
        L.orc_theta_solve.argtypes = [i64, i64p, i32p, dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_theta_solve_p.restype = ctypes.c_int
        L.orc_theta_solve_p.argtypes = [i64, i64p, i32p, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, dp, dp, dp, i64, i32p, i32p, dp, dp, dp, dp]
        L.orc_periodic.restype = i64
        L.orc_periodic.argtypes = [i64, dp, i64, i32p, i32p, i32p, i32p, ctypes.c_double, i64p, i32p, dp, i32p,
                                   i32p, dp, dp, i64]
        L.orc_spmv.restype = None
        L.orc_spmv.argtypes = [i64, i64p, i32p, dp, dp, ctypes.c_double, dp, dp]
        L.orc_assemble_sbc.restype = ctypes.c_int
        L.orc_assemble_sbc.argtypes = [i64, dp, i64, i32p, i32p, i32p, dp, i64p, i32p, i64, dp, dp, dp, dp, dp, dp]
        L.orc_signal.restype = None
        L.orc_signal.argtypes = [i64, dp, dp, dp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ------------------------------------------------------------------------------------
# time sequences
# ------------------------------------------------------------------------------------
def pgse_piecewise(delta, Delta):
    """sympy Piecewise of GCloudDmriSolver.py:188-193 evaluated in floats."""
    T = Delta + delta

    def f(s):
        if s < delta:
            return 1.0
        if s < Delta:
            return 0.0
        if s < T:
            return -1.0
        return 0.0

    return f


def pgse_comri(delta, Delta):
    """comri FT (main.cpp:87-92)."""

    def f(t):
        return 1.0 * (0 <= t < delta) - 1.0 * (Delta <= t <= Delta + delta)

    return f


def pgse_integral(delta, Delta):
    """F(t) = int_0^t f for the Piecewise PGSE of GCloudDmriSolver.py:188-193 (closed form of
    MRI_parameters.itime_profile_sym, DmriFemLib.py:829-832)."""
    T = Delta + delta

    def F(t):
        if t < delta:
            return float(t)
        if t < Delta:
            return float(delta)
        if t < T:
            return float(delta - (t - Delta))
        return 0.0

    return F


def q_from_b_pgse(b, delta, Delta):
    return math.sqrt(b) / math.sqrt(delta * delta * (Delta - delta / 3.0))


def cos_ogse(nperiod, delta, Delta, t0=100.0, seq_ext=100.0):
    """cos-OGSE, profile 3 of ArbitraryTimeSequence.ipynb:380-392, evaluated in floats:
    omega = 2 pi nperiod / delta, tau = (Delta + delta)/2, T = Delta + delta + t0 + seq_ext and the
    Piecewise branches (0, s < t0), (cos(omega (s - t0)), s <= delta + t0), (0, s <= tau + t0),
    (-cos(omega (s - t0 - tau)), s <= delta + tau + t0), (0, True).  Returns (f, T)."""
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


def cos_ogse_integral(nperiod, delta, Delta, t0=100.0):
    """F(t) = int_0^t f of cos_ogse (MRI_parameters.itime_profile_sym, DmriFemLib.py:829-832, for
    that profile): sin(omega (t - t0))/omega on the first lobe, held at its end value until the
    second lobe, which subtracts sin(omega (t - t0 - tau))/omega."""
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


def int_F2_cos_ogse(nperiod, delta, Delta):
    """int_0^T F(s)^2 ds of cos_ogse (MRI_parameters.integral_term_for_gb, DmriFemLib.py:826-827) for
    whole periods (F returns to 0 after each lobe and the lobes do not overlap, tau >= delta):
    2 int_0^delta sin^2(omega u)/omega^2 du = delta / omega^2."""
    om = 2.0 * nperiod * math.pi / delta
    return delta / (om * om)


def gcloud_sequence(f, T, k):
    """(fA, fB) of MRI_simulation.solve: loop while t < T + k, t += k, RHS lagged
    (tp starts at 0), DmriFemLib.py:886-911."""
    fA, fB = [], []
    t, tp = 0.0, 0.0
    while t < T + k:
        fA.append(f(t))
        fB.append(f(tp))
        tp = t
        t += k
    return np.asarray(fA), np.asarray(fB)


def lagged_sequence_until(f, T, k):
    """The inline loop of ConvergenceTest.ipynb cell 10: while t < T, lagged RHS
    starting from f(0)."""
    fA, fB = [], []
    t = 0.0
    fprev = f(t)
    while t < T:
        ft = f(t)
        fA.append(ft)
        fB.append(fprev)
        fprev = ft
        t += k
    return np.asarray(fA), np.asarray(fB)


def comri_sequence(f, T, dt, until_T=False):
    """comri loops: same f on both sides; while (t < T + dt) (one-/two-comp main.cpp:296,854)
    or while (t < T) (multilayer main.cpp:301)."""
    fA = []
    t = 0.0
    end = T if until_T else T + dt
    while t < end:
        fA.append(f(t))
        t += dt
    fA = np.asarray(fA)
    return fA, fA.copy()


# ------------------------------------------------------------------------------------
# assembled system
# ------------------------------------------------------------------------------------
class System:
    """Oracle-assembled matrices on the DOF pattern of a mesh.

    D: scalar or per-compartment pair; invT2 likewise; kappa scalar (interface)."""

    def __init__(self, xyz, tets, phase=None, D=3e-3, invT2=1e-16, kappa=0.0, Dcell=None, invT2cell=None,
                 kappa_facet=None, periodic=None):
        L = lib()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32)
        self.phase = None if phase is None else np.ascontiguousarray(phase, dtype=np.int32)
        self.ncomp = 1 if self.phase is None else 2
        nv, nc = self.xyz.shape[0], self.tets.shape[0]
        self.nv, self.nc = nv, nc
        self.dof_of = np.zeros(self.ncomp * nv, dtype=np.int32)
        ndof = L.orc_dof_map(nv, nc, _p(self.tets, ctypes.c_int32), _p(self.phase, ctypes.c_int32), self.ncomp,
                             _p(self.dof_of, ctypes.c_int32))
        if ndof < 0:
            raise ValueError("bad mesh / phase")
        self.ndof = int(ndof)
        if self.phase is not None:
            nf = L.orc_interface_facets(nc, _p(self.tets, ctypes.c_int32), _p(self.phase, ctypes.c_int32), None,
                                        None, 0)
            self.fv = np.zeros((max(nf, 1), 3), dtype=np.int32)
            self.fc = np.zeros((max(nf, 1), 2), dtype=np.int32)
            L.orc_interface_facets(nc, _p(self.tets, ctypes.c_int32), _p(self.phase, ctypes.c_int32),
                                   _p(self.fv, ctypes.c_int32), _p(self.fc, ctypes.c_int32), nf)
            self.fv, self.fc = self.fv[:nf], self.fc[:nf]
        else:
            self.fv = np.zeros((0, 3), dtype=np.int32)
            self.fc = np.zeros((0, 2), dtype=np.int32)
        nf = self.fv.shape[0]
        self.rowptr = np.zeros(self.ndof + 1, dtype=np.int64)
        fvp = _p(self.fv, ctypes.c_int32) if nf else None
        nnz = L.orc_pattern(self.ndof, nv, nc, _p(self.tets, ctypes.c_int32), _p(self.phase, ctypes.c_int32),
                            _p(self.dof_of, ctypes.c_int32), nf, fvp, _p(self.rowptr, ctypes.c_int64), None)
        if nnz < 0:
            raise MemoryError("oracle pattern")
        self.col = np.zeros(nnz, dtype=np.int32)
        L.orc_pattern(self.ndof, nv, nc, _p(self.tets, ctypes.c_int32), _p(self.phase, ctypes.c_int32),
                      _p(self.dof_of, ctypes.c_int32), nf, fvp, _p(self.rowptr, ctypes.c_int64),
                      _p(self.col, ctypes.c_int32))
        self.nnz = int(nnz)
        Dc = np.ascontiguousarray(np.broadcast_to(np.asarray(D, dtype=np.float64), (2,)))
        iT = np.ascontiguousarray(np.broadcast_to(np.asarray(invT2, dtype=np.float64), (2,)))
        Dcell = None if Dcell is None else np.ascontiguousarray(Dcell, dtype=np.float64).reshape(nc, 9)
        invT2cell = None if invT2cell is None else np.ascontiguousarray(invT2cell, dtype=np.float64)
        kappa_facet = None if kappa_facet is None else np.ascontiguousarray(kappa_facet, dtype=np.float64)
        self.mats = {k: np.zeros(nnz) for k in ("m", "k", "r", "gi", "jx", "jy", "jz")}
        M = self.mats
        rc = L.orc_assemble(nv, _p(self.xyz, ctypes.c_double), nc, _p(self.tets, ctypes.c_int32),
                            _p(self.phase, ctypes.c_int32), _p(self.dof_of, ctypes.c_int32), _p(Dc, ctypes.c_double),
                            _p(Dcell, ctypes.c_double), _p(iT, ctypes.c_double), _p(invT2cell, ctypes.c_double),
                            float(kappa), _p(kappa_facet, ctypes.c_double), nf, fvp,
                            _p(self.rowptr, ctypes.c_int64), _p(self.col, ctypes.c_int32), nnz,
                            *[_p(M[k], ctypes.c_double) for k in ("m", "k", "r", "gi", "jx", "jy", "jz")])
        if rc != 0:
            raise RuntimeError(f"oracle assembly failed ({rc})")
        # weak pseudo-periodic boundary mass and mirror entries
        self.periodic = None
        if periodic is not None and any(periodic):
            X = self.xyz[self.tets]                                      # nc x 4 x 3
            ii, jj = np.triu_indices(4, 1)
            self.hmin = float(np.linalg.norm(X[:, ii] - X[:, jj], axis=2).max(axis=1).min())
            pdir = np.ascontiguousarray([1 if p else 0 for p in periodic], dtype=np.int32)
            pv = np.zeros(nnz)
            args = [nv, _p(self.xyz, ctypes.c_double), nc, _p(self.tets, ctypes.c_int32),
                    _p(self.phase, ctypes.c_int32), _p(self.dof_of, ctypes.c_int32), _p(pdir, ctypes.c_int32),
                    self.hmin, _p(self.rowptr, ctypes.c_int64), _p(self.col, ctypes.c_int32)]
            ne = L.orc_periodic(*args, _p(np.zeros(nnz), ctypes.c_double), None, None, None, None, 0)
            if ne < 0:
                raise RuntimeError(f"oracle periodic set-up failed ({ne})")
            ed = np.zeros(max(ne, 1), np.int32)
            es = np.zeros((max(ne, 1), 3), np.int32)
            ew = np.zeros((max(ne, 1), 3))
            el = np.zeros((max(ne, 1), 3))
            L.orc_periodic(*args, _p(pv, ctypes.c_double), _p(ed, ctypes.c_int32), _p(es, ctypes.c_int32),
                           _p(ew, ctypes.c_double), _p(el, ctypes.c_double), ne)
            self.mats["p"] = pv
            self.periodic = {"dir": pdir, "dof": ed[:ne], "src": es[:ne], "w": ew[:ne], "ell": el[:ne],
                             "kappa_e": 3e-3 / self.hmin}
        # signal weights w = M 1 (row sums of the mass matrix on live DOFs)
        self.w = np.add.reduceat(M["m"], self.rowptr[:-1]) if nnz else np.zeros(self.ndof)
        self.volume = float(self.w.sum())

    # -- helpers ------------------------------------------------------------------
    def l(self):
        return self.mats["k"] + self.mats["r"] + self.mats["gi"]

    def jg(self, g):
        g = np.asarray(g, dtype=np.float64)
        g = g / np.linalg.norm(g)
        return g[0] * self.mats["jx"] + g[1] * self.mats["jy"] + g[2] * self.mats["jz"]

    def dense(self, name_or_vals):
        vals = self.mats[name_or_vals] if isinstance(name_or_vals, str) else name_or_vals
        A = np.zeros((self.ndof, self.ndof))
        rows = np.repeat(np.arange(self.ndof), np.diff(self.rowptr))
        A[rows, self.col] = vals
        return A

    def initial_condition(self):
        u = np.zeros(self.ndof, dtype=np.complex128)
        u[:] = 1.0
        return u

    def solve(self, gdir, q, dt, fA, fB, theta=0.5, rtol=1e-13, atol=0.0, maxit=100000, u0=None, l=None, FB=None):
        """FB: F(t) = int_0^t f at the explicit side's time of every step (periodic systems only)."""
        L = lib()
        fA = np.ascontiguousarray(fA, dtype=np.float64)
        fB = np.ascontiguousarray(fB, dtype=np.float64)
        u = self.initial_condition() if u0 is None else np.array(u0, dtype=np.complex128)
        uv = u.view(np.float64)
        lv = np.ascontiguousarray(self.l() if l is None else l)
        jg = np.ascontiguousarray(self.jg(gdir))
        stats = np.zeros(4)
        base = [self.ndof, _p(self.rowptr, ctypes.c_int64), _p(self.col, ctypes.c_int32),
                _p(self.mats["m"], ctypes.c_double), _p(lv, ctypes.c_double), _p(jg, ctypes.c_double),
                float(dt), float(theta), float(q), len(fA), _p(fA, ctypes.c_double),
                _p(fB, ctypes.c_double), float(rtol), float(atol), int(maxit), _p(uv, ctypes.c_double),
                _p(stats, ctypes.c_double)]
        if self.periodic is None:
            rc = L.orc_theta_solve(*base)
        else:
            P = self.periodic
            FB = np.ascontiguousarray(np.zeros(len(fA)) if FB is None else FB, dtype=np.float64)
            assert FB.shape[0] == len(fA)
            g = np.asarray(gdir, dtype=np.float64)
            g = np.ascontiguousarray(g / np.linalg.norm(g))
            rc = L.orc_theta_solve_p(*base, _p(self.mats["p"], ctypes.c_double), int(P["dof"].shape[0]),
                                     _p(P["dof"], ctypes.c_int32), _p(P["src"], ctypes.c_int32),
                                     _p(P["w"], ctypes.c_double), _p(P["ell"], ctypes.c_double),
                                     _p(g, ctypes.c_double), _p(FB, ctypes.c_double))
        if rc != 0:
            raise RuntimeError("oracle Krylov solve did not converge")
        sig = np.zeros(2)
        L.orc_signal(self.ndof, _p(self.w, ctypes.c_double), _p(uv, ctypes.c_double), _p(sig, ctypes.c_double))
        return sig[0] / self.volume, u, {"iters": stats[0], "max_iters": stats[1], "max_relres": stats[2],
                                         "restarts": stats[3], "signal_raw": sig.copy()}


# ------------------------------------------------------------------------------------
# strong pseudo-periodic formulation (IsDomainPeriodic = True)
# ------------------------------------------------------------------------------------
def periodic_masters(xyz, pdir, tol):
    """Master vertex of every vertex under the periodic constraint of PeriodicBD (DmriFemLib.py:
    327-375): a vertex within tol of the max face of a periodic axis is the image of the vertex at
    the same point moved to the min face (map() moves x by -L along that axis); the chain of maps is
    followed to its end, so a vertex on several max faces takes the vertex at the corner of the min
    faces.  Brute force over the vertices; tol = 1e-2 hmin as in PeriodicBD."""
    xyz = np.asarray(xyz, dtype=np.float64)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    canon = xyz.copy()
    for d in range(3):
        if pdir[d]:
            canon[np.abs(xyz[:, d] - hi[d]) < tol, d] = lo[d]
    master = np.arange(xyz.shape[0])
    for v in np.nonzero(np.any(canon != xyz, axis=1))[0]:
        cand = np.nonzero(np.abs(xyz - canon[v]).max(axis=1) < tol)[0]
        if cand.shape[0] != 1:
            raise ValueError(f"vertex {v}: {cand.shape[0]} vertices at its periodic image (the mesh is not periodic)")
        master[v] = cand[0]
    return master


class StrongPeriodicSystem:
    """The strong pseudo-periodic scheme on the constrained (periodic) P1 space, built from the
    unconstrained oracle matrices by P^T X P (P maps the constrained DOFs onto the vertices they
    stand for), with scalar D per compartment.  Per step (ThetaMethodF/L_sBC1c/2c, DmriFemLib.py:
    167-238, theta on both sides as there, F = int_0^t f, FA = F(t_n), FB = F(t_{n-1})):
      A_n = M/k + theta (K + R + kappa I) + theta q^2 FA^2 D_c M + i theta q FA (C - B).g
      B_n = M/k - theta (K + R + kappa I) - theta q^2 FB^2 D_c M - i theta q FB (C - B).g
    with C, B the orc_assemble_sbc terms (the 1e-16 the reference adds to q F in outer_interface is
    left out: it moves the signal by about 1e-17).  Solved by a sparse LU per step (exact discrete
    scheme); signal = sum (M 1) Re u / |Omega|."""

    def __init__(self, xyz, tets, phase=None, D=3e-3, invT2=1e-16, kappa=0.0, pdir=(1, 1, 1)):
        import scipy.sparse as sps

        self.base = System(xyz, tets, phase, D=D, invT2=invT2, kappa=kappa)
        S = self.base
        X = S.xyz[S.tets]
        ii, jj = np.triu_indices(4, 1)
        hmin = float(np.linalg.norm(X[:, ii] - X[:, jj], axis=2).max(axis=1).min())
        self.master = periodic_masters(S.xyz, pdir, 1e-2 * hmin)
        L = lib()
        Dc = np.ascontiguousarray(np.broadcast_to(np.asarray(D, dtype=np.float64), (2,)))
        ext = {k: np.zeros(S.nnz) for k in ("cx", "cy", "cz", "bx", "by", "bz")}
        rc = L.orc_assemble_sbc(S.nv, _p(S.xyz, ctypes.c_double), S.nc, _p(S.tets, ctypes.c_int32),
                                _p(S.phase, ctypes.c_int32), _p(S.dof_of, ctypes.c_int32), _p(Dc, ctypes.c_double),
                                _p(S.rowptr, ctypes.c_int64), _p(S.col, ctypes.c_int32), S.nnz,
                                *[_p(ext[k], ctypes.c_double) for k in ("cx", "cy", "cz", "bx", "by", "bz")])
        if rc != 0:
            raise RuntimeError(f"oracle sBC assembly failed ({rc})")
        self.ext = ext
        # constrained DOFs: the DOF of (compartment, master vertex)
        nv, ncomp = S.nv, S.ncomp
        key = np.full(S.ndof, -1, dtype=np.int64)
        for c in range(ncomp):
            d = S.dof_of[c * nv:(c + 1) * nv]
            live = np.nonzero(d >= 0)[0]
            dm = S.dof_of[c * nv + self.master[live]]
            if np.any(dm < 0):
                raise ValueError("a periodic image vertex has no DOF in the compartment of its master")
            key[d[live]] = dm
        uniq, self.merged_of = np.unique(key, return_inverse=True)
        self.n = uniq.shape[0]
        rows = np.repeat(np.arange(S.ndof), np.diff(S.rowptr))
        P = sps.csr_matrix((np.ones(S.ndof), (np.arange(S.ndof), self.merged_of)), shape=(S.ndof, self.n))

        def merged(vals):
            A = sps.csr_matrix((vals, (rows, S.col)), shape=(S.ndof, S.ndof))
            return (P.T @ A @ P).tocsr()

        self.pattern = merged(np.ones(S.nnz))  # the constrained sparsity (every product entry)
        self.M = merged(S.mats["m"])
        self.L = merged(S.mats["k"] + S.mats["r"] + S.mats["gi"])
        comp = np.zeros(S.ndof)
        if ncomp == 2:
            comp[S.dof_of[nv:][S.dof_of[nv:] >= 0]] = 1.0
        dcomp = np.where(comp == 1.0, Dc[1], Dc[0])
        self.Mgg = merged(S.mats["m"] * dcomp[rows])
        self.G = [merged(ext["c" + a] - ext["b" + a]) for a in "xyz"]
        self.w = np.asarray(self.M.sum(axis=1)).ravel()
        self.volume = float(self.w.sum())

    def solve(self, gdir, q, dt, FA, FB, theta=0.5, u0=None):
        import scipy.sparse.linalg as spl

        g = np.asarray(gdir, dtype=np.float64)
        g = g / np.linalg.norm(g)
        Gg = g[0] * self.G[0] + g[1] * self.G[1] + g[2] * self.G[2]
        u = np.ones(self.n, dtype=np.complex128) if u0 is None else np.asarray(u0, dtype=np.complex128)
        base_a = (self.M / dt + theta * self.L).astype(np.complex128)
        base_b = (self.M / dt - theta * self.L).astype(np.complex128)
        for fa, fb in zip(FA, FB):
            A = base_a + (theta * q * q * fa * fa) * self.Mgg + (1j * theta * q * fa) * Gg
            B = base_b - (theta * q * q * fb * fb) * self.Mgg - (1j * theta * q * fb) * Gg
            u = spl.splu(A.tocsc()).solve(B @ u)
        return float((self.w * u.real).sum() / self.volume), u
