"""ctypes binding of libdmri (include/dmri.h).

This is the only route from Python to the solver: there is no CPU fallback.  If
libdmri.so is missing the import of this module raises immediately.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DMRI_LIB", os.path.join(_PKG, "lib", "libdmri.so"))

DMRI_MAT = {"m": 0, "k": 1, "r": 2, "gi": 3, "jx": 4, "jy": 5, "jz": 6, "p": 7, "gx": 8, "gy": 9, "gz": 10}
ERRORS = {-1: "ARG", -2: "HIP", -3: "STATE", -4: "NOCONV", -5: "MESH", -6: "ALLOC", -7: "LIMIT"}


class DmriError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libdmri error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class _Coeffs(ctypes.Structure):
    _fields_ = [("D", ctypes.c_double * 2), ("D_cell", ctypes.POINTER(ctypes.c_double)),
                ("invT2", ctypes.c_double * 2), ("invT2_cell", ctypes.POINTER(ctypes.c_double)),
                ("kappa", ctypes.c_double), ("kappa_facet", ctypes.POINTER(ctypes.c_double))]


class _Info(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("nc", ctypes.c_int64), ("ncomp", ctypes.c_int64), ("ndof", ctypes.c_int64),
                ("nnz", ctypes.c_int64), ("n_iface", ctypes.c_int64), ("n_bfacet", ctypes.c_int64),
                ("ncolors_cells", ctypes.c_int32), ("ncolors_iface", ctypes.c_int32), ("volume", ctypes.c_double),
                ("hmin", ctypes.c_double), ("hmax", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("n_periodic", ctypes.c_int64), ("kappa_e", ctypes.c_double)]


class _Seq(ctypes.Structure):
    _fields_ = [("nsteps", ctypes.c_int32), ("f_implicit", ctypes.POINTER(ctypes.c_double)),
                ("f_explicit", ctypes.POINTER(ctypes.c_double)), ("dt", ctypes.c_double),
                ("theta", ctypes.c_double), ("q", ctypes.c_double), ("gdir", ctypes.c_double * 3),
                ("F_explicit", ctypes.POINTER(ctypes.c_double)), ("u0_vertex", ctypes.POINTER(ctypes.c_double)),
                ("snap_every", ctypes.c_int32), ("snap_out", ctypes.POINTER(ctypes.c_double)),
                ("F_implicit", ctypes.POINTER(ctypes.c_double))]


class _Opts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("maxit", ctypes.c_int32),
                ("profile", ctypes.c_int32), ("lanes_per_row", ctypes.c_int32), ("initial_guess", ctypes.c_int32),
                ("mapping", ctypes.c_int32), ("preconditioner", ctypes.c_int32), ("krylov", ctypes.c_int32),
                ("restart", ctypes.c_int32), ("graphs", ctypes.c_int32), ("pre_fp32", ctypes.c_int32)]


class _Stats(ctypes.Structure):
    _fields_ = [("iters_total", ctypes.c_int64), ("iters_max", ctypes.c_int32), ("restarts", ctypes.c_int32),
                ("relres_max", ctypes.c_double), ("solve_ms", ctypes.c_double), ("spmv_ms", ctypes.c_double),
                ("spmv_launches", ctypes.c_int64), ("spmv_bytes", ctypes.c_double),
                ("precond_used", ctypes.c_int32), ("graph_launches", ctypes.c_int32),
                ("kern_ms", ctypes.c_double * 7), ("kern_launches", ctypes.c_int64 * 7),
                ("kern_bytes", ctypes.c_double * 7)]

KERNELS = ("bicg_p", "spmv_v_pre", "spmv_v", "spass", "spmv_t_pre", "spmv_t", "bicg_update")


def _stats_dict(st) -> dict:
    out = {}
    for k, _ in _Stats._fields_:
        v = getattr(st, k)
        out[k] = list(v) if k.startswith("kern_") else v
    return out


class _AdjOpts(ctypes.Structure):
    _fields_ = [("checkpoint_stride", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class _Grads(ctypes.Structure):
    _fields_ = [("dS_dD", ctypes.c_double * 2), ("dS_dinvT2", ctypes.c_double * 2), ("dS_dkappa", ctypes.c_double),
                ("dS_dq", ctypes.c_double), ("dS_df_implicit", ctypes.POINTER(ctypes.c_double)),
                ("dS_df_explicit", ctypes.POINTER(ctypes.c_double)), ("ms_forward", ctypes.c_double),
                ("ms_adjoint", ctypes.c_double), ("checkpoints", ctypes.c_int32),
                ("recomputed_steps", ctypes.c_int32), ("iters_forward", ctypes.c_int64),
                ("iters_recompute", ctypes.c_int64), ("iters_backward", ctypes.c_int64)]


EXPORTED = ["dmri_create", "dmri_destroy", "dmri_last_error", "dmri_version", "dmri_mesh_set", "dmri_coeffs_set",
            "dmri_assemble", "dmri_get_info", "dmri_export_pattern", "dmri_export_values", "dmri_export_facets",
            "dmri_periodic_set", "dmri_periodic_strong_set", "dmri_export_periodic",
            "dmri_solve", "dmri_solve_batched", "dmri_tune", "dmri_set_tuning", "dmri_time_spmv", "dmri_copy_bandwidth",
            "dmri_tune_db", "dmri_adjoint", "dmri_get_solution", "dmri_get_solution_noalloc", "dmri_batch_begin",
            "dmri_batch_advance", "dmri_batch_signals", "dmri_batch_end"]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libdmri.so not found at {LIB_PATH}: build it with "
                          f"`python dmri-fem-cloud_amd/build_native.py` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, i64, i32p, dp = ctypes.c_void_p, ctypes.c_int64, P(ctypes.c_int32), P(ctypes.c_double)
    L.dmri_create.argtypes = [P(vp), ctypes.c_int]
    L.dmri_destroy.argtypes = [vp]
    L.dmri_last_error.restype = ctypes.c_char_p
    L.dmri_version.restype = ctypes.c_char_p
    L.dmri_mesh_set.argtypes = [vp, i64, dp, i64, i32p, i32p]
    L.dmri_coeffs_set.argtypes = [vp, P(_Coeffs)]
    L.dmri_assemble.argtypes = [vp]
    L.dmri_get_info.argtypes = [vp, P(_Info)]
    L.dmri_export_pattern.argtypes = [vp, P(ctypes.c_int64), i32p, i32p]
    L.dmri_export_values.argtypes = [vp, ctypes.c_int, dp]
    L.dmri_export_facets.argtypes = [vp, i32p]
    L.dmri_periodic_set.argtypes = [vp, i32p]
    L.dmri_periodic_strong_set.argtypes = [vp, i32p]
    L.dmri_export_periodic.argtypes = [vp, i32p, i32p, dp, dp]
    L.dmri_solve.argtypes = [vp, P(_Seq), P(_Opts), dp, P(_Stats)]
    L.dmri_solve_batched.argtypes = [vp, ctypes.c_int32, dp, dp, P(_Seq), P(_Opts), dp, P(_Stats)]
    L.dmri_get_solution.argtypes = [vp, ctypes.c_int32, dp]
    L.dmri_get_solution_noalloc.argtypes = [vp, ctypes.c_int32, dp]
    L.dmri_tune.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, i32p, i32p, dp]
    L.dmri_set_tuning.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.dmri_tune_db.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, i32p, i32p, i32p]
    L.dmri_time_spmv.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, dp, dp, dp]
    L.dmri_copy_bandwidth.argtypes = [vp, i64, ctypes.c_int32, dp]
    L.dmri_adjoint.argtypes = [vp, P(_Seq), P(_Opts), P(_AdjOpts), dp, P(_Grads), P(_Stats)]
    L.dmri_batch_begin.argtypes = [vp, ctypes.c_int32, dp, dp, P(_Seq), P(_Opts), ctypes.c_int32, P(vp)]
    L.dmri_batch_advance.argtypes = [vp, ctypes.c_int32, P(_Stats)]
    L.dmri_batch_signals.argtypes = [vp, dp, i32p]
    L.dmri_batch_end.argtypes = [vp]
    for name in EXPORTED:
        if name not in ("dmri_last_error", "dmri_version"):
            getattr(L, name).restype = ctypes.c_int
    return L


LIB = _load()


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ct))


def _check(rc):
    if rc != 0:
        raise DmriError(rc, LIB.dmri_last_error().decode())


def version() -> str:
    return LIB.dmri_version().decode()


class Context:
    """One solver context on one GPU (see include/dmri.h)."""

    KRYLOV = {"bicgstab": 0, "gmres": 1, "cocg": 2}

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        self.has_fields = False
        _check(LIB.dmri_create(ctypes.byref(self._h), int(device)))

    def close(self):
        if self._h:
            LIB.dmri_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up -------------------------------------------------------------------
    def set_mesh(self, xyz, tets, phase=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.int32)
        _check(LIB.dmri_mesh_set(self._h, xyz.shape[0], _ptr(xyz, ctypes.c_double), tets.shape[0],
                                 _ptr(tets, ctypes.c_int32), _ptr(ph, ctypes.c_int32)))

    def set_coeffs(self, D=3e-3, invT2=1e-16, kappa=0.0, D_cell=None, invT2_cell=None, kappa_facet=None):
        c = _Coeffs()
        Dv = np.broadcast_to(np.asarray(D, dtype=np.float64), (2,))
        Tv = np.broadcast_to(np.asarray(invT2, dtype=np.float64), (2,))
        c.D[0], c.D[1] = float(Dv[0]), float(Dv[1])
        c.invT2[0], c.invT2[1] = float(Tv[0]), float(Tv[1])
        c.kappa = float(kappa)
        self._keep = []
        self.has_fields = any(a is not None for a in (D_cell, invT2_cell, kappa_facet))
        for name, arr in (("D_cell", D_cell), ("invT2_cell", invT2_cell), ("kappa_facet", kappa_facet)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
                self._keep.append(a)
                setattr(c, name, _ptr(a, ctypes.c_double))
        _check(LIB.dmri_coeffs_set(self._h, ctypes.byref(c)))

    def assemble(self):
        _check(LIB.dmri_assemble(self._h))

    def info(self) -> dict:
        i = _Info()
        _check(LIB.dmri_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _Info._fields_}

    # -- parity hooks ---------------------------------------------------------------
    def export_pattern(self):
        inf = self.info()
        rowptr = np.zeros(inf["ndof"] + 1, dtype=np.int64)
        col = np.zeros(inf["nnz"], dtype=np.int32)
        dof_of = np.zeros(inf["ncomp"] * inf["nv"], dtype=np.int32)
        _check(LIB.dmri_export_pattern(self._h, _ptr(rowptr, ctypes.c_int64), _ptr(col, ctypes.c_int32),
                                       _ptr(dof_of, ctypes.c_int32)))
        return rowptr, col, dof_of

    def export_values(self, which: str) -> np.ndarray:
        out = np.zeros(self.info()["nnz"])
        _check(LIB.dmri_export_values(self._h, DMRI_MAT[which], _ptr(out, ctypes.c_double)))
        return out

    def export_facets(self) -> np.ndarray:
        n = self.info()["n_iface"]
        out = np.zeros((max(n, 1), 3), dtype=np.int32)
        _check(LIB.dmri_export_facets(self._h, _ptr(out, ctypes.c_int32)))
        return out[:n]

    def set_periodic(self, dirs):
        """Weak pseudo-periodic faces normal to the axes with dirs[d] != 0 (next assemble)."""
        d = np.ascontiguousarray([1 if x else 0 for x in dirs], dtype=np.int32)
        _check(LIB.dmri_periodic_set(self._h, _ptr(d, ctypes.c_int32)))

    def set_periodic_strong(self, dirs):
        """Strong pseudo-periodic condition on the faces normal to the axes with dirs[d] != 0 (next
        assemble): max-face vertices share their min-face image's DOF (dmri_periodic_strong_set)."""
        d = np.ascontiguousarray([1 if x else 0 for x in dirs], dtype=np.int32)
        _check(LIB.dmri_periodic_strong_set(self._h, _ptr(d, ctypes.c_int32)))

    def export_periodic(self):
        """Mirror entries of the last assemble: dict(dof, src (ne,3), w (ne,3), ell (ne,3))."""
        ne = self.info()["n_periodic"]
        dof = np.zeros(max(ne, 1), np.int32)
        src = np.zeros((max(ne, 1), 3), np.int32)
        w = np.zeros((max(ne, 1), 3))
        ell = np.zeros((max(ne, 1), 3))
        _check(LIB.dmri_export_periodic(self._h, _ptr(dof, ctypes.c_int32), _ptr(src, ctypes.c_int32),
                                        _ptr(w, ctypes.c_double), _ptr(ell, ctypes.c_double)))
        return {"dof": dof[:ne], "src": src[:ne], "w": w[:ne], "ell": ell[:ne]}

    # -- solve --------------------------------------------------------------------------
    @staticmethod
    def _seq_opts(f_implicit, f_explicit, dt, q, gdir, theta, rtol, atol, maxit, profile, lanes, guess=0,
                  mapping=0, F_explicit=None, precond=0, u0=None, nv=None, F_implicit=None, krylov=0, restart=0,
                  graphs=0, pre_fp32=0):
        fA = np.ascontiguousarray(f_implicit, dtype=np.float64)
        fB = np.ascontiguousarray(f_explicit, dtype=np.float64)
        if fA.shape != fB.shape:
            raise ValueError("f_implicit and f_explicit differ in length")
        FB = None if F_explicit is None else np.ascontiguousarray(F_explicit, dtype=np.float64)
        if FB is not None and FB.shape != fA.shape:
            raise ValueError("F_explicit and the time profile differ in length")
        s = _Seq()
        s.nsteps = fA.shape[0]
        s.f_implicit = _ptr(fA, ctypes.c_double)
        s.f_explicit = _ptr(fB, ctypes.c_double)
        s.F_explicit = _ptr(FB, ctypes.c_double) if FB is not None else None
        FA = None if F_implicit is None else np.ascontiguousarray(F_implicit, dtype=np.float64)
        if FA is not None and FA.shape != fA.shape:
            raise ValueError("F_implicit and the time profile differ in length")
        s.F_implicit = _ptr(FA, ctypes.c_double) if FA is not None else None
        U0 = None if u0 is None else np.ascontiguousarray(u0, dtype=np.float64).ravel()
        if U0 is not None and nv is not None and U0.shape[0] != nv:
            raise ValueError(f"u0 needs one value per mesh vertex ({nv}), got {U0.shape[0]}")
        s.u0_vertex = _ptr(U0, ctypes.c_double) if U0 is not None else None
        s.dt, s.theta, s.q = float(dt), float(theta), float(q)
        for i in range(3):
            s.gdir[i] = float(gdir[i])
        o = _Opts()
        o.rtol, o.atol, o.maxit, o.profile = float(rtol), float(atol), int(maxit), int(profile)
        o.lanes_per_row = int(lanes)
        o.initial_guess = int(guess)
        o.mapping = int(mapping)
        o.preconditioner = int(precond)
        o.krylov = Context.KRYLOV[krylov] if isinstance(krylov, str) else int(krylov)
        o.restart = int(restart)
        o.graphs = int(graphs)
        o.pre_fp32 = int(pre_fp32)
        return s, o, (fA, fB, FB, U0, FA)

    def _nv(self, u0):
        return None if u0 is None else int(self.info()["nv"])

    def _snapshots(self, s, snap_every, nsig):
        """Attach a snapshot buffer to the sequence: (nsnap, nsig, ndof) complex, or None."""
        snap_every = int(snap_every)
        if snap_every < 0:
            raise ValueError("snap_every must be >= 0")
        if snap_every == 0:
            return None
        nsnap = (s.nsteps + snap_every - 1) // snap_every
        buf = np.zeros((nsnap, nsig, int(self.info()["ndof"])), dtype=np.complex128)
        s.snap_every = snap_every
        if buf.size:
            s.snap_out = _ptr(buf.view(np.float64), ctypes.c_double)
        return buf

    def solve(self, f_implicit, f_explicit, dt, q, gdir, theta=0.5, rtol=1e-13, atol=0.0, maxit=10000,
              profile=False, lanes=0, guess=0, mapping=0, F_explicit=None, precond=0, u0=None, snap_every=0,
              F_implicit=None, krylov=0, restart=0, graphs=0, pre_fp32=0):
        """precond: 0 = Jacobi + degree-1 polynomial (default), 1 = Jacobi only.  guess: Krylov start
        of each step, 0 = residual-minimising projection on the last states (default), 1 = previous
        state, 2 = linear, 3 = quadratic extrapolation.  u0: initial state per mesh vertex (real part
        of every compartment's DOF there; None = 1).  snap_every > 0: stats["snapshots"] holds u
        (complex, DOF order) after every step whose 0-based index is a multiple of it, shape
        (nsnap, ndof)."""
        s, o, keep = self._seq_opts(f_implicit, f_explicit, dt, q, gdir, theta, rtol, atol, maxit, profile, lanes,
                                    guess, mapping, F_explicit, precond, u0, self._nv(u0), F_implicit, krylov, restart,
                                    graphs, pre_fp32)
        snaps = self._snapshots(s, snap_every, 1)
        sig = np.zeros(2)
        st = _Stats()
        _check(LIB.dmri_solve(self._h, ctypes.byref(s), ctypes.byref(o), _ptr(sig, ctypes.c_double),
                              ctypes.byref(st)))
        stats = _stats_dict(st)
        if snaps is not None:
            stats["snapshots"] = snaps[:, 0, :]
        return complex(sig[0], sig[1]), stats

    def solve_batched(self, f_implicit, f_explicit, dt, qs, gdirs, theta=0.5, rtol=1e-13, atol=0.0, maxit=10000,
                      profile=False, lanes=0, guess=0, mapping=0, F_explicit=None, precond=0, u0=None,
                      snap_every=0, F_implicit=None, krylov=0, restart=0, graphs=0, pre_fp32=0):
        """Signals for every (q[s], gdir[s]) on the shared pattern; returns complex array (nsig,).
        snap_every > 0: stats["snapshots"] has shape (nsnap, nsig, ndof), as in solve."""
        qs = np.ascontiguousarray(np.atleast_1d(qs), dtype=np.float64)
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gdirs, dtype=np.float64).reshape(-1, 3),
                                                 (qs.shape[0], 3)))
        s, o, keep = self._seq_opts(f_implicit, f_explicit, dt, 0.0, (1, 0, 0), theta, rtol, atol, maxit, profile,
                                    lanes, guess, mapping, F_explicit, precond, u0, self._nv(u0), F_implicit, krylov,
                                    restart, graphs, pre_fp32)
        snaps = self._snapshots(s, snap_every, qs.shape[0])
        sig = np.zeros(2 * qs.shape[0])
        st = _Stats()
        _check(LIB.dmri_solve_batched(self._h, qs.shape[0], _ptr(qs, ctypes.c_double), _ptr(g, ctypes.c_double),
                                      ctypes.byref(s), ctypes.byref(o), _ptr(sig, ctypes.c_double),
                                      ctypes.byref(st)))
        stats = _stats_dict(st)
        if snaps is not None:
            stats["snapshots"] = snaps
        return sig.view(np.complex128), stats

    def batch(self, f_implicit, f_explicit, dt, qs, gdirs, theta=0.5, rtol=1e-13, atol=0.0, maxit=10000,
              profile=False, lanes=0, guess=0, mapping=0, F_explicit=None, precond=0, u0=None, group=0,
              F_implicit=None, krylov=0, restart=0, graphs=0, pre_fp32=0):
        """Resumable batched solve (dmri_batch_begin): returns a Batch whose advance(k) runs k more
        theta steps of every signal and whose signals() gives the signals of the current states."""
        qs = np.ascontiguousarray(np.atleast_1d(qs), dtype=np.float64)
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gdirs, dtype=np.float64).reshape(-1, 3),
                                                 (qs.shape[0], 3)))
        s, o, keep = self._seq_opts(f_implicit, f_explicit, dt, 0.0, (1, 0, 0), theta, rtol, atol, maxit, profile,
                                    lanes, guess, mapping, F_explicit, precond, u0, self._nv(u0), F_implicit, krylov,
                                    restart, graphs, pre_fp32)
        h = ctypes.c_void_p()
        _check(LIB.dmri_batch_begin(self._h, qs.shape[0], _ptr(qs, ctypes.c_double), _ptr(g, ctypes.c_double),
                                    ctypes.byref(s), ctypes.byref(o), int(group), ctypes.byref(h)))
        return Batch(self, h, qs.shape[0], s.nsteps)

    MAPPINGS = {1: "spread", 2: "spread-staged", 3: "registers", 4: "wave-per-row", 5: "tile-compressed"}

    def tune(self, nsig: int, per_gdir: bool = False):
        """Time every (mapping, lanes) candidate; returns ((mapping, lanes), {(mapping, lanes): us})."""
        bl, bm = ctypes.c_int32(), ctypes.c_int32()
        us = np.zeros(36)
        _check(LIB.dmri_tune(self._h, int(nsig), 1 if per_gdir else 0, ctypes.byref(bl), ctypes.byref(bm),
                             _ptr(us, ctypes.c_double)))
        times = {(m, l): us[(m - 1) * 6 + i] for m in range(1, 7) for i, l in enumerate((1, 2, 4, 8, 16, 32))
                 if us[(m - 1) * 6 + i] >= 0}
        return (bm.value, bl.value), times

    def tune_db(self, path, nsig: int, per_gdir: bool = False):
        """Stored choice from the tuning database at path, or tune and append; ((mapping, lanes), hit)."""
        bl, bm, hit = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(LIB.dmri_tune_db(self._h, None if path is None else str(path).encode(), int(nsig),
                                1 if per_gdir else 0, ctypes.byref(bl), ctypes.byref(bm), ctypes.byref(hit)))
        return (bm.value, bl.value), bool(hit.value)

    def set_tuning(self, nsig: int, per_gdir: bool, lanes: int, mapping: int):
        _check(LIB.dmri_set_tuning(self._h, int(nsig), 1 if per_gdir else 0, int(lanes), int(mapping)))

    def adjoint(self, f_implicit, f_explicit, dt, q, gdir, theta=0.5, rtol=1e-13, atol=0.0, maxit=10000,
                checkpoint_stride=0, lanes=0, guess=0, mapping=0, precond=0, u0=None, krylov=0, restart=0,
                graphs=0, pre_fp32=0):
        """Signal and its adjoint gradients (see dmri_adjoint): returns (signal, grads dict, stats)."""
        s, o, keep = self._seq_opts(f_implicit, f_explicit, dt, q, gdir, theta, rtol, atol, maxit, False, lanes,
                                    guess, mapping, None, precond, u0, self._nv(u0), None, krylov, restart, graphs,
                                    pre_fp32)
        ao = _AdjOpts()
        ao.checkpoint_stride = int(checkpoint_stride)
        n = s.nsteps
        dfa, dfb = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        g = _Grads()
        g.dS_df_implicit = _ptr(dfa, ctypes.c_double)
        g.dS_df_explicit = _ptr(dfb, ctypes.c_double)
        sig = np.zeros(2)
        st = _Stats()
        _check(LIB.dmri_adjoint(self._h, ctypes.byref(s), ctypes.byref(o), ctypes.byref(ao), _ptr(sig, ctypes.c_double),
                                ctypes.byref(g), ctypes.byref(st)))
        grads = {"D": np.array(g.dS_dD[:]), "invT2": np.array(g.dS_dinvT2[:]), "kappa": g.dS_dkappa, "q": g.dS_dq,
                 "f_implicit": dfa[:n].copy(), "f_explicit": dfb[:n].copy(), "ms_forward": g.ms_forward,
                 "ms_adjoint": g.ms_adjoint, "checkpoints": g.checkpoints, "recomputed_steps": g.recomputed_steps,
                 "iters_forward": g.iters_forward, "iters_recompute": g.iters_recompute,
                 "iters_backward": g.iters_backward}
        return complex(sig[0], sig[1]), grads, _stats_dict(st)

    def time_spmv(self, nsig: int, per_gdir: bool = False, reps: int = 50):
        """(us per spmv_v launch, us per spmv_t launch, algorithmic bytes per launch) of back-to-back
        replays of the two SpMV kernels with the mapping a solve of nsig signals would use."""
        uv, ut, nb = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(LIB.dmri_time_spmv(self._h, int(nsig), int(bool(per_gdir)), int(reps), ctypes.byref(uv),
                                  ctypes.byref(ut), ctypes.byref(nb)))
        return uv.value, ut.value, nb.value

    def copy_bandwidth(self, nbytes: int, reps: int = 10) -> float:
        """GB/s of the library's plain copy kernel (read + write bytes)."""
        g = ctypes.c_double()
        _check(LIB.dmri_copy_bandwidth(self._h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    def solution(self, sig: int = 0) -> np.ndarray:
        n = self.info()["ndof"]
        u = np.zeros(2 * n)
        _check(LIB.dmri_get_solution(self._h, int(sig), _ptr(u, ctypes.c_double)))
        return u.view(np.complex128)


class Batch:
    """A resumable batched solve of a Context (see dmri_batch_begin in include/dmri.h)."""

    def __init__(self, ctx: Context, handle, nsig: int, nsteps: int):
        self.ctx, self._h, self.nsig, self.nsteps = ctx, handle, nsig, nsteps

    def advance(self, nsteps: int = 1) -> dict:
        st = _Stats()
        _check(LIB.dmri_batch_advance(self._h, int(nsteps), ctypes.byref(st)))
        return _stats_dict(st)

    def signals(self):
        """(complex signals of the current states, theta steps done)."""
        sig = np.zeros(2 * self.nsig)
        done = ctypes.c_int32()
        _check(LIB.dmri_batch_signals(self._h, _ptr(sig, ctypes.c_double), ctypes.byref(done)))
        return sig.view(np.complex128), done.value

    def close(self):
        if self._h:
            LIB.dmri_batch_end(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
